// Device form of fused passes: gate descriptors, the register-tiled gate
// bodies and generator overlaps, shared by the single-state kernels
// (lgh_fused.hip) and the batched kernels (lgh_batch.hip).  Internal to
// liblgh.so.
#pragma once

#include <vector>

#include "lgh_fused.hpp"

namespace lgh {
namespace fz {

// A wavefront's tile: the 64 lanes carry six index bits (the lowest ones, or
// the lowest three or four and others chosen per pass), and each lane holds
// NR = 2^RB registers, one per combination of the pass's RB register bits.
// RB is a compile-time parameter of the kernels (3 or 4): four register bits
// fuse more gates per pass, three leave room for two resident wavefronts per
// SIMD in the adjoint kernel, which holds two states.
constexpr int kMaxRB = 4;
constexpr int kMaxNR = 1 << kMaxRB;
constexpr int kLaneBits = 6;
// Trainable gates per pass.  The adjoint kernels keep their overlap sums in
// LDS, 48 KiB per workgroup of 256 lanes (three workgroups per CU): the
// batched kernel one double per slot and lane (kMaxSlots), the single-state
// kernel one per slot and quad of lanes (kMaxSlotsSv), the quad's four values
// added with two lane exchanges first.  A pass holds as many trainable gates
// as it has slots, so twice the slots means fewer passes over the two states;
// a quarter of a column per lane keeps the workgroup's LDS at 26 KiB, which
// lets four workgroups (four waves per SIMD) share a CU.
constexpr int kMaxSlots = 24;
constexpr int kAccLanes = 4;                     // lanes per accumulator column (single-state kernel)
constexpr int kAccCols = 256 / kAccLanes;        // columns per workgroup
constexpr int kMaxSlotsSv = 2 * kMaxSlots;  // three register bits: 26 KiB per workgroup
// Four register bits run at most three workgroups per CU (register-bound), so
// their passes can hold twice the trainable gates in 50 KiB of LDS.
constexpr int kMaxSlotsSv4 = 4 * kMaxSlots;
__host__ __device__ constexpr int adj_slot_cap(int rb) { return rb >= 4 ? kMaxSlotsSv4 : kMaxSlotsSv; }
constexpr int kRBForward = 4;  // forward passes (one state)
constexpr int kRBAdjoint = 3;  // adjoint passes (lambda and one observable state)
// Complex128 adjoint passes from this many qubits on take four register bits
// by default: a pass then holds more gates, and the sweep of the 28-qubit
// brick circuit takes 8 passes of 4.50 ms instead of 10 of 4.14 ms.
constexpr int kAdjointWideQubits = 24;

// Device description of one fused gate (22 dwords, loaded whole and
// prefetched one gate ahead).  Controls and parity bits among the register
// bits become two 16-bit masks over the registers (bit k of rok: the
// register-bit controls hold in register k; bit k of rpar: odd parity of
// register k's offset), so each register's test is a compile-time bit of a
// scalar.  The remaining bits are tested once per lane on the tile base.
struct FOp {
    uint32_t meta; // kind | loc << 4 | (slot + 1) << 8 | gkind << 16 | dneg << 20 | mreal << 21 | dsym << 22
    uint32_t rbits; // rok | rpar << 16
    uint64_t cm_rest, pm_rest;
    uint64_t pad;
    cx<double> m[4]; // U (dense) or (d_even, d_odd) in m[0], m[1]
};
// generator of a trainable gate, indexed by its slot
struct FGen {
    cx<double> g[4]; // dense generator, or (g_even, g_odd) in g[0].re, g[1].re
};
enum : uint32_t { kLaneDense = 0, kRegDense = 1, kDiag = 2, kDiagRun = 3 };
enum : uint32_t { kGenDense = 0, kGenDiag = 1, kGenX = 2, kGenY = 3 };
__device__ __forceinline__ uint32_t f_kind(const FOp &o) { return o.meta & 15u; }
__device__ __forceinline__ int f_loc(const FOp &o) { return int((o.meta >> 4) & 15u); }
__device__ __forceinline__ int f_slot(const FOp &o) { return int((o.meta >> 8) & 255u) - 1; }
__device__ __forceinline__ uint32_t f_gkind(const FOp &o) { return (o.meta >> 16) & 15u; }
__device__ __forceinline__ bool f_dneg(const FOp &o) { return (o.meta >> 20) & 1u; }
__device__ __forceinline__ bool f_real(const FOp &o) { return (o.meta >> 21) & 1u; }
// diagonal with d_odd = conj(d_even) (RZ, CRZ, IsingZZ, MultiRZ)
__device__ __forceinline__ bool f_dsym(const FOp &o) { return (o.meta >> 22) & 1u; }
// dense 2x2 equal to [[0, 1], [1, 0]] (X, CNOT, Toffoli)
__device__ __forceinline__ bool f_isx(const FOp &o) { return (o.meta >> 23) & 1u; }
// rotation with a factor moved by the host into a scale of the whole pass:
// 1: [[1, -t], [t, 1]] (t in m[2]), 2: [[t, -1], [1, t]] (t in m[0]), 0: neither
__device__ __forceinline__ int f_rot(const FOp &o) { return int((o.meta >> 28) & 3u); }
// every register satisfies the register-bit controls (the controls, if any,
// are on lane or higher bits)
__device__ __forceinline__ bool f_regs_ok(const FOp &o) { return (o.rbits & 0xffffu) == 0xffffu; }

// Diagonal runs.  Diagonal gates commute, so a run of consecutive diagonal
// gates of a pass is applied as one phase per amplitude: the product of the
// members whose bits are all outside the registers is a per-lane factor q,
// the product of the members whose bits are all register bits is a constant
// P[k] per register (computed on the host), members that negate where their
// controls hold (CZ) with bits on both sides flip a per-register sign mask,
// and the other members are applied as they come.  At the end of the run
// every register is multiplied once by q * P[k] (with its sign).  In the
// adjoint sweep both states get the same diagonal unitaries, which leave
// conj(b_x) a_x unchanged, so every member's generator overlap can be formed
// from w_k = Im(conj(b_k) a_k) at any point of the run: T = sum_k w_k and
// E_B = sum of w_k over the registers with register bit B clear give the
// overlaps of generators that do not depend on the registers or depend on
// one register bit.
//
// The run is stored as a header op (kind kDiagRun; pad = members |
// data ops << 16 | flags << 24), then data ops holding P[k] in their m[]
// (four registers per op, only with kRunHasP), then the members, whose meta
// carries their class (bits 24-25) and overlap form (bits 26-27).
enum : uint32_t { kMemQ = 0, kMemR = 1, kMemS = 2, kMemX = 3 };
enum : uint32_t { kOvT = 0, kOvB = 1, kOvF = 2 };
enum : uint32_t { kRunHasP = 1, kRunNeedsTE = 2, kRunHasS = 4, kRunHasQ = 8 };
__device__ __forceinline__ uint32_t f_mclass(const FOp &o) { return (o.meta >> 24) & 3u; }
__device__ __forceinline__ uint32_t f_oclass(const FOp &o) { return (o.meta >> 26) & 3u; }
__device__ __forceinline__ int run_nmem(const FOp &o) { return int(o.pad & 0xffffu); }
__device__ __forceinline__ int run_ndata(const FOp &o) { return int((o.pad >> 16) & 0xffu); }
__device__ __forceinline__ uint32_t run_flags(const FOp &o) { return uint32_t(o.pad >> 24) & 0xffu; }

// tile bits above the fixed lane bits: at most 5 lane bits and the register bits
constexpr int kMaxFree = kLaneBits - 1 + kMaxRB;

struct PassDesc {
    int32_t nops;
    int32_t nslots;
    int32_t rb;                  // register bits of the pass
    int32_t nfix;                // lane positions 0..nfix-1 carry index bits 0..nfix-1
    int32_t reg_bits[kMaxRB];
    int32_t reg_sorted[kMaxRB];  // register bits in ascending order
    uint64_t reg_off[kMaxNR];    // index offset of each register
    int32_t lane_bits[kLaneBits]; // index bit of each lane position
    int32_t nfree;               // tile bits besides the fixed lane bits
    int32_t pad;
    int32_t free_sorted[kMaxFree]; // those bits in ascending order
    int32_t pad2;
    uint64_t tile_mask;          // all tile bits (lane and register bits)
};

// first index of the tile after the one at g (the next index with zeros at
// every tile bit)
__host__ __device__ __forceinline__ uint64_t next_tile(uint64_t g, uint64_t tile_mask) {
    return ((g | tile_mask) + 1) & ~tile_mask;
}

// First index of tile t when the lanes carry index bits 0..5: t's bits go
// around the lane bits and the register bits, which are all >= 6.
template <int RB> __device__ __forceinline__ uint64_t tile_base(uint64_t t, const int (&rs)[RB]) {
    uint64_t g = t << kLaneBits;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int p = rs[r];
        g = ((g >> p) << (p + 1)) | (g & ((1ull << p) - 1ull));
    }
    return g;
}

// First index of tile t for any lane bits: t's bits go around the fixed lane
// bits (0..nfix-1) and the nfree other tile bits fs[] (ascending).
__host__ __device__ __forceinline__ uint64_t tile_base_free(uint64_t t, int nfix, const int (&fs)[kMaxFree], int nfree) {
    uint64_t g = t << nfix;
#pragma unroll
    for (int r = 0; r < kMaxFree; ++r)
        if (r < nfree) {
            const int p = fs[r];
            g = ((g >> p) << (p + 1)) | (g & ((1ull << p) - 1ull));
        }
    return g;
}

// index offset of a lane: bit p of the lane id goes to index bit lane_bits[p]
__host__ __device__ __forceinline__ uint64_t lane_offset(int lane, const PassDesc *pd) {
    uint64_t off = 0;
#pragma unroll
    for (int p = 0; p < kLaneBits; ++p)
        if ((lane >> p) & 1) off |= 1ull << pd->lane_bits[p];
    return off;
}

template <class R> __device__ __forceinline__ cx<R> lo(cx<double> z) { return cx<R>{R(z.re), R(z.im)}; }
template <class R> __device__ __forceinline__ cx<R> neg(cx<R> a) { return cx<R>{-a.re, -a.im}; }

__device__ __forceinline__ bool reg_ok(const FOp &op, int k) { return (op.rbits >> k) & 1u; }
__device__ __forceinline__ bool reg_par(const FOp &op, int k) { return (op.rbits >> (16 + k)) & 1u; }

// Dense 2x2 on register bit B.
template <class R> __device__ __forceinline__ cx<R> rmul(R s, cx<R> a) { return cx<R>{s * a.re, s * a.im}; }
template <class R> __device__ __forceinline__ cx<R> rfma(R s, cx<R> a, cx<R> c) {
    return cx<R>{fma(s, a.re, c.re), fma(s, a.im, c.im)};
}
template <class R> __device__ __forceinline__ cx<R> sel(bool c, cx<R> x, cx<R> y) {
    return cx<R>{c ? x.re : y.re, c ? x.im : y.im};
}

// A gate without controls updates every amplitude of the tile; the masked
// variants select per register and lane instead of branching.
// (rok has its bits for registers beyond NR set, so the test is the same for any NR)
__device__ __forceinline__ bool f_plain(const FOp &o) { return (o.rbits & 0xffffu) == 0xffffu && o.cm_rest == 0; }

// 2x2 matrix in the working precision; Real keeps only the real parts.
template <class R> struct M2 {
    cx<R> m00, m01, m10, m11;
};
template <class R> __device__ __forceinline__ M2<R> m2(const FOp &op) {
    return M2<R>{lo<R>(op.m[0]), lo<R>(op.m[1]), lo<R>(op.m[2]), lo<R>(op.m[3])};
}
template <class R, bool Real, int Rot = 0>
__device__ __forceinline__ void mat2(const M2<R> &m, cx<R> &x0, cx<R> &x1) {
    cx<R> y0, y1;
    if constexpr (Rot == 1) {
        y0 = rfma(m.m01.re, x1, x0);
        y1 = rfma(m.m10.re, x0, x1);
    } else if constexpr (Rot == 2) {
        y0 = rfma(m.m00.re, x0, neg(x1));
        y1 = rfma(m.m11.re, x1, x0);
    } else if constexpr (Real) {
        y0 = rfma(m.m01.re, x1, rmul(m.m00.re, x0));
        y1 = rfma(m.m11.re, x1, rmul(m.m10.re, x0));
    } else {
        y0 = cfma(m.m01, x1, cmul(m.m00, x0));
        y1 = cfma(m.m11, x1, cmul(m.m10, x0));
    }
    x0 = y0;
    x1 = y1;
}
// lane bit L <= 3: each lane updates its own amplitude from the partner's
// (form 2 on lane bits takes the real path: the partner's sign would differ by lane)
template <class R, bool Real, int Rot = 0>
__device__ __forceinline__ cx<R> row2(cx<R> cs, cx<R> cp, cx<R> own, cx<R> partner) {
    if constexpr (Rot == 1)
        return rfma(cp.re, partner, own);
    else if constexpr (Real)
        return rfma(cp.re, partner, rmul(cs.re, own));
    else
        return cfma(cp, partner, cmul(cs, own));
}

// --- uncontrolled gates -------------------------------------------------
template <class R, int NR, int B, bool Real, int Rot = 0>
__device__ __forceinline__ void reg_plain(const M2<R> &m, cx<R> (&v)[NR]) {
#pragma unroll
    for (int k = 0; k < NR; ++k)
        if (!(k & (1 << B))) mat2<R, Real, Rot>(m, v[k], v[k | (1 << B)]);
}
template <class R, int NR, int L, bool Real, int Rot = 0>
__device__ __forceinline__ void lane_plain(const M2<R> &m, int lane, cx<R> (&v)[NR]) {
    if constexpr (L >= 4) {
#pragma unroll
        for (int k = 0; k < NR; k += 2) {
            row_swap_cx<L>(v[k], v[k + 1]);
            mat2<R, Real, Rot>(m, v[k], v[k + 1]);
            row_swap_cx<L>(v[k], v[k + 1]);
        }
    } else {
        const bool hi = (lane >> L) & 1;
        const cx<R> cs = hi ? m.m11 : m.m00, cp = hi ? m.m10 : m.m01;
#pragma unroll
        for (int k = 0; k < NR; ++k) v[k] = row2<R, Real, Rot == 1 ? 1 : 0>(cs, cp, v[k], xor_lane_cx<L>(v[k], lane));
    }
}

// --- controlled gates (controls on register bits and/or other bits) ------
// AllReg: every register satisfies the register-bit controls; IsX: the
// matrix is X, so the gate swaps the pair where the controls hold.
template <class R, int NR, int B, bool AllReg, bool IsX>
__device__ __forceinline__ void reg_masked(const FOp &op, const M2<R> &m, bool lane_ok, cx<R> (&v)[NR]) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (k & (1 << B)) continue;
        const int k1 = k | (1 << B);
        cx<R> x = v[k], y = v[k1];
        if constexpr (IsX) {
            x = v[k1];
            y = v[k];
        } else {
            mat2<R, false>(m, x, y);
        }
        const bool ok = (AllReg || reg_ok(op, k)) && lane_ok;
        v[k] = sel(ok, x, v[k]);
        v[k1] = sel(ok, y, v[k1]);
    }
}
template <class R, int NR, int L, bool AllReg, bool IsX>
__device__ __forceinline__ void lane_masked(const FOp &op, const M2<R> &m, int lane, bool lane_ok, cx<R> (&v)[NR]) {
    const bool hi = (lane >> L) & 1;
    const cx<R> cs = hi ? m.m11 : m.m00, cp = hi ? m.m10 : m.m01;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const cx<R> p = xor_lane_cx<L>(v[k], lane);
        cx<R> r;
        if constexpr (IsX)
            r = p;
        else
            r = row2<R, false>(cs, cp, v[k], p);
        v[k] = sel((AllReg || reg_ok(op, k)) && lane_ok, r, v[k]);
    }
}

// x with its sign bit xor-ed by m (0 or 0x80000000)
__device__ __forceinline__ float flip(float x, uint32_t m) { return __uint_as_float(__float_as_uint(x) ^ m); }
__device__ __forceinline__ double flip(double x, uint32_t m) {
    return __hiloint2double(int(uint32_t(__double2hiint(x)) ^ m), __double2loint(x));
}
constexpr uint32_t kSignBit = 0x80000000u;
// location code of a diagonal gate whose parity mask has no register bit
constexpr int kDiagNoReg = 7;

// phase (c, s) where the parity of the mask bits is odd, (c, -s) where even;
// the register part of the parity is register bit B (B < 0: none), so each
// register's sign is known at compile time
template <class R> __device__ __forceinline__ cx<R> phase_cs(cx<R> x, R c, R s) {
    return cx<R>{c * x.re - s * x.im, c * x.im + s * x.re};
}
template <class R, int NR, int B>
__device__ __forceinline__ void diag_sym(const FOp &op, bool lpar, cx<R> (&v)[NR]) {
    const R c = R(op.m[1].re), s = R(op.m[1].im);
    const R s0 = flip(s, lpar ? 0u : kSignBit); // registers with even register parity
    const R s1 = flip(s, lpar ? kSignBit : 0u);
#pragma unroll
    for (int k = 0; k < NR; ++k) v[k] = phase_cs(v[k], c, (B >= 0 && ((k >> B) & 1)) ? s1 : s0);
}

template <class R, int NR>
__device__ __forceinline__ void diag_gate(const FOp &op, uint64_t base, bool lane_ok, cx<R> (&v)[NR]) {
    if (f_dneg(op)) {
        // -1 where the controls hold (CZ, Z): flip the sign bits
        const uint32_t ml = lane_ok ? kSignBit : 0u;
        if (f_regs_ok(op)) {
#pragma unroll
            for (int k = 0; k < NR; ++k) v[k] = cx<R>{flip(v[k].re, ml), flip(v[k].im, ml)};
        } else {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const uint32_t m = reg_ok(op, k) ? ml : 0u;
                v[k] = cx<R>{flip(v[k].re, m), flip(v[k].im, m)};
            }
        }
        return;
    }
    const bool lpar = __popcll(base & op.pm_rest) & 1;
    if (f_dsym(op) && f_plain(op)) {
        switch (f_loc(op)) {
        case 0: diag_sym<R, NR, 0>(op, lpar, v); return;
        case 1: diag_sym<R, NR, 1>(op, lpar, v); return;
        case 2: diag_sym<R, NR, 2>(op, lpar, v); return;
        case 3:
            if constexpr (NR > 8) {
                diag_sym<R, NR, 3>(op, lpar, v);
                return;
            }
            break;
        case kDiagNoReg: diag_sym<R, NR, -1>(op, lpar, v); return;
        default: break;
        }
        // parity over several register bits
        const R c = R(op.m[1].re), s = R(op.m[1].im);
        const uint32_t ml = lpar ? 0u : kSignBit;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const R sk = flip(s, reg_par(op, k) ? (ml ^ kSignBit) : ml);
            const cx<R> x = v[k];
            v[k] = cx<R>{c * x.re - sk * x.im, c * x.im + sk * x.re};
        }
        return;
    }
    cx<R> pe = lo<R>(lpar ? op.m[1] : op.m[0]); // phase when the register part is even
    cx<R> po = lo<R>(lpar ? op.m[0] : op.m[1]);
    if (!lane_ok) pe = po = cx<R>{R(1), R(0)};
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const cx<R> ph = reg_ok(op, k) ? (reg_par(op, k) ? po : pe) : cx<R>{R(1), R(0)};
        v[k] = cmul(v[k], ph);
    }
}

template <class R, int NR, int B> __device__ __forceinline__ void reg_gate(const FOp &op, bool lane_ok, cx<R> (&v)[NR]) {
    const M2<R> m = m2<R>(op);
    if (!f_plain(op)) {
        if (f_isx(op)) {
            if (f_regs_ok(op))
                reg_masked<R, NR, B, true, true>(op, m, lane_ok, v);
            else
                reg_masked<R, NR, B, false, true>(op, m, lane_ok, v);
        } else {
            if (f_regs_ok(op))
                reg_masked<R, NR, B, true, false>(op, m, lane_ok, v);
            else
                reg_masked<R, NR, B, false, false>(op, m, lane_ok, v);
        }
    }
    else if (f_rot(op) == 1)
        reg_plain<R, NR, B, true, 1>(m, v);
    else if (f_rot(op) == 2)
        reg_plain<R, NR, B, true, 2>(m, v);
    else if (f_real(op))
        reg_plain<R, NR, B, true>(m, v);
    else
        reg_plain<R, NR, B, false>(m, v);
}
template <class R, int NR, int L>
__device__ __forceinline__ void lane_gate(const FOp &op, int lane, bool lane_ok, cx<R> (&v)[NR]) {
    const M2<R> m = m2<R>(op);
    if (!f_plain(op)) {
        if (f_isx(op)) {
            if (f_regs_ok(op))
                lane_masked<R, NR, L, true, true>(op, m, lane, lane_ok, v);
            else
                lane_masked<R, NR, L, false, true>(op, m, lane, lane_ok, v);
        } else {
            if (f_regs_ok(op))
                lane_masked<R, NR, L, true, false>(op, m, lane, lane_ok, v);
            else
                lane_masked<R, NR, L, false, false>(op, m, lane, lane_ok, v);
        }
    }
    else if (f_rot(op) == 1)
        lane_plain<R, NR, L, true, 1>(m, lane, v);
    else if (L >= 4 && f_rot(op) == 2)
        lane_plain<R, NR, L, true, 2>(m, lane, v);
    else if (f_real(op))
        lane_plain<R, NR, L, true>(m, lane, v);
    else
        lane_plain<R, NR, L, false>(m, lane, v);
}

template <class R, int NR>
__device__ __forceinline__ void apply_fop(const FOp &op, int lane, uint64_t base, cx<R> (&v)[NR]) {
    const bool lane_ok = (base & op.cm_rest) == op.cm_rest;
    const uint32_t kind = f_kind(op);
    const int loc = f_loc(op);
    if (kind == kDiag) {
        diag_gate<R, NR>(op, base, lane_ok, v);
    } else if (kind == kRegDense) {
        switch (loc) {
        case 0: reg_gate<R, NR, 0>(op, lane_ok, v); break;
        case 1: reg_gate<R, NR, 1>(op, lane_ok, v); break;
        case 2: reg_gate<R, NR, 2>(op, lane_ok, v); break;
        default:
            if constexpr (NR > 8) reg_gate<R, NR, 3>(op, lane_ok, v);
            break;
        }
    } else {
        switch (loc) {
        case 0: lane_gate<R, NR, 0>(op, lane, lane_ok, v); break;
        case 1: lane_gate<R, NR, 1>(op, lane, lane_ok, v); break;
        case 2: lane_gate<R, NR, 2>(op, lane, lane_ok, v); break;
        case 3: lane_gate<R, NR, 3>(op, lane, lane_ok, v); break;
        case 4: lane_gate<R, NR, 4>(op, lane, lane_ok, v); break;
        default: lane_gate<R, NR, 5>(op, lane, lane_ok, v); break;
        }
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <class R> __device__ __forceinline__ cx<double> up(cx<R> a) { return cx<double>{double(a.re), double(a.im)}; }
template <class R> __device__ __forceinline__ double im_cdot(cx<R> b, cx<double> q) {
    return double(b.re) * q.im - double(b.im) * q.re;
}
template <class R> __device__ __forceinline__ double re_cdot(cx<R> b, cx<R> a) {
    return double(b.re) * a.re + double(b.im) * a.im;
}
template <class R> __device__ __forceinline__ double im_cdot2(cx<R> b, cx<R> a) {
    return double(b.re) * a.im - double(b.im) * a.re;
}

// Im(conj(b0) (G a)_0 + conj(b1) (G a)_1) for a 2x2 generator.  G = X and
// G = Y avoid the general complex products:
//   X: Im(conj(b0) a1 + conj(b1) a0);   Y: Re(conj(b1) a0) - Re(conj(b0) a1).
template <class R>
__device__ __forceinline__ double pair_dot(uint32_t gk, const FGen &g, cx<R> a0, cx<R> a1, cx<R> b0, cx<R> b1) {
    if (gk == kGenX) return im_cdot2(b0, a1) + im_cdot2(b1, a0);
    if (gk == kGenY) return re_cdot(b1, a0) - re_cdot(b0, a1);
    const cx<double> x0 = up(a0), x1 = up(a1);
    return im_cdot(b0, cfma(g.g[1], x1, cmul(g.g[0], x0))) + im_cdot(b1, cfma(g.g[3], x1, cmul(g.g[2], x0)));
}
// s plus that overlap; for G = X and G = Y one chain of four FMAs
template <class R>
__device__ __forceinline__ double pair_dot_acc(uint32_t gk, const FGen &g, cx<R> a0, cx<R> a1, cx<R> b0, cx<R> b1,
                                               double s) {
    if (gk == kGenX) { // b0.re a1.im - b0.im a1.re + b1.re a0.im - b1.im a0.re
        s = fma(double(b0.re), double(a1.im), s);
        s = fma(-double(b0.im), double(a1.re), s);
        s = fma(double(b1.re), double(a0.im), s);
        return fma(-double(b1.im), double(a0.re), s);
    }
    if (gk == kGenY) { // b1.re a0.re + b1.im a0.im - b0.re a1.re - b0.im a1.im
        s = fma(double(b1.re), double(a0.re), s);
        s = fma(double(b1.im), double(a0.im), s);
        s = fma(-double(b0.re), double(a1.re), s);
        return fma(-double(b0.im), double(a1.im), s);
    }
    return s + pair_dot<R>(gk, g, a0, a1, b0, b1);
}
// the same with the generator kind fixed at compile time (GK < 0: read gk)
template <int GK, class R>
__device__ __forceinline__ double pair_dot_acc_k(uint32_t gk, const FGen &g, cx<R> a0, cx<R> a1, cx<R> b0,
                                                 cx<R> b1, double s) {
    if constexpr (GK >= 0) gk = uint32_t(GK);
    return pair_dot_acc<R>(gk, g, a0, a1, b0, b1, s);
}
// lane bit L <= 3: one lane's row of the pair sum; own = its amplitude,
// partner = the amplitude in lane ^ 2^L, hi = the lane's bit L
template <class R>
__device__ __forceinline__ double row_dot(uint32_t gk, const FGen &g, bool hi, cx<R> own, cx<R> partner, cx<R> b) {
    if (gk == kGenX) return im_cdot2(b, partner);
    if (gk == kGenY) return hi ? re_cdot(b, partner) : -re_cdot(b, partner);
    const cx<double> gs = hi ? g.g[3] : g.g[0], gp = hi ? g.g[2] : g.g[1];
    return im_cdot(b, cfma(gp, up(partner), cmul(gs, up(own))));
}
// diagonal generator: eigenvalue g_even / g_odd by the parity of the mask bits
template <class R, int NR>
__device__ __forceinline__ double diag_dot(const FOp &op, const FGen &g, uint64_t base, bool lane_ok,
                                           const cx<R> (&a)[NR], const cx<R> (&b)[NR]) {
    const bool lpar = __popcll(base & op.pm_rest) & 1;
    const double ge = lpar ? g.g[1].re : g.g[0].re, go = lpar ? g.g[0].re : g.g[1].re;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (!reg_ok(op, k)) continue;
        s += (reg_par(op, k) ? go : ge) * im_cdot2(b[k], a[k]);
    }
    return lane_ok ? s : 0.0;
}

// Generator overlap of a controlled gate (either state unchanged).
template <class R, int NR, int B>
__device__ __forceinline__ double reg_dot_masked(const FOp &op, const FGen &g, bool lane_ok, const cx<R> (&a)[NR],
                                                 const cx<R> (&b)[NR]) {
    const uint32_t gk = f_gkind(op);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (k & (1 << B)) continue;
        if (!reg_ok(op, k)) continue;
        const int k1 = k | (1 << B);
        s += pair_dot<R>(gk, g, a[k], a[k1], b[k], b[k1]);
    }
    return lane_ok ? s : 0.0;
}
template <class R, int NR, int L>
__device__ __forceinline__ double lane_dot_masked(const FOp &op, const FGen &g, int lane, bool lane_ok,
                                                  const cx<R> (&a)[NR], const cx<R> (&b)[NR]) {
    const bool hi = (lane >> L) & 1;
    const uint32_t gk = f_gkind(op);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (!reg_ok(op, k)) continue;
        s += row_dot<R>(gk, g, hi, a[k], xor_lane_cx<L>(a[k], lane), b[k]);
    }
    return lane_ok ? s : 0.0;
}

// One reverse step of an uncontrolled dense gate on (lambda, bra): the
// generator overlap (with Gen) and U^dagger on both states, sharing the lane
// exchanges between them.
// GK >= 0: the generator kind is known at compile time.
template <class R, int NR, int B, bool Real, bool Gen, int Rot = 0, int GK = -1>
__device__ __forceinline__ double adj_reg_plain(const M2<R> &m, uint32_t gk, const FGen *g, cx<R> (&a)[NR],
                                                cx<R> (&b)[NR]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (k & (1 << B)) continue;
        const int k1 = k | (1 << B);
        if constexpr (Gen) s = pair_dot_acc_k<GK, R>(gk, *g, a[k], a[k1], b[k], b[k1], s);
        mat2<R, Real, Rot>(m, a[k], a[k1]);
        mat2<R, Real, Rot>(m, b[k], b[k1]);
    }
    return s;
}
template <class R, int NR, int L, bool Real, bool Gen, int Rot = 0, int GK = -1>
__device__ __forceinline__ double adj_lane_plain(const M2<R> &m, uint32_t gk, const FGen *g, int lane,
                                                 cx<R> (&a)[NR], cx<R> (&b)[NR]) {
    double s = 0.0;
    if constexpr (L >= 4) {
#pragma unroll
        for (int k = 0; k < NR; k += 2) {
            row_swap_cx<L>(a[k], a[k + 1]);
            row_swap_cx<L>(b[k], b[k + 1]);
            if constexpr (Gen) s = pair_dot_acc_k<GK, R>(gk, *g, a[k], a[k + 1], b[k], b[k + 1], s);
            mat2<R, Real, Rot>(m, a[k], a[k + 1]);
            mat2<R, Real, Rot>(m, b[k], b[k + 1]);
            row_swap_cx<L>(a[k], a[k + 1]);
            row_swap_cx<L>(b[k], b[k + 1]);
        }
    } else {
        const bool hi = (lane >> L) & 1;
        const cx<R> cs = hi ? m.m11 : m.m00, cp = hi ? m.m10 : m.m01;
        auto step = [&](int k, const cx<R> &pa) {
            a[k] = row2<R, Real, Rot == 1 ? 1 : 0>(cs, cp, a[k], pa);
            b[k] = row2<R, Real, Rot == 1 ? 1 : 0>(cs, cp, b[k], xor_lane_cx<L>(b[k], lane));
        };
        if constexpr (Gen) {
            // X and Y generators: one chain of two FMAs per register; the Y
            // row's sign depends only on the lane and is applied once
            if (gk == kGenY) {
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const cx<R> pa = xor_lane_cx<L>(a[k], lane);
                    s = fma(double(b[k].re), double(pa.re), s);
                    s = fma(double(b[k].im), double(pa.im), s);
                    step(k, pa);
                }
                return hi ? s : -s;
            }
            if (gk == kGenX) {
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const cx<R> pa = xor_lane_cx<L>(a[k], lane);
                    s = fma(double(b[k].re), double(pa.im), s);
                    s = fma(-double(b[k].im), double(pa.re), s);
                    step(k, pa);
                }
                return s;
            }
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const cx<R> pa = xor_lane_cx<L>(a[k], lane);
            if constexpr (Gen) s += row_dot<R>(gk, *g, hi, a[k], pa, b[k]);
            step(k, pa);
        }
    }
    return s;
}

// the three matrix forms of an uncontrolled dense step
template <class R, int NR, int B, bool Gen>
__device__ __forceinline__ double adj_reg_any(const M2<R> &m, int rot, bool real, uint32_t gk, const FGen *g,
                                              cx<R> (&a)[NR], cx<R> (&b)[NR]) {
    // the rotation forms of RY (Y generator) get the overlap without a
    // per-pair test of the generator kind
    if (rot == 1) {
        if (Gen && gk == kGenY) return adj_reg_plain<R, NR, B, true, Gen, 1, kGenY>(m, gk, g, a, b);
        return adj_reg_plain<R, NR, B, true, Gen, 1>(m, gk, g, a, b);
    }
    if (rot == 2) {
        if (Gen && gk == kGenY) return adj_reg_plain<R, NR, B, true, Gen, 2, kGenY>(m, gk, g, a, b);
        return adj_reg_plain<R, NR, B, true, Gen, 2>(m, gk, g, a, b);
    }
    if (real) return adj_reg_plain<R, NR, B, true, Gen>(m, gk, g, a, b);
    return adj_reg_plain<R, NR, B, false, Gen>(m, gk, g, a, b);
}
template <class R, int NR, int L, bool Gen>
__device__ __forceinline__ double adj_lane_any(const M2<R> &m, int rot, bool real, uint32_t gk, const FGen *g,
                                               int lane, cx<R> (&a)[NR], cx<R> (&b)[NR]) {
    if (rot == 1) {
        if constexpr (L >= 4)
            if (Gen && gk == kGenY) return adj_lane_plain<R, NR, L, true, Gen, 1, kGenY>(m, gk, g, lane, a, b);
        return adj_lane_plain<R, NR, L, true, Gen, 1>(m, gk, g, lane, a, b);
    }
    if constexpr (L >= 4)
        if (rot == 2) {
            if (Gen && gk == kGenY) return adj_lane_plain<R, NR, L, true, Gen, 2, kGenY>(m, gk, g, lane, a, b);
            return adj_lane_plain<R, NR, L, true, Gen, 2>(m, gk, g, lane, a, b);
        }
    if (real) return adj_lane_plain<R, NR, L, true, Gen>(m, gk, g, lane, a, b);
    return adj_lane_plain<R, NR, L, false, Gen>(m, gk, g, lane, a, b);
}

template <class R, int NR, bool Gen>
__device__ __forceinline__ double adj_plain_dense(const FOp &op, const FGen *g, int lane, cx<R> (&a)[NR],
                                                  cx<R> (&b)[NR]) {
    const M2<R> m = m2<R>(op);
    const uint32_t gk = f_gkind(op);
    const bool real = f_real(op);
    const int rot = f_rot(op);
    const int loc = f_loc(op);
    if (f_kind(op) == kRegDense) {
        switch (loc) {
        case 0: return adj_reg_any<R, NR, 0, Gen>(m, rot, real, gk, g, a, b);
        case 1: return adj_reg_any<R, NR, 1, Gen>(m, rot, real, gk, g, a, b);
        case 2: return adj_reg_any<R, NR, 2, Gen>(m, rot, real, gk, g, a, b);
        default:
            if constexpr (NR > 8) return adj_reg_any<R, NR, 3, Gen>(m, rot, real, gk, g, a, b);
            return 0.0;
        }
    }
    switch (loc) {
    case 0: return adj_lane_any<R, NR, 0, Gen>(m, rot, real, gk, g, lane, a, b);
    case 1: return adj_lane_any<R, NR, 1, Gen>(m, rot, real, gk, g, lane, a, b);
    case 2: return adj_lane_any<R, NR, 2, Gen>(m, rot, real, gk, g, lane, a, b);
    case 3: return adj_lane_any<R, NR, 3, Gen>(m, rot, real, gk, g, lane, a, b);
    case 4: return adj_lane_any<R, NR, 4, Gen>(m, rot, real, gk, g, lane, a, b);
    default: return adj_lane_any<R, NR, 5, Gen>(m, rot, real, gk, g, lane, a, b);
    }
}

// Adjoint: a lean pass's dense gate (uncontrolled: rotation form 1 or 2, a
// plain real matrix (F = 0) or a complex one (F = 3); trainable ones with an
// X or Y generator).
template <class R, int NR, int B, bool Gen, int F>
__device__ __forceinline__ double adj_reg_lean(const M2<R> &m, uint32_t gk, cx<R> (&a)[NR], cx<R> (&b)[NR]) {
    const FGen g{};
    constexpr bool Real = F != 3;
    constexpr int Rot = F == 3 ? 0 : F;
    if constexpr (Gen) {
        if (gk == kGenY) return adj_reg_plain<R, NR, B, Real, true, Rot, kGenY>(m, gk, &g, a, b);
        return adj_reg_plain<R, NR, B, Real, true, Rot, kGenX>(m, gk, &g, a, b);
    }
    return adj_reg_plain<R, NR, B, Real, false, Rot>(m, gk, &g, a, b);
}
template <class R, int NR, int L, bool Gen, int F>
__device__ __forceinline__ double adj_lane_lean(const M2<R> &m, uint32_t gk, int lane, cx<R> (&a)[NR],
                                                cx<R> (&b)[NR]) {
    const FGen g{};
    constexpr bool Real = F != 3;
    constexpr int Rot = F == 3 ? 0 : F;
    if constexpr (L >= 4 && Gen) {
        if (gk == kGenY) return adj_lane_plain<R, NR, L, Real, true, Rot, kGenY>(m, gk, &g, lane, a, b);
        return adj_lane_plain<R, NR, L, Real, true, Rot, kGenX>(m, gk, &g, lane, a, b);
    }
    // lane bits 0-3 take form 2 through the real path
    return adj_lane_plain<R, NR, L, Real, Gen, (L < 4 && Rot == 2) ? 0 : Rot>(m, gk, &g, lane, a, b);
}
// `sink` takes a trainable gate's overlap inside the gate's own branch: with
// one call after the dispatch, every branch's tile results met in a common
// block, where the register allocator gave the tile a second register set
// and copied it in and out (32-64 moves per trainable gate).
template <class R, int NR, bool Gen, int Lv, class Sink>
__device__ __forceinline__ void adj_plain_lean(const FOp &op, int lane, cx<R> (&a)[NR], cx<R> (&b)[NR],
                                               const Sink &sink) {
    const M2<R> m = m2<R>(op);
    const uint32_t gk = f_gkind(op);
    const int rot = f_rot(op), loc = f_loc(op);
    const int form = rot ? rot : ((Lv == 1 || f_real(op)) ? 0 : 3);
#define LGH_REG3(B)                                                                                                    \
    (form == 1   ? sink(adj_reg_lean<R, NR, B, Gen, 1>(m, gk, a, b))                                                 \
     : form == 2 ? sink(adj_reg_lean<R, NR, B, Gen, 2>(m, gk, a, b))                                                 \
     : form == 0 ? sink(adj_reg_lean<R, NR, B, Gen, 0>(m, gk, a, b))                                                 \
                 : sink(adj_reg_lean<R, NR, B, Gen, Lv == 2 ? 3 : 0>(m, gk, a, b)))
#define LGH_LANE3(L)                                                                                                   \
    (form == 1   ? sink(adj_lane_lean<R, NR, L, Gen, 1>(m, gk, lane, a, b))                                          \
     : form == 2 ? sink(adj_lane_lean<R, NR, L, Gen, 2>(m, gk, lane, a, b))                                          \
     : form == 0 ? sink(adj_lane_lean<R, NR, L, Gen, 0>(m, gk, lane, a, b))                                          \
                 : sink(adj_lane_lean<R, NR, L, Gen, Lv == 2 ? 3 : 0>(m, gk, lane, a, b)))
    if (f_kind(op) == kRegDense) {
        switch (loc) {
        case 0: LGH_REG3(0); return;
        case 1: LGH_REG3(1); return;
        case 2: LGH_REG3(2); return;
        default:
            if constexpr (NR > 8) LGH_REG3(3);
            return;
        }
    }
    switch (loc) {
    case 0: LGH_LANE3(0); return;
    case 1: LGH_LANE3(1); return;
    case 2: LGH_LANE3(2); return;
    case 3: LGH_LANE3(3); return;
    case 4: LGH_LANE3(4); return;
    default: LGH_LANE3(5); return;
    }
#undef LGH_REG3
#undef LGH_LANE3
}

// One reverse step of an uncontrolled RZ-type diagonal gate on (lambda, bra):
// the overlap with a diagonal generator (eigenvalue g_even / g_odd by the
// parity of the mask bits) and the phase on both states, the lane parity
// computed once.
template <class R, int NR, int B, bool Gen>
__device__ __forceinline__ double adj_diag_sym(const FOp &op, double ge, double go, uint64_t base, cx<R> (&a)[NR],
                                               cx<R> (&b)[NR]) {
    const bool lpar = __popcll(base & op.pm_rest) & 1;
    const R c = R(op.m[1].re), s = R(op.m[1].im);
    const R s0 = flip(s, lpar ? 0u : kSignBit), s1 = flip(s, lpar ? kSignBit : 0u);
    const double g0 = lpar ? go : ge; // registers with even register parity
    const double g1 = lpar ? ge : go;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const bool odd = B >= 0 && ((k >> B) & 1);
        if constexpr (Gen) acc += (odd ? g1 : g0) * im_cdot2(b[k], a[k]);
        a[k] = phase_cs(a[k], c, odd ? s1 : s0);
        b[k] = phase_cs(b[k], c, odd ? s1 : s0);
    }
    return acc;
}

// true when adj_diag_step handles the gate
__device__ __forceinline__ bool diag_fast(const FOp &op) {
    const int loc = f_loc(op);
    return f_kind(op) == kDiag && f_dsym(op) && f_plain(op) && (loc < 4 || loc == kDiagNoReg) &&
           (f_slot(op) < 0 || f_gkind(op) == kGenDiag);
}
// ge, go: generator eigenvalues for even / odd parity (unused without Gen)
template <class R, int NR, bool Gen>
__device__ __forceinline__ double adj_diag_step(const FOp &op, double ge, double go, uint64_t base, cx<R> (&a)[NR],
                                                cx<R> (&b)[NR]) {
    switch (f_loc(op)) {
    case 0: return adj_diag_sym<R, NR, 0, Gen>(op, ge, go, base, a, b);
    case 1: return adj_diag_sym<R, NR, 1, Gen>(op, ge, go, base, a, b);
    case 2: return adj_diag_sym<R, NR, 2, Gen>(op, ge, go, base, a, b);
    case 3:
        if constexpr (NR > 8) return adj_diag_sym<R, NR, 3, Gen>(op, ge, go, base, a, b);
        return 0.0;
    default: return adj_diag_sym<R, NR, -1, Gen>(op, ge, go, base, a, b);
    }
}

// Generator overlap before the gate for the general (controlled or
// diagonal) case.
template <class R, int NR>
__device__ __forceinline__ double gen_dot(const FOp &op, const FGen &g, int lane, uint64_t base,
                                          const cx<R> (&a)[NR], const cx<R> (&b)[NR]) {
    const bool lane_ok = (base & op.cm_rest) == op.cm_rest;
    if (f_gkind(op) == kGenDiag) return diag_dot<R, NR>(op, g, base, lane_ok, a, b);
    const int loc = f_loc(op);
    if (f_kind(op) == kRegDense) {
        switch (loc) {
        case 0: return reg_dot_masked<R, NR, 0>(op, g, lane_ok, a, b);
        case 1: return reg_dot_masked<R, NR, 1>(op, g, lane_ok, a, b);
        case 2: return reg_dot_masked<R, NR, 2>(op, g, lane_ok, a, b);
        default:
            if constexpr (NR > 8) return reg_dot_masked<R, NR, 3>(op, g, lane_ok, a, b);
            return 0.0;
        }
    }
    switch (loc) {
    case 0: return lane_dot_masked<R, NR, 0>(op, g, lane, lane_ok, a, b);
    case 1: return lane_dot_masked<R, NR, 1>(op, g, lane, lane_ok, a, b);
    case 2: return lane_dot_masked<R, NR, 2>(op, g, lane, lane_ok, a, b);
    case 3: return lane_dot_masked<R, NR, 3>(op, g, lane, lane_ok, a, b);
    case 4: return lane_dot_masked<R, NR, 4>(op, g, lane, lane_ok, a, b);
    default: return lane_dot_masked<R, NR, 5>(op, g, lane, lane_ok, a, b);
    }
}

// Gate descriptors are read through the constant address space: the loads
// are then scalar, and everything decided from them stays uniform.
typedef __attribute__((address_space(4))) const FOp CFOp;
typedef __attribute__((address_space(4))) const FGen CFGen;
__device__ __forceinline__ FOp load_op(const CFOp *p) {
    FOp f;
    f.meta = p->meta;
    f.rbits = p->rbits;
    f.cm_rest = p->cm_rest;
    f.pm_rest = p->pm_rest;
    f.pad = p->pad;
#pragma unroll
    for (int q = 0; q < 4; ++q) f.m[q] = cx<double>{p->m[q].re, p->m[q].im};
    return f;
}
// Makes the compiler wait for the descriptor op (an empty asm reading its
// first and last dwords).  Called on a descriptor loaded before a loop:
// without it the wait for that load lands at the loop header, where it also
// waits for the loop's own one-ahead prefetch, so the prefetch would hide
// nothing.
__device__ __forceinline__ void wait_op(const FOp &op) {
    asm volatile("" ::"s"(op.meta), "s"(op.rbits), "s"(__double2hiint(op.m[3].im)), "s"(op.pad));
}

__device__ __forceinline__ FGen load_gen(const CFGen *p) {
    FGen g;
#pragma unroll
    for (int q = 0; q < 4; ++q) g.g[q] = cx<double>{p->g[q].re, p->g[q].im};
    return g;
}

// Trainable diagonal gates carry their generator eigenvalues (even / odd
// parity) in m[2].re and m[3].re, next to their phases, so the adjoint sweep
// reads no separate generator record for them.
__device__ __forceinline__ FGen diag_gen(const FOp &op) {
    FGen g{};
    g.g[0].re = op.m[2].re;
    g.g[1].re = op.m[3].re;
    return g;
}

// --- lean passes -------------------------------------------------------------
// A pass whose dense gates are all uncontrolled 2x2 matrices (the two
// rotation forms, plain real or complex matrices; trainable ones with an X
// or Y generator) or controlled X gates (not trainable), and whose diagonal gates are all members of diagonal runs
// without general members (or a pass's closing real scale), runs in kernels
// compiled with only those cases.  With the controlled and complex cases compiled in,
// the compiler keeps the loop-carried tile in one set of registers and each
// gate's results in another, and copies the whole tile back after every
// gate; without them the gates update the tile where it lies.
// a real factor on every amplitude (the closing scale gate of a pass
// without diagonal runs)
inline bool pure_scale(const FOp &f) {
    return (f.meta & 15u) == kDiag && int((f.meta >> 4) & 15u) == kDiagNoReg && f.pm_rest == 0 &&
           f.cm_rest == 0 && (f.rbits & 0xffffu) == 0xffffu && f.m[0].im == 0.0 && f.m[1].re == f.m[0].re &&
           f.m[1].im == 0.0 && int((f.meta >> 8) & 255u) == 0;
}
// Lean level of a built op: 1 for the forms of level-1 kernels (uncontrolled
// real 2x2 gates, diagonal runs, a pass's closing real scale), 2 for the
// further forms of level-2 kernels (uncontrolled complex 2x2 gates,
// controlled X gates, phases symmetric under conjugation on at most one
// register bit), 0 for anything else.  Trainable dense gates need an X or Y
// generator, trainable phases a diagonal one.  Each level is a kernel of its
// own: the level-2 cases cost the level-1 passes a little.
inline int lean_level_op(const FOp &f, bool adjoint) {
    const uint32_t kind = f.meta & 15u;
    const bool trainable = adjoint && int((f.meta >> 8) & 255u) > 0;
    if (kind == kDiagRun) return 1; // members are checked separately
    if (kind == kDiag) {
        if (pure_scale(f)) return 1;
        const int loc = int((f.meta >> 4) & 15u);
        const bool dsym = (f.meta >> 22) & 1u, plain = (f.rbits & 0xffffu) == 0xffffu && f.cm_rest == 0;
        return dsym && plain && (loc < 4 || loc == kDiagNoReg) && (!trainable || ((f.meta >> 16) & 15u) == kGenDiag)
                   ? 2
                   : 0;
    }
    if (kind != kRegDense && kind != kLaneDense) return 0;
    if ((f.rbits & 0xffffu) != 0xffffu || f.cm_rest != 0) // controls: only a controlled X
        return ((f.meta >> 23) & 1u) && !trainable ? 2 : 0;
    if (trainable) {
        const uint32_t gk = (f.meta >> 16) & 15u;
        if (gk != kGenX && gk != kGenY) return 0;
    }
    const int rot = int((f.meta >> 28) & 3u);
    const bool real = (f.meta >> 21) & 1u;
    return rot != 0 || real ? 1 : 2;
}
inline bool lean_member(const FOp &m, bool adjoint) {
    const uint32_t mc = (m.meta >> 24) & 3u;
    if (mc == kMemX) return false;
    if (adjoint && int((m.meta >> 8) & 255u) > 0 && ((m.meta >> 26) & 3u) == kOvF) return false;
    return true;
}
// the lowest lean level whose kernels cover every op of a pass (0: none)
inline int lean_level(const std::vector<FOp> &ops, bool adjoint) {
    int level = 1;
    for (size_t j = 0; j < ops.size();) {
        const FOp &f = ops[j];
        const int l = lean_level_op(f, adjoint);
        if (l == 0) return 0;
        level = l > level ? l : level;
        if ((f.meta & 15u) == kDiagRun) {
            const size_t nm = size_t(f.pad & 0xffffu), nd = size_t((f.pad >> 16) & 0xffu);
            for (size_t q = 0; q < nm; ++q)
                if (!lean_member(ops[j + 1 + nd + q], adjoint)) return 0;
            j += 1 + nd + nm;
        } else {
            ++j;
        }
    }
    return level;
}

// a controlled X (CNOT, Toffoli): the pair swaps where the controls hold
template <class R, int NR>
__device__ __forceinline__ void cx_lean(const FOp &op, const M2<R> &m, int lane, uint64_t base, cx<R> (&v)[NR]) {
    const bool lane_ok = (base & op.cm_rest) == op.cm_rest;
    const int loc = f_loc(op);
    if (f_kind(op) == kRegDense) {
        switch (loc) {
        case 0: reg_masked<R, NR, 0, false, true>(op, m, lane_ok, v); break;
        case 1: reg_masked<R, NR, 1, false, true>(op, m, lane_ok, v); break;
        case 2: reg_masked<R, NR, 2, false, true>(op, m, lane_ok, v); break;
        default:
            if constexpr (NR > 8) reg_masked<R, NR, 3, false, true>(op, m, lane_ok, v);
            break;
        }
        return;
    }
    switch (loc) {
    case 0: lane_masked<R, NR, 0, false, true>(op, m, lane, lane_ok, v); break;
    case 1: lane_masked<R, NR, 1, false, true>(op, m, lane, lane_ok, v); break;
    case 2: lane_masked<R, NR, 2, false, true>(op, m, lane, lane_ok, v); break;
    case 3: lane_masked<R, NR, 3, false, true>(op, m, lane, lane_ok, v); break;
    case 4: lane_masked<R, NR, 4, false, true>(op, m, lane, lane_ok, v); break;
    default: lane_masked<R, NR, 5, false, true>(op, m, lane, lane_ok, v); break;
    }
}

// forward: a lean pass's dense gate
template <class R, int NR> __device__ __forceinline__ void scale_lean(const FOp &op, cx<R> (&v)[NR]) {
    const R c = R(op.m[0].re);
#pragma unroll
    for (int k = 0; k < NR; ++k) v[k] = rmul(c, v[k]);
}
template <class R, int NR, int Lv>
__device__ __forceinline__ void apply_lean(const FOp &op, int lane, uint64_t base, cx<R> (&v)[NR]) {
    if (f_kind(op) == kDiag) {
        if (Lv == 1 || !f_dsym(op)) {
            scale_lean<R, NR>(op, v); // a pure scale
            return;
        }
        const bool lpar = __popcll(base & op.pm_rest) & 1;
        switch (f_loc(op)) {
        case 0: diag_sym<R, NR, 0>(op, lpar, v); return;
        case 1: diag_sym<R, NR, 1>(op, lpar, v); return;
        case 2: diag_sym<R, NR, 2>(op, lpar, v); return;
        case 3:
            if constexpr (NR > 8) diag_sym<R, NR, 3>(op, lpar, v);
            return;
        default: diag_sym<R, NR, -1>(op, lpar, v); return;
        }
    }
    const M2<R> m = m2<R>(op);
    const int loc = f_loc(op), rot = f_rot(op);
    const bool real = Lv == 1 || f_real(op);
    if constexpr (Lv == 2) {
        if (!f_plain(op)) {
            cx_lean<R, NR>(op, m, lane, base, v);
            return;
        }
    }
    if (f_kind(op) == kRegDense) {
#define LGH_REG_ARM(B)                                                                                                 \
    if (rot == 1)                                                                                                      \
        reg_plain<R, NR, B, true, 1>(m, v);                                                                            \
    else if (rot == 2)                                                                                                 \
        reg_plain<R, NR, B, true, 2>(m, v);                                                                            \
    else if (real)                                                                                                     \
        reg_plain<R, NR, B, true>(m, v);                                                                               \
    else                                                                                                               \
        reg_plain<R, NR, B, false>(m, v);
        switch (loc) {
        case 0: LGH_REG_ARM(0) break;
        case 1: LGH_REG_ARM(1) break;
        case 2: LGH_REG_ARM(2) break;
        default:
            if constexpr (NR > 8) { LGH_REG_ARM(3) }
            break;
        }
#undef LGH_REG_ARM
        return;
    }
#define LGH_LANE_ARM(L)                                                                                                \
    if (rot == 1)                                                                                                      \
        lane_plain<R, NR, L, true, 1>(m, lane, v);                                                                     \
    else if (L >= 4 && rot == 2)                                                                                       \
        lane_plain<R, NR, L, true, 2>(m, lane, v);                                                                     \
    else if (real)                                                                                                     \
        lane_plain<R, NR, L, true>(m, lane, v);                                                                        \
    else                                                                                                               \
        lane_plain<R, NR, L, false>(m, lane, v);
    switch (loc) {
    case 0: LGH_LANE_ARM(0) break;
    case 1: LGH_LANE_ARM(1) break;
    case 2: LGH_LANE_ARM(2) break;
    case 3: LGH_LANE_ARM(3) break;
    case 4: LGH_LANE_ARM(4) break;
    default: LGH_LANE_ARM(5) break;
    }
#undef LGH_LANE_ARM
}

// --- diagonal runs ---------------------------------------------------------
// per-lane factor of a member without register bits
template <class R> __device__ __forceinline__ cx<R> run_q(const FOp &m, uint64_t base, cx<R> q) {
    bool lane_ok = true;
    if (m.cm_rest) lane_ok = (base & m.cm_rest) == m.cm_rest; // uniform test first
    if (f_dneg(m)) {
        const uint32_t fm = lane_ok ? kSignBit : 0u;
        return cx<R>{flip(q.re, fm), flip(q.im, fm)};
    }
    const bool lpar = __popcll(base & m.pm_rest) & 1;
    cx<R> d = lo<R>(lpar ? m.m[1] : m.m[0]);
    if (!lane_ok) d = cx<R>{R(1), R(0)};
    return cmul(q, d);
}
// sign mask update of a negating member with register and other bits
__device__ __forceinline__ uint32_t run_s(const FOp &m, uint64_t base, uint32_t sm) {
    return sm ^ (((base & m.cm_rest) == m.cm_rest) ? (m.rbits & 0xffffu) : 0u);
}
// The end of a run: register k's factor is the lane factor q (HasQ: the run
// has per-lane members), times its constant P[k] (HasP: the data ops exist
// only then), with its sign from bit k of sm (HasS).  The flags are uniform and
// fixed per variant, so each variant is straight-line code that forms a factor
// and applies it register by register (with per-register branches the factors
// of all registers were formed first and held at once: 64 more VGPRs at four
// register bits).  A run with signs only flips signs.
template <class R, bool HasQ, bool HasP, bool HasS>
__device__ __forceinline__ cx<R> run_factor(const CFOp *hp, int k, cx<R> q, uint32_t sm) {
    cx<R> f = HasQ ? q : cx<R>{R(1), R(0)};
    if constexpr (HasP) {
        const cx<R> p = lo<R>(cx<double>{hp[1 + k / 4].m[k % 4].re, hp[1 + k / 4].m[k % 4].im});
        f = HasQ ? cmul(f, p) : p;
    }
    if constexpr (HasS) {
        const uint32_t fm = (sm << (31 - k)) & kSignBit;
        f = cx<R>{flip(f.re, fm), flip(f.im, fm)};
    }
    return f;
}
template <class R, int NR, int NS, bool HasQ, bool HasP, bool HasS>
__device__ __forceinline__ void run_finish_t(const CFOp *hp, cx<R> q, uint32_t sm, cx<R> (&a)[NR], cx<R> (&b)[NR]) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if constexpr (!HasQ && !HasP) {
            const uint32_t fm = (sm << (31 - k)) & kSignBit;
            a[k] = cx<R>{flip(a[k].re, fm), flip(a[k].im, fm)};
            if constexpr (NS == 2) b[k] = cx<R>{flip(b[k].re, fm), flip(b[k].im, fm)};
        } else {
            const cx<R> f = run_factor<R, HasQ, HasP, HasS>(hp, k, q, sm);
            a[k] = cmul(a[k], f);
            if constexpr (NS == 2) b[k] = cmul(b[k], f);
        }
    }
}
template <class R, int NR, int NS>
__device__ __forceinline__ void run_finish_any(const CFOp *hp, uint32_t flags, cx<R> q, uint32_t sm, cx<R> (&a)[NR],
                                               cx<R> (&b)[NR]) {
    switch (flags & (kRunHasQ | kRunHasP | kRunHasS)) {
    case kRunHasS: run_finish_t<R, NR, NS, false, false, true>(hp, q, sm, a, b); break;
    case kRunHasP: run_finish_t<R, NR, NS, false, true, false>(hp, q, sm, a, b); break;
    case kRunHasP | kRunHasS: run_finish_t<R, NR, NS, false, true, true>(hp, q, sm, a, b); break;
    case kRunHasQ: run_finish_t<R, NR, NS, true, false, false>(hp, q, sm, a, b); break;
    case kRunHasQ | kRunHasS: run_finish_t<R, NR, NS, true, false, true>(hp, q, sm, a, b); break;
    case kRunHasQ | kRunHasP: run_finish_t<R, NR, NS, true, true, false>(hp, q, sm, a, b); break;
    case kRunHasQ | kRunHasP | kRunHasS: run_finish_t<R, NR, NS, true, true, true>(hp, q, sm, a, b); break;
    default: break; // no factor
    }
}
template <class R, int NR>
__device__ __forceinline__ void run_finish(const CFOp *hp, uint32_t flags, cx<R> q, uint32_t sm, cx<R> (&a)[NR]) {
    run_finish_any<R, NR, 1>(hp, flags, q, sm, a, a);
}
template <class R, int NR>
__device__ __forceinline__ void run_finish2(const CFOp *hp, uint32_t flags, cx<R> q, uint32_t sm, cx<R> (&a)[NR],
                                            cx<R> (&b)[NR]) {
    run_finish_any<R, NR, 2>(hp, flags, q, sm, a, b);
}

// Forward: the run whose header hdr is at hp (Lean: no general members).
template <class R, int NR, bool Lean = false>
__device__ __forceinline__ void diag_run(const CFOp *hp, const FOp &hdr, uint64_t base, cx<R> (&v)[NR]) {
    const int nmem = run_nmem(hdr);
    const uint32_t flags = run_flags(hdr);
    const CFOp *mp = hp + 1 + run_ndata(hdr);
    cx<R> q{R(1), R(0)};
    uint32_t sm = 0;
    FOp m = load_op(mp);
    wait_op(m);
    for (int j = 0; j < nmem; ++j) {
        const FOp mn = load_op(mp + (j + 1 < nmem ? j + 1 : j));
        switch (f_mclass(m)) {
        case kMemQ: q = run_q(m, base, q); break;
        case kMemR: break;
        case kMemS: sm = run_s(m, base, sm); break;
        default:
            if constexpr (!Lean) diag_gate<R, NR>(m, base, (base & m.cm_rest) == m.cm_rest, v);
            break;
        }
        m = mn;
    }
    run_finish<R, NR>(hp, flags, q, sm, v);
}

// Overlap sums of the single-state adjoint kernel: acc[slot][column], one
// column per quad of lanes, and a sink row (one entry per thread, never read)
// that takes the other three lanes' adds, so that adding stays free of
// divergent branches.
struct AccCols {
    double (*acc)[kAccCols];
    double *sink;
};
// Adds a lane's overlap to its slot: the quad's values are summed first and
// its first lane adds the sum to the quad's column (its own column: no
// contention, program order).
__device__ __forceinline__ void acc_add(const AccCols &c, int slot, double v) {
    static_assert(kAccLanes == 4, "two exchanges sum a lane quad");
    const int lane = threadIdx.x & 63, col = threadIdx.x >> 2;
    v += xor_lane<0>(v, lane);
    v += xor_lane<1>(v, lane);
    atomicAdd((lane & 3) ? &c.sink[threadIdx.x] : &c.acc[slot][col], v);
}

// Adjoint: the run applied to lambda (a) and the observable state (b); the
// overlap of every trainable member is added to its slot of acc.
template <class R, int NR, int RB, bool Lean = false>
__device__ __forceinline__ void diag_run_adj(const CFOp *hp, const FOp &hdr, uint64_t base, cx<R> (&a)[NR],
                                             cx<R> (&b)[NR], const AccCols &acc) {
    const int nmem = run_nmem(hdr);
    const uint32_t flags = run_flags(hdr);
    const CFOp *mp = hp + 1 + run_ndata(hdr);
    double T = 0.0, E[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) E[r] = 0.0;
    if (flags & kRunNeedsTE) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double w = im_cdot2(b[k], a[k]);
            T += w;
#pragma unroll
            for (int r = 0; r < RB; ++r)
                if (!((k >> r) & 1)) E[r] += w;
        }
    }
    cx<R> q{R(1), R(0)};
    uint32_t sm = 0;
    FOp m = load_op(mp);
    wait_op(m);
    for (int j = 0; j < nmem; ++j) {
        const FOp mn = load_op(mp + (j + 1 < nmem ? j + 1 : j));
        const int slot = f_slot(m);
        if (slot >= 0) {
            const uint32_t oc = f_oclass(m);
            double val;
            if (oc == kOvB && m.pm_rest == 0 && m.cm_rest == 0) {
                // a phase on register bits only (RZ on a register bit): the
                // registers with the bit clear take g_even, the others g_odd
                const double ge = m.m[2].re, go = m.m[3].re;
                const double gt = go * T, d = ge - go;
                switch (f_loc(m)) { // uniform: each case reads its own E
                case 0: val = fma(d, E[0], gt); break;
                case 1:
                    if constexpr (RB > 1) {
                        val = fma(d, E[1], gt);
                        break;
                    }
                    [[fallthrough]];
                case 2:
                    if constexpr (RB > 2) {
                        val = fma(d, E[2], gt);
                        break;
                    }
                    [[fallthrough]];
                default:
                    if constexpr (RB > 3) {
                        val = fma(d, E[RB - 1], gt);
                        break;
                    }
                    val = fma(d, E[0], gt);
                    break;
                }
            } else {
                bool lane_ok = true;
                if (m.cm_rest) lane_ok = (base & m.cm_rest) == m.cm_rest;
                if (!Lean && oc == kOvF) {
                    if constexpr (!Lean) val = diag_dot<R, NR>(m, diag_gen(m), base, lane_ok, a, b);
                } else {
                    const double ge = m.m[2].re, go = m.m[3].re;
                    const bool lpar = __popcll(base & m.pm_rest) & 1;
                    const double g0 = lpar ? go : ge, g1 = lpar ? ge : go;
                    if (oc == kOvT) {
                        val = g0 * T;
                    } else {
                        double e = E[0];
#pragma unroll
                        for (int r = 1; r < RB; ++r)
                            if (f_loc(m) == r) e = E[r];
                        val = g0 * e + g1 * (T - e);
                    }
                    val = lane_ok ? val : 0.0;
                }
            }
            acc_add(acc, slot, val);
        }
        switch (f_mclass(m)) {
        case kMemQ: q = run_q(m, base, q); break;
        case kMemR: break;
        case kMemS: sm = run_s(m, base, sm); break;
        default: {
            if constexpr (!Lean) {
                const bool lane_ok = (base & m.cm_rest) == m.cm_rest;
                diag_gate<R, NR>(m, base, lane_ok, a);
                diag_gate<R, NR>(m, base, lane_ok, b);
            }
            break;
        }
        }
        m = mn;
    }
    run_finish2<R, NR>(hp, flags, q, sm, a, b);
}

// Host: device description of one pass (descriptor, gates, and for the
// reverse direction the generator of every trainable gate by slot).
struct Built {
    PassDesc pd;
    std::vector<FOp> ops;
    std::vector<FGen> gens; // one per slot
    std::vector<int> slot_ops; // per slot: the index into kops of its gate
    std::vector<double> slot_scale; // per slot: factor of its Jacobian entry (deferred rotation scales)
    int scale_members = 0;          // 1 when the pass's closing scale gate is a diagonal-run member
};

// reverse: gates in reverse order (the adjoint sweep); gens / slot_of mark
// trainable gates (nullptr in the forward direction, which instead merges
// consecutive single-qubit gates on one bit unless merge is false; without
// merging, ops[j] is the j-th gate of the pass in execution order).
// diag_runs: group consecutive diagonal gates into diagonal runs and take
// the factors out of real rotations (the single-state kernels; the batched
// kernels take plain gates).  single: the pass runs on complex64 states,
// which bounds how far the rotation factors may shrink the registers' scale.
Built build_pass(const Pass &p, const std::vector<KOp> &kops, int n, bool reverse, const std::vector<KGen> *gens,
                 const std::vector<int> *slot_of, int rb, bool merge = true, bool diag_runs = false,
                 bool single = false);

} // namespace fz
} // namespace lgh
