// Gate fusion: runs of gates applied in one HBM pass.
//
// A pass holds a set of "register bits" (RB of the state's index bits, chosen
// per pass) besides the 6 lowest index bits.  Each wavefront loads a tile of
// 2^(6+RB) amplitudes -- lane l holds the amplitudes whose low 6 index bits
// are l, one register per combination of the register bits -- applies every
// gate of the pass in registers, and stores the tile back.  A gate fits a pass
// when its target bits are lane or register bits; controls and diagonal gates
// (RZ, CZ, PhaseShift, MultiRZ, ...) fit any pass because they only need the
// amplitude's global index.  Targets on lane bits exchange values with
// cross-lane shuffles; targets on register bits are plain register math.
//
// The host scheduler builds passes greedily over the gate dependency graph
// (two gates commute when every qubit they share is used diagonally by both),
// so a layer of single-qubit rotations plus an entangling brick costs a few
// passes instead of one pass per gate.  The adjoint Jacobian runs the same
// passes in reverse order on lambda and one observable state, accumulating
// Im<H lambda | G lambda> for every trainable gate of the pass.
//
// Reference behaviour: applyOperation sequences (pennylane_lightning_gpu/src/
// simulator/StateVectorCudaManaged.hpp:279-315) and the reverse sweep of
// AdjointJacobianGPU::adjointJacobian (algorithms/AdjointDiffGPU.hpp:499-596);
// the reference applies one cuStateVec kernel per gate and has no fusion.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>

#include "lgh_fused.hpp"

namespace lgh {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int RB = 4;              // register bits per pass
constexpr int NREG = 1 << RB;      // amplitudes per lane
constexpr int kLaneBits = 6;
constexpr int kMaxSlots = 32;      // trainable gates per pass

std::atomic<int> g_fusion{1};

struct Use {
    uint64_t diag = 0;    // bits used diagonally (controls, phases)
    uint64_t nondiag = 0; // bits used as dense targets
};

Use uses_of(const KOp &k) {
    Use u;
    if (k.kind == KOp::DIAG) {
        u.diag = k.cmask | k.pmask;
    } else if (k.kind == KOp::DENSE) {
        u.diag = k.cmask;
        for (int i = 0; i < k.k; ++i) u.nondiag |= 1ull << k.tbits[i];
    }
    return u;
}

} // namespace

void set_fusion(int enabled) { g_fusion.store(enabled ? 1 : 0); }
int fusion_enabled() { return g_fusion.load(); }

bool fusable(const KOp &k, int n) {
    if (n < kLaneBits + RB) return false;
    if (k.kind == KOp::DIAG) return true;
    return k.kind == KOp::DENSE && k.k == 1;
}

// Greedy list scheduling of `ops` (already compiled) into passes.
std::vector<Pass> schedule_passes(const std::vector<KOp> &ops, int n, const std::vector<int> &trainable_slot) {
    const int N = static_cast<int>(ops.size());
    std::vector<Use> use(N);
    for (int i = 0; i < N; ++i) use[i] = uses_of(ops[i]);
    // dependency edges via per-bit "last non-diagonal user" and "diagonal users since"
    std::vector<std::vector<int>> succ(N);
    std::vector<int> indeg(N, 0);
    std::vector<int> last_nd(n, -1);
    std::vector<std::vector<int>> diag_since(n);
    auto edge = [&](int a, int b) {
        if (a < 0 || a == b) return;
        succ[a].push_back(b);
        indeg[b]++;
    };
    for (int j = 0; j < N; ++j) {
        std::vector<int> deps;
        for (int b = 0; b < n; ++b) {
            const uint64_t m = 1ull << b;
            if (use[j].nondiag & m) {
                deps.push_back(last_nd[b]);
                for (int d : diag_since[b]) deps.push_back(d);
            } else if (use[j].diag & m) {
                deps.push_back(last_nd[b]);
            }
        }
        std::sort(deps.begin(), deps.end());
        deps.erase(std::unique(deps.begin(), deps.end()), deps.end());
        for (int d : deps) edge(d, j);
        for (int b = 0; b < n; ++b) {
            const uint64_t m = 1ull << b;
            if (use[j].nondiag & m) {
                last_nd[b] = j;
                diag_since[b].clear();
            } else if (use[j].diag & m) {
                diag_since[b].push_back(j);
            }
        }
    }
    std::vector<Pass> passes;
    std::vector<char> done(N, 0);
    std::vector<int> ready;
    for (int i = 0; i < N; ++i)
        if (indeg[i] == 0) ready.push_back(i);
    int remaining = N;
    auto release = [&](int i) {
        done[i] = 1;
        --remaining;
        for (int s : succ[i])
            if (--indeg[s] == 0) ready.push_back(s);
    };
    const uint64_t lane_mask = (1ull << kLaneBits) - 1;
    while (remaining > 0) {
        std::sort(ready.begin(), ready.end());
        // a ready gate that cannot be fused runs on its own
        if (!fusable(ops[ready.front()], n)) {
            const int i = ready.front();
            ready.erase(ready.begin());
            Pass p;
            p.standalone = true;
            p.ops.push_back(i);
            passes.push_back(p);
            release(i);
            continue;
        }
        // register bits: targets of the earliest ready dense gates outside the lane bits
        Pass p;
        uint64_t reg = 0;
        int nreg = 0;
        for (int i : ready) {
            if (!fusable(ops[i], n) || ops[i].kind != KOp::DENSE) continue;
            const uint64_t t = use[i].nondiag & ~lane_mask;
            if (t && !(reg & t) && nreg < RB) {
                reg |= t;
                ++nreg;
            }
        }
        // pad with the lowest unused bits so that every pass has RB register bits
        for (int b = kLaneBits; b < n && nreg < RB; ++b)
            if (!(reg & (1ull << b))) {
                reg |= 1ull << b;
                ++nreg;
            }
        int slots = 0;
        bool progress = true;
        while (progress) {
            progress = false;
            std::sort(ready.begin(), ready.end());
            for (size_t r = 0; r < ready.size(); ++r) {
                const int i = ready[r];
                if (!fusable(ops[i], n)) continue;
                const uint64_t t = use[i].nondiag & ~lane_mask;
                if (t & ~reg) continue;
                const bool tr = !trainable_slot.empty() && trainable_slot[i] >= 0;
                if (tr && slots >= kMaxSlots) continue;
                slots += tr;
                p.ops.push_back(i);
                ready.erase(ready.begin() + r);
                release(i);
                progress = true;
                break;
            }
        }
        for (int b = 0, k = 0; b < n; ++b)
            if (reg & (1ull << b)) p.reg_bits[k++] = b;
        passes.push_back(p);
    }
    return passes;
}

} // namespace lgh

using namespace lgh;

namespace {

// Device description of one fused gate.
struct FOp {
    int32_t kind;  // 0 = dense on lane bit, 1 = dense on register bit, 2 = diagonal
    int32_t loc;   // lane bit or register bit
    int32_t slot;  // trainable slot in the pass, -1 if none
    int32_t gkind; // generator: 0 dense 2x2, 1 diagonal
    uint64_t cmask, pmask;
    cx<double> m[4]; // U (dense) or (d_even, d_odd) in m[0], m[1]
    cx<double> g[4]; // generator (dense) or (g_even, g_odd) in g[0].re, g[1].re
};

struct PassDesc {
    int32_t nops;
    int32_t nslots;
    int32_t reg_bits[RB];
    uint64_t reg_off[NREG]; // index offset of each register
    Fixed tile_fx;          // zero-insertion positions of lane and register bits
};

template <class R> __device__ __forceinline__ cx<R> lo(cx<double> z) { return cx<R>{R(z.re), R(z.im)}; }

// Apply a dense 2x2 on register bit B to all NREG registers.
template <class R, int B>
__device__ __forceinline__ void reg_gate(cx<R> (&v)[NREG], const uint64_t (&idx)[NREG], uint64_t cmask, cx<R> m00,
                                         cx<R> m01, cx<R> m10, cx<R> m11) {
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        if (k & (1 << B)) continue;
        const int k1 = k | (1 << B);
        if ((idx[k] & cmask) != cmask) continue;
        const cx<R> a = v[k], b = v[k1];
        v[k] = cfma(m01, b, cmul(m00, a));
        v[k1] = cfma(m11, b, cmul(m10, a));
    }
}

template <class R>
__device__ __forceinline__ void reg_gate_dispatch(int b, cx<R> (&v)[NREG], const uint64_t (&idx)[NREG], uint64_t cm,
                                                  cx<R> m00, cx<R> m01, cx<R> m10, cx<R> m11) {
    switch (b) {
    case 0: reg_gate<R, 0>(v, idx, cm, m00, m01, m10, m11); break;
    case 1: reg_gate<R, 1>(v, idx, cm, m00, m01, m10, m11); break;
    case 2: reg_gate<R, 2>(v, idx, cm, m00, m01, m10, m11); break;
    default: reg_gate<R, 3>(v, idx, cm, m00, m01, m10, m11); break;
    }
}

// Dense 2x2 on lane bit l: partner in lane ^ 2^l.
template <class R>
__device__ __forceinline__ void lane_gate(int l, int lane, cx<R> (&v)[NREG], const uint64_t (&idx)[NREG],
                                          uint64_t cmask, cx<R> m00, cx<R> m01, cx<R> m10, cx<R> m11) {
    const int mask = 1 << l;
    const bool hi = (lane & mask) != 0;
    const cx<R> cs = hi ? m11 : m00, cp = hi ? m10 : m01;
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        const cx<R> p = shfl_xor_cx(v[k], mask);
        if ((idx[k] & cmask) == cmask) v[k] = cfma(cp, p, cmul(cs, v[k]));
    }
}

template <class R>
__device__ __forceinline__ void diag_gate(cx<R> (&v)[NREG], const uint64_t (&idx)[NREG], uint64_t cmask,
                                          uint64_t pmask, cx<R> de, cx<R> dodd) {
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        if ((idx[k] & cmask) != cmask) continue;
        v[k] = cmul(v[k], (__popcll(idx[k] & pmask) & 1) ? dodd : de);
    }
}

template <class R>
__device__ __forceinline__ void apply_fop(const FOp &op, int lane, cx<R> (&v)[NREG], const uint64_t (&idx)[NREG],
                                          const cx<double> (&m)[4]) {
    if (op.kind == 2) {
        diag_gate<R>(v, idx, op.cmask, op.pmask, lo<R>(m[0]), lo<R>(m[1]));
    } else if (op.kind == 1) {
        reg_gate_dispatch<R>(op.loc, v, idx, op.cmask, lo<R>(m[0]), lo<R>(m[1]), lo<R>(m[2]), lo<R>(m[3]));
    } else {
        lane_gate<R>(op.loc, lane, v, idx, op.cmask, lo<R>(m[0]), lo<R>(m[1]), lo<R>(m[2]), lo<R>(m[3]));
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Im sum_i conj(b_i) (G a)_i for a dense generator on register bit B.
template <class R, int B>
__device__ __forceinline__ double gen_dot_reg(const FOp &op, const cx<R> (&a)[NREG], const cx<R> (&b)[NREG],
                                              const uint64_t (&idx)[NREG]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        if ((idx[k] & op.cmask) != op.cmask) continue;
        const bool hi = (k >> B) & 1;
        const cx<R> p = a[k ^ (1 << B)];
        const cx<double> gs = hi ? op.g[3] : op.g[0], gp = hi ? op.g[2] : op.g[1];
        const cx<double> ga = cfma(gp, cx<double>{double(p.re), double(p.im)},
                                   cmul(gs, cx<double>{double(a[k].re), double(a[k].im)}));
        s += double(b[k].re) * ga.im - double(b[k].im) * ga.re;
    }
    return s;
}

// Im sum_i conj(b_i) (G a)_i over the tile for one trainable gate.
template <class R>
__device__ __forceinline__ double gen_dot(const FOp &op, int lane, const cx<R> (&a)[NREG], const cx<R> (&b)[NREG],
                                          const uint64_t (&idx)[NREG]) {
    double s = 0.0;
    if (op.gkind == 1) { // diagonal generator
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            if ((idx[k] & op.cmask) != op.cmask) continue;
            const double gv = (__popcll(idx[k] & op.pmask) & 1) ? op.g[1].re : op.g[0].re;
            s += gv * (double(b[k].re) * a[k].im - double(b[k].im) * a[k].re);
        }
        return s;
    }
    if (op.kind == 1) { // register bit
        switch (op.loc) {
        case 0: return gen_dot_reg<R, 0>(op, a, b, idx);
        case 1: return gen_dot_reg<R, 1>(op, a, b, idx);
        case 2: return gen_dot_reg<R, 2>(op, a, b, idx);
        default: return gen_dot_reg<R, 3>(op, a, b, idx);
        }
    }
    // lane bit
    const int mask = 1 << op.loc;
    const bool hi = (lane & mask) != 0;
    const cx<double> gs = hi ? op.g[3] : op.g[0], gp = hi ? op.g[2] : op.g[1];
#pragma unroll
    for (int k = 0; k < NREG; ++k) {
        const cx<R> p = shfl_xor_cx(a[k], mask);
        if ((idx[k] & op.cmask) != op.cmask) continue;
        const cx<double> ga =
            cfma(gp, cx<double>{double(p.re), double(p.im)}, cmul(gs, cx<double>{double(a[k].re), double(a[k].im)}));
        s += double(b[k].re) * ga.im - double(b[k].im) * ga.re;
    }
    return s;
}

// Forward pass: every wavefront processes tiles [w*chunk, (w+1)*chunk).
template <class R>
__global__ __launch_bounds__(kBlock) void k_fused_pass(cx<R> *__restrict__ sv, uint64_t ntiles, uint64_t chunk,
                                                       const PassDesc *__restrict__ pd_g,
                                                       const FOp *__restrict__ ops_g) {
    const int lane = threadIdx.x & 63;
    const uint64_t wave = xcd_block() * kWavesPerBlock + (threadIdx.x >> 6);
    const PassDesc pd = *pd_g;
    const uint64_t t0 = wave * chunk, t1 = min(t0 + chunk, ntiles);
    for (uint64_t t = t0; t < t1; ++t) {
        const uint64_t base = expand(t, pd.tile_fx) | uint64_t(lane);
        uint64_t idx[NREG];
        cx<R> v[NREG];
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            idx[k] = base | pd.reg_off[k];
            v[k] = ldnt(sv + idx[k]);
        }
        for (int o = 0; o < pd.nops; ++o) {
            const FOp &op = ops_g[o];
            const cx<double> m[4] = {op.m[0], op.m[1], op.m[2], op.m[3]};
            apply_fop<R>(op, lane, v, idx, m);
        }
#pragma unroll
        for (int k = 0; k < NREG; ++k) stnt(sv + idx[k], v[k]);
    }
}

// Reverse pass of the adjoint sweep on lambda and one observable state.  The
// ops of the pass are stored in reverse order with U^dagger matrices.
template <class R>
__global__ __launch_bounds__(kBlock) void k_fused_adj(cx<R> *__restrict__ lam, cx<R> *__restrict__ bra,
                                                      uint64_t ntiles, uint64_t chunk,
                                                      const PassDesc *__restrict__ pd_g,
                                                      const FOp *__restrict__ ops_g, double *__restrict__ partials) {
    __shared__ double acc[kWavesPerBlock][kMaxSlots];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int s = lane; s < kMaxSlots; s += 64) acc[w][s] = 0.0;
    const uint64_t wave = xcd_block() * kWavesPerBlock + w;
    const PassDesc pd = *pd_g;
    const uint64_t t0 = wave * chunk, t1 = min(t0 + chunk, ntiles);
    for (uint64_t t = t0; t < t1; ++t) {
        const uint64_t base = expand(t, pd.tile_fx) | uint64_t(lane);
        uint64_t idx[NREG];
        cx<R> a[NREG], b[NREG];
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            idx[k] = base | pd.reg_off[k];
            a[k] = ldnt(lam + idx[k]);
            b[k] = ldnt(bra + idx[k]);
        }
        for (int o = 0; o < pd.nops; ++o) {
            const FOp &op = ops_g[o];
            if (op.slot >= 0) {
                const double d = wave_sum(gen_dot<R>(op, lane, a, b, idx));
                if (lane == 0) acc[w][op.slot] += d;
            }
            const cx<double> m[4] = {op.m[0], op.m[1], op.m[2], op.m[3]};
            apply_fop<R>(op, lane, a, idx, m);
            apply_fop<R>(op, lane, b, idx, m);
        }
#pragma unroll
        for (int k = 0; k < NREG; ++k) {
            stnt(lam + idx[k], a[k]);
            stnt(bra + idx[k], b[k]);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < pd.nslots; s += kBlock) {
        double v = 0.0;
        for (int q = 0; q < kWavesPerBlock; ++q) v += acc[q][s];
        partials[size_t(s) * gridDim.x + blockIdx.x] = v;
    }
}

// slot sums -> jac[col[s]] * factor[s]
__global__ __launch_bounds__(kBlock) void k_slot_store(const double *__restrict__ partials, int nparts, int nslots,
                                                       const int64_t *__restrict__ cols,
                                                       const double *__restrict__ factors, double *__restrict__ jac) {
    __shared__ double sh[kWavesPerBlock];
    const int s = blockIdx.x;
    if (s >= nslots) return;
    double v = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) v += partials[size_t(s) * nparts + b];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kWavesPerBlock; ++q) t += sh[q];
        jac[cols[s]] = factors[s] * t;
    }
}

// Host: device description of a pass.
struct Built {
    PassDesc pd;
    std::vector<FOp> ops;
};

Built build_pass(const Pass &p, const std::vector<KOp> &kops, int n, bool reverse,
                 const std::vector<KGen> *gens, const std::vector<int> *slot_of) {
    Built b{};
    std::vector<int> tile_bits;
    for (int l = 0; l < kLaneBits; ++l) tile_bits.push_back(l);
    for (int r = 0; r < RB; ++r) {
        b.pd.reg_bits[r] = p.reg_bits[r];
        tile_bits.push_back(p.reg_bits[r]);
    }
    b.pd.tile_fx = make_fixed(tile_bits);
    for (int k = 0; k < NREG; ++k) {
        uint64_t o = 0;
        for (int r = 0; r < RB; ++r)
            if ((k >> r) & 1) o |= 1ull << p.reg_bits[r];
        b.pd.reg_off[k] = o;
    }
    std::vector<int> order = p.ops;
    if (reverse) std::reverse(order.begin(), order.end());
    int nslots = 0;
    for (int i : order) {
        const KOp &k = kops[size_t(i)];
        FOp f{};
        f.slot = -1;
        if (k.kind == KOp::DIAG) {
            f.kind = 2;
            f.cmask = k.cmask;
            f.pmask = k.pmask;
            f.m[0] = to_cx<double>(k.d_even);
            f.m[1] = to_cx<double>(k.d_odd);
        } else {
            const int t = k.tbits[0];
            f.cmask = k.cmask;
            if (t < kLaneBits) {
                f.kind = 0;
                f.loc = t;
            } else {
                f.kind = 1;
                f.loc = -1;
                for (int r = 0; r < RB; ++r)
                    if (p.reg_bits[r] == t) f.loc = r;
                LGH_REQUIRE(f.loc >= 0, "fused pass: target is not a register bit");
            }
            for (int q = 0; q < 4; ++q) f.m[q] = to_cx<double>(k.mat[size_t(q)]);
        }
        if (slot_of && (*slot_of)[size_t(i)] >= 0) {
            const KGen &g = (*gens)[size_t(i)];
            f.slot = nslots++;
            if (g.mat.empty()) {
                f.gkind = 1;
                f.g[0] = cx<double>{g.g_even, 0.0};
                f.g[1] = cx<double>{g.g_odd, 0.0};
            } else {
                f.gkind = 0;
                for (int q = 0; q < 4; ++q) f.g[q] = to_cx<double>(g.mat[size_t(q)]);
            }
        }
        b.ops.push_back(f);
    }
    b.pd.nops = static_cast<int32_t>(b.ops.size());
    b.pd.nslots = nslots;
    return b;
}

struct Geometry {
    uint64_t ntiles, chunk;
    unsigned blocks;
};
Geometry geometry(int n, int device) {
    Geometry g{};
    g.ntiles = 1ull << (n - kLaneBits - RB);
    // enough wavefronts to fill the chip several times over
    const uint64_t max_waves = uint64_t(std::max(1, sm_count(device))) * 32;
    uint64_t waves = std::min<uint64_t>(g.ntiles, max_waves);
    g.chunk = (g.ntiles + waves - 1) / waves;
    waves = (g.ntiles + g.chunk - 1) / g.chunk;
    g.blocks = unsigned((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (g.blocks >= 8) g.blocks = (g.blocks + 7) / 8 * 8; // whole XCD rounds for the remap
    return g;
}

template <class R> void run_forward(lgh_sv *sv, const std::vector<KOp> &kops, const std::vector<Pass> &passes) {
    StreamCtx *ctx = sv->ctx;
    const Geometry geo = geometry(sv->n, sv->device);
    for (const Pass &p : passes) {
        if (p.standalone) {
            launch_kop(sv, kops[size_t(p.ops[0])]);
            continue;
        }
        Built b = build_pass(p, kops, sv->n, false, nullptr, nullptr);
        const size_t pb = (sizeof(PassDesc) + 255) & ~size_t(255);
        std::vector<unsigned char> blob(pb + b.ops.size() * sizeof(FOp));
        std::memcpy(blob.data(), &b.pd, sizeof(PassDesc));
        std::memcpy(blob.data() + pb, b.ops.data(), b.ops.size() * sizeof(FOp));
        unsigned char *d = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
        hipLaunchKernelGGL(k_fused_pass<R>, dim3(geo.blocks), dim3(kBlock), 0, ctx->stream,
                           static_cast<cx<R> *>(sv->data), geo.ntiles, geo.chunk,
                           reinterpret_cast<const PassDesc *>(d), reinterpret_cast<const FOp *>(d + pb));
        LGH_LAUNCH_CHECK();
    }
}

} // namespace

namespace lgh {

void fused_apply(lgh_sv *sv, const std::vector<KOp> &kops) {
    const std::vector<Pass> passes = schedule_passes(kops, sv->n, {});
    if (sv->dtype == LGH_C128)
        run_forward<double>(sv, kops, passes);
    else
        run_forward<float>(sv, kops, passes);
}

int count_passes(const std::vector<KOp> &kops, int n) {
    return static_cast<int>(schedule_passes(kops, n, {}).size());
}

template <class R>
static void adjoint_one(lgh_sv *lam, lgh_sv *bra, const std::vector<KOp> &udag, const std::vector<KGen> &gens,
                        const std::vector<int> &slot_col, const std::vector<double> &factor, double *jac_row_dev,
                        double *partials, StandaloneStep standalone) {
    // `udag` lists the ops in forward order with U^dagger matrices; the passes
    // are scheduled on that order and executed backwards.
    StreamCtx *ctx = lam->ctx;
    const int n = lam->n;
    const std::vector<Pass> passes = schedule_passes(udag, n, slot_col);
    const Geometry geo = geometry(n, lam->device);
    for (auto it = passes.rbegin(); it != passes.rend(); ++it) {
        const Pass &p = *it;
        if (p.standalone) {
            const int i = p.ops[0];
            standalone(udag[size_t(i)], slot_col[size_t(i)] >= 0 ? &gens[size_t(i)] : nullptr, slot_col[size_t(i)],
                       factor[size_t(i)]);
            continue;
        }
        Built b = build_pass(p, udag, n, true, &gens, &slot_col);
        std::vector<int64_t> cols;
        std::vector<double> facs;
        for (auto oi = p.ops.rbegin(); oi != p.ops.rend(); ++oi)
            if (slot_col[size_t(*oi)] >= 0) {
                cols.push_back(slot_col[size_t(*oi)]);
                facs.push_back(factor[size_t(*oi)]);
            }
        const size_t pb = (sizeof(PassDesc) + 255) & ~size_t(255);
        const size_t ob = (b.ops.size() * sizeof(FOp) + 255) & ~size_t(255);
        const size_t cb = (cols.size() * 8 + 255) & ~size_t(255);
        std::vector<unsigned char> blob(pb + ob + 2 * cb + 256);
        std::memcpy(blob.data(), &b.pd, sizeof(PassDesc));
        std::memcpy(blob.data() + pb, b.ops.data(), b.ops.size() * sizeof(FOp));
        std::memcpy(blob.data() + pb + ob, cols.data(), cols.size() * 8);
        std::memcpy(blob.data() + pb + ob + cb, facs.data(), facs.size() * 8);
        unsigned char *d = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
        hipLaunchKernelGGL(k_fused_adj<R>, dim3(geo.blocks), dim3(kBlock), 0, ctx->stream,
                           static_cast<cx<R> *>(lam->data), static_cast<cx<R> *>(bra->data), geo.ntiles, geo.chunk,
                           reinterpret_cast<const PassDesc *>(d), reinterpret_cast<const FOp *>(d + pb), partials);
        LGH_LAUNCH_CHECK();
        if (!cols.empty()) {
            hipLaunchKernelGGL(k_slot_store, dim3(unsigned(cols.size())), dim3(kBlock), 0, ctx->stream, partials,
                               int(geo.blocks), int(cols.size()), reinterpret_cast<const int64_t *>(d + pb + ob),
                               reinterpret_cast<const double *>(d + pb + ob + cb), jac_row_dev);
            LGH_LAUNCH_CHECK();
        }
    }
}

void fused_adjoint_one(lgh_sv *lam, lgh_sv *bra, const std::vector<KOp> &udag, const std::vector<KGen> &gens,
                       const std::vector<int> &slot_col, const std::vector<double> &factor, double *jac_row_dev,
                       double *partials, StandaloneStep standalone) {
    if (lam->dtype == LGH_C128)
        adjoint_one<double>(lam, bra, udag, gens, slot_col, factor, jac_row_dev, partials, standalone);
    else
        adjoint_one<float>(lam, bra, udag, gens, slot_col, factor, jac_row_dev, partials, standalone);
}

size_t fused_partials_bytes(int n, int device) {
    if (n < kLaneBits + RB) return 0;
    return sizeof(double) * size_t(kMaxSlots) * geometry(n, device).blocks;
}

} // namespace lgh

extern "C" lgh_status lgh_set_fusion(int enabled) {
    return guard([&] { set_fusion(enabled); });
}
