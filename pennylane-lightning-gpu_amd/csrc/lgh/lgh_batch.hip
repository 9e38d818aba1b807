// Batched circuits: many small circuits in one persistent launch.
//
// A template is the gate structure shared by a family of circuits that
// differ only in their parameters (a training batch, a parameter sweep, the
// requests of one model in a serving queue).  The host compiles it once: the
// forward gates, the adjoint sweep (U^dagger per gate, generators, Jacobian
// factors) and the observable are scheduled into register-tiled passes (the
// same passes and gate bodies as lgh_fused.hip), and every gate descriptor
// records how its matrix follows from the circuit's parameter vector.
//
// A run takes any mix of templates and parameter vectors.  A small kernel
// writes every circuit's gate descriptors from its parameters; then one
// persistent launch processes whole circuits per workgroup, largest first,
// pulled from an atomic counter: |0..0>, the forward passes, bra = P psi
// with <P> = <psi|bra>, and the reverse passes on (psi, bra) that accumulate
// d<P>/d theta for every trainable gate.  All passes of a circuit run inside
// one workgroup, separated by a workgroup barrier and a device-scope fence,
// so states of up to 2^20 amplitudes stay in L2 / MALL between passes and no
// circuit waits for a launch.  Circuits below 9 qubits (one 512-amplitude
// tile) run on 9 qubits with idle low-order wires; their results are those of
// the smaller circuit.
//
// The reference has no batched form; each circuit is one sequence of device
// calls (StateVectorCudaManaged::applyOperation, SVCM:279-315, expval, and
// AdjointJacobianGPU::adjointJacobian, AdjointDiffGPU.hpp:499-596), which is
// what lgh_batch_run_serial reproduces for comparison.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>

#include "lgh_fused_dev.hpp"

using namespace lgh;
using namespace lgh::fz;

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
// lane positions 0-3 stay on index bits 0-3 (runs of 256 B for complex128,
// 128 B for complex64); the forward and reverse passes choose the other two
constexpr int kFixedLanes = 4;
constexpr int kMaxBatchQubits = 20;
// One kernel holds the forward, observable and adjoint passes; three
// register bits keep its register use low enough for two wavefronts per SIMD.
constexpr int RB = 3;
constexpr int NR = 1 << RB;
constexpr int kTileBits = kLaneBits + RB;

// How a gate's matrix follows from theta' = sign * theta (sign = -1 for U^dagger
// and for inverse gates): the forms the parametric template gates compile to
// (compile_op in lgh_gates.hip).
enum Form : int32_t {
    kFixed = 0,   // matrix fixed at template creation
    kPhase = 1,   // diagonal, d_even = d_odd = e^{i theta'} (PhaseShift, ControlledPhaseShift)
    kRotDiag = 2, // diagonal, d_even = e^{-i theta'/2}, d_odd = e^{+i theta'/2} (RZ, CRZ, IsingZZ, MultiRZ)
    kRX = 3,      // [[c, -is], [-is, c]], c = cos(theta'/2), s = sin(theta'/2) (RX, CRX)
    kRY = 4,      // [[c, -s], [s, c]] (RY, CRY)
};

struct Recipe {
    FOp base;     // masks and meta; the matrix for fixed gates
    int32_t pidx; // parameter index, -1 for fixed gates
    int32_t form;
    double sign;
};
struct TPass {
    PassDesc pd;
    int32_t op_begin;   // first FOp of the pass in a circuit's FOp array
    int32_t nops;
    int32_t slot_begin; // first slot in the template's slot table
    int32_t nslots;
};
struct TSlot {
    int32_t pidx;
    int32_t pad;
    double factor; // -2 * generator scale * (inverse ? -1 : 1)
};
struct TmplDev {
    int32_t nq, nfwd, nrev, nparams;
    int32_t nfops, pad;
    uint64_t ntiles;
    const TPass *fwd;
    const TPass *rev;
    const TPass *obs;
    const FGen *gens;
    const TSlot *slots;
    const Recipe *recipes;
};
struct CircDev {
    const TmplDev *t;
    FOp *fops;
    const double *params;
    void *psi;
    void *bra;
    double *jac; // nparams entries, or null without gradient
    double *expval;
};

Form form_of(int gate) {
    switch (gate) {
    case LGH_GATE_PHASESHIFT: case LGH_GATE_CONTROLLEDPHASESHIFT: return kPhase;
    case LGH_GATE_RZ: case LGH_GATE_CRZ: case LGH_GATE_ISINGZZ: case LGH_GATE_MULTIRZ: return kRotDiag;
    case LGH_GATE_RX: case LGH_GATE_CRX: return kRX;
    case LGH_GATE_RY: case LGH_GATE_CRY: return kRY;
    default: return kFixed;
    }
}

// ------------------------------------------------------------------ device

typedef __attribute__((address_space(4))) const CircDev CCirc;
typedef __attribute__((address_space(4))) const TmplDev CTmpl;
typedef __attribute__((address_space(4))) const TPass CTPass;
typedef __attribute__((address_space(4))) const TSlot CTSlot;

__global__ __launch_bounds__(kBlock) void k_batch_prep(const CircDev *__restrict__ circs, uint32_t *counter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    const CircDev &cd = circs[blockIdx.x];
    const TmplDev *t = cd.t;
    for (int i = threadIdx.x; i < t->nfops; i += kBlock) {
        const Recipe &r = t->recipes[i];
        FOp f = r.base;
        if (r.form != kFixed) {
            const double th = r.sign * cd.params[r.pidx];
            double s, c;
            sincos(0.5 * th, &s, &c);
            switch (r.form) {
            case kPhase: {
                double s1, c1;
                sincos(th, &s1, &c1);
                f.m[0] = f.m[1] = cx<double>{c1, s1};
                break;
            }
            case kRotDiag:
                f.m[0] = cx<double>{c, -s};
                f.m[1] = cx<double>{c, s};
                break;
            case kRX:
                f.m[0] = f.m[3] = cx<double>{c, 0.0};
                f.m[1] = f.m[2] = cx<double>{0.0, -s};
                break;
            default: // kRY
                f.m[0] = f.m[3] = cx<double>{c, 0.0};
                f.m[1] = cx<double>{-s, 0.0};
                f.m[2] = cx<double>{s, 0.0};
                break;
            }
            bool real = (f.meta & 15u) != kDiag;
            for (int q = 0; q < 4; ++q) real = real && f.m[q].im == 0.0;
            const bool dsym = (f.meta & 15u) == kDiag && f.m[1].re == f.m[0].re && f.m[1].im == -f.m[0].im;
            // (parametric gates are never exactly X: bit 23 stays as compiled, clear)
            f.meta = (f.meta & ~(7u << 21)) | (uint32_t(real) << 21) | (uint32_t(dsym) << 22);
        }
        cd.fops[i] = f;
    }
}

// Stores of this workgroup become visible to its other wavefronts' loads.
__device__ __forceinline__ void block_fence() {
    __threadfence();
    __syncthreads();
}

template <class R>
__device__ __forceinline__ void load_tile(const cx<R> *p, const uint64_t (&roff)[NR], cx<R> (&v)[NR]) {
#pragma unroll
    for (int k = 0; k < NR; ++k) v[k] = ldnt(p + roff[k]);
}
template <class R>
__device__ __forceinline__ void store_tile(cx<R> *p, const uint64_t (&roff)[NR], const cx<R> (&v)[NR]) {
#pragma unroll
    for (int k = 0; k < NR; ++k) stnt(p + roff[k], v[k]);
}

// Tile geometry of one pass: register offsets, all tile bits, and this
// lane's index offset (lane positions carry the pass's lane bits).  A
// wavefront visits tiles w, w + kWaves, ...: tile_start gives the first
// index of tile w, tile_step moves kWaves tiles on.
struct PassGeo {
    uint64_t roff[NR];
    uint64_t tmask;
    uint64_t loff;
};
__device__ __forceinline__ PassGeo pass_geo(const CTPass *tp, int lane) {
    PassGeo g;
#pragma unroll
    for (int k = 0; k < NR; ++k) g.roff[k] = tp->pd.reg_off[k];
    g.tmask = tp->pd.tile_mask;
    g.loff = 0;
#pragma unroll
    for (int q = 0; q < kLaneBits; ++q)
        if ((lane >> q) & 1) g.loff |= 1ull << tp->pd.lane_bits[q];
    return g;
}
__device__ __forceinline__ uint64_t tile_start(const CTPass *tp, uint64_t tile) {
    int fs[kMaxFree];
#pragma unroll
    for (int r = 0; r < kMaxFree; ++r) fs[r] = tp->pd.free_sorted[r];
    return tile_base_free(tile, tp->pd.nfix, fs, tp->pd.nfree);
}
__device__ __forceinline__ uint64_t tile_step(uint64_t g, uint64_t tmask) {
#pragma unroll
    for (int k = 0; k < kWaves; ++k) g = next_tile(g, tmask);
    return g;
}

// Lv > 0: every gate of the pass has a form of that lean level (see lean_level)
template <class R, int Lv>
__device__ __forceinline__ void fwd_pass(const CTPass *tp, const CFOp *fops, cx<R> *psi, uint64_t ntiles, int w, int lane) {
    const PassGeo g = pass_geo(tp, lane);
    const int nops = tp->nops;
    const CFOp *ops = fops + tp->op_begin;
    uint64_t gt = tile_start(tp, uint64_t(w));
    for (uint64_t tile = w; tile < ntiles; tile += kWaves, gt = tile_step(gt, g.tmask)) {
        const uint64_t base = gt | g.loff;
        cx<R> *p = psi + base;
        cx<R> v[NR];
        load_tile(p, g.roff, v);
        if (nops > 0) {
            FOp cur = load_op(ops);
            for (int o = 0; o < nops; ++o) {
                const FOp nxt = load_op(ops + (o + 1 < nops ? o + 1 : o));
                if constexpr (Lv > 0)
                    apply_lean<R, NR, Lv>(cur, lane, base, v);
                else
                    apply_fop<R, NR>(cur, lane, base, v);
                cur = nxt;
            }
        }
        asm volatile("" : "+v"(p));
        store_tile(p, g.roff, v);
    }
}

// bra = P psi over the tile; returns this lane's share of Re <psi|bra>.
template <class R>
__device__ __forceinline__ double obs_pass(const CTPass *tp, const CFOp *fops, const cx<R> *psi, cx<R> *bra, uint64_t ntiles, int w,
                           int lane) {
    const PassGeo g = pass_geo(tp, lane);
    const int nops = tp->nops;
    const CFOp *ops = fops + tp->op_begin;
    double e = 0.0;
    uint64_t gt = tile_start(tp, uint64_t(w));
    for (uint64_t tile = w; tile < ntiles; tile += kWaves, gt = tile_step(gt, g.tmask)) {
        const uint64_t base = gt | g.loff;
        cx<R> v[NR];
        load_tile(psi + base, g.roff, v);
        cx<R> u[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) u[k] = v[k];
        for (int o = 0; o < nops; ++o) apply_fop<R, NR>(load_op(ops + o), lane, base, u);
#pragma unroll
        for (int k = 0; k < NR; ++k) e += double(v[k].re) * u[k].re + double(v[k].im) * u[k].im;
        store_tile(bra + base, g.roff, u);
    }
    return e;
}

template <class R, int Lv>
__device__ __forceinline__ void rev_pass(const CTPass *tp, const CFOp *fops, const CFGen *gens, cx<R> *lam, cx<R> *bra,
                         uint64_t ntiles, int w, int lane, double (*acc)[kBlock]) {
    const PassGeo g = pass_geo(tp, lane);
    const int nops = tp->nops;
    const CFOp *ops = fops + tp->op_begin;
    const int tid = threadIdx.x;
    uint64_t gt = tile_start(tp, uint64_t(w));
    for (uint64_t tile = w; tile < ntiles; tile += kWaves, gt = tile_step(gt, g.tmask)) {
        const uint64_t base = gt | g.loff;
        cx<R> *pa = lam + base, *pb = bra + base;
        cx<R> a[NR], b[NR];
        load_tile(pa, g.roff, a);
        load_tile(pb, g.roff, b);
        if (nops > 0) {
            FOp cur = load_op(ops);
            for (int o = 0; o < nops; ++o) {
                const FOp nxt = load_op(ops + (o + 1 < nops ? o + 1 : o));
                const int slot = f_slot(cur);
                if constexpr (Lv > 0) {
                    if (f_kind(cur) == kDiag) { // a phase on at most one register bit
                        if (slot >= 0) {
                            const CFGen *gp = gens + slot;
                            acc[slot][tid] += adj_diag_step<R, NR, true>(cur, gp->g[0].re, gp->g[1].re, base, a, b);
                        } else {
                            adj_diag_step<R, NR, false>(cur, 0.0, 0.0, base, a, b);
                        }
                    } else if (Lv == 2 && !f_plain(cur)) { // a controlled X, not trainable
                        if constexpr (Lv == 2) {
                            const M2<R> m = m2<R>(cur);
                            cx_lean<R, NR>(cur, m, lane, base, a);
                            cx_lean<R, NR>(cur, m, lane, base, b);
                        }
                    } else if (slot >= 0) {
                        // one add after the dispatch measured faster here than an add
                        // in every branch
                        double s = 0.0;
                        adj_plain_lean<R, NR, true, Lv>(cur, lane, a, b, [&](double v) { s = v; });
                        acc[slot][tid] += s;
                    } else {
                        adj_plain_lean<R, NR, false, Lv>(cur, lane, a, b, [](double) {});
                    }
                } else if (f_kind(cur) != kDiag && f_plain(cur)) {
                    if (slot >= 0) {
                        const FGen gg = load_gen(gens + slot);
                        acc[slot][tid] += adj_plain_dense<R, NR, true>(cur, &gg, lane, a, b);
                    } else {
                        adj_plain_dense<R, NR, false>(cur, nullptr, lane, a, b);
                    }
                } else if (diag_fast(cur)) {
                    if (slot >= 0) {
                        const CFGen *gp = gens + slot;
                        acc[slot][tid] += adj_diag_step<R, NR, true>(cur, gp->g[0].re, gp->g[1].re, base, a, b);
                    } else {
                        adj_diag_step<R, NR, false>(cur, 0.0, 0.0, base, a, b);
                    }
                } else {
                    if (slot >= 0) {
                        const FGen gg = load_gen(gens + slot);
                        acc[slot][tid] += gen_dot<R, NR>(cur, gg, lane, base, a, b);
                    }
                    apply_fop<R, NR>(cur, lane, base, a);
                    apply_fop<R, NR>(cur, lane, base, b);
                }
                cur = nxt;
            }
        }
        asm volatile("" : "+v"(pa), "+v"(pb));
        store_tile(pa, g.roff, a);
        store_tile(pb, g.roff, b);
    }
}

// (two wavefronts per SIMD: forcing three spills registers inside the gate
// loop and halves throughput)
// Lv: lean level of every forward and reverse pass of the batch's templates
// (0: any fusable gate); the observable pass always takes the general form.
template <class R, int Lv = 0>
__global__ __launch_bounds__(kBlock) void k_batch(const CircDev *__restrict__ circs, int ncirc,
                                                  uint32_t *__restrict__ counter, int grad) {
    __shared__ double acc[kMaxSlots][kBlock];
    __shared__ double red[kMaxSlots];
    __shared__ double red_e[kWaves];
    __shared__ int s_c;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const CCirc *cc = (const CCirc *)circs;
    for (;;) {
        if (tid == 0) s_c = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int c = s_c;
        if (c >= ncirc) break;
        const CCirc *cd = cc + c;
        const CTmpl *t = (const CTmpl *)cd->t;
        const CFOp *fops = (const CFOp *)cd->fops;
        const int nq = t->nq, nfwd = t->nfwd, nrev = t->nrev, nparams = t->nparams;
        const uint64_t ntiles = t->ntiles;
        cx<R> *psi = static_cast<cx<R> *>(cd->psi);
        cx<R> *bra = static_cast<cx<R> *>(cd->bra);
        double *jac = cd->jac;
        // |0..0>
        const uint64_t N = 1ull << nq;
        for (uint64_t i = tid; i < N; i += kBlock) psi[i] = cx<R>{R(i == 0 ? 1 : 0), R(0)};
        if (grad && jac)
            for (int j = tid; j < nparams; j += kBlock) jac[j] = 0.0;
        block_fence();
        for (int p = 0; p < nfwd; ++p) {
            fwd_pass<R, Lv>((const CTPass *)t->fwd + p, fops, psi, ntiles, w, lane);
            block_fence();
        }
        const double e = wave_sum(obs_pass<R>((const CTPass *)t->obs, fops, psi, bra, ntiles, w, lane));
        if (lane == 0) red_e[w] = e;
        block_fence();
        if (tid == 0) {
            double s = 0.0;
            for (int q = 0; q < kWaves; ++q) s += red_e[q];
            *cd->expval = s;
        }
        if (grad && jac) {
            const CTSlot *slots = (const CTSlot *)t->slots;
            const CFGen *gens = (const CFGen *)t->gens;
            for (int p = 0; p < nrev; ++p) {
                const CTPass *tp = (const CTPass *)t->rev + p;
                const int ns = tp->nslots, s0 = tp->slot_begin;
                for (int s = 0; s < ns; ++s) acc[s][tid] = 0.0;
                rev_pass<R, Lv>(tp, fops, gens + s0, psi, bra, ntiles, w, lane, acc);
                __threadfence();
                __syncthreads();
                for (int s = w; s < ns; s += kWaves) {
                    double v = acc[s][lane] + acc[s][lane + 64] + acc[s][lane + 128] + acc[s][lane + 192];
                    v = wave_sum(v);
                    if (lane == 0) red[s] = v;
                }
                __syncthreads();
                if (tid == 0)
                    for (int s = 0; s < ns; ++s) {
                        const int pi = slots[s0 + s].pidx;
                        jac[pi] += slots[s0 + s].factor * red[s];
                    }
                block_fence();
            }
        }
        __syncthreads(); // s_c is rewritten by the next iteration
    }
}

} // namespace

// ------------------------------------------------------------------ host

struct lgh_tmpl {
    int n = 0, nq = 0, device = 0, nparams = 0, nfops = 0;
    std::vector<TPass> fwd, rev;
    TPass obs{};
    std::vector<FGen> gens;
    std::vector<TSlot> slots;
    std::vector<Recipe> recipes;
    uint64_t cost = 0; // tiles x passes, for ordering a batch
    int lean = 0;      // lean level of every forward and reverse pass (0: general kernels)
    // the circuit as given, for the serial form
    std::vector<lgh_op> ops;
    std::vector<std::vector<int32_t>> wires;
    std::vector<std::vector<double>> mats; // copies of explicit matrices
    std::vector<int32_t> pidx;
    std::vector<int32_t> obs_paulis, obs_wires;
    // device copy of the compiled template (uploaded on first use)
    std::mutex upload_mu;
    void *dev = nullptr;
    size_t dev_bytes = 0;
    ~lgh_tmpl() {
        if (dev) {
            (void)hipDeviceSynchronize();
            pool_free(device, dev, dev_bytes);
        }
    }
};

namespace {

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// The copy runs on the caller's stream: a copy on the null stream would wait
// for the work of every other blocking stream of the device (a serving
// engine's large circuits, for one).
void upload_template(lgh_tmpl *t, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(t->upload_mu);
    if (t->dev) return;
    const size_t o_hdr = 0;
    const size_t o_fwd = al256(sizeof(TmplDev));
    const size_t o_rev = o_fwd + al256(sizeof(TPass) * std::max<size_t>(1, t->fwd.size()));
    const size_t o_obs = o_rev + al256(sizeof(TPass) * std::max<size_t>(1, t->rev.size()));
    const size_t o_gen = o_obs + al256(sizeof(TPass));
    const size_t o_slot = o_gen + al256(sizeof(FGen) * std::max<size_t>(1, t->gens.size()));
    const size_t o_rec = o_slot + al256(sizeof(TSlot) * std::max<size_t>(1, t->slots.size()));
    const size_t total = o_rec + al256(sizeof(Recipe) * std::max<size_t>(1, t->recipes.size()));
    DeviceGuard dg(t->device);
    char *d = static_cast<char *>(pool_alloc(t->device, total));
    std::vector<char> h(total, 0);
    TmplDev hd{};
    hd.nq = t->nq;
    hd.nfwd = int32_t(t->fwd.size());
    hd.nrev = int32_t(t->rev.size());
    hd.nparams = t->nparams;
    hd.nfops = t->nfops;
    hd.ntiles = 1ull << (t->nq - kTileBits);
    hd.fwd = reinterpret_cast<const TPass *>(d + o_fwd);
    hd.rev = reinterpret_cast<const TPass *>(d + o_rev);
    hd.obs = reinterpret_cast<const TPass *>(d + o_obs);
    hd.gens = reinterpret_cast<const FGen *>(d + o_gen);
    hd.slots = reinterpret_cast<const TSlot *>(d + o_slot);
    hd.recipes = reinterpret_cast<const Recipe *>(d + o_rec);
    std::memcpy(h.data() + o_hdr, &hd, sizeof(hd));
    std::memcpy(h.data() + o_fwd, t->fwd.data(), sizeof(TPass) * t->fwd.size());
    std::memcpy(h.data() + o_rev, t->rev.data(), sizeof(TPass) * t->rev.size());
    std::memcpy(h.data() + o_obs, &t->obs, sizeof(TPass));
    std::memcpy(h.data() + o_gen, t->gens.data(), sizeof(FGen) * t->gens.size());
    std::memcpy(h.data() + o_slot, t->slots.data(), sizeof(TSlot) * t->slots.size());
    std::memcpy(h.data() + o_rec, t->recipes.data(), sizeof(Recipe) * t->recipes.size());
    hipError_t e = hipMemcpyAsync(d, h.data(), total, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream); // h is a host temporary
    if (e != hipSuccess) {
        pool_free(t->device, d, total);
        LGH_HIP(e);
    }
    t->dev = d;
    t->dev_bytes = total;
}

lgh_op op_with(const lgh_tmpl &t, size_t i, const double *params) {
    lgh_op o = t.ops[i];
    o.wires = t.wires[i].data();
    if (t.pidx[i] >= 0) {
        o.params[0] = params[t.pidx[i]];
        o.num_params = 1;
    }
    return o;
}

void check_batch_args(lgh_tmpl *const *tmpls, int32_t ncirc, lgh_dtype dtype, const double *params,
                      const int64_t *param_offset, double *expval, double *jac, const int64_t *jac_offset) {
    LGH_REQUIRE(ncirc >= 0, "negative circuit count");
    LGH_REQUIRE(dtype == LGH_C64 || dtype == LGH_C128, "unknown dtype");
    if (ncirc == 0) return;
    LGH_REQUIRE(tmpls && expval, "null pointer");
    LGH_REQUIRE(!jac || jac_offset, "jac needs jac_offset");
    const int dev = tmpls[0] ? tmpls[0]->device : 0;
    for (int32_t c = 0; c < ncirc; ++c) {
        LGH_REQUIRE(tmpls[c], "null template");
        LGH_REQUIRE(tmpls[c]->device == dev, "all templates of a batch must live on one device");
        if (tmpls[c]->nparams > 0) LGH_REQUIRE(params && param_offset, "parameters missing");
        if (param_offset) LGH_REQUIRE(param_offset[c] >= 0, "negative parameter offset");
        if (jac) LGH_REQUIRE(jac_offset[c] >= 0, "negative jacobian offset");
    }
}

// Device layout of one batch: one block with the circuit table, parameters,
// descriptors, Jacobian rows, expectation values and the work counter; a
// second block with every circuit's psi and bra.  Circuits are ordered
// largest first (the tail of the persistent launch is then made of the
// cheapest ones); p_off / j_off / expval slots stay indexed by the caller's
// circuit order.
struct BatchLayout {
    int dev = 0;
    int32_t ncirc = 0;
    bool grad = false;
    std::vector<int64_t> p_off, f_off, j_off;
    std::vector<uint64_t> s_off;
    int64_t np = 0, nf = 0, nj = 0;
    uint64_t ns = 0;
    size_t b_circ = 0, b_par = 0, b_fop = 0, b_jac = 0, b_exp = 0, total = 0, sbytes = 0;
    std::vector<int32_t> order;
    int lean = 0; // the kernel's lean level: the highest of the templates', 0 if any has none
    char *d = nullptr;
    void *states = nullptr;

    double *d_par() const { return reinterpret_cast<double *>(d + b_circ); }
    FOp *d_fop() const { return reinterpret_cast<FOp *>(d + b_circ + b_par); }
    double *d_jac() const { return reinterpret_cast<double *>(d + b_circ + b_par + b_fop); }
    double *d_exp() const { return reinterpret_cast<double *>(d + b_circ + b_par + b_fop + b_jac); }
    uint32_t *d_cnt() const { return reinterpret_cast<uint32_t *>(d + b_circ + b_par + b_fop + b_jac + b_exp); }

    template <class R> void plan(lgh_tmpl *const *tmpls, int32_t n, bool with_grad) {
        dev = tmpls[0]->device;
        ncirc = n;
        grad = with_grad;
        lean = lean_enabled() ? 1 : 0;
        for (int32_t i = 0; i < n && lean > 0; ++i) {
            if (tmpls[i]->lean == 0) {
                lean = 0;
                break;
            }
            lean = std::max(lean, tmpls[i]->lean);
        }
        p_off.resize(size_t(n));
        f_off.resize(size_t(n));
        j_off.resize(size_t(n));
        s_off.resize(size_t(n));
        for (int32_t c = 0; c < n; ++c) {
            const lgh_tmpl *t = tmpls[c];
            p_off[size_t(c)] = np;
            np += t->nparams;
            f_off[size_t(c)] = nf;
            nf += t->nfops;
            j_off[size_t(c)] = nj;
            nj += grad ? t->nparams : 0;
            s_off[size_t(c)] = ns;
            ns += 2ull << t->nq; // psi and bra
        }
        b_circ = al256(sizeof(CircDev) * size_t(n));
        b_par = al256(sizeof(double) * size_t(std::max<int64_t>(1, np)));
        b_fop = al256(sizeof(FOp) * size_t(std::max<int64_t>(1, nf)));
        b_jac = al256(sizeof(double) * size_t(std::max<int64_t>(1, nj)));
        b_exp = al256(sizeof(double) * size_t(n));
        total = b_circ + b_par + b_fop + b_jac + b_exp + 256;
        sbytes = sizeof(cx<R>) * ns;
        order.resize(size_t(n));
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return tmpls[a]->cost > tmpls[b]->cost; });
    }
    void allocate() {
        d = static_cast<char *>(pool_alloc(dev, total));
        try {
            states = pool_alloc(dev, sbytes);
        } catch (...) {
            pool_free(dev, d, total);
            d = nullptr;
            throw;
        }
    }
    void release() {
        if (d) pool_free(dev, d, total);
        if (states) pool_free(dev, states, sbytes);
        d = nullptr;
        states = nullptr;
    }
    // circuit table in launch order
    template <class R> std::vector<CircDev> table(lgh_tmpl *const *tmpls) const {
        std::vector<CircDev> hc(static_cast<size_t>(ncirc));
        cx<R> *st = static_cast<cx<R> *>(states);
        for (int32_t r = 0; r < ncirc; ++r) {
            const int32_t c = order[size_t(r)];
            const lgh_tmpl *t = tmpls[c];
            CircDev &cd = hc[size_t(r)];
            cd.t = static_cast<const TmplDev *>(t->dev);
            cd.fops = d_fop() + f_off[size_t(c)];
            cd.params = d_par() + p_off[size_t(c)];
            cd.psi = st + s_off[size_t(c)];
            cd.bra = st + s_off[size_t(c)] + (1ull << t->nq);
            cd.jac = grad ? d_jac() + j_off[size_t(c)] : nullptr;
            cd.expval = d_exp() + c;
        }
        return hc;
    }
};

// Resident workgroups per CU of the persistent launch (the kernel runs at
// two waves per SIMD: one or two 256-thread workgroups per CU).
unsigned batch_grid(int dev, int32_t ncirc, int per_cu) {
    return unsigned(std::min<int64_t>(ncirc, int64_t(per_cu) * int64_t(std::max(1, sm_count(dev)))));
}

typedef void (*BatchKernel)(const CircDev *, int, uint32_t *, int);
template <class R> BatchKernel batch_kernel(int lean) {
    return lean == 1 ? k_batch<R, 1> : lean == 2 ? k_batch<R, 2> : k_batch<R, 0>;
}

template <class R>
void launch_batch(const BatchLayout &L, unsigned grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_prep, dim3(unsigned(L.ncirc)), dim3(kBlock), 0, stream,
                       reinterpret_cast<const CircDev *>(L.d), L.d_cnt());
    LGH_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_kernel<R>(L.lean), dim3(grid), dim3(kBlock), 0, stream,
                       reinterpret_cast<const CircDev *>(L.d), int(L.ncirc), L.d_cnt(), int(L.grad));
    LGH_LAUNCH_CHECK();
}

template <class R>
void batch_run_t(lgh_tmpl *const *tmpls, int32_t ncirc, const double *params, const int64_t *param_offset,
                 double *expval, double *jac, const int64_t *jac_offset, void *stream) {
    const int dev = tmpls[0]->device;
    DeviceGuard dg(dev);
    StreamCtx *ctx = acquire_ctx(dev, static_cast<hipStream_t>(stream));
    struct Rel {
        StreamCtx *c;
        ~Rel() { release_ctx(c); }
    } rel{ctx};
    for (int32_t c = 0; c < ncirc; ++c) upload_template(tmpls[c], ctx->stream);
    BatchLayout L;
    L.plan<R>(tmpls, ncirc, jac != nullptr);
    L.allocate();
    struct FreeAll {
        StreamCtx *ctx;
        BatchLayout &L;
        ~FreeAll() {
            (void)hipStreamSynchronize(ctx->stream);
            L.release();
        }
    } fa{ctx, L};
    const std::vector<CircDev> hc = L.table<R>(tmpls);
    std::vector<double> hp(size_t(std::max<int64_t>(1, L.np)), 0.0);
    for (int32_t c = 0; c < ncirc; ++c)
        if (tmpls[c]->nparams > 0)
            std::memcpy(hp.data() + L.p_off[size_t(c)], params + param_offset[c],
                        sizeof(double) * size_t(tmpls[c]->nparams));
    LGH_HIP(hipMemcpyAsync(L.d, hc.data(), sizeof(CircDev) * hc.size(), hipMemcpyHostToDevice, ctx->stream));
    LGH_HIP(hipMemcpyAsync(L.d_par(), hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, ctx->stream));
    {
        ProfScope prof(kProfBatch, ctx->stream);
        launch_batch<R>(L, batch_grid(dev, ncirc, 2), ctx->stream);
    }
    LGH_HIP(hipMemcpyAsync(expval, L.d_exp(), sizeof(double) * size_t(ncirc), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> hj;
    if (L.grad && L.nj > 0) {
        hj.resize(size_t(L.nj));
        LGH_HIP(hipMemcpyAsync(hj.data(), L.d_jac(), sizeof(double) * size_t(L.nj), hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    LGH_HIP(hipStreamSynchronize(ctx->stream));
    if (L.grad)
        for (int32_t c = 0; c < ncirc; ++c)
            if (tmpls[c]->nparams > 0)
                std::memcpy(jac + jac_offset[c], hj.data() + L.j_off[size_t(c)],
                            sizeof(double) * size_t(tmpls[c]->nparams));
}

} // namespace

// A batch of fixed composition prepared once (lgh_batch_plan_*): device
// buffers, circuit table, pinned staging of parameters and results, and the
// launch sequence -- parameter upload, descriptor preparation, persistent
// launch, result download -- captured in a HIP graph that each run replays.
// The graph holds copies and kernels only (the work counter is cleared by
// the preparation kernel, not by a memset node).
struct lgh_batch_plan {
    int dev = 0;
    lgh_dtype dtype = LGH_C128;
    std::vector<lgh_tmpl *> tmpls;
    BatchLayout L;
    hipStream_t stream = nullptr;
    double *h_par = nullptr, *h_exp = nullptr, *h_jac = nullptr; // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int grid_per_cu = 2;
    double tuned_ms[2] = {0.0, 0.0}; // per candidate grid (1 and 2 workgroups per CU)

    ~lgh_batch_plan() {
        DeviceGuard g(dev);
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        L.release();
        if (h_par) (void)hipHostFree(h_par);
        if (h_exp) (void)hipHostFree(h_exp);
        if (h_jac) (void)hipHostFree(h_jac);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class R> void capture(int per_cu, hipGraph_t *g, hipGraphExec_t *e) {
        LGH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t err = hipMemcpyAsync(L.d_par(), h_par, sizeof(double) * size_t(std::max<int64_t>(1, L.np)),
                                        hipMemcpyHostToDevice, stream);
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_batch_prep, dim3(unsigned(L.ncirc)), dim3(kBlock), 0, stream,
                               reinterpret_cast<const CircDev *>(L.d), L.d_cnt());
            hipLaunchKernelGGL(batch_kernel<R>(L.lean), dim3(batch_grid(dev, L.ncirc, per_cu)), dim3(kBlock), 0,
                               stream, reinterpret_cast<const CircDev *>(L.d), int(L.ncirc), L.d_cnt(), int(L.grad));
            err = hipGetLastError();
        }
        if (err == hipSuccess)
            err = hipMemcpyAsync(h_exp, L.d_exp(), sizeof(double) * size_t(L.ncirc), hipMemcpyDeviceToHost, stream);
        if (err == hipSuccess && L.grad && L.nj > 0)
            err = hipMemcpyAsync(h_jac, L.d_jac(), sizeof(double) * size_t(L.nj), hipMemcpyDeviceToHost, stream);
        hipGraph_t graph_out = nullptr;
        const hipError_t end = hipStreamEndCapture(stream, &graph_out);
        if (err == hipSuccess) err = end;
        if (err != hipSuccess) {
            if (graph_out) (void)hipGraphDestroy(graph_out);
            LGH_HIP(err);
        }
        err = hipGraphInstantiate(e, graph_out, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            (void)hipGraphDestroy(graph_out);
            LGH_HIP(err);
        }
        *g = graph_out;
    }

    double time_graph(hipGraphExec_t e, int reps) {
        hipEvent_t a = nullptr, b = nullptr;
        LGH_HIP(hipEventCreate(&a));
        hipError_t err = hipEventCreate(&b);
        float best = 1e30f;
        for (int r = 0; r < reps && err == hipSuccess; ++r) {
            err = hipEventRecord(a, stream);
            if (err == hipSuccess) err = hipGraphLaunch(e, stream);
            if (err == hipSuccess) err = hipEventRecord(b, stream);
            if (err == hipSuccess) err = hipEventSynchronize(b);
            float ms = 0.0f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
            if (err == hipSuccess) best = std::min(best, ms);
        }
        (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        LGH_HIP(err);
        return double(best);
    }

    // Autotuning of the persistent launch's grid: both candidates replayed
    // on the plan's own circuits (parameters of a fixed pseudo-random
    // sequence), the faster kept.  Each circuit is computed by one
    // workgroup either way, so the choice does not change any result.
    template <class R> void build_and_tune() {
        for (int64_t k = 0; k < L.np; ++k) h_par[k] = std::fmod(0.1 + 0.7548776662 * double(k), 6.283185307179586);
        hipGraph_t g[2] = {nullptr, nullptr};
        hipGraphExec_t e[2] = {nullptr, nullptr};
        try {
            for (int q = 0; q < 2; ++q) {
                capture<R>(q + 1, &g[q], &e[q]);
                time_graph(e[q], 1); // warm-up
                tuned_ms[q] = time_graph(e[q], 2);
            }
        } catch (...) {
            for (int q = 0; q < 2; ++q) {
                if (e[q]) (void)hipGraphExecDestroy(e[q]);
                if (g[q]) (void)hipGraphDestroy(g[q]);
            }
            throw;
        }
        const int best = tuned_ms[0] < tuned_ms[1] ? 0 : 1;
        grid_per_cu = best + 1;
        graph = g[best];
        exec = e[best];
        (void)hipGraphExecDestroy(e[1 - best]);
        (void)hipGraphDestroy(g[1 - best]);
    }
};

extern "C" {

lgh_status lgh_batch_plan_create(lgh_batch_plan **out, lgh_tmpl *const *tmpls, int32_t num_circuits,
                                 lgh_dtype dtype, int32_t with_grad) {
    return guard([&] {
        LGH_REQUIRE(out, "null pointer");
        *out = nullptr;
        LGH_REQUIRE(num_circuits >= 1, "a batch plan needs at least one circuit");
        LGH_REQUIRE(dtype == LGH_C64 || dtype == LGH_C128, "unknown dtype");
        LGH_REQUIRE(tmpls, "null pointer");
        const int dev = tmpls[0] ? tmpls[0]->device : 0;
        for (int32_t c = 0; c < num_circuits; ++c) {
            LGH_REQUIRE(tmpls[c], "null template");
            LGH_REQUIRE(tmpls[c]->device == dev, "all templates of a batch must live on one device");
        }
        DeviceGuard dg(dev);
        auto p = std::make_unique<lgh_batch_plan>();
        p->dev = dev;
        p->dtype = dtype;
        p->tmpls.assign(tmpls, tmpls + num_circuits);
        LGH_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)); // default priority
        for (int32_t c = 0; c < num_circuits; ++c) upload_template(tmpls[c], p->stream);
        auto setup = [&](auto r) {
            using R = decltype(r);
            p->L.plan<R>(tmpls, num_circuits, with_grad != 0);
            p->L.allocate();
            LGH_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->h_par),
                                  sizeof(double) * size_t(std::max<int64_t>(1, p->L.np)), hipHostMallocDefault));
            LGH_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->h_exp), sizeof(double) * size_t(num_circuits),
                                  hipHostMallocDefault));
            LGH_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->h_jac),
                                  sizeof(double) * size_t(std::max<int64_t>(1, p->L.nj)), hipHostMallocDefault));
            const std::vector<CircDev> hc = p->L.table<R>(tmpls);
            LGH_HIP(hipMemcpyAsync(p->L.d, hc.data(), sizeof(CircDev) * hc.size(), hipMemcpyHostToDevice, p->stream));
            LGH_HIP(hipStreamSynchronize(p->stream));
            p->build_and_tune<R>();
        };
        if (dtype == LGH_C128)
            setup(double{});
        else
            setup(float{});
        *out = p.release();
    });
}

lgh_status lgh_batch_plan_run(lgh_batch_plan *plan, const double *params, const int64_t *param_offset,
                              double *expval, double *jac, const int64_t *jac_offset) {
    return guard([&] {
        Range rg("lgh_batch_plan_run");
        LGH_REQUIRE(plan && expval, "null pointer");
        const BatchLayout &L = plan->L;
        LGH_REQUIRE(!jac || L.grad, "the plan was created without gradients");
        LGH_REQUIRE(!jac || jac_offset, "jac needs jac_offset");
        DeviceGuard dg(plan->dev);
        for (int32_t c = 0; c < L.ncirc; ++c) {
            const int32_t np = plan->tmpls[size_t(c)]->nparams;
            if (np == 0) continue;
            LGH_REQUIRE(params && param_offset && param_offset[c] >= 0, "parameters missing");
            std::memcpy(plan->h_par + L.p_off[size_t(c)], params + param_offset[c], sizeof(double) * size_t(np));
        }
        {
            ProfScope prof(kProfBatch, plan->stream);
            LGH_HIP(hipGraphLaunch(plan->exec, plan->stream));
        }
        LGH_HIP(hipStreamSynchronize(plan->stream));
        std::memcpy(expval, plan->h_exp, sizeof(double) * size_t(L.ncirc));
        if (jac)
            for (int32_t c = 0; c < L.ncirc; ++c) {
                const int32_t np = plan->tmpls[size_t(c)]->nparams;
                if (np > 0)
                    std::memcpy(jac + jac_offset[c], plan->h_jac + L.j_off[size_t(c)], sizeof(double) * size_t(np));
            }
    });
}

lgh_status lgh_batch_plan_info(const lgh_batch_plan *plan, int32_t *grid_per_cu, double *tuned_ms) {
    return guard([&] {
        LGH_REQUIRE(plan, "null plan");
        if (grid_per_cu) *grid_per_cu = plan->grid_per_cu;
        if (tuned_ms) {
            tuned_ms[0] = plan->tuned_ms[0];
            tuned_ms[1] = plan->tuned_ms[1];
        }
    });
}

lgh_status lgh_batch_plan_destroy(lgh_batch_plan *plan) {
    return guard([&] { delete plan; });
}



lgh_status lgh_tmpl_create(lgh_tmpl **out, int32_t num_qubits, const lgh_op *ops, int32_t num_ops,
                           const int32_t *param_index, const int32_t *obs_paulis, const int32_t *obs_wires,
                           int32_t obs_len, int32_t device) {
    return guard([&] {
        LGH_REQUIRE(out, "null pointer");
        *out = nullptr;
        LGH_REQUIRE(num_qubits >= 1 && num_qubits <= kMaxBatchQubits,
                    "batched circuits have 1 to " + std::to_string(kMaxBatchQubits) + " qubits");
        LGH_REQUIRE(num_ops >= 0 && (num_ops == 0 || ops), "bad op list");
        LGH_REQUIRE(obs_len >= 0 && (obs_len == 0 || (obs_paulis && obs_wires)), "bad observable");
        auto t = std::make_unique<lgh_tmpl>();
        t->n = num_qubits;
        t->nq = std::max(num_qubits, kTileBits);
        t->device = device;
        const int nq = t->nq;
        // the circuit as given (validated against its own width)
        std::vector<int> gate_ops; // indices of non-identity ops
        for (int32_t i = 0; i < num_ops; ++i) {
            validate_op(ops[i], num_qubits);
            const int32_t pi = param_index ? param_index[i] : -1;
            LGH_REQUIRE(pi >= -1, "bad parameter index");
            if (pi >= 0) {
                LGH_REQUIRE(form_of(ops[i].gate) != kFixed,
                            "gate " + std::to_string(ops[i].gate) +
                                " cannot take a circuit parameter (supported: RX, RY, RZ, PhaseShift, CRX, CRY, "
                                "CRZ, ControlledPhaseShift, IsingZZ, MultiRZ)");
                t->nparams = std::max(t->nparams, pi + 1);
            }
            LGH_REQUIRE(ops[i].gate != LGH_GATE_STATEPREP, "state preparation is not part of a template");
            t->ops.push_back(ops[i]);
            t->wires.emplace_back(ops[i].wires, ops[i].wires + ops[i].num_wires);
            if (ops[i].gate == LGH_GATE_MATRIX) {
                const size_t dim = size_t(1) << (ops[i].num_wires - ops[i].num_ctrl);
                t->mats.emplace_back(ops[i].matrix, ops[i].matrix + 2 * dim * dim);
            } else {
                t->mats.emplace_back();
            }
            t->pidx.push_back(pi);
            if (ops[i].gate != LGH_GATE_IDENTITY) gate_ops.push_back(i);
        }
        for (size_t i = 0; i < t->ops.size(); ++i) {
            t->ops[i].wires = t->wires[i].data();
            if (!t->mats[i].empty()) t->ops[i].matrix = t->mats[i].data();
        }
        for (int32_t k = 0; k < obs_len; ++k) {
            const int32_t g = obs_paulis[k];
            LGH_REQUIRE(g == LGH_GATE_PAULIX || g == LGH_GATE_PAULIY || g == LGH_GATE_PAULIZ,
                        "observable factors are PauliX, PauliY or PauliZ");
            LGH_REQUIRE(obs_wires[k] >= 0 && obs_wires[k] < num_qubits, "observable wire out of range");
            for (int32_t j = 0; j < k; ++j) LGH_REQUIRE(obs_wires[j] != obs_wires[k], "repeated observable wire");
            t->obs_paulis.push_back(g);
            t->obs_wires.push_back(obs_wires[k]);
        }
        // forward program
        std::vector<KOp> fk;
        std::vector<int> fsrc;
        for (int i : gate_ops) {
            fk.push_back(compile_op(t->ops[size_t(i)], nq, false));
            fsrc.push_back(i);
            LGH_REQUIRE(fusable(fk.back(), nq, RB),
                        "gate " + std::to_string(t->ops[size_t(i)].gate) +
                            " is not supported in a batched circuit (single-target dense or diagonal gates only)");
        }
        auto recipe = [&](const FOp &f, int src, bool dagger) {
            Recipe r{};
            r.base = f;
            r.pidx = t->pidx[size_t(src)];
            r.form = r.pidx >= 0 ? form_of(t->ops[size_t(src)].gate) : kFixed;
            const bool inv = (t->ops[size_t(src)].inverse != 0) != dagger;
            r.sign = inv ? -1.0 : 1.0;
            if (r.form != kFixed) r.base.meta &= ~(1u << 20); // the -1 shortcut only for fixed gates
            return r;
        };
        int nf = 0;
        {
            const std::vector<Pass> passes = schedule_passes(fk, nq, {}, RB, kFixedLanes, kMaxSlots);
            for (const Pass &p : passes) {
                LGH_REQUIRE(!p.standalone, "gate does not fit a fused pass");
                Built b = build_pass(p, fk, nq, false, nullptr, nullptr, RB, false);
                LGH_REQUIRE(b.ops.size() == p.ops.size(), "fused pass: unexpected gate count");
                TPass tp{};
                tp.pd = b.pd;
                tp.op_begin = nf;
                tp.nops = int32_t(b.ops.size());
                for (size_t j = 0; j < b.ops.size(); ++j)
                    t->recipes.push_back(recipe(b.ops[j], fsrc[size_t(p.ops[j])], false));
                nf += tp.nops;
                t->fwd.push_back(tp);
            }
        }
        // adjoint program: U^dagger of every gate, generators of the trainable ones
        if (t->nparams > 0) {
            std::vector<KOp> uk;
            std::vector<KGen> gk;
            std::vector<int> slot_of;
            std::vector<double> fac;
            for (int i : gate_ops) {
                const lgh_op &o = t->ops[size_t(i)];
                uk.push_back(compile_op(o, nq, true));
                if (t->pidx[size_t(i)] >= 0) {
                    gk.push_back(compile_generator(o, nq, uk.back()));
                    fac.push_back(-2.0 * gk.back().scale * (o.inverse ? -1.0 : 1.0));
                    slot_of.push_back(t->pidx[size_t(i)]);
                } else {
                    gk.emplace_back();
                    fac.push_back(0.0);
                    slot_of.push_back(-1);
                }
            }
            const std::vector<Pass> passes = schedule_passes(uk, nq, slot_of, RB, kFixedLanes, kMaxSlots);
            for (auto it = passes.rbegin(); it != passes.rend(); ++it) {
                const Pass &p = *it;
                LGH_REQUIRE(!p.standalone, "gate does not fit a fused pass");
                Built b = build_pass(p, uk, nq, true, &gk, &slot_of, RB, false);
                LGH_REQUIRE(b.ops.size() == p.ops.size(), "fused pass: unexpected gate count");
                TPass tp{};
                tp.pd = b.pd;
                tp.op_begin = nf;
                tp.nops = int32_t(b.ops.size());
                tp.slot_begin = int32_t(t->slots.size());
                tp.nslots = b.pd.nslots;
                const size_t m = p.ops.size();
                for (size_t j = 0; j < m; ++j) {
                    const int q = p.ops[m - 1 - j]; // the pass runs its gates backwards
                    t->recipes.push_back(recipe(b.ops[j], fsrc[size_t(q)], true));
                    if (slot_of[size_t(q)] >= 0) {
                        TSlot s{};
                        s.pidx = slot_of[size_t(q)];
                        s.factor = fac[size_t(q)];
                        t->slots.push_back(s);
                    }
                }
                LGH_REQUIRE(int32_t(t->slots.size()) - tp.slot_begin == tp.nslots, "slot bookkeeping");
                t->gens.insert(t->gens.end(), b.gens.begin(), b.gens.end());
                nf += tp.nops;
                t->rev.push_back(tp);
            }
        }
        // observable: one pass
        {
            std::vector<KOp> ok;
            for (size_t k = 0; k < t->obs_paulis.size(); ++k) {
                lgh_op o{};
                o.gate = t->obs_paulis[k];
                o.num_wires = 1;
                o.wires = &t->obs_wires[k];
                ok.push_back(compile_op(o, nq, false));
            }
            Pass p;
            if (!ok.empty()) {
                const std::vector<Pass> passes = schedule_passes(ok, nq, {}, RB, 6, kMaxSlots);
                LGH_REQUIRE(passes.size() == 1 && !passes[0].standalone,
                            "observable needs more than one pass (at most 4 X/Y factors on wires 0..n-7)");
                p = passes[0];
            } else {
                for (int r = 0; r < RB; ++r) p.reg_bits[r] = kLaneBits + r;
            }
            Built b = build_pass(p, ok, nq, false, nullptr, nullptr, RB, false);
            t->obs.pd = b.pd;
            t->obs.op_begin = nf;
            t->obs.nops = int32_t(b.ops.size());
            for (const FOp &f : b.ops) {
                Recipe r{};
                r.base = f;
                r.pidx = -1;
                r.form = kFixed;
                r.sign = 1.0;
                t->recipes.push_back(r);
            }
            nf += t->obs.nops;
        }
        t->nfops = nf;
        t->cost = (1ull << (nq - kTileBits)) * (t->fwd.size() + 2 * t->rev.size() + 2);
        // lean level of the forward and reverse passes, for every parameter value: the
        // prep kernel recomputes the real and symmetric-phase flags of parametrised gates
        auto recipe_level = [](const Recipe &r, bool rev) -> int {
            const FOp &f = r.base;
            const uint32_t kind = f.meta & 15u;
            const bool dense = kind == kRegDense || kind == kLaneDense;
            const bool plain = (f.rbits & 0xffffu) == 0xffffu && f.cm_rest == 0;
            const int loc = int((f.meta >> 4) & 15u);
            switch (r.form) {
            case kRY: return dense && plain ? 1 : 0;  // real for every angle; Y generator
            case kRX: return dense && plain ? 2 : 0;  // X generator
            case kRotDiag: // d_odd = conj(d_even) for every angle; diagonal generator
                return kind == kDiag && plain && (loc < 4 || loc == kDiagNoReg) ? 2 : 0;
            case kPhase: return 0;
            default: return lean_level_op(f, rev); // fixed: as built, never trainable
            }
        };
        t->lean = 1;
        auto take = [&](const TPass &tp, bool rev) {
            for (int32_t j = 0; j < tp.nops && t->lean > 0; ++j) {
                const int l = recipe_level(t->recipes[size_t(tp.op_begin + j)], rev);
                t->lean = l == 0 ? 0 : std::max(t->lean, l);
            }
        };
        for (const TPass &tp : t->fwd) take(tp, false);
        for (const TPass &tp : t->rev) take(tp, true);
        *out = t.release();
    });
}

lgh_status lgh_tmpl_destroy(lgh_tmpl *tmpl) {
    return guard([&] { delete tmpl; });
}

lgh_status lgh_tmpl_info(const lgh_tmpl *tmpl, int32_t *num_params, int32_t *num_qubits) {
    return guard([&] {
        LGH_REQUIRE(tmpl, "null template");
        if (num_params) *num_params = tmpl->nparams;
        if (num_qubits) *num_qubits = tmpl->n;
    });
}

lgh_status lgh_tmpl_passes(const lgh_tmpl *tmpl, int32_t *forward_passes, int32_t *reverse_passes,
                           int32_t *state_qubits) {
    return guard([&] {
        LGH_REQUIRE(tmpl, "null template");
        if (forward_passes) *forward_passes = int32_t(tmpl->fwd.size());
        if (reverse_passes) *reverse_passes = int32_t(tmpl->rev.size());
        if (state_qubits) *state_qubits = tmpl->nq;
    });
}

lgh_status lgh_batch_run(lgh_tmpl *const *tmpls, int32_t num_circuits, lgh_dtype dtype, const double *params,
                         const int64_t *param_offset, double *expval, double *jac, const int64_t *jac_offset,
                         void *stream) {
    return guard([&] {
        Range rg("lgh_batch_run");
        check_batch_args(tmpls, num_circuits, dtype, params, param_offset, expval, jac, jac_offset);
        if (num_circuits == 0) return;
        if (dtype == LGH_C128)
            batch_run_t<double>(tmpls, num_circuits, params, param_offset, expval, jac, jac_offset, stream);
        else
            batch_run_t<float>(tmpls, num_circuits, params, param_offset, expval, jac, jac_offset, stream);
    });
}

lgh_status lgh_batch_run_serial(lgh_tmpl *const *tmpls, int32_t num_circuits, lgh_dtype dtype,
                                const double *params, const int64_t *param_offset, double *expval, double *jac,
                                const int64_t *jac_offset, void *stream) {
    return guard([&] {
        check_batch_args(tmpls, num_circuits, dtype, params, param_offset, expval, jac, jac_offset);
        auto ok = [](lgh_status s) {
            if (s != LGH_OK) throw Error(s, lgh_last_error());
        };
        for (int32_t c = 0; c < num_circuits; ++c) {
            const lgh_tmpl &t = *tmpls[c];
            const double *pc = t.nparams > 0 ? params + param_offset[c] : nullptr;
            std::vector<lgh_op> ops;
            for (size_t i = 0; i < t.ops.size(); ++i) ops.push_back(op_with(t, i, pc));
            lgh_sv *psi = nullptr, *bra = nullptr;
            ok(lgh_sv_create(&psi, t.n, dtype, t.device, stream));
            struct Del {
                lgh_sv *&a, *&b;
                ~Del() {
                    if (a) lgh_sv_destroy(a);
                    if (b) lgh_sv_destroy(b);
                }
            } del{psi, bra};
            ok(lgh_sv_reset(psi));
            ok(lgh_apply(psi, ops.data(), int64_t(ops.size())));
            std::string word(static_cast<size_t>(t.n), 'I');
            const char *names = "IXYZ";
            std::vector<int32_t> wl(static_cast<size_t>(t.n));
            std::iota(wl.begin(), wl.end(), 0);
            for (size_t k = 0; k < t.obs_paulis.size(); ++k)
                word[size_t(t.obs_wires[k])] = names[t.obs_paulis[k] - LGH_GATE_PAULIX + 1];
            const int64_t offs[2] = {0, t.n};
            const double coef[2] = {1.0, 0.0};
            ok(lgh_expval_pauli_words(psi, 1, word.data(), wl.data(), offs, coef, nullptr, &expval[c]));
            if (!jac || t.nparams == 0) continue;
            ok(lgh_sv_create(&bra, t.n, dtype, t.device, stream));
            ok(lgh_sv_copy_d2d(bra, psi));
            std::vector<lgh_op> pops;
            for (size_t k = 0; k < t.obs_paulis.size(); ++k) {
                lgh_op o{};
                o.gate = t.obs_paulis[k];
                o.num_wires = 1;
                o.wires = &t.obs_wires[k];
                pops.push_back(o);
            }
            if (!pops.empty()) ok(lgh_apply(bra, pops.data(), int64_t(pops.size())));
            // trainable = the parametrised ops driven by a circuit parameter
            std::vector<int64_t> tp;
            std::vector<int32_t> tp_pidx;
            int64_t k = 0;
            for (size_t i = 0; i < ops.size(); ++i) {
                if (ops[i].num_params <= 0) continue;
                if (t.pidx[i] >= 0) {
                    tp.push_back(k);
                    tp_pidx.push_back(t.pidx[i]);
                }
                ++k;
            }
            std::vector<double> row(tp.size());
            lgh_sv *bl[1] = {bra};
            ok(lgh_adjoint_jacobian(psi, bl, 1, ops.data(), int64_t(ops.size()), tp.data(), int64_t(tp.size()), 0,
                                    row.data()));
            double *jr = jac + jac_offset[c];
            for (int32_t j = 0; j < t.nparams; ++j) jr[j] = 0.0;
            for (size_t j = 0; j < tp.size(); ++j) jr[tp_pidx[j]] += row[j];
        }
    });
}

} // extern "C"
