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
#include <complex>
#include <atomic>
#include <cstring>
#include <array>
#include <chrono>
#include <map>
#include <tuple>
#include <unordered_map>

#include "lgh_fused_dev.hpp"

using namespace lgh::fz;

namespace lgh {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

std::atomic<int> g_fusion{1};

// Lane positions fixed to the lowest index bits: a load instruction then
// covers runs of 128 consecutive bytes (8 complex doubles or 16 complex
// floats), whole cache lines; the other lane positions are chosen per pass.
int fixed_lanes(lgh_dtype dt) { return dt == LGH_C128 ? 3 : 4; }
// adjoint 0: by state size and dtype (adjoint_rb_default)
std::atomic<int> g_rb_fwd{kRBForward}, g_rb_adj{0};
int adjoint_rb_default(int n, lgh_dtype dt) { return dt == LGH_C128 && n >= kAdjointWideQubits ? 4 : kRBAdjoint; }
// passes with only lean gate forms run in the lean kernels (lgh_set_lean_passes)
std::atomic<bool> g_lean{true};
// caps on the wavefronts of a pass's grid (0: the per-CU defaults below)
std::atomic<int64_t> g_waves_fwd{0}, g_waves_adj{0};

// Wavefronts per CU of a pass's grid: many more than fit at once (12 per CU
// for the default kernels), each taking a short run of consecutive tiles, so
// that the dispatcher keeps every CU fed to the end of the pass.  Measured at
// 28 qubits: 192-256 per CU (8-11 tiles each) beat 32 (64 tiles each) by
// 6-7 % for the adjoint pass, whose per-workgroup overlap reduction makes
// still shorter runs cost more; the forward pass is fastest with one tile
// per wavefront (5 % over 32 tiles).  The autotuner may pick other values
// per state size (at most kMaxWavesPerCUAdjoint for the adjoint).
constexpr int kWavesPerCUForward = 4096; // one tile per wavefront up to 30 qubits
constexpr int kWavesPerCUAdjoint = 256;
constexpr int kMaxWavesPerCUAdjoint = 512;

// Tile height and grid of the fused passes per (device, qubits, dtype,
// direction): chosen by lgh_autotune_fused, else the defaults (tile bits of
// lgh_set_tile_bits, the wavefronts per CU above).
struct FusedCfg {
    int rb;
    int waves_per_cu;
    int slots; // trainable gates per adjoint pass (<= adj_slot_cap(rb)); unused forward
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, FusedCfg> g_tuned;
// the autotuner's candidate, for the thread that times it
thread_local const FusedCfg *t_force[2] = {nullptr, nullptr};

void clear_tuned() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

// dir 0: forward passes, 1: adjoint passes
FusedCfg fused_cfg(int device, int n, lgh_dtype dt, int dir) {
    if (t_force[dir]) return *t_force[dir];
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(std::make_tuple(device, n, int(dt), dir));
        if (it != g_tuned.end()) return it->second;
    }
    const int rb_adj = g_rb_adj.load() ? g_rb_adj.load() : adjoint_rb_default(n, dt);
    return dir == 0 ? FusedCfg{g_rb_fwd.load(), kWavesPerCUForward, kMaxSlots}
                    : FusedCfg{rb_adj, kWavesPerCUAdjoint, adj_slot_cap(rb_adj)};
}

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
void set_tile_bits(int forward, int adjoint) {
    LGH_REQUIRE((forward == 3 || forward == 4) && (adjoint == 0 || adjoint == 3 || adjoint == 4),
                "tile bits are 3 or 4 (adjoint 0: the default by state size)");
    g_rb_fwd.store(forward);
    g_rb_adj.store(adjoint);
    clear_tuned();
}
int fusion_enabled() { return g_fusion.load(); }
void set_pass_waves(int64_t forward, int64_t adjoint) {
    LGH_REQUIRE(forward >= 0 && adjoint >= 0, "wavefront caps are >= 0 (0: default)");
    g_waves_fwd.store(forward);
    g_waves_adj.store(adjoint);
    clear_tuned();
}

bool fusable(const KOp &k, int n, int rb) {
    if (n < kLaneBits + rb) return false;
    if (k.kind == KOp::DIAG) return true;
    return k.kind == KOp::DENSE && k.k == 1;
}

namespace {
// Greedy list scheduling of `ops` (already compiled) into passes.  depth: how
// many tile bits ahead a choice of tile bit is scored (1: the gates it lets in
// alone; 2, 3: plus the most the next one or two choices let in).
std::vector<Pass> schedule_greedy(const std::vector<KOp> &ops, int n, const std::vector<int> &trainable_slot, int rb,
                                  int fixed_lanes, int max_slots, int depth) {
    // index bits the scheduler chooses per pass: the free lane positions and the registers
    const int nfree = kLaneBits - fixed_lanes + rb;
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
    // list-scheduler state; copied to try out a register bit before taking it
    struct State {
        std::vector<int> indeg, ready;
        int slots = 0;
    };
    State st;
    st.indeg = indeg;
    for (int i = 0; i < N; ++i)
        if (indeg[i] == 0) st.ready.push_back(i);
    int remaining = N;
    const uint64_t lane_mask = (1ull << fixed_lanes) - 1;
    std::vector<char> fus(N), dense(N), train(N);
    for (int i = 0; i < N; ++i) {
        fus[i] = fusable(ops[i], n, rb);
        dense[i] = ops[i].kind == KOp::DENSE;
        train[i] = !trainable_slot.empty() && trainable_slot[i] >= 0;
    }
    auto release = [&](State &s, int i) {
        for (int t : succ[i])
            if (--s.indeg[t] == 0) s.ready.push_back(t);
    };
    // Absorbs every ready gate whose targets are lane or register bits (and
    // the gates this makes ready); returns the number of dense gates taken.
    auto absorb = [&](State &s, uint64_t reg, std::vector<int> *taken) {
        int nd = 0;
        for (bool progress = true; progress;) {
            progress = false;
            for (size_t r = 0; r < s.ready.size();) {
                const int i = s.ready[r];
                if (!fus[i] || (use[i].nondiag & ~lane_mask & ~reg) || (train[i] && s.slots >= max_slots)) {
                    ++r;
                    continue;
                }
                s.slots += train[i];
                nd += dense[i];
                if (taken) taken->push_back(i);
                s.ready[r] = s.ready.back();
                s.ready.pop_back();
                release(s, i);
                progress = true;
            }
        }
        return nd;
    };
    // targets (outside the lane bits and `reg`) of ready dense gates
    auto blocked_targets = [&](const State &s, uint64_t reg) {
        uint64_t t = 0;
        for (int i : s.ready)
            if (fus[i] && dense[i] && !(train[i] && s.slots >= max_slots)) t |= use[i].nondiag & ~lane_mask & ~reg;
        return t;
    };
    while (remaining > 0) {
        // a ready gate that cannot be fused runs on its own
        int solo = -1;
        for (int i : st.ready)
            if (!fus[i] && (solo < 0 || i < solo)) solo = i;
        if (solo >= 0) {
            Pass p;
            p.standalone = true;
            p.ops.push_back(solo);
            passes.push_back(p);
            st.ready.erase(std::find(st.ready.begin(), st.ready.end(), solo));
            release(st, solo);
            --remaining;
            continue;
        }
        // The free tile bits (free lane positions and registers) are chosen
        // while the pass grows: absorb every ready gate whose targets are
        // tile bits; while a tile bit is free, try each target of a blocked
        // dense gate as that bit and keep the one that lets the most dense
        // gates in (ties: the target of the earliest gate), then absorb
        // again.  A chain of gates (each one's target the next one's
        // control) then fills a pass instead of starting a new one per gate.
        Pass p;
        uint64_t reg = 0;
        int nreg = 0;
        st.slots = 0;
        for (;;) {
            absorb(st, reg, &p.ops);
            if (nreg >= nfree) break;
            const uint64_t cand = blocked_targets(st, reg);
            if (!cand) break;
            uint64_t best = 0;
            long best_score = -1;
            for (uint64_t c = cand; c; c &= c - 1) {
                const uint64_t bit = c & (~c + 1);
                State t = st;
                long score = absorb(t, reg | bit, nullptr);
                if (depth >= 2 && nreg + 1 < nfree) {
                    long best2 = 0;
                    const uint64_t cand2 = blocked_targets(t, reg | bit);
                    for (uint64_t c2 = cand2; c2; c2 &= c2 - 1) {
                        const uint64_t bit2 = c2 & (~c2 + 1);
                        State t2 = t;
                        long s2 = absorb(t2, reg | bit | bit2, nullptr);
                        if (depth >= 3 && nreg + 2 < nfree) {
                            long best3 = 0;
                            const uint64_t cand3 = blocked_targets(t2, reg | bit | bit2);
                            for (uint64_t c3 = cand3; c3; c3 &= c3 - 1) {
                                State t3 = t2;
                                best3 = std::max(best3, long(absorb(t3, reg | bit | bit2 | (c3 & (~c3 + 1)), nullptr)));
                            }
                            s2 += best3;
                        }
                        best2 = std::max(best2, s2);
                    }
                    score += best2;
                }
                int first = N;
                for (int i : st.ready)
                    if (dense[i] && (use[i].nondiag & bit)) first = std::min(first, i);
                score = score * (N + 1) + (N - first);
                if (score > best_score) {
                    best_score = score;
                    best = bit;
                }
            }
            reg |= best;
            ++nreg;
        }
        remaining -= static_cast<int>(p.ops.size());
        // gates only depend on earlier gates: index order is a valid order
        std::sort(p.ops.begin(), p.ops.end());
        // pad with the lowest unused bits so that every pass has all its tile bits
        for (int b = fixed_lanes; b < n && nreg < nfree; ++b)
            if (!(reg & (1ull << b))) {
                reg |= 1ull << b;
                ++nreg;
            }
        // The bits with the most dense gates become register bits (plain
        // register math); the others go to the free lane positions (cross-lane
        // exchanges).  Ties: the lower bit is a lane bit.
        std::vector<int> chosen, weight(size_t(n), 0);
        for (int b = 0; b < n; ++b)
            if (reg & (1ull << b)) chosen.push_back(b);
        for (int i : p.ops)
            if (dense[i]) weight[size_t(ops[size_t(i)].tbits[0])]++;
        std::stable_sort(chosen.begin(), chosen.end(), [&](int x, int y) { return weight[size_t(x)] < weight[size_t(y)]; });
        std::vector<int> lanes(chosen.begin(), chosen.begin() + (kLaneBits - fixed_lanes));
        std::vector<int> regs(chosen.begin() + (kLaneBits - fixed_lanes), chosen.end());
        // free lane positions in order of cost: the ones inside a 16-lane
        // row exchange with one DPP move per dword, positions 4 and 5 swap
        // rows; the lane bit with the most dense gates takes the cheapest
        std::stable_sort(lanes.begin(), lanes.end(), [&](int x, int y) {
            return weight[size_t(x)] != weight[size_t(y)] ? weight[size_t(x)] > weight[size_t(y)] : x < y;
        });
        std::sort(regs.begin(), regs.end());
        for (int q = 0; q < kLaneBits; ++q) p.lane_bits[q] = q < fixed_lanes ? q : lanes[size_t(q - fixed_lanes)];
        for (int r = 0; r < rb; ++r) p.reg_bits[r] = regs[size_t(r)];
        passes.push_back(p);
    }
    return passes;
}
} // namespace

// The greedy schedule is sensitive to which tile bit it takes next: scoring
// each choice one, two or three choices ahead gives different partitions, none
// always the shortest (28-qubit brick circuit, 498 gates: 11, 11 and 10
// adjoint passes, 8, 8 and 7 forward passes).  The schedule with the fewest
// passes is kept (ties: the shallower lookahead).  Three ahead is skipped for
// very long circuits, where it takes more than a few tenths of a second.
std::vector<Pass> schedule_passes(const std::vector<KOp> &ops, int n, const std::vector<int> &trainable_slot,
                                  int rb, int fixed_lanes, int max_slots) {
    LGH_REQUIRE(max_slots >= 1 && max_slots <= 254, "fused pass: bad number of slots");
    LGH_REQUIRE(fixed_lanes >= 1 && fixed_lanes <= kLaneBits, "fused pass: bad number of fixed lane bits");
    const int max_depth = ops.size() <= 8192 ? 3 : 2;
    std::vector<Pass> best = schedule_greedy(ops, n, trainable_slot, rb, fixed_lanes, max_slots, 1);
    for (int depth = 2; depth <= max_depth; ++depth) {
        std::vector<Pass> p = schedule_greedy(ops, n, trainable_slot, rb, fixed_lanes, max_slots, depth);
        if (p.size() < best.size()) best = std::move(p);
    }
    return best;
}

// A schedule depends only on the gates' structure -- kind, targets,
// controls, parity mask, which gates are trainable -- and on n, the register
// bits and the fixed lane positions, never on angles or matrix entries.
// Repeated executions of one circuit shape (optimisation loops, a serving
// stream of one circuit family, the forward re-execution before each
// adjoint) reuse the schedule instead of rebuilding the dependency graph.
namespace {
struct SchedCache {
    std::mutex mu;
    std::unordered_map<std::string, std::shared_ptr<const std::vector<Pass>>> map;
};
SchedCache &sched_cache() {
    static SchedCache *c = new SchedCache();
    return *c;
}
constexpr size_t kSchedCacheMax = 512;
} // namespace

bool lean_enabled() { return g_lean.load(); }

std::shared_ptr<const std::vector<Pass>> schedule_passes_cached(const std::vector<KOp> &ops, int n,
                                                                const std::vector<int> &trainable_slot, int rb,
                                                                int fixed_lanes, int max_slots) {
    std::string key;
    key.reserve(24 + ops.size() * 24);
    auto put = [&](uint64_t v) { key.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    put(uint64_t(n) | uint64_t(rb) << 8 | uint64_t(fixed_lanes) << 16 | uint64_t(ops.size()) << 24);
    put(uint64_t(max_slots));
    for (size_t i = 0; i < ops.size(); ++i) {
        const KOp &k = ops[i];
        const bool train = !trainable_slot.empty() && trainable_slot[i] >= 0;
        const Use u = uses_of(k);
        put(uint64_t(k.kind) | uint64_t(k.k) << 4 | uint64_t(train) << 12 |
            uint64_t(k.kind == KOp::DENSE ? k.tbits[0] : 0) << 16);
        put(u.diag);
        put(u.nondiag);
    }
    SchedCache &c = sched_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.map.find(key);
        if (it != c.map.end()) return it->second;
    }
    auto sched =
        std::make_shared<const std::vector<Pass>>(schedule_passes(ops, n, trainable_slot, rb, fixed_lanes, max_slots));
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.map.size() >= kSchedCacheMax) c.map.clear();
    c.map.emplace(std::move(key), sched);
    return sched;
}

} // namespace lgh

using namespace lgh;

namespace {

// Forward pass: every wavefront processes tiles [w*chunk, (w+1)*chunk).
// Lv 1, 2: the pass holds only the gate forms of that lean level
// (lean_level); 0: any fusable gate.
template <class R, int RB, int Lv = 0>
__global__ __launch_bounds__(kBlock) void k_fused_pass(cx<R> *__restrict__ sv, uint64_t ntiles, uint64_t chunk,
                                                       const PassDesc *__restrict__ pd,
                                                       const FOp *__restrict__ ops) {
    const int lane = threadIdx.x & 63;
    const CFOp *cops = (const CFOp *)ops;
    const uint64_t wave = xcd_block() * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t t0 = wave * chunk, t1 = min(t0 + chunk, ntiles);
    const int nops = pd->nops;
    constexpr int NR = 1 << RB;
    uint64_t roff[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) roff[k] = pd->reg_off[k];
    int fs[kMaxFree];
#pragma unroll
    for (int r = 0; r < kMaxFree; ++r) fs[r] = pd->free_sorted[r];
    const uint64_t tmask = pd->tile_mask, loff = lane_offset(lane, pd);
    // first index of the wavefront's tile, stepped from tile to tile
    uint64_t g = tile_base_free(t0, pd->nfix, fs, pd->nfree);
    for (uint64_t t = t0; t < t1; ++t, g = next_tile(g, tmask)) {
        const uint64_t base = g | loff;
        cx<R> v[NR];
        cx<R> *p = sv + base; // the tile bits are disjoint: | equals +
#pragma unroll
        for (int k = 0; k < NR; ++k) v[k] = ldnt(p + roff[k]);
        FOp cur = load_op(cops);
        wait_op(cur);
        for (int o = 0; o < nops;) {
            if (f_kind(cur) == kDiagRun) {
                diag_run<R, NR, (Lv > 0)>(cops + o, cur, base, v);
                o += 1 + run_ndata(cur) + run_nmem(cur);
                cur = load_op(cops + (o < nops ? o : 0));
                wait_op(cur);
                continue;
            }
            const FOp nxt = load_op(cops + (o + 1 < nops ? o + 1 : o)); // prefetch the next descriptor
            if constexpr (Lv > 0)
                apply_lean<R, NR, Lv>(cur, lane, base, v);
            else
                apply_fop<R, NR>(cur, lane, base, v);
            cur = nxt;
            ++o;
        }
        // recompute the store addresses instead of keeping NR of them live
        asm volatile("" : "+v"(p));
#pragma unroll
        for (int k = 0; k < NR; ++k) stnt(p + roff[k], v[k]);
    }
}

// Reverse pass of the adjoint sweep on lambda and one observable state.  The
// ops of the pass are stored in reverse order with U^dagger matrices.  Each
// lane accumulates its overlap contributions in its own LDS column; the
// reduction over lanes and wavefronts happens once at the end.
// lam_in: where lambda is read (lam itself, or the forward state for the
// first pass of a sweep, which then needs no copy of it); lam: where it is
// written.  Each thread reads and writes its own amplitudes only, so the two
// may be the same array.
// bra_zmask != 0 with bra_from_lam: the observable state is not read but
// formed from lambda's loaded values as Z-word(bra_zmask) lambda (the sign
// (-1)^popcount(index & bra_zmask) per amplitude), then written like any
// pass's output: the first reverse pass of a sweep with a Z-word observable
// needs no separate pass that forms it.
template <class R, int RB, int Lv>
__device__ __forceinline__ void fused_adj_body(const cx<R> *lam_in, cx<R> *lam, cx<R> *__restrict__ bra,
                                               uint64_t ntiles, uint64_t chunk, const PassDesc *__restrict__ pd,
                                               const FOp *__restrict__ ops, const FGen *__restrict__ gens,
                                               double *__restrict__ partials, uint64_t bra_zmask, int bra_from_lam) {
    static_assert(kAccCols * kAccLanes == kBlock, "one accumulator column per lane quad");
    __shared__ double acc[adj_slot_cap(RB)][kAccCols];
    __shared__ double sink[kBlock];
    const AccCols cols{acc, sink};
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const CFOp *cops = (const CFOp *)ops;
    const CFGen *cgens = (const CFGen *)gens;
    const int nslots = pd->nslots;
    // each column is cleared by the lane quad that adds to it (all four
    // lanes store the same zero)
    for (int s = 0; s < nslots; ++s) acc[s][threadIdx.x >> 2] = 0.0;
    const uint64_t wave = xcd_block() * kWavesPerBlock + w;
    const uint64_t t0 = wave * chunk, t1 = min(t0 + chunk, ntiles);
    const int nops = pd->nops;
    constexpr int NR = 1 << RB;
    uint64_t roff[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) roff[k] = pd->reg_off[k];
    int fs[kMaxFree];
#pragma unroll
    for (int r = 0; r < kMaxFree; ++r) fs[r] = pd->free_sorted[r];
    const uint64_t tmask = pd->tile_mask, loff = lane_offset(lane, pd);
    // first index of the wavefront's tile, stepped from tile to tile
    uint64_t g = tile_base_free(t0, pd->nfix, fs, pd->nfree);
    for (uint64_t t = t0; t < t1; ++t, g = next_tile(g, tmask)) {
        const uint64_t base = g | loff;
        cx<R> a[NR], b[NR];
        cx<R> *pa = lam + base, *pb = bra + base; // the tile bits are disjoint
        const cx<R> *qa = lam_in + base;
        if (bra_from_lam) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                a[k] = ldnt(qa + roff[k]);
                const uint32_t m = (__popcll((base | roff[k]) & bra_zmask) & 1) ? kSignBit : 0u;
                b[k] = cx<R>{flip(a[k].re, m), flip(a[k].im, m)};
            }
        } else {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                a[k] = ldnt(qa + roff[k]);
                b[k] = ldnt(pb + roff[k]);
            }
        }
        FOp cur = load_op(cops);
        wait_op(cur);
        for (int o = 0; o < nops;) {
            if (f_kind(cur) == kDiagRun) {
                diag_run_adj<R, NR, RB, (Lv > 0)>(cops + o, cur, base, a, b, cols);
                o += 1 + run_ndata(cur) + run_nmem(cur);
                cur = load_op(cops + (o < nops ? o : 0));
                wait_op(cur);
                continue;
            }
            const FOp nxt = load_op(cops + (o + 1 < nops ? o + 1 : o));
            const int slot = f_slot(cur);
            // overlaps go to the lane pair's LDS column (no contention; the
            // additions keep program order)
            if constexpr (Lv > 0) {
                if (f_kind(cur) == kDiag) {
                    if (Lv == 1 || !f_dsym(cur)) { // the pass's closing scale
                        scale_lean<R, NR>(cur, a);
                        scale_lean<R, NR>(cur, b);
                    } else if constexpr (Lv == 2) {
                        if (slot >= 0)
                            acc_add(cols, slot, adj_diag_step<R, NR, true>(cur, cur.m[2].re, cur.m[3].re, base, a, b));
                        else
                            adj_diag_step<R, NR, false>(cur, 0.0, 0.0, base, a, b);
                    }
                } else if (Lv == 2 && !f_plain(cur)) { // a controlled X, not trainable
                    if constexpr (Lv == 2) {
                        const M2<R> m = m2<R>(cur);
                        cx_lean<R, NR>(cur, m, lane, base, a);
                        cx_lean<R, NR>(cur, m, lane, base, b);
                    }
                } else if (slot >= 0) {
                    adj_plain_lean<R, NR, true, Lv>(cur, lane, a, b, [&](double v) { acc_add(cols, slot, v); });
                } else {
                    adj_plain_lean<R, NR, false, Lv>(cur, lane, a, b, [](double) {});
                }
            } else if (f_kind(cur) != kDiag && f_plain(cur)) {
                if (slot >= 0) {
                    // X and Y generators need no generator record
                    FGen g{};
                    if (f_gkind(cur) == kGenDense) g = load_gen(cgens + slot);
                    acc_add(cols, slot, adj_plain_dense<R, NR, true>(cur, &g, lane, a, b));
                } else {
                    adj_plain_dense<R, NR, false>(cur, nullptr, lane, a, b);
                }
            } else if (diag_fast(cur)) {
                if (slot >= 0)
                    acc_add(cols, slot, adj_diag_step<R, NR, true>(cur, cur.m[2].re, cur.m[3].re, base, a, b));
                else
                    adj_diag_step<R, NR, false>(cur, 0.0, 0.0, base, a, b);
            } else {
                if (slot >= 0) {
                    const FGen g = f_kind(cur) == kDiag ? diag_gen(cur) : load_gen(cgens + slot);
                    acc_add(cols, slot, gen_dot<R, NR>(cur, g, lane, base, a, b));
                }
                apply_fop<R, NR>(cur, lane, base, a);
                apply_fop<R, NR>(cur, lane, base, b);
            }
            cur = nxt;
            ++o;
        }
        asm volatile("" : "+v"(pa), "+v"(pb));
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            stnt(pa + roff[k], a[k]);
            stnt(pb + roff[k], b[k]);
        }
    }
    __syncthreads();
    // slot s: wavefront w sums the 64 columns in a fixed order
    static_assert(kAccCols == 64, "one column per lane of the summing wavefront");
    for (int s = w; s < nslots; s += kWavesPerBlock) {
        double v = acc[s][lane];
        v = wave_sum(v);
        if (lane == 0) partials[size_t(s) * gridDim.x + blockIdx.x] = v;
    }
}

// Three register bits fit three wavefronts per SIMD in the general kernel
// (at most 168 VGPRs) and four in the lean ones (103-117 VGPRs; 26 KiB of LDS
// per workgroup).  The complex128
// lean kernels are held to 96 VGPRs for five wavefronts per SIMD (a few spills,
// mostly outside the gate loop; measured 1.3 % faster per pass at level 1 and
// 2.2 % at level 2).
// Lv: lean level of the pass (0: any fusable gate).
template <class R, int Lv = 0>
__global__ __launch_bounds__(kBlock, (Lv >= 1 && sizeof(R) == 8) ? 5 : 3) void k_fused_adj3(
    const cx<R> *lam_in, cx<R> *lam, cx<R> *__restrict__ bra, uint64_t ntiles, uint64_t chunk,
    const PassDesc *__restrict__ pd, const FOp *__restrict__ ops, const FGen *__restrict__ gens,
    double *__restrict__ partials, uint64_t bra_zmask, int bra_from_lam) {
    fused_adj_body<R, 3, Lv>(lam_in, lam, bra, ntiles, chunk, pd, ops, gens, partials, bra_zmask, bra_from_lam);
}
// Four register bits: the complex128 lean kernels are held to 168 VGPRs for
// three wavefronts per SIMD (39 dwords spilled to scratch at level 1; 4.50
// against 4.88 ms per 28-qubit pass at two wavefronts).
template <class R, int Lv = 0>
__global__ __launch_bounds__(kBlock, (Lv >= 1 && sizeof(R) == 8) ? 3 : 1) void k_fused_adj4(const cx<R> *lam_in, cx<R> *lam, cx<R> *__restrict__ bra,
                                                    uint64_t ntiles, uint64_t chunk, const PassDesc *__restrict__ pd,
                                                    const FOp *__restrict__ ops, const FGen *__restrict__ gens,
                                                    double *__restrict__ partials, uint64_t bra_zmask,
                                                    int bra_from_lam) {
    fused_adj_body<R, 4, Lv>(lam_in, lam, bra, ntiles, chunk, pd, ops, gens, partials, bra_zmask, bra_from_lam);
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

// host form of FOp before its fields are packed into meta
struct HOp {
    uint32_t kind = 0, gkind = 0, dneg = 0;
    uint32_t rot = 0; // 1: matrix [[1, -t], [t, 1]], 2: [[t, -1], [1, t]] (a rotation with a factor taken out)
    int loc = 0, slot = -1;
    uint64_t cm_reg = 0, pm_reg = 0; // control / parity bits among the register bits
    FOp f{};
};

FOp pack(const HOp &h, const uint64_t (&reg_off)[kMaxNR], int nr) {
    FOp f = h.f;
    bool real = h.kind != kDiag;
    for (int q = 0; q < 4 && real; ++q) real = f.m[q].im == 0.0;
    const bool dsym = h.kind == kDiag && f.m[1].re == f.m[0].re && f.m[1].im == -f.m[0].im;
    const bool isx = h.kind != kDiag && f.m[0].re == 0.0 && f.m[0].im == 0.0 && f.m[1].re == 1.0 &&
                     f.m[1].im == 0.0 && f.m[2].re == 1.0 && f.m[2].im == 0.0 && f.m[3].re == 0.0 && f.m[3].im == 0.0;
    f.meta = h.kind | uint32_t(h.loc) << 4 | uint32_t(h.slot + 1) << 8 | h.gkind << 16 | h.dneg << 20 |
             uint32_t(real) << 21 | uint32_t(dsym) << 22 | uint32_t(isx) << 23 | h.rot << 28;
    // (bits 28-29: the rotation form)
    uint32_t rok = 0, rpar = 0;
    for (int k = 0; k < kMaxNR; ++k) {
        // registers beyond nr do not exist; their bits read as "controls hold"
        if (k >= nr || (reg_off[k] & h.cm_reg) == h.cm_reg) rok |= 1u << k;
        if (k >= nr) continue;
        if (__builtin_popcountll(reg_off[k] & h.pm_reg) & 1) rpar |= 1u << k;
    }
    f.rbits = rok | rpar << 16;
    return f;
}

bool is_neg_diag(const KOp &k) {
    return k.kind == KOp::DIAG && std::abs(k.d_even + 1.0) < 1e-300 && std::abs(k.d_odd + 1.0) < 1e-300;
}

// Reorders the gates of a pass (`order`, a valid execution order) so that the
// diagonal gates come in as few groups as possible: take every diagonal gate
// whose predecessors are placed, then every other gate whose predecessors are
// placed, and repeat.  Only commuting gates trade places (two gates commute
// when every qubit they share is used diagonally by both), so the pass's
// product is unchanged, and so is every adjoint overlap: a gate's derivative
// commutes with whatever the gate commutes with.  Consecutive diagonal gates
// become one diagonal run, and single-qubit gates on one bit stay mergeable
// (nothing that touches the bit moves between them).
void cluster_diagonals(std::vector<int> &order, const std::vector<KOp> &kops, int n) {
    const int m = static_cast<int>(order.size());
    if (m < 3) return;
    std::vector<Use> use(size_t(m), Use{});
    std::vector<char> diag(static_cast<size_t>(m));
    for (int i = 0; i < m; ++i) {
        use[size_t(i)] = uses_of(kops[size_t(order[size_t(i)])]);
        diag[size_t(i)] = kops[size_t(order[size_t(i)])].kind == KOp::DIAG;
    }
    // dependency edges as in schedule_passes: per bit, the last non-diagonal
    // user and the diagonal users since
    std::vector<std::vector<int>> succ(static_cast<size_t>(m));
    std::vector<int> indeg(size_t(m), 0), last_nd(size_t(n), -1), deps;
    std::vector<std::vector<int>> diag_since(static_cast<size_t>(n));
    for (int j = 0; j < m; ++j) {
        deps.clear();
        for (int b = 0; b < n; ++b) {
            const uint64_t bit = 1ull << b;
            if (use[size_t(j)].nondiag & bit) {
                if (last_nd[size_t(b)] >= 0) deps.push_back(last_nd[size_t(b)]);
                deps.insert(deps.end(), diag_since[size_t(b)].begin(), diag_since[size_t(b)].end());
                last_nd[size_t(b)] = j;
                diag_since[size_t(b)].clear();
            } else if (use[size_t(j)].diag & bit) {
                if (last_nd[size_t(b)] >= 0) deps.push_back(last_nd[size_t(b)]);
                diag_since[size_t(b)].push_back(j);
            }
        }
        std::sort(deps.begin(), deps.end());
        deps.erase(std::unique(deps.begin(), deps.end()), deps.end());
        for (int d : deps) {
            succ[size_t(d)].push_back(j);
            ++indeg[size_t(j)];
        }
    }
    std::vector<int> out;
    out.reserve(size_t(m));
    std::vector<char> placed(size_t(m), 0);
    bool want_diag = diag[0] != 0;
    while (static_cast<int>(out.size()) < m) {
        // one sweep in index order places a gate's successors later in the
        // same sweep once they are ready, so the sweep reaches the closure
        for (int i = 0; i < m; ++i) {
            if (placed[size_t(i)] || indeg[size_t(i)] != 0 || (diag[size_t(i)] != 0) != want_diag) continue;
            placed[size_t(i)] = 1;
            out.push_back(order[size_t(i)]);
            for (int s : succ[size_t(i)]) --indeg[size_t(s)];
        }
        want_diag = !want_diag;
    }
    order.swap(out);
}

} // namespace

Built lgh::fz::build_pass(const Pass &p, const std::vector<KOp> &kops, int n, bool reverse,
                          const std::vector<KGen> *gens, const std::vector<int> *slot_of, int rb, bool merge_forward,
                          bool diag_runs, bool single) {
    LGH_REQUIRE(rb >= 1 && rb <= kMaxRB, "fused pass: bad register-bit count");
    const int nr = 1 << rb;
    Built b{};
    b.pd.rb = rb;
    uint64_t regmask = 0, lanemask = 0;
    for (int q = 0; q < kLaneBits; ++q) {
        LGH_REQUIRE(p.lane_bits[q] >= 0 && p.lane_bits[q] < n && !(lanemask >> p.lane_bits[q] & 1),
                    "fused pass: bad lane bits");
        lanemask |= 1ull << p.lane_bits[q];
        b.pd.lane_bits[q] = p.lane_bits[q];
    }
    for (int r = 0; r < rb; ++r) {
        LGH_REQUIRE(p.reg_bits[r] >= 0 && p.reg_bits[r] < n && !((lanemask | regmask) >> p.reg_bits[r] & 1),
                    "fused pass: bad register bits");
        b.pd.reg_bits[r] = p.reg_bits[r];
        b.pd.reg_sorted[r] = p.reg_bits[r];
        regmask |= 1ull << p.reg_bits[r];
    }
    std::sort(b.pd.reg_sorted, b.pd.reg_sorted + rb);
    const uint64_t tilemask = regmask | lanemask;
    // lane positions that carry the lowest index bits stay out of the tile id
    int nfix = 0;
    while (nfix < kLaneBits && p.lane_bits[nfix] == nfix) ++nfix;
    b.pd.nfix = nfix;
    b.pd.nfree = 0;
    for (int bit = nfix; bit < n; ++bit)
        if (tilemask >> bit & 1) b.pd.free_sorted[b.pd.nfree++] = bit;
    LGH_REQUIRE(b.pd.nfree <= kMaxFree, "fused pass: too many free tile bits");
    b.pd.tile_mask = tilemask;
    for (int k = 0; k < nr; ++k) {
        uint64_t o = 0;
        for (int r = 0; r < rb; ++r)
            if ((k >> r) & 1) o |= 1ull << p.reg_bits[r];
        b.pd.reg_off[k] = o;
    }
    auto reg_loc = [&](int bit) {
        for (int r = 0; r < rb; ++r)
            if (p.reg_bits[r] == bit) return r;
        return -1;
    };
    auto lane_loc = [&](int bit) {
        for (int q = 0; q < kLaneBits; ++q)
            if (p.lane_bits[q] == bit) return q;
        return -1;
    };
    std::vector<int> order = p.ops;
    if (reverse) std::reverse(order.begin(), order.end());
    if (diag_runs) cluster_diagonals(order, kops, n);
    // forward passes merge consecutive single-qubit gates on the same tile bit
    // into one 2x2 (untrained passes only)
    const bool merge = merge_forward && slot_of == nullptr;
    std::vector<HOp> hops;
    std::vector<int> last_single(64, -1); // per bit: index in hops of a mergeable gate
    int nslots = 0;
    for (int i : order) {
        const KOp &k = kops[size_t(i)];
        HOp h;
        FOp &f = h.f;
        int single_bit = -1; // uncontrolled gate acting on exactly one tile bit
        if (k.kind == KOp::DIAG) {
            h.kind = kDiag;
            h.cm_reg = k.cmask & regmask;
            f.cm_rest = k.cmask & ~regmask;
            h.pm_reg = k.pmask & regmask;
            f.pm_rest = k.pmask & ~regmask;
            f.m[0] = to_cx<double>(k.d_even);
            f.m[1] = to_cx<double>(k.d_odd);
            h.dneg = is_neg_diag(k) ? 1 : 0;
            // register bit of the parity mask (diagonal fast paths)
            h.loc = h.pm_reg == 0 ? kDiagNoReg
                  : __builtin_popcountll(h.pm_reg) == 1 ? reg_loc(__builtin_ctzll(h.pm_reg)) : 15;
            if (k.cmask == 0 && __builtin_popcountll(k.pmask) == 1 && (k.pmask & tilemask))
                single_bit = __builtin_ctzll(k.pmask);
        } else {
            const int t = k.tbits[0];
            h.cm_reg = k.cmask & regmask;
            f.cm_rest = k.cmask & ~regmask;
            if (lane_loc(t) >= 0) {
                h.kind = kLaneDense;
                h.loc = lane_loc(t);
            } else {
                h.kind = kRegDense;
                h.loc = reg_loc(t);
                LGH_REQUIRE(h.loc >= 0, "fused pass: target is not a register bit");
            }
            for (int q = 0; q < 4; ++q) f.m[q] = to_cx<double>(k.mat[size_t(q)]);
            if (k.cmask == 0) single_bit = t;
        }
        const bool trained = slot_of && (*slot_of)[size_t(i)] >= 0;
        if (trained) {
            const KGen &g = (*gens)[size_t(i)];
            FGen fg{};
            h.slot = nslots++;
            b.slot_ops.push_back(i);
            if (g.mat.empty()) {
                h.gkind = kGenDiag;
                fg.g[0] = cx<double>{g.g_even, 0.0};
                fg.g[1] = cx<double>{g.g_odd, 0.0};
                f.m[2] = fg.g[0]; // read in place by the adjoint kernels
                f.m[3] = fg.g[1];
            } else {
                h.gkind = kGenDense;
                for (int q = 0; q < 4; ++q) fg.g[q] = to_cx<double>(g.mat[size_t(q)]);
                const auto &G = g.mat;
                const std::complex<double> i1(0, 1);
                if (G[0] == 0.0 && G[3] == 0.0 && G[1] == 1.0 && G[2] == 1.0) h.gkind = kGenX;
                else if (G[0] == 0.0 && G[3] == 0.0 && G[1] == -i1 && G[2] == i1) h.gkind = kGenY;
            }
            b.gens.push_back(fg);
        }
        // bits this gate touches invalidate pending merges on them
        const uint64_t touched = k.kind == KOp::DIAG ? (k.cmask | k.pmask) : (k.cmask | (1ull << k.tbits[0]));
        // With diagonal runs a diagonal gate costs less as a run member (on
        // a register bit it is folded into the run's per-register constant)
        // than as a factor that turns a real 2x2 complex, and two diagonal
        // gates stay diagonal in a run: such pairs are not merged.
        auto real_dense = [](const HOp &x) {
            if (x.kind == kDiag) return false;
            for (int q = 0; q < 4; ++q)
                if (x.f.m[q].im != 0.0) return false;
            return true;
        };
        auto keep_apart = [&](const HOp &x, const HOp &y) {
            if (!diag_runs || (x.kind != kDiag && y.kind != kDiag)) return false;
            const HOp &other = x.kind == kDiag ? y : x;
            return other.kind == kDiag || real_dense(other);
        };
        if (merge && single_bit >= 0 && last_single[size_t(single_bit)] >= 0 &&
            !keep_apart(h, hops[size_t(last_single[size_t(single_bit)])])) {
            // fold into the earlier gate on the same bit: M <- M_new * M_old
            HOp &prev = hops[size_t(last_single[size_t(single_bit)])];
            const cx<double> zero{0, 0};
            cx<double> mn[4], mo[4];
            if (h.kind == kDiag) {
                mn[0] = f.m[0]; mn[1] = zero; mn[2] = zero; mn[3] = f.m[1];
            } else {
                for (int q = 0; q < 4; ++q) mn[q] = f.m[q];
            }
            if (prev.kind == kDiag) {
                mo[0] = prev.f.m[0]; mo[1] = zero; mo[2] = zero; mo[3] = prev.f.m[1];
            } else {
                for (int q = 0; q < 4; ++q) mo[q] = prev.f.m[q];
            }
            auto mul = [](cx<double> x, cx<double> y) { return cx<double>{x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re}; };
            auto add = [](cx<double> x, cx<double> y) { return cx<double>{x.re + y.re, x.im + y.im}; };
            const cx<double> r[4] = {add(mul(mn[0], mo[0]), mul(mn[1], mo[2])), add(mul(mn[0], mo[1]), mul(mn[1], mo[3])),
                                     add(mul(mn[2], mo[0]), mul(mn[3], mo[2])), add(mul(mn[2], mo[1]), mul(mn[3], mo[3]))};
            const int bit = single_bit;
            prev = HOp{};
            if (lane_loc(bit) >= 0) {
                prev.kind = kLaneDense;
                prev.loc = lane_loc(bit);
            } else {
                prev.kind = kRegDense;
                prev.loc = reg_loc(bit);
            }
            for (int q = 0; q < 4; ++q) prev.f.m[q] = r[q];
            continue;
        }
        for (int bit = 0; bit < 64; ++bit)
            if (touched & (1ull << bit)) last_single[size_t(bit)] = -1;
        hops.push_back(h);
        if (merge && single_bit >= 0) last_single[size_t(single_bit)] = int(hops.size()) - 1;
    }
    // Rotations [[c, -s], [s, c]] (RY, its adjoint, products of them) become
    // c [[1, -t], [t, 1]] with t = s / c when |c| >= |s|, else
    // s [[t, -1], [1, t]] with t = c / s: one FMA per output component instead
    // of a multiply and an FMA.  The factors taken out multiply
    // into one scale C for the whole pass, applied by a diagonal gate at its
    // end.  Until then the registers hold the states divided by the running
    // product, so the overlap of a gate taken there is divided by its square:
    // slot_scale carries that square into the Jacobian factor.
    b.slot_scale.assign(size_t(nslots), 1.0);
    // a diagonal member needs a diagonal generator when it is trainable
    auto member = [&](const HOp &h) { return h.kind == kDiag && (h.slot < 0 || h.gkind == kGenDiag); };
    // end of the diagonal run that starts at hops[i], or i when hops[i] is
    // not the start of one (a run needs two members applied by the run)
    auto run_end = [&](size_t i) {
        size_t j = i;
        while (diag_runs && j < hops.size() && member(hops[j])) ++j;
        int deferred = 0;
        for (size_t t = i; t < j; ++t) {
            const HOp &h = hops[t];
            const bool reg = (h.cm_reg | h.pm_reg) != 0, rest = (h.f.cm_rest | h.f.pm_rest) != 0;
            deferred += !reg || !rest || h.dneg;
        }
        return deferred < 2 ? i : j;
    };
    size_t scale_hop = hops.size() + 1;           // index of a closing scale gate, if any
    size_t fold_i = hops.size() + 1, fold_j = 0;  // the run that applies the scale, if any
    double scale = 1.0;
    // smallest running product: 2^-20 for complex64 registers (values up to
    // 2^20 times the amplitudes, far inside float range), 2^-100 for
    // complex128 (the squared ratios of the Jacobian factors stay normal)
    const double scale_floor = single ? 0x1p-20 : 0x1p-100;
    if (diag_runs) {
        std::vector<double> before(hops.size(), 1.0); // running scale before each gate
        for (size_t x = 0; x < hops.size(); ++x) {
            HOp &h = hops[x];
            before[x] = scale;
            if (h.kind == kDiag || h.cm_reg != 0 || h.f.cm_rest != 0) continue;
            const cx<double> *m = h.f.m;
            if (m[0].im != 0.0 || m[1].im != 0.0 || m[2].im != 0.0 || m[3].im != 0.0) continue;
            const double c = m[0].re, s = m[2].re;
            if (m[3].re != c || m[1].re != -s || (c == 0.0 && s == 0.0)) continue;
            // The factor taken out is the larger of |c|, |s| (at least 2^-1/2),
            // so the running product only shrinks.  The registers hold the
            // states divided by it: past the floor they could overflow (and
            // the closing factor underflow) in the working precision, so
            // the rest of the pass keeps its rotations whole.
            if (std::abs(scale * (std::abs(c) >= std::abs(s) ? c : s)) < scale_floor) continue;
            if (std::abs(c) >= std::abs(s)) { // c [[1, -s/c], [s/c, 1]]
                const double t = s / c;
                h.f.m[0] = cx<double>{1.0, 0.0};
                h.f.m[1] = cx<double>{-t, 0.0};
                h.f.m[2] = cx<double>{t, 0.0};
                h.f.m[3] = cx<double>{1.0, 0.0};
                h.rot = 1;
                scale *= c;
            } else { // s [[c/s, -1], [1, c/s]]
                const double t = c / s;
                h.f.m[0] = cx<double>{t, 0.0};
                h.f.m[1] = cx<double>{-1.0, 0.0};
                h.f.m[2] = cx<double>{1.0, 0.0};
                h.f.m[3] = cx<double>{t, 0.0};
                h.rot = 2;
                scale *= s;
            }
        }
        if (scale != 1.0) {
            // The scale is a global factor, so any point of the pass may
            // apply it: the last diagonal run takes it into its per-register
            // constants at no extra cost.  Gates after that run see states
            // multiplied by the whole scale and divided by their running one.
            for (size_t i = 0; i < hops.size();) {
                const size_t j = run_end(i);
                if (j == i) {
                    ++i;
                    continue;
                }
                fold_i = i;
                fold_j = j;
                i = j;
            }
            if (fold_j == 0) {
                HOp fin; // no run: a diagonal gate closes the pass
                fin.kind = kDiag;
                fin.loc = kDiagNoReg;
                fin.f.m[0] = fin.f.m[1] = cx<double>{scale, 0.0};
                scale_hop = hops.size();
                hops.push_back(fin);
                before.push_back(scale);
            }
        }
        for (size_t x = 0; x < before.size() && x < hops.size(); ++x) {
            if (hops[x].slot < 0) continue;
            const double r = x < fold_j || fold_j == 0 ? before[x] : before[x] / scale;
            b.slot_scale[size_t(hops[x].slot)] = r * r;
        }
    }
    for (size_t i = 0; i < hops.size();) {
        const size_t j = run_end(i);
        if (j == i) {
            b.ops.push_back(pack(hops[i], b.pd.reg_off, nr));
            ++i;
            continue;
        }
        // diagonal run over hops[i, j): header, P[k] data ops, members
        if (scale_hop >= i && scale_hop < j) b.scale_members = 1;
        std::vector<std::complex<double>> P(size_t(nr), i == fold_i ? scale : 1.0);
        uint32_t flags = i == fold_i ? kRunHasP : 0u;
        std::vector<FOp> members;
        for (size_t t = i; t < j; ++t) {
            const HOp &h = hops[t];
            const bool reg = (h.cm_reg | h.pm_reg) != 0, rest = (h.f.cm_rest | h.f.pm_rest) != 0;
            uint32_t cls, ov = kOvF;
            if (!reg) {
                cls = kMemQ;
                flags |= kRunHasQ;
            } else if (!rest) {
                cls = kMemR;
                flags |= kRunHasP;
                const std::complex<double> de(h.f.m[0].re, h.f.m[0].im), dodd(h.f.m[1].re, h.f.m[1].im);
                for (int k = 0; k < nr; ++k) {
                    if ((b.pd.reg_off[k] & h.cm_reg) != h.cm_reg) continue;
                    P[size_t(k)] *= (__builtin_popcountll(b.pd.reg_off[k] & h.pm_reg) & 1) ? dodd : de;
                }
            } else {
                cls = h.dneg ? kMemS : kMemX;
                if (h.dneg) flags |= kRunHasS;
            }
            if (h.slot >= 0 && h.cm_reg == 0) {
                if (h.pm_reg == 0) ov = kOvT;
                else if (__builtin_popcountll(h.pm_reg) == 1 && h.loc < rb) ov = kOvB;
                if (ov != kOvF) flags |= kRunNeedsTE;
            }
            FOp f = pack(h, b.pd.reg_off, nr);
            f.meta |= cls << 24 | ov << 26;
            members.push_back(f);
        }
        const int ndata = (flags & kRunHasP) ? (nr + 3) / 4 : 0;
        FOp hdr{};
        hdr.meta = kDiagRun;
        hdr.pad = uint64_t(members.size()) | uint64_t(ndata) << 16 | uint64_t(flags) << 24;
        b.ops.push_back(hdr);
        for (int d = 0; d < ndata; ++d) {
            FOp data{};
            for (int q = 0; q < 4; ++q) {
                const int k = 4 * d + q;
                if (k < nr) data.m[q] = cx<double>{P[size_t(k)].real(), P[size_t(k)].imag()};
            }
            b.ops.push_back(data);
        }
        b.ops.insert(b.ops.end(), members.begin(), members.end());
        i = j;
    }
    b.pd.nops = static_cast<int32_t>(b.ops.size());
    b.pd.nslots = nslots;
    return b;
}

namespace {

struct Geometry {
    uint64_t ntiles, chunk;
    unsigned blocks;
};
// cap: at most that many wavefronts (0: no cap beyond waves_per_cu)
Geometry geometry(int n, int device, int rb, int waves_per_cu, int64_t cap = 0) {
    Geometry g{};
    g.ntiles = 1ull << (n - kLaneBits - rb);
    uint64_t max_waves = uint64_t(std::max(1, sm_count(device))) * uint64_t(waves_per_cu);
    if (cap > 0) max_waves = std::min<uint64_t>(max_waves, uint64_t(cap));
    uint64_t waves = std::min<uint64_t>(g.ntiles, max_waves);
    g.chunk = (g.ntiles + waves - 1) / waves;
    waves = (g.ntiles + g.chunk - 1) / g.chunk;
    g.blocks = unsigned((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    if (g.blocks >= 8) g.blocks = (g.blocks + 7) / 8 * 8; // whole XCD rounds for the remap
    return g;
}

template <class R>
void run_forward(lgh_sv *sv, const std::vector<KOp> &kops, const std::vector<Pass> &passes, FusedCfg cfg) {
    StreamCtx *ctx = sv->ctx;
    const int rb = cfg.rb;
    const Geometry geo = geometry(sv->n, sv->device, rb, cfg.waves_per_cu, g_waves_fwd.load());
    for (const Pass &p : passes) {
        if (p.standalone) {
            launch_kop(sv, kops[size_t(p.ops[0])]);
            continue;
        }
        Built b = build_pass(p, kops, sv->n, false, nullptr, nullptr, rb, true, true, sv->dtype == LGH_C64);
        const size_t pb = (sizeof(PassDesc) + 255) & ~size_t(255);
        std::vector<unsigned char> blob(pb + b.ops.size() * sizeof(FOp));
        std::memcpy(blob.data(), &b.pd, sizeof(PassDesc));
        std::memcpy(blob.data() + pb, b.ops.data(), b.ops.size() * sizeof(FOp));
        unsigned char *d = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
        ProfScope prof(kProfFusedPass, ctx->stream);
        const int lv = g_lean.load() ? lean_level(b.ops, false) : 0;
        auto kern = rb == 4 ? (lv == 1 ? k_fused_pass<R, 4, 1> : lv == 2 ? k_fused_pass<R, 4, 2> : k_fused_pass<R, 4>)
                            : (lv == 1 ? k_fused_pass<R, 3, 1> : lv == 2 ? k_fused_pass<R, 3, 2> : k_fused_pass<R, 3>);
        hipLaunchKernelGGL(kern, dim3(geo.blocks), dim3(kBlock), 0, ctx->stream, static_cast<cx<R> *>(sv->data),
                           geo.ntiles, geo.chunk, reinterpret_cast<const PassDesc *>(d),
                           reinterpret_cast<const FOp *>(d + pb));
        LGH_LAUNCH_CHECK();
    }
}

} // namespace

namespace lgh {

void fused_apply(lgh_sv *sv, const std::vector<KOp> &kops) {
    const FusedCfg cfg = fused_cfg(sv->device, sv->n, sv->dtype, 0);
    const auto passes = schedule_passes_cached(kops, sv->n, {}, cfg.rb, fixed_lanes(sv->dtype), kMaxSlots);
    if (sv->dtype == LGH_C128)
        run_forward<double>(sv, kops, *passes, cfg);
    else
        run_forward<float>(sv, kops, *passes, cfg);
}

FusionPlan plan_passes(const lgh_op *ops, int64_t num_ops, int n, bool adjoint, lgh_dtype dtype) {
    std::vector<KOp> kops;
    std::vector<KGen> gens;
    std::vector<int> slot;
    int col = 0;
    for (int64_t i = 0; i < num_ops; ++i) {
        const lgh_op &o = ops[i];
        if (o.gate < 0) continue; // state preparation
        kops.push_back(compile_op(o, n, adjoint));
        if (adjoint && o.num_params > 0) {
            gens.push_back(compile_generator(o, n, kops.back()));
            slot.push_back(col++);
        } else {
            gens.emplace_back();
            slot.push_back(-1);
        }
    }
    const FusedCfg cfg = fused_cfg(0, n, dtype, adjoint ? 1 : 0);
    const int rb = cfg.rb;
    const auto passes = schedule_passes_cached(kops, n, adjoint ? slot : std::vector<int>{}, rb, fixed_lanes(dtype),
                                               adjoint ? std::min(cfg.slots, adj_slot_cap(rb)) : kMaxSlots);
    FusionPlan plan{};
    for (const Pass &p : *passes) {
        ++plan.passes;
        if (p.standalone) {
            ++plan.standalone;
            continue;
        }
        const bool single = dtype == LGH_C64;
        const Built b = adjoint ? build_pass(p, kops, n, true, &gens, &slot, rb, true, true, single)
                                : build_pass(p, kops, n, false, nullptr, nullptr, rb, true, true, single);
        plan.lean += lean_level(b.ops, adjoint) > 0 ? 1 : 0;
        for (size_t j = 0; j < b.ops.size();) {
            const FOp &f = b.ops[j];
            if ((f.meta & 15u) != kDiagRun) {
                ++j;
                continue;
            }
            const int nmem = int(f.pad & 0xffffu), ndata = int((f.pad >> 16) & 0xffu);
            ++plan.runs;
            plan.run_members += nmem;
            j += size_t(1 + ndata + nmem);
        }
        plan.run_members -= b.scale_members; // the closing scale gate is no gate of the list
    }
    return plan;
}

template <class R>
static void adjoint_one(lgh_sv *lam, const void *lam_src, lgh_sv *bra, const std::vector<KOp> &udag,
                        const std::vector<KGen> &gens, const std::vector<int> &slot_col,
                        const std::vector<double> &factor, double *jac_row_dev, double *partials,
                        StandaloneStep standalone, const BraFromZ *bra_z) {
    // `udag` lists the ops in forward order with U^dagger matrices; the passes
    // are scheduled on that order and executed backwards.
    StreamCtx *ctx = lam->ctx;
    const int n = lam->n;
    const FusedCfg cfg = fused_cfg(lam->device, n, lam->dtype, 1);
    const int rb = cfg.rb;
    // the kernel's LDS holds adj_slot_cap(rb) slots
    const auto passes = schedule_passes_cached(udag, n, slot_col, rb, fixed_lanes(lam->dtype),
                                               std::min(cfg.slots, adj_slot_cap(rb)));
    const Geometry geo = geometry(n, lam->device, rb, cfg.waves_per_cu, g_waves_adj.load());
    // lambda still to be read from lam_src (until the first pass has run)
    const void *src = lam_src;
    // the observable state still to be formed from lambda (first pass only)
    bool bra_pending = bra_z != nullptr;
    for (auto it = passes->rbegin(); it != passes->rend(); ++it) {
        const Pass &p = *it;
        if (p.standalone) {
            if (bra_pending) {
                bra_z->materialize(); // before lambda's first change
                bra_pending = false;
            }
            if (src) {
                LGH_HIP(hipMemcpyAsync(lam->data, src, lam->bytes, hipMemcpyDeviceToDevice, ctx->stream));
                src = nullptr;
            }
            const int i = p.ops[0];
            standalone(udag[size_t(i)], slot_col[size_t(i)] >= 0 ? &gens[size_t(i)] : nullptr, slot_col[size_t(i)],
                       factor[size_t(i)]);
            continue;
        }
        Built b = build_pass(p, udag, n, true, &gens, &slot_col, rb, true, true, lam->dtype == LGH_C64);
        LGH_REQUIRE(b.pd.nslots <= adj_slot_cap(rb), "fused adjoint pass: more trainable gates than LDS slots");
        std::vector<int64_t> cols;
        std::vector<double> facs;
        for (size_t s = 0; s < b.slot_ops.size(); ++s) {
            const int oi = b.slot_ops[s];
            cols.push_back(slot_col[size_t(oi)]);
            facs.push_back(factor[size_t(oi)] * b.slot_scale[s]);
        }
        const size_t pb = (sizeof(PassDesc) + 255) & ~size_t(255);
        const size_t ob = (b.ops.size() * sizeof(FOp) + 255) & ~size_t(255);
        const size_t gb = (b.gens.size() * sizeof(FGen) + 255) & ~size_t(255);
        const size_t cb = (cols.size() * 8 + 255) & ~size_t(255);
        std::vector<unsigned char> blob(pb + ob + gb + 2 * cb + 256);
        std::memcpy(blob.data(), &b.pd, sizeof(PassDesc));
        std::memcpy(blob.data() + pb, b.ops.data(), b.ops.size() * sizeof(FOp));
        std::memcpy(blob.data() + pb + ob, b.gens.data(), b.gens.size() * sizeof(FGen));
        std::memcpy(blob.data() + pb + ob + gb, cols.data(), cols.size() * 8);
        std::memcpy(blob.data() + pb + ob + gb + cb, facs.data(), facs.size() * 8);
        unsigned char *d = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
        {
            ProfScope prof(kProfFusedAdj, ctx->stream);
            const int lv = g_lean.load() ? lean_level(b.ops, true) : 0;
            auto kern = rb == 4 ? (lv == 1 ? k_fused_adj4<R, 1> : lv == 2 ? k_fused_adj4<R, 2> : k_fused_adj4<R>)
                                : (lv == 1 ? k_fused_adj3<R, 1> : lv == 2 ? k_fused_adj3<R, 2> : k_fused_adj3<R>);
            const cx<R> *lam_in = static_cast<const cx<R> *>(src ? src : lam->data);
            src = nullptr;
            const uint64_t zmask = bra_pending ? bra_z->zmask : 0ull;
            hipLaunchKernelGGL(kern, dim3(geo.blocks), dim3(kBlock), 0, ctx->stream, lam_in,
                               static_cast<cx<R> *>(lam->data), static_cast<cx<R> *>(bra->data), geo.ntiles,
                               geo.chunk, reinterpret_cast<const PassDesc *>(d), reinterpret_cast<const FOp *>(d + pb),
                               reinterpret_cast<const FGen *>(d + pb + ob), partials, zmask, int(bra_pending));
            bra_pending = false;
            LGH_LAUNCH_CHECK();
        }
        if (!cols.empty()) {
            hipLaunchKernelGGL(k_slot_store, dim3(unsigned(cols.size())), dim3(kBlock), 0, ctx->stream, partials,
                               int(geo.blocks), int(cols.size()), reinterpret_cast<const int64_t *>(d + pb + ob + gb),
                               reinterpret_cast<const double *>(d + pb + ob + gb + cb), jac_row_dev);
            LGH_LAUNCH_CHECK();
        }
    }
}

void fused_adjoint_one(lgh_sv *lam, const void *lam_src, lgh_sv *bra, const std::vector<KOp> &udag,
                       const std::vector<KGen> &gens, const std::vector<int> &slot_col,
                       const std::vector<double> &factor, double *jac_row_dev, double *partials,
                       StandaloneStep standalone, const BraFromZ *bra_z) {
    if (lam->dtype == LGH_C128)
        adjoint_one<double>(lam, lam_src, bra, udag, gens, slot_col, factor, jac_row_dev, partials, standalone,
                            bra_z);
    else
        adjoint_one<float>(lam, lam_src, bra, udag, gens, slot_col, factor, jac_row_dev, partials, standalone,
                           bra_z);
}

size_t fused_partials_bytes(int n, int device) {
    // sized for the smaller tile, which has more workgroups
    if (n < kLaneBits + 3) return 0;
    // the most slots of any tile height times the most blocks (three register bits)
    return sizeof(double) * size_t(kMaxSlotsSv4) * geometry(n, device, 3, kMaxWavesPerCUAdjoint).blocks;
}

} // namespace lgh

extern "C" lgh_status lgh_set_fusion(int enabled) {
    return guard([&] { set_fusion(enabled); });
}

extern "C" lgh_status lgh_set_tile_bits(int forward, int adjoint) {
    return guard([&] { set_tile_bits(forward, adjoint); });
}

extern "C" lgh_status lgh_set_pass_waves(int64_t forward, int64_t adjoint) {
    return guard([&] { set_pass_waves(forward, adjoint); });
}

extern "C" lgh_status lgh_fusion_lean_passes(int32_t num_qubits, const lgh_op *ops, int64_t num_ops, int32_t adjoint,
                                             lgh_dtype dtype, int32_t *lean_passes) {
    int32_t p = 0, s = 0, r = 0, m = 0;
    const lgh_status st = lgh_fusion_plan(num_qubits, ops, num_ops, adjoint, dtype, &p, &s, &r, &m);
    if (st != LGH_OK) return st;
    return guard([&] {
        LGH_REQUIRE(lean_passes, "null output");
        *lean_passes = plan_passes(ops, num_ops, num_qubits, adjoint != 0, dtype).lean;
    });
}

extern "C" lgh_status lgh_set_lean_passes(int enabled) {
    return guard([&] { g_lean.store(enabled != 0); });
}

extern "C" lgh_status lgh_fusion_plan(int32_t num_qubits, const lgh_op *ops, int64_t num_ops, int32_t adjoint,
                                      lgh_dtype dtype, int32_t *passes, int32_t *standalone, int32_t *runs,
                                      int32_t *run_members) {
    return guard([&] {
        LGH_REQUIRE(num_qubits >= 1 && num_qubits <= 63, "bad number of qubits");
        LGH_REQUIRE(num_ops >= 0 && (num_ops == 0 || ops), "bad op list");
        LGH_REQUIRE(dtype == LGH_C64 || dtype == LGH_C128, "bad dtype");
        LGH_REQUIRE(passes && standalone && runs && run_members, "null output");
        for (int64_t i = 0; i < num_ops; ++i) {
            if (ops[i].gate < 0) continue;
            LGH_REQUIRE(!adjoint || ops[i].num_params <= 1,
                        "The operation is not supported using the adjoint differentiation method");
            validate_op(ops[i], num_qubits);
        }
        const FusionPlan plan = plan_passes(ops, num_ops, num_qubits, adjoint != 0, dtype);
        *passes = plan.passes;
        *standalone = plan.standalone;
        *runs = plan.runs;
        *run_members = plan.run_members;
    });
}

// ---------------------------------------------------------------- autotuner
//
// Times the fused passes of one circuit under every candidate tile height
// (register bits 3 and 4), grid (wavefronts per CU) and, for the adjoint,
// number of trainable gates per pass, on states of the requested size,
// forward and adjoint separately, and keeps the fastest per (device, qubits,
// dtype, direction).  Results do not depend on the grid;
// other tile heights schedule the gates into other passes, which changes the
// rounding of the results but not their values beyond it.

namespace {

struct ForceCfg {
    int dir;
    ForceCfg(int d, const FusedCfg *c) : dir(d) { t_force[d] = c; }
    ~ForceCfg() { t_force[dir] = nullptr; }
};

void ok(lgh_status s) {
    if (s != LGH_OK) throw Error(s, lgh_last_error());
}

struct Owned {
    lgh_sv *sv = nullptr;
    ~Owned() {
        if (sv) lgh_sv_destroy(sv);
    }
};

// milliseconds of fn() on `stream`, best of `reps` after one warm-up
template <class F> double time_on(hipStream_t stream, int reps, F &&fn) {
    hipEvent_t a = nullptr, b = nullptr;
    LGH_HIP(hipEventCreate(&a));
    struct Ev {
        hipEvent_t &a, &b;
        ~Ev() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev{a, b};
    LGH_HIP(hipEventCreate(&b));
    fn();
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        LGH_HIP(hipEventRecord(a, stream));
        fn();
        LGH_HIP(hipEventRecord(b, stream));
        LGH_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        LGH_HIP(hipEventElapsedTime(&ms, a, b));
        best = std::min(best, ms);
    }
    return double(best);
}

} // namespace

extern "C" lgh_status lgh_autotune_fused(int device, int32_t num_qubits, lgh_dtype dtype, const lgh_op *ops,
                                         int64_t num_ops, lgh_tune_result *out) {
    return guard([&] {
        LGH_REQUIRE(out, "null pointer");
        LGH_REQUIRE(dtype == LGH_C64 || dtype == LGH_C128, "unknown dtype");
        LGH_REQUIRE(num_qubits >= kLaneBits + 4 && num_qubits <= 34, "autotuning needs 10 to 34 qubits");
        LGH_REQUIRE(num_ops >= 0 && (num_ops == 0 || ops), "bad op list");
        *out = lgh_tune_result{};
        const int n = num_qubits;
        // the circuit: as given, or 2 layers of RY, RZ on every wire and a CZ brick
        std::vector<lgh_op> circ;
        std::vector<std::array<int32_t, 2>> wire_store;
        if (num_ops > 0) {
            for (int64_t i = 0; i < num_ops; ++i)
                if (ops[i].gate >= 0) circ.push_back(ops[i]);
        } else {
            wire_store.reserve(size_t(4 * n));
            for (int layer = 0; layer < 2; ++layer) {
                for (int w = 0; w < n; ++w)
                    for (int g : {int(LGH_GATE_RY), int(LGH_GATE_RZ)}) {
                        wire_store.push_back({w, 0});
                        lgh_op o{};
                        o.gate = g;
                        o.num_wires = 1;
                        o.wires = wire_store.back().data();
                        o.params[0] = 0.37 + 0.61 * double(circ.size());
                        o.num_params = 1;
                        circ.push_back(o);
                    }
                for (int s0 = 0; s0 < 2; ++s0)
                    for (int w = s0; w + 1 < n; w += 2) {
                        wire_store.push_back({w, w + 1});
                        lgh_op o{};
                        o.gate = LGH_GATE_CZ;
                        o.num_wires = 2;
                        o.wires = wire_store.back().data();
                        circ.push_back(o);
                    }
            }
        }
        for (const lgh_op &o : circ) validate_op(o, n);
        int64_t npar = 0;
        for (const lgh_op &o : circ) npar += o.num_params > 0;
        LGH_REQUIRE(npar > 0, "the circuit has no parametrised gate to differentiate");
        std::vector<int64_t> tp(static_cast<size_t>(npar));
        for (int64_t k = 0; k < npar; ++k) tp[size_t(k)] = k;
        std::vector<double> jac(static_cast<size_t>(npar));
        DeviceGuard g(device);
        Owned psi, bra;
        ok(lgh_sv_create(&psi.sv, n, dtype, device, nullptr));
        ok(lgh_sv_create(&bra.sv, n, dtype, device, nullptr));
        hipStream_t st = psi.sv->ctx->stream;
        const int32_t zw = 0;
        const int64_t ncirc = int64_t(circ.size());

        // forward: tile height x grid; adjoint: tile height x grid x trainable
        // gates per pass (more per pass: fewer passes over the two states,
        // more work per pass)
        const FusedCfg fwd_c[] = {{3, 512, kMaxSlots}, {3, 4096, kMaxSlots}, {4, 512, kMaxSlots}, {4, 4096, kMaxSlots}};
        const FusedCfg adj_c[] = {{3, 128, kMaxSlotsSv}, {3, 256, kMaxSlotsSv}, {3, 512, kMaxSlotsSv},
                                  {3, 256, kMaxSlotsSv / 2}, {4, 256, kMaxSlotsSv}, {4, 128, kMaxSlotsSv4},
                                  {4, 256, kMaxSlotsSv4}, {4, 512, kMaxSlotsSv4}};
        const FusedCfg def_f = fused_cfg(device, n, dtype, 0), def_a = fused_cfg(device, n, dtype, 1);
        FusedCfg best_f = def_f, best_a = def_a;
        double tbest_f = 1e300, tbest_a = 1e300;
        ok(lgh_sv_reset(psi.sv));
        for (const FusedCfg &c : fwd_c) {
            ForceCfg fc(0, &c);
            const double ms = time_on(st, 2, [&] { ok(lgh_apply(psi.sv, circ.data(), ncirc)); });
            const int q = out->candidates++;
            out->cand_dir[q] = 0;
            out->cand_tile_bits[q] = c.rb;
            out->cand_waves_per_cu[q] = c.waves_per_cu;
            out->cand_slots[q] = 0;
            out->cand_ms[q] = ms;
            if (c.rb == def_f.rb && c.waves_per_cu == def_f.waves_per_cu) out->default_forward_ms = ms;
            if (ms < tbest_f) {
                tbest_f = ms;
                best_f = c;
            }
        }
        // adjoint of <Z_0> on the circuit's state; the observable state is
        // rebuilt before every timed sweep (the sweep overwrites it)
        ok(lgh_sv_reset(psi.sv));
        ok(lgh_apply(psi.sv, circ.data(), ncirc));
        for (const FusedCfg &c : adj_c) {
            ForceCfg fc(1, &c);
            double ms = 1e300;
            for (int r = 0; r < 3; ++r) {
                ok(lgh_sv_pauli_word_from(bra.sv, psi.sv, "Z", &zw, 1));
                ok(lgh_sv_synchronize(bra.sv));
                const auto t0 = std::chrono::steady_clock::now();
                lgh_sv *bl[1] = {bra.sv};
                ok(lgh_adjoint_jacobian(psi.sv, bl, 1, circ.data(), ncirc, tp.data(), npar, 0, jac.data()));
                const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (r > 0) ms = std::min(ms, t); // the first sweep warms up
            }
            const int q = out->candidates++;
            out->cand_dir[q] = 1;
            out->cand_tile_bits[q] = c.rb;
            out->cand_waves_per_cu[q] = c.waves_per_cu;
            out->cand_slots[q] = c.slots;
            out->cand_ms[q] = ms;
            if (c.rb == def_a.rb && c.waves_per_cu == def_a.waves_per_cu && c.slots == def_a.slots)
                out->default_adjoint_ms = ms;
            if (ms < tbest_a) {
                tbest_a = ms;
                best_a = c;
            }
        }
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[std::make_tuple(device, n, int(dtype), 0)] = best_f;
            g_tuned[std::make_tuple(device, n, int(dtype), 1)] = best_a;
        }
        out->forward_tile_bits = best_f.rb;
        out->forward_waves_per_cu = best_f.waves_per_cu;
        out->adjoint_tile_bits = best_a.rb;
        out->adjoint_waves_per_cu = best_a.waves_per_cu;
        out->adjoint_slots = best_a.slots;
        out->forward_ms = tbest_f;
        out->adjoint_ms = tbest_a;
    });
}

extern "C" lgh_status lgh_autotune_clear(void) {
    return guard([&] { clear_tuned(); });
}
