// Gate fusion (lgh_fused.hip): pass scheduling and fused forward / adjoint
// passes.  Internal to liblgh.so.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "lgh_internal.hpp"

namespace lgh {

// One fused pass: gates (indices into the compiled op list, in execution
// order) whose dense targets are lane bits or register bits.  Lane position
// p (bit p of the lane id) carries index bit lane_bits[p]; the lowest lane
// positions always carry the lowest index bits, so that every load covers
// whole runs of consecutive amplitudes.  A standalone pass holds one gate
// that the fused kernels cannot run.
struct Pass {
    bool standalone = false;
    int lane_bits[6] = {0, 1, 2, 3, 4, 5};
    int reg_bits[4] = {0, 0, 0, 0}; // the first rb entries are used
    std::vector<int> ops;
};

void set_fusion(int enabled);
// register bits of forward and adjoint passes (3 or 4 each)
void set_tile_bits(int forward, int adjoint);
int fusion_enabled();
// caps on the wavefronts of the forward and adjoint passes' grids (0: default)
void set_pass_waves(int64_t forward, int64_t adjoint);
// rb: register bits of the passes (tile = 2^(6 + rb) amplitudes per wavefront)
bool fusable(const KOp &k, int n, int rb);
// trainable_slot[i] >= 0 marks gates whose generator overlap is accumulated
// (at most max_slots per pass: kMaxSlots for the batched kernels,
// kMaxSlotsSv for the single-state adjoint kernel); empty for the forward
// direction.
// fixed_lanes: the lane positions 0..fixed_lanes-1 carry index bits
// 0..fixed_lanes-1 in every pass; the scheduler chooses the index bits of the
// other lane positions per pass, like the register bits (6: all lane bits
// fixed).  max_slots does not matter without trainable gates.
std::vector<Pass> schedule_passes(const std::vector<KOp> &ops, int n, const std::vector<int> &trainable_slot,
                                  int rb, int fixed_lanes, int max_slots);
// schedule_passes through a cache keyed by the gates' structure (the
// schedule does not depend on parameter values).
std::shared_ptr<const std::vector<Pass>> schedule_passes_cached(const std::vector<KOp> &ops, int n,
                                                                const std::vector<int> &trainable_slot, int rb,
                                                                int fixed_lanes, int max_slots);
// lgh_set_lean_passes: passes of lean gate forms run in the lean kernels
bool lean_enabled();
// Host-side plan of the fused passes a gate sequence gets (no device use):
// forward passes, or adjoint passes with every parametric gate trainable.
struct FusionPlan {
    int32_t passes = 0, standalone = 0, runs = 0, run_members = 0, lean = 0;
};
FusionPlan plan_passes(const lgh_op *ops, int64_t num_ops, int n, bool adjoint, lgh_dtype dtype);

// Forward application of a compiled op list through fused passes.
void fused_apply(lgh_sv *sv, const std::vector<KOp> &kops);

// Applies U^dagger of one non-fusable gate to lambda and the observable
// state; with a generator, stores factor * Im<bra|G lambda> into jac_row[col].
using StandaloneStep = std::function<void(const KOp &udag, const KGen *gen, int64_t col, double factor)>;

// An observable state that the sweep forms itself: bra = Z-word(zmask)
// lambda, inside the first reverse pass when that pass is fused; otherwise
// `materialize` forms it before lambda first changes.
struct BraFromZ {
    uint64_t zmask;
    std::function<void()> materialize;
};

// Reverse sweep of one observable state through fused passes.  udag: the
// ops in forward order compiled as U^dagger; slot_col[i]: Jacobian column of
// a trainable gate or -1; factor[i] = -2 s (inverse ? -1 : 1).  lam_src: if
// not null, lambda's starting values, read by the first pass (or copied
// before a first standalone step) instead of lam's own contents.
void fused_adjoint_one(lgh_sv *lam, const void *lam_src, lgh_sv *bra, const std::vector<KOp> &udag,
                       const std::vector<KGen> &gens, const std::vector<int> &slot_col,
                       const std::vector<double> &factor, double *jac_row_dev, double *partials,
                       StandaloneStep standalone, const BraFromZ *bra_z = nullptr);
size_t fused_partials_bytes(int n, int device);

} // namespace lgh
