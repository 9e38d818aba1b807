/*
 * lgh.h -- C ABI of the MI355X (gfx950) statevector library behind the
 * lightning.gpu device.
 *
 * The reference crosses from its pybind11 module into closed NVIDIA
 * libraries (cuStateVec, cuBLAS, cuSPARSE) inside StateVectorCudaManaged;
 * this header is the equivalent seam of the MI355X build: the C++ host class
 * StateVectorHIPManaged and the pybind11 module `lightning_gpu_qubit_ops`
 * call only these functions, and every function is implemented with
 * hand-written HIP kernels.  Each declaration cites the reference interface
 * it replaces (paths relative to the reference repository root; "SVCM" is
 * pennylane_lightning_gpu/src/simulator/StateVectorCudaManaged.hpp).
 *
 * Conventions
 *  - Wire w of an n-qubit state is amplitude-index bit (n-1-w) (SVCM:1349-1356).
 *  - Matrices are row-major, interleaved (re, im) doubles; the first listed
 *    wire is the most significant bit of the matrix index (PennyLane order).
 *  - Complex host buffers of states are interleaved (re, im) in the state's
 *    precision: float for LGH_C64, double for LGH_C128.
 *  - Every function returns LGH_OK or an error code; lgh_last_error() returns
 *    the message of the calling thread's last failure.
 *  - Work on one state is ordered on that state's HIP stream.  Functions that
 *    return host values synchronise that stream.  A handle must not be used
 *    from two threads at once.
 */
#ifndef LGH_H
#define LGH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGH_ABI_VERSION 1

typedef enum lgh_status {
    LGH_OK = 0,
    LGH_ERR_INVALID_ARGUMENT = 1,
    LGH_ERR_UNSUPPORTED = 2,
    LGH_ERR_OUT_OF_MEMORY = 3,
    LGH_ERR_DEVICE = 4,
    LGH_ERR_NO_DEVICE = 5
} lgh_status;

typedef enum lgh_dtype { LGH_C64 = 0, LGH_C128 = 1 } lgh_dtype;

/* Gate identifiers: the named methods of LightningGPU_C{64,128}
 * (pennylane_lightning_gpu/src/bindings/Bindings.cpp:120-427) plus an
 * explicit-matrix gate (Bindings.cpp:258-262, SVCM:226-245). */
typedef enum lgh_gate {
    LGH_GATE_STATEPREP = -1, /* state preparation record: skipped by the adjoint sweep */
    LGH_GATE_IDENTITY = 0,
    LGH_GATE_PAULIX,
    LGH_GATE_PAULIY,
    LGH_GATE_PAULIZ,
    LGH_GATE_HADAMARD,
    LGH_GATE_S,
    LGH_GATE_T,
    LGH_GATE_CNOT,
    LGH_GATE_SWAP,
    LGH_GATE_CSWAP,
    LGH_GATE_TOFFOLI,
    LGH_GATE_CY,
    LGH_GATE_CZ,
    LGH_GATE_PHASESHIFT,
    LGH_GATE_CONTROLLEDPHASESHIFT,
    LGH_GATE_RX,
    LGH_GATE_RY,
    LGH_GATE_RZ,
    LGH_GATE_ROT,
    LGH_GATE_CRX,
    LGH_GATE_CRY,
    LGH_GATE_CRZ,
    LGH_GATE_CROT,
    LGH_GATE_ISINGXX,
    LGH_GATE_ISINGYY,
    LGH_GATE_ISINGZZ,
    LGH_GATE_SINGLEEXCITATION,
    LGH_GATE_SINGLEEXCITATIONMINUS,
    LGH_GATE_SINGLEEXCITATIONPLUS,
    LGH_GATE_DOUBLEEXCITATION,
    LGH_GATE_DOUBLEEXCITATIONMINUS,
    LGH_GATE_DOUBLEEXCITATIONPLUS,
    LGH_GATE_MULTIRZ,
    LGH_GATE_MATRIX, /* op.matrix on the last (num_wires - num_ctrl) wires */
    LGH_GATE_COUNT
} lgh_gate;

/* One operation.  `wires` lists controls first, as the reference's
 * ctrl_map does (pennylane_lightning_gpu/src/simulator/StateVectorCudaBase.hpp:288-312).
 * Named gates read params[0] (Rot/CRot read all three). */
typedef struct lgh_op {
    int32_t gate;          /* lgh_gate */
    int32_t num_wires;
    int32_t num_ctrl;      /* LGH_GATE_MATRIX only: leading control wires */
    int32_t inverse;       /* apply the adjoint */
    const int32_t *wires;  /* host pointer, num_wires entries */
    double params[3];
    const double *matrix;  /* LGH_GATE_MATRIX: 2^k x 2^k complex, k = num_wires - num_ctrl */
    int32_t num_params;    /* parameters the op carries (adjoint bookkeeping) */
    int32_t reserved;
} lgh_op;

typedef struct lgh_sv lgh_sv;   /* device-resident state vector */
typedef struct lgh_csr lgh_csr; /* device-resident CSR matrix */

/* ---- errors / version ------------------------------------------------ */
const char *lgh_last_error(void);
int lgh_abi_version(void);

/* ---- devices (Bindings.cpp:1736-1777; cuda_helpers.hpp:627-631) ------ */
/* GetNumGPUs / DevPool.getTotalDevices */
lgh_status lgh_device_count(int *count);
/* get_gpu_arch: writes the gfx target name, e.g. "gfx950" */
lgh_status lgh_device_arch(int device, char *name, size_t len);
/* is_gpu_supported: 1 for gfx950 */
lgh_status lgh_device_supported(int device, int *supported);
lgh_status lgh_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);
lgh_status lgh_device_synchronize(int device);
/* device_reset: frees cached buffers, then resets the device */
lgh_status lgh_device_reset(void);
/* Returns cached device buffers of `device` to the runtime. */
lgh_status lgh_pool_trim(int device);
/* Gate fusion switch for lgh_apply sequences and the adjoint sweep (on by
 * default).  No counterpart in the reference, which applies one kernel per
 * gate; exposed for comparisons and tests. */
lgh_status lgh_set_fusion(int enabled);
/* Tile height of the fused passes: register bits (3 or 4) of the forward and
 * of the adjoint passes; a wavefront holds 2^(6 + bits) amplitudes per state.
 * Defaults: forward 4; adjoint 4 for complex128 states of 24 qubits and more,
 * else 3 (adjoint 0 restores that default).  No counterpart in the
 * reference; a tuning knob. */
lgh_status lgh_set_tile_bits(int forward, int adjoint);
/* Caps on the wavefronts of the fused passes' grids, forward and adjoint
 * (0: the defaults, up to 4096 and 256 wavefronts per CU).  Each wavefront
 * steps through ceil(tiles / wavefronts) consecutive tiles of 2^(6 + bits)
 * amplitudes, so a small cap exercises the multi-tile path at any size.  No
 * counterpart in the reference; a tuning knob used by tests. */
lgh_status lgh_set_pass_waves(int64_t forward, int64_t adjoint);
/* Lean fused passes (on by default; single-state passes and the batched
 * kernel): a pass whose dense gates are all
 * uncontrolled 2x2 matrices (trainable ones with an X or Y generator) or
 * controlled X gates, and whose diagonal gates all sit in diagonal runs
 * without general members, runs in kernels compiled for just those forms.  Results agree with the general kernels within rounding.  No
 * counterpart in the reference; a switch for comparisons and tests. */
lgh_status lgh_set_lean_passes(int enabled);
/* Autotuner of the fused passes (no counterpart in the reference): times
 * `ops` (or, with num_ops = 0, two layers of RY and RZ on every wire plus a
 * CZ brick) on fresh states of num_qubits (10..34) qubits under every
 * candidate tile height (register bits 3 and 4) and grid (forward: 512 or
 * 4096 wavefronts per CU; adjoint: 128, 256 or 512), the forward passes
 * (lgh_apply) and the adjoint sweep of <Z_0> (lgh_adjoint_jacobian, every
 * parametrised gate trainable) separately, and keeps the fastest of each for
 * (device, num_qubits, dtype): later lgh_apply / lgh_adjoint_jacobian calls
 * on such states use it.  Other tile heights group the gates into other
 * passes: results change within rounding.  lgh_set_tile_bits,
 * lgh_set_pass_waves and lgh_autotune_clear drop every tuned choice. */
typedef struct lgh_tune_result {
    int32_t forward_tile_bits, forward_waves_per_cu, adjoint_tile_bits, adjoint_waves_per_cu;
    double forward_ms, adjoint_ms;                 /* the chosen candidates */
    double default_forward_ms, default_adjoint_ms; /* the defaults in force before */
    int32_t candidates;                            /* timed candidates, in order */
    int32_t cand_dir[16];                          /* 0: forward, 1: adjoint */
    int32_t cand_tile_bits[16];
    int32_t cand_waves_per_cu[16];
    double cand_ms[16];
    int32_t adjoint_slots;  /* trainable gates per adjoint pass of the chosen candidate */
    int32_t cand_slots[16]; /* per candidate (0 for forward candidates) */
} lgh_tune_result;
lgh_status lgh_autotune_fused(int device, int32_t num_qubits, lgh_dtype dtype, const lgh_op *ops, int64_t num_ops,
                              lgh_tune_result *out);
lgh_status lgh_autotune_clear(void);
/* Plan of the fused passes that `ops` get on `num_qubits` qubits, computed on
 * the host without touching a device: passes in total, of those the
 * standalone ones (gates that fit no fused pass), and over the fused passes
 * the diagonal runs and the diagonal gates applied as their members.
 * adjoint 0: the forward passes of lgh_apply; 1: the reverse passes of
 * lgh_adjoint_jacobian with every parametric gate trainable.  State
 * preparation ops (gate < 0) are skipped.  No counterpart in the reference,
 * which applies one kernel per gate; for inspecting fusion and for tests. */
lgh_status lgh_fusion_plan(int32_t num_qubits, const lgh_op *ops, int64_t num_ops, int32_t adjoint, lgh_dtype dtype,
                           int32_t *passes, int32_t *standalone, int32_t *runs, int32_t *run_members);
/* Of the fused passes lgh_fusion_plan describes, how many run in the lean
 * kernels (lgh_set_lean_passes): passes of uncontrolled 2x2 gates,
 * controlled X gates and diagonal runs only.  Host only; no counterpart in
 * the reference. */
lgh_status lgh_fusion_lean_passes(int32_t num_qubits, const lgh_op *ops, int64_t num_ops, int32_t adjoint,
                                  lgh_dtype dtype, int32_t *lean_passes);
/* Launch timing of the hot kernels with HIP events recorded on their
 * launch stream (off by default; no counterpart in the reference).  Enabling
 * or disabling clears what was recorded.  kind 0: fused forward passes,
 * 1: fused adjoint passes, 2: batched-circuit launches.  lgh_profile_read
 * waits for the recorded events and returns their count and summed time. */
lgh_status lgh_profile_enable(int enabled);
lgh_status lgh_profile_read(int kind, int64_t *launches, double *total_ms);
/* Streams for callers that overlap independent states (serving). */
lgh_status lgh_stream_create(int device, void **stream);
lgh_status lgh_stream_destroy(void *stream);
lgh_status lgh_stream_synchronize(void *stream);

/* ---- state vectors (SVCM:84-185, StateVectorCudaBase.hpp:104-240) ---- */
/* StateVectorCudaManaged(num_qubits[, DevTag]); stream NULL = library stream
 * of `device`.  The state is left uninitialised. */
lgh_status lgh_sv_create(lgh_sv **out, int num_qubits, lgh_dtype dtype, int device, void *stream);
lgh_status lgh_sv_destroy(lgh_sv *sv);
lgh_status lgh_sv_info(const lgh_sv *sv, int *num_qubits, lgh_dtype *dtype, int *device,
                       void **stream, void **device_data);
/* initSV / resetGPU: |0...0> (StateVectorCudaBase.hpp:234-240) */
lgh_status lgh_sv_reset(lgh_sv *sv);
/* setBasisState: zero, then amplitude `index` = value (SVCM:144-152) */
lgh_status lgh_sv_set_basis(lgh_sv *sv, uint64_t index, double re, double im);
/* setStateVector: zero, then sv[indices[i]] = values[i] (SVCM:164-185,
 * initSV.cu:64-97).  values: complex in the state's precision. */
lgh_status lgh_sv_set_indexed(lgh_sv *sv, const int64_t *indices, const void *values, uint64_t count);
/* CopyHostDataToGpu (StateVectorCudaBase.hpp:104-130) */
lgh_status lgh_sv_copy_h2d(lgh_sv *sv, const void *host, uint64_t count);
/* CopyGpuDataToHost (StateVectorCudaBase.hpp:131-170) */
lgh_status lgh_sv_copy_d2h(const lgh_sv *sv, void *host, uint64_t count);
/* DeviceToDevice / updateData (StateVectorCudaBase.hpp:205-219) */
lgh_status lgh_sv_copy_d2d(lgh_sv *dst, const lgh_sv *src);
lgh_status lgh_sv_synchronize(const lgh_sv *sv);

/* ---- gates (SVCM:198-688) ------------------------------------------- */
/* applyOperation for one op or a sequence (SVCM:198-247, :279-315). */
lgh_status lgh_apply(lgh_sv *sv, const lgh_op *ops, int64_t num_ops);
/* Host-only utility (no device needed): the 2^w x 2^w matrix, controls
 * included and wires in op order, that lgh_apply applies for `op`; complex
 * (re, im) row-major into `out`. */
lgh_status lgh_gate_matrix(const lgh_op *op, double *out);
/* Generator of a differentiable gate, unscaled (GateGenerators.hpp:58-320,
 * SVCM:562-688).  Writes the generator's scaling factor to *scale. */
lgh_status lgh_apply_generator(lgh_sv *sv, const lgh_op *op, double *scale);

/* ---- BLAS-like helpers (cuda_helpers.hpp:456-560) -------------------- */
/* innerProdC_CUDA: <a|b> */
lgh_status lgh_inner(const lgh_sv *a, const lgh_sv *b, double *re, double *im);
/* scaleAndAddC_CUDA: y += alpha * x */
lgh_status lgh_axpy(lgh_sv *y, double alpha_re, double alpha_im, const lgh_sv *x);
/* scaleC_CUDA: x *= alpha */
lgh_status lgh_scale(lgh_sv *x, double alpha_re, double alpha_im);
/* dst = P src for the Pauli word P = paulis[k] ('I', 'X', 'Y', 'Z') on
 * wires[k]; dst and src are different states of the same size.  One pass.
 * It replaces the reference's copy of the state followed by the observable's
 * gates when the adjoint Jacobian forms H|lambda> for a Pauli-word
 * observable (AdjointDiffGPU.hpp:228-268, ObservablesGPU.hpp:88-128,
 * :209-260); no single reference function corresponds to it. */
lgh_status lgh_sv_pauli_word_from(lgh_sv *dst, const lgh_sv *src, const char *paulis, const int32_t *wires,
                                  int32_t num_wires);

/* ---- measurements (SVCM:702-1148) ------------------------------------ */
/* expval(wires, matrix) / getExpectationValueDeviceMatrix (SVCM:755-780, :1649-1713) */
lgh_status lgh_expval_matrix(const lgh_sv *sv, const int32_t *wires, int32_t num_wires,
                             const double *matrix, double *re, double *im);
/* getExpectationValuePauliWords (SVCM:1071-1148).  Word w is
 * paulis[offsets[w]..offsets[w+1]) over {'I','X','Y','Z'} on the wires at the
 * same positions; coeffs are complex (re, im).  *result = Re(sum_w c_w <P_w>);
 * per_word (nullable) receives <P_w>. */
lgh_status lgh_expval_pauli_words(const lgh_sv *sv, int64_t num_words, const char *paulis,
                                  const int32_t *wires, const int64_t *offsets,
                                  const double *coeffs, double *per_word, double *result);
/* Host-only plan of lgh_expval_pauli_words (no device needed): tile passes
 * (words grouped by the index bits their flips touch; each pass reads the
 * state once per launch of up to 32 words) and the pair-kernel launches of
 * words whose flips span more bits than a tile holds (one pass per flip
 * mask, up to 8 words per launch).  No counterpart in the reference; for
 * byte accounting and tests. */
lgh_status lgh_pauli_plan(int32_t num_qubits, int64_t num_words, const char *paulis, const int32_t *wires,
                          const int64_t *offsets, int32_t *tile_passes, int32_t *tile_launches,
                          int32_t *pair_launches);
/* lgh_expval_pauli_words without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results (bit for bit) as lgh_expval_pauli_words; at most 1024
 * words, and the reduction partials of a pass plus 16 B per word must fit a
 * fixed 512 KiB device scratch array (16 B * 16 * 8 * CUs + 16 B * words;
 * 256 KiB of partials on a 256-CU device), else LGH_ERR_INVALID_ARGUMENT.
 * The scratch array is shared by every call on a device: calls that overlap
 * on one device race and must be serialised by the caller.  The first call
 * on a device looks up the array's address; later calls allocate nothing. */
lgh_status lgh_expval_pauli_words_noalloc(const lgh_sv *sv, int64_t num_words, const char *paulis,
                                          const int32_t *wires, const int64_t *offsets, const double *coeffs,
                                          double *per_word, double *result);
/* CSR kept resident on the device (replaces the per-call upload of
 * getExpectationValueOnSparseSpMV, SVCM:795-922).  values: complex128. */
lgh_status lgh_csr_create(lgh_csr **out, int device, lgh_dtype dtype, int64_t num_rows,
                          const int64_t *indptr, const int64_t *indices, const double *values,
                          int64_t nnz);
lgh_status lgh_csr_destroy(lgh_csr *csr);
/* <psi|H|psi> with SpMV and dot fused (SVCM:854-919) */
lgh_status lgh_expval_csr(const lgh_sv *sv, const lgh_csr *csr, double *re, double *im);
/* out = H in (SparseHamiltonianGPU::applyInPlace, ObservablesGPU.hpp:469-587) */
lgh_status lgh_apply_csr(lgh_sv *out, const lgh_csr *csr, const lgh_sv *in);
/* probability(wires): wires[0] is the least significant output bit, as the
 * reference returns it (SVCM:931-970).  out: 2^num_wires doubles. */
lgh_status lgh_probs(const lgh_sv *sv, const int32_t *wires, int32_t num_wires, double *out);
/* generate_samples (SVCM:982-1061) with a seeded Philox4x32-10 stream:
 * writes `shots` basis indices in ascending order. */
lgh_status lgh_sample(const lgh_sv *sv, uint64_t shots, uint64_t seed, uint64_t *indices);
/* Same, expanded to a (shots x num_qubits) 0/1 array: column (n-1-j) holds
 * bit j of the index (SVCM:1037-1055). */
lgh_status lgh_sample_bits(const lgh_sv *sv, uint64_t shots, uint64_t seed, uint64_t *bits);

/* ---- adjoint Jacobian (AdjointDiffGPU.hpp:499-596) -------------------- */
/* psi: forward state (read only; copied).  bras[j] = O_j psi on entry,
 * overwritten.  trainable: sorted indices into the parametrised ops.
 * jac: num_obs x num_trainable, row-major.  If apply_operations, psi is the
 * initial state and the ops are applied first. */
lgh_status lgh_adjoint_jacobian(const lgh_sv *psi, lgh_sv *const *bras, int32_t num_obs,
                                const lgh_op *ops, int64_t num_ops, const int64_t *trainable,
                                int64_t num_trainable, int32_t apply_operations, double *jac);

/* lgh_adjoint_jacobian for observables that are words of Z (and identity)
 * factors, the common case of PennyLane's expval(qml.PauliZ(i) @ ...): the
 * reference forms every H lambda with the observable's gates before the
 * sweep (applyObservables, AdjointDiffGPU.hpp:228-268, ObservablesGPU.hpp:88-
 * 128, :209-260).  Here observable j is given by zmasks[j], the
 * amplitude-index bits (bit n-1-w for wire w) of its Z factors, and bras[j]
 * is scratch of psi's size that the sweep fills with Z_j psi itself: in its
 * first reverse pass when that pass is fused, which saves a read and a write
 * of the state per observable.  Results equal lgh_adjoint_jacobian's with
 * bras[j] = Z_j psi, bit for bit. */
lgh_status lgh_adjoint_jacobian_zwords(const lgh_sv *psi, lgh_sv *const *bras, const uint64_t *zmasks,
                                       int32_t num_obs, const lgh_op *ops, int64_t num_ops, const int64_t *trainable,
                                       int64_t num_trainable, int32_t apply_operations, double *jac);

/* ---- batched circuits ---------------------------------------------------
 * No counterpart in the reference, which runs one circuit per device call
 * (the closest form is the per-GPU observable batching of batch_obs,
 * AdjointDiffGPU.hpp:392-473).  A template is the gate structure shared by a
 * family of circuits that differ only in their parameters, with one Pauli
 * word as observable; a batch runs many circuits (any mix of templates) in
 * one persistent launch: forward pass from |0..0>, <P>, and optionally the
 * adjoint gradient of <P> with respect to every circuit parameter. */
typedef struct lgh_tmpl lgh_tmpl;

/* ops: the circuit (num_qubits <= 20; every gate single-target dense or
 * diagonal, e.g. RX/RY/RZ/PhaseShift/CRX/CRY/CRZ/CPhase/IsingZZ/MultiRZ and
 * fixed 1- and controlled gates).  param_index[i]: index into the circuit's
 * parameter vector that drives ops[i].params[0] (parameter sharing allowed),
 * or -1 for a gate whose params are fixed as given.  The observable is the
 * tensor product of obs_paulis[k] (LGH_GATE_PAULIX/Y/Z) on obs_wires[k]. */
lgh_status lgh_tmpl_create(lgh_tmpl **out, int32_t num_qubits, const lgh_op *ops, int32_t num_ops,
                           const int32_t *param_index, const int32_t *obs_paulis, const int32_t *obs_wires,
                           int32_t obs_len, int32_t device);
lgh_status lgh_tmpl_destroy(lgh_tmpl *tmpl);
/* number of parameters (1 + the largest param_index) and of qubits */
lgh_status lgh_tmpl_info(const lgh_tmpl *tmpl, int32_t *num_params, int32_t *num_qubits);
/* Passes a batched run makes over the circuit's states: forward passes,
 * reverse (adjoint) passes, and the qubits of the states it runs on (at
 * least 9; smaller circuits run with idle low-order wires).  For byte
 * accounting of the batched kernel. */
lgh_status lgh_tmpl_passes(const lgh_tmpl *tmpl, int32_t *forward_passes, int32_t *reverse_passes,
                           int32_t *state_qubits);

/* Circuit c (0 <= c < num_circuits) is tmpls[c] with parameters
 * params[param_offset[c] ...] (host memory, float64).  expval[c] = <P>; if
 * jac is not NULL, jac[jac_offset[c] + j] = d<P>/d param_j for the template's
 * parameters.  dtype selects complex64 or complex128 states; stream as for
 * lgh_sv_create (NULL: the library's stream).  Blocks until the results are
 * on the host. */
lgh_status lgh_batch_run(lgh_tmpl *const *tmpls, int32_t num_circuits, lgh_dtype dtype, const double *params,
                         const int64_t *param_offset, double *expval, double *jac, const int64_t *jac_offset,
                         void *stream);
/* Same results through one state per circuit and the single-circuit entry
 * points (lgh_apply, lgh_expval_pauli_words, lgh_adjoint_jacobian), one
 * circuit after another: the per-circuit launch form the batch replaces. */
lgh_status lgh_batch_run_serial(lgh_tmpl *const *tmpls, int32_t num_circuits, lgh_dtype dtype,
                                const double *params, const int64_t *param_offset, double *expval, double *jac,
                                const int64_t *jac_offset, void *stream);

/* A batch of fixed composition (the same templates, in the same order, run
 * again and again with new parameters: a training loop, a parameter sweep)
 * prepared once.  Creation uploads the templates, allocates the device and
 * pinned host buffers and captures the launch sequence (parameter upload,
 * descriptor preparation, persistent launch, result download) in a HIP
 * graph on a stream of the plan's own; it also autotunes the persistent
 * launch's grid (one or two resident workgroups per CU), timing both on the
 * plan's circuits and keeping the faster.  Each run replays the graph: the
 * results equal lgh_batch_run's bit for bit.  No counterpart in the
 * reference (see lgh_batch_run).  A plan must not be run from two threads at
 * once, and its templates must outlive it. */
typedef struct lgh_batch_plan lgh_batch_plan;
lgh_status lgh_batch_plan_create(lgh_batch_plan **out, lgh_tmpl *const *tmpls, int32_t num_circuits,
                                 lgh_dtype dtype, int32_t with_grad);
/* params / param_offset / expval / jac / jac_offset as for lgh_batch_run;
 * jac must be NULL for a plan created without gradients. */
lgh_status lgh_batch_plan_run(lgh_batch_plan *plan, const double *params, const int64_t *param_offset,
                              double *expval, double *jac, const int64_t *jac_offset);
/* The tuned grid (workgroups per CU) and the timed candidates (ms, one and
 * two workgroups per CU). */
lgh_status lgh_batch_plan_info(const lgh_batch_plan *plan, int32_t *grid_per_cu, double *tuned_ms);
lgh_status lgh_batch_plan_destroy(lgh_batch_plan *plan);

/* ---- serving ------------------------------------------------------------
 * No counterpart in the reference, where each PennyLane execution calls the
 * device synchronously (lightning_gpu.py:519-555, :638-752); one request is
 * one such execution.  Worker threads with one stream each take requests
 * from a queue: circuits of at most batch_max_qubits qubits are batched
 * together into persistent launches (continuous batching), larger ones run
 * on states of their own; all work is admitted against a device-memory
 * budget. */
typedef struct lgh_server lgh_server;
typedef struct lgh_request {
    int32_t num_qubits;
    const lgh_op *ops;          /* bound parameters; copied at submission */
    int32_t num_ops;
    const int32_t *obs_paulis;  /* LGH_GATE_PAULIX/Y/Z */
    const int32_t *obs_wires;
    int32_t obs_len;
    int32_t want_grad;          /* d<P>/d params of every parametrised op, in op order */
} lgh_request;

/* memory_budget: device bytes the server may hold (0: 90% of free memory);
 * num_streams: worker threads / streams, 1..16; batch_max_qubits: 0..20. */
lgh_status lgh_server_create(lgh_server **out, int32_t device, uint64_t memory_budget, int32_t num_streams,
                             int32_t batch_max_qubits);
/* Stops the workers after their current work; requests still queued are dropped. */
lgh_status lgh_server_destroy(lgh_server *server);
/* Queues a request; rejects one that can never fit the budget. */
lgh_status lgh_server_submit(lgh_server *server, const lgh_request *request, uint64_t *ticket);
/* Blocks until the request is done.  grad (nullable) receives grad_len >=
 * number of parametrised ops entries; timing_ms (nullable, 3 entries):
 * queueing delay, run time, total latency. */
lgh_status lgh_server_wait(lgh_server *server, uint64_t ticket, double *expval, double *grad, int64_t grad_len,
                           double *timing_ms);
lgh_status lgh_server_stats(lgh_server *server, uint64_t *batches, uint64_t *batched_requests,
                            uint64_t *single_requests, uint64_t *budget);
/* Batches that may run at once (default: num_streams - 1, at least 1) and
 * circuits per batch (default 4096). */
lgh_status lgh_server_set_limits(lgh_server *server, int32_t max_concurrent_batches, int32_t max_batch_circuits);
/* Batch formation under load: while other work runs, a batch of fewer than
 * min_batch circuits waits up to wait_us microseconds after its oldest
 * circuit arrived for more of its width class; an idle engine launches at
 * once.  Defaults 64 and 1000; wait_us = 0 launches every batch at once. */
lgh_status lgh_server_set_batching(lgh_server *server, int32_t min_batch, int32_t wait_us);
/* Failure injection for tests: the next `count` pieces of work (a batch or a
 * single circuit) fail with LGH_ERR_DEVICE before launching anything; their
 * requests report the error from lgh_server_wait, all other requests run
 * normally. */
lgh_status lgh_server_inject_failures(lgh_server *server, uint64_t count);

#ifdef __cplusplus
}
#endif
#endif /* LGH_H */
