/*
 * oracle.h -- CPU restatement of the lightning.gpu statevector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against: only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links,
 * loads or falls back to it.
 *
 * Conventions (PennyLane, as used by the reference):
 *   - amplitude index bit (n-1-w) <-> wire w (wire 0 is the MSB)
 *     [pennylane_lightning_gpu/src/simulator/StateVectorCudaManaged.hpp:1349-1356]
 *   - gate matrices are row-major and the first listed wire is the MSB of the
 *     matrix index [pennylane_lightning_gpu/lightning_gpu.py:545-551]
 *   - adjoint Jacobian: jac = -2 * s * Im<H lambda | G mu>
 *     [pennylane_lightning_gpu/src/algorithms/AdjointDiffGPU.hpp:132-162, :499-596]
 *
 * Every entry point returns 0 on success and a negative value on error,
 * unless documented otherwise.
 */
#ifndef LGH_ORACLE_H
#define LGH_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fill `out` (2^nw x 2^nw, row-major) with the matrix of a named gate
 * (its adjoint when `inverse`).  Returns the number of wires the gate acts
 * on, or -1 for an unknown name / wrong arity. */
int orc_gate_matrix(const char *name, int nw, const double *params,
                    int inverse, double complex *out);

/* Generator G and scaling factor s with U(p) = exp(i*s*p*G) for the
 * differentiable gates.  Returns -1 when the gate has no generator. */
int orc_generator(const char *name, int nw, double complex *out,
                  double *scale);

/* Apply one operation in place.  `matrix` (2^nw x 2^nw, row-major, may be
 * NULL) is used when `name` is not a known gate. */
int orc_apply_op(int n, double complex *psi, const char *name,
                 const int *wires, int nw, int inverse, const double *params,
                 const double complex *matrix);

/* Apply a list of operations.  wires_flat holds the wires of op i at
 * [woff[i], woff[i+1]); params holds 3 doubles per op; matrices[i] may be
 * NULL. */
int orc_apply_ops(int n, double complex *psi, int nops,
                  const char *const *names, const int *wires_flat,
                  const int *woff, const int *inverses, const double *params,
                  const double complex *const *matrices);

/* Apply the generator of op (name, wires) in place (no scaling). */
int orc_apply_generator(int n, double complex *psi, const char *name,
                        const int *wires, int nw);

/* <psi| M |psi> for a dense M on `wires` (PennyLane wire order). */
int orc_expval_matrix(int n, const double complex *psi, const int *wires,
                      int k, const double complex *M, double *re,
                      double *im);

/* Pauli-word expectation values.  paulis[offsets[w] .. offsets[w+1]) are
 * characters in {I,X,Y,Z} acting on wires_flat[same range].  per_word
 * receives <P_w> (may be NULL); the return value in *total is
 * Re(sum_w coeffs[w] * <P_w>). */
int orc_expval_pauli_words(int n, const double complex *psi, int nwords,
                           const char *paulis, const int *wires_flat,
                           const int64_t *offsets,
                           const double complex *coeffs, double *per_word,
                           double *total);

/* y = H x for a CSR matrix (0-based). */
int orc_spmv(int64_t nrows, const int64_t *indptr, const int64_t *indices,
             const double complex *data, const double complex *x,
             double complex *y);

/* <psi| H |psi> for a CSR matrix. */
int orc_expval_csr(int n, const double complex *psi, const int64_t *indptr,
                   const int64_t *indices, const double complex *data,
                   double *re, double *im);

/* Marginal probabilities over `wires`, in the reference binding's order:
 * wires[0] is the least significant bit of the output index
 * [pennylane_lightning_gpu/src/simulator/StateVectorCudaManaged.hpp:931-970]. */
int orc_probs(int n, const double complex *psi, const int *wires, int k,
              double *out);

/* Philox4x32-10 block for a full 128-bit counter and 64-bit key. */
void orc_philox4x32_10_raw(const uint32_t ctr[4], const uint32_t key[2],
                           uint32_t out[4]);
/* Philox4x32-10 block for counter (counter lo, hi, 0, 0), key (seed lo, hi). */
void orc_philox4x32_10(uint64_t seed, uint64_t counter, uint32_t out[4]);

/* Seeded inverse-CDF sampler.  Writes `shots` basis indices sorted in
 * ascending order.  The CDF is built in 2^-60 fixed point so that the
 * result does not depend on summation order. */
int orc_sample(int n, const double complex *psi, uint64_t shots,
               uint64_t seed, uint64_t *out);
int orc_sample_c64(int n, const float complex *psi, uint64_t shots,
                   uint64_t seed, uint64_t *out);

/* Adjoint Jacobian.  psi is the forward state U_N..U_1|init>; bras[j] is
 * O_j psi (input, consumed as scratch copies internally).  trainable holds
 * sorted indices into the list of parametrised ops.  jac is
 * nobs x ntp row-major. */
int orc_adjoint_jacobian(int n, const double complex *psi, int nobs,
                         const double complex *const *bras, int nops,
                         const char *const *names, const int *wires_flat,
                         const int *woff, const int *inverses,
                         const double *params, const int *nparams,
                         const double complex *const *matrices,
                         const int64_t *trainable, int ntp, double *jac);

/* Number of OpenMP threads the oracle uses. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
