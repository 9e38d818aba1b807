/*
 * oracle.c -- CPU restatement of the lightning.gpu statevector hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the gate
 * mathematics with PennyLane conventions; the reference's own CPU path
 * (pennylane-lightning, fetched from GitHub at configure time,
 * CMakeLists.txt:109-114) is not available offline, and its GPU path needs
 * cuStateVec/cuBLAS/cuSPARSE, so this file restates the semantics of the
 * call sites cited below.
 *
 * Build: oracle/build.sh (gcc -O3 -fopenmp -ffp-contract=off).
 * -ffp-contract=off matters for orc_sample: the probability of an amplitude
 * must be re*re + im*im with two roundings, exactly as the HIP sampler
 * computes it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex dc;

#define ORC_MAX_K 12

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Gate matrices (PennyLane definitions, first wire = MSB of matrix index).  */
/* Reference semantics: cuGates_host.hpp (LSB-first there; re-derived here). */
/* ------------------------------------------------------------------------ */

static void mat_zero(dc *m, int dim) { memset(m, 0, sizeof(dc) * dim * dim); }
static void mat_eye(dc *m, int dim) {
    mat_zero(m, dim);
    for (int i = 0; i < dim; i++) m[i * dim + i] = 1.0;
}

/* out = controlled(U) with `nc` leading control wires; U is dim_t x dim_t. */
static void controlled(const dc *u, int kt, int nc, dc *out) {
    int dim = 1 << (kt + nc);
    int dt = 1 << kt;
    mat_eye(out, dim);
    int off = dim - dt; /* all controls = 1 block sits at the bottom-right */
    for (int r = 0; r < dt; r++)
        for (int c = 0; c < dt; c++) out[(off + r) * dim + off + c] = u[r * dt + c];
}

static void rot_matrix(double phi, double theta, double omega, dc *m) {
    /* Rot(phi,theta,omega) = RZ(omega) RY(theta) RZ(phi)
       [cuGates_host.hpp:390-422; _serialize.py:311-312] */
    double c = cos(theta / 2), s = sin(theta / 2);
    m[0] = cexp(-I * (phi + omega) / 2) * c;
    m[1] = -cexp(I * (phi - omega) / 2) * s;
    m[2] = cexp(-I * (phi - omega) / 2) * s;
    m[3] = cexp(I * (phi + omega) / 2) * c;
}

static int one_qubit(const char *name, const double *p, dc *m) {
    double c, s;
    if (!strcmp(name, "Identity")) { mat_eye(m, 2); return 1; }
    if (!strcmp(name, "PauliX")) { m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; return 1; }
    if (!strcmp(name, "PauliY")) { m[0] = 0; m[1] = -I; m[2] = I; m[3] = 0; return 1; }
    if (!strcmp(name, "PauliZ")) { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; return 1; }
    if (!strcmp(name, "Hadamard")) {
        double r = 1.0 / sqrt(2.0);
        m[0] = r; m[1] = r; m[2] = r; m[3] = -r; return 1;
    }
    if (!strcmp(name, "S")) { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = I; return 1; }
    if (!strcmp(name, "T")) { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cexp(I * M_PI / 4); return 1; }
    if (!strcmp(name, "PhaseShift")) { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cexp(I * p[0]); return 1; }
    if (!strcmp(name, "RX")) {
        c = cos(p[0] / 2); s = sin(p[0] / 2);
        m[0] = c; m[1] = -I * s; m[2] = -I * s; m[3] = c; return 1;
    }
    if (!strcmp(name, "RY")) {
        c = cos(p[0] / 2); s = sin(p[0] / 2);
        m[0] = c; m[1] = -s; m[2] = s; m[3] = c; return 1;
    }
    if (!strcmp(name, "RZ")) {
        m[0] = cexp(-I * p[0] / 2); m[1] = 0; m[2] = 0; m[3] = cexp(I * p[0] / 2); return 1;
    }
    if (!strcmp(name, "Rot")) { rot_matrix(p[0], p[1], p[2], m); return 1; }
    return 0;
}

static void excitation2(double phi, dc diag, dc *m) {
    double c = cos(phi / 2), s = sin(phi / 2);
    mat_zero(m, 4);
    m[0 * 4 + 0] = diag; m[3 * 4 + 3] = diag;
    m[1 * 4 + 1] = c; m[1 * 4 + 2] = -s;
    m[2 * 4 + 1] = s; m[2 * 4 + 2] = c;
}

static void excitation4(double phi, dc diag, dc *m) {
    double c = cos(phi / 2), s = sin(phi / 2);
    mat_zero(m, 16);
    for (int i = 0; i < 16; i++) m[i * 16 + i] = diag;
    m[3 * 16 + 3] = c; m[3 * 16 + 12] = -s;
    m[12 * 16 + 3] = s; m[12 * 16 + 12] = c;
}

static int gate_matrix_fwd(const char *name, int nw, const double *p, dc *m) {
    dc u[4];
    double c, s;
    if (nw == 1 && one_qubit(name, p, m)) return 1;
    if (nw == 2) {
        if (!strcmp(name, "CNOT")) { one_qubit("PauliX", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CY")) { one_qubit("PauliY", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CZ")) { one_qubit("PauliZ", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CRX")) { one_qubit("RX", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CRY")) { one_qubit("RY", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CRZ")) { one_qubit("RZ", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "CRot")) { rot_matrix(p[0], p[1], p[2], u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "ControlledPhaseShift")) { one_qubit("PhaseShift", p, u); controlled(u, 1, 1, m); return 2; }
        if (!strcmp(name, "SWAP")) {
            mat_zero(m, 4); m[0] = 1; m[1 * 4 + 2] = 1; m[2 * 4 + 1] = 1; m[15] = 1; return 2;
        }
        if (!strcmp(name, "IsingXX")) {
            c = cos(p[0] / 2); s = sin(p[0] / 2);
            mat_zero(m, 4);
            for (int i = 0; i < 4; i++) { m[i * 4 + i] = c; m[i * 4 + (3 - i)] = -I * s; }
            return 2;
        }
        if (!strcmp(name, "IsingYY")) {
            /* c*I - i*s*(Y (x) Y); Y(x)Y = antidiag(-1, 1, 1, -1) */
            c = cos(p[0] / 2); s = sin(p[0] / 2);
            mat_zero(m, 4);
            for (int i = 0; i < 4; i++) m[i * 4 + i] = c;
            m[0 * 4 + 3] = I * s; m[1 * 4 + 2] = -I * s; m[2 * 4 + 1] = -I * s; m[3 * 4 + 0] = I * s;
            return 2;
        }
        if (!strcmp(name, "IsingZZ")) {
            mat_zero(m, 4);
            m[0] = cexp(-I * p[0] / 2); m[5] = cexp(I * p[0] / 2);
            m[10] = cexp(I * p[0] / 2); m[15] = cexp(-I * p[0] / 2);
            return 2;
        }
        if (!strcmp(name, "SingleExcitation")) { excitation2(p[0], 1.0, m); return 2; }
        if (!strcmp(name, "SingleExcitationMinus")) { excitation2(p[0], cexp(-I * p[0] / 2), m); return 2; }
        if (!strcmp(name, "SingleExcitationPlus")) { excitation2(p[0], cexp(I * p[0] / 2), m); return 2; }
    }
    if (nw == 3) {
        if (!strcmp(name, "Toffoli")) { one_qubit("PauliX", p, u); controlled(u, 1, 2, m); return 3; }
        if (!strcmp(name, "CSWAP")) {
            dc sw[16];
            gate_matrix_fwd("SWAP", 2, p, sw);
            controlled(sw, 2, 1, m);
            return 3;
        }
    }
    if (nw == 4) {
        if (!strcmp(name, "DoubleExcitation")) { excitation4(p[0], 1.0, m); return 4; }
        if (!strcmp(name, "DoubleExcitationMinus")) { excitation4(p[0], cexp(-I * p[0] / 2), m); return 4; }
        if (!strcmp(name, "DoubleExcitationPlus")) { excitation4(p[0], cexp(I * p[0] / 2), m); return 4; }
    }
    if (!strcmp(name, "MultiRZ") && nw >= 1 && nw <= ORC_MAX_K) {
        int dim = 1 << nw;
        mat_zero(m, dim);
        for (int i = 0; i < dim; i++) {
            int par = __builtin_popcount((unsigned)i) & 1;
            m[i * dim + i] = cexp(-I * p[0] / 2 * (par ? -1.0 : 1.0));
        }
        return nw;
    }
    return -1;
}

static void dagger(dc *m, int dim) {
    for (int r = 0; r < dim; r++)
        for (int c = r; c < dim; c++) {
            dc a = m[r * dim + c], b = m[c * dim + r];
            m[r * dim + c] = conj(b);
            m[c * dim + r] = conj(a);
        }
}

int orc_gate_matrix(const char *name, int nw, const double *params,
                    int inverse, double complex *out) {
    double zero[3] = {0, 0, 0};
    const double *p = params ? params : zero;
    int k = gate_matrix_fwd(name, nw, p, out);
    if (k < 0) return -1;
    if (inverse) dagger(out, 1 << k);
    return k;
}

/* Generators [GateGenerators.hpp:58-320; scaling AdjointDiffGPU.hpp:96-114] */
int orc_generator(const char *name, int nw, double complex *g, double *scale) {
    dc u[4];
    if (nw == 1) {
        if (!strcmp(name, "RX")) { one_qubit("PauliX", NULL, g); *scale = -0.5; return 1; }
        if (!strcmp(name, "RY")) { one_qubit("PauliY", NULL, g); *scale = -0.5; return 1; }
        if (!strcmp(name, "RZ")) { one_qubit("PauliZ", NULL, g); *scale = -0.5; return 1; }
        if (!strcmp(name, "PhaseShift")) { mat_zero(g, 2); g[3] = 1; *scale = 1.0; return 1; }
    }
    if (nw == 2) {
        if (!strcmp(name, "CRX") || !strcmp(name, "CRY") || !strcmp(name, "CRZ")) {
            const char *pn = name[2] == 'X' ? "PauliX" : (name[2] == 'Y' ? "PauliY" : "PauliZ");
            one_qubit(pn, NULL, u);
            mat_zero(g, 4);
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) g[(2 + r) * 4 + 2 + c] = u[r * 2 + c];
            *scale = -0.5;
            return 2;
        }
        if (!strcmp(name, "ControlledPhaseShift")) { mat_zero(g, 4); g[15] = 1; *scale = 1.0; return 2; }
        if (!strncmp(name, "Ising", 5)) {
            const char *pn = name[5] == 'X' ? "PauliX" : (name[5] == 'Y' ? "PauliY" : "PauliZ");
            if (strlen(name) != 7 || name[5] != name[6]) return -1;
            one_qubit(pn, NULL, u);
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) g[r * 4 + c] = u[(r >> 1) * 2 + (c >> 1)] * u[(r & 1) * 2 + (c & 1)];
            *scale = -0.5;
            return 2;
        }
        if (!strncmp(name, "SingleExcitation", 16)) {
            dc d = 0;
            if (!strcmp(name, "SingleExcitationMinus")) d = 1;
            else if (!strcmp(name, "SingleExcitationPlus")) d = -1;
            else if (strcmp(name, "SingleExcitation")) return -1;
            mat_zero(g, 4);
            g[0] = d; g[15] = d;
            g[1 * 4 + 2] = -I; g[2 * 4 + 1] = I;
            *scale = -0.5;
            return 2;
        }
    }
    if (nw == 4 && !strncmp(name, "DoubleExcitation", 16)) {
        dc d = 0;
        if (!strcmp(name, "DoubleExcitationMinus")) d = 1;
        else if (!strcmp(name, "DoubleExcitationPlus")) d = -1;
        else if (strcmp(name, "DoubleExcitation")) return -1;
        mat_zero(g, 16);
        for (int i = 0; i < 16; i++) g[i * 16 + i] = d;
        g[3 * 16 + 3] = 0; g[12 * 16 + 12] = 0;
        g[3 * 16 + 12] = -I; g[12 * 16 + 3] = I;
        *scale = -0.5;
        return 4;
    }
    if (!strcmp(name, "MultiRZ") && nw >= 1 && nw <= ORC_MAX_K) {
        int dim = 1 << nw;
        mat_zero(g, dim);
        for (int i = 0; i < dim; i++) g[i * dim + i] = (__builtin_popcount((unsigned)i) & 1) ? -1.0 : 1.0;
        *scale = -0.5;
        return nw;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Dense application of a k-wire matrix.                                    */
/* ------------------------------------------------------------------------ */

/* insert a zero bit at each (ascending) position in `pos` */
static inline uint64_t insert_zeros(uint64_t g, const int *pos_sorted, int k) {
    for (int j = 0; j < k; j++) {
        uint64_t p = pos_sorted[j];
        uint64_t lo = g & ((1ULL << p) - 1);
        g = ((g >> p) << (p + 1)) | lo;
    }
    return g;
}

static int check_wires(int n, const int *wires, int k) {
    if (k < 0 || k > n || k > ORC_MAX_K) return -1;
    for (int i = 0; i < k; i++) {
        if (wires[i] < 0 || wires[i] >= n) return -1;
        for (int j = 0; j < i; j++)
            if (wires[i] == wires[j]) return -1;
    }
    return 0;
}

static int apply_dense(int n, dc *psi, const int *wires, int k, const dc *M) {
    if (check_wires(n, wires, k)) return -1;
    int dim = 1 << k;
    int bitpos[ORC_MAX_K], sorted[ORC_MAX_K];
    for (int j = 0; j < k; j++) bitpos[j] = n - 1 - wires[j];
    memcpy(sorted, bitpos, sizeof(int) * k);
    for (int a = 0; a < k; a++)
        for (int b = a + 1; b < k; b++)
            if (sorted[b] < sorted[a]) { int t = sorted[a]; sorted[a] = sorted[b]; sorted[b] = t; }
    /* offset of matrix index m: bit (k-1-j) of m <-> wire j */
    uint64_t offs[1 << ORC_MAX_K];
    for (int m = 0; m < dim; m++) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if ((m >> (k - 1 - j)) & 1) o |= 1ULL << bitpos[j];
        offs[m] = o;
    }
    int64_t ngroups = (int64_t)1 << (n - k);
    if (k == 1) {
        dc m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
        uint64_t o1 = offs[1];
#pragma omp parallel for schedule(static) if (ngroups > 4096)
        for (int64_t g = 0; g < ngroups; g++) {
            uint64_t b = insert_zeros((uint64_t)g, sorted, 1);
            dc a0 = psi[b], a1 = psi[b | o1];
            psi[b] = m00 * a0 + m01 * a1;
            psi[b | o1] = m10 * a0 + m11 * a1;
        }
        return 0;
    }
#pragma omp parallel if (ngroups > 1024)
    {
        dc v[1 << ORC_MAX_K], w[1 << ORC_MAX_K];
#pragma omp for schedule(static)
        for (int64_t g = 0; g < ngroups; g++) {
            uint64_t b = insert_zeros((uint64_t)g, sorted, k);
            for (int m = 0; m < dim; m++) v[m] = psi[b | offs[m]];
            for (int r = 0; r < dim; r++) {
                dc acc = 0;
                for (int c = 0; c < dim; c++) acc += M[r * dim + c] * v[c];
                w[r] = acc;
            }
            for (int m = 0; m < dim; m++) psi[b | offs[m]] = w[m];
        }
    }
    return 0;
}

/* MultiRZ on many wires: diagonal phase by parity. */
static int apply_multirz(int n, dc *psi, const int *wires, int k, double phi, int inverse) {
    if (k < 1 || k > n) return -1;
    uint64_t mask = 0;
    for (int j = 0; j < k; j++) {
        if (wires[j] < 0 || wires[j] >= n) return -1;
        mask |= 1ULL << (n - 1 - wires[j]);
    }
    double a = inverse ? -phi : phi;
    dc even = cexp(-I * a / 2), odd = cexp(I * a / 2);
    int64_t N = (int64_t)1 << n;
#pragma omp parallel for schedule(static) if (N > 4096)
    for (int64_t i = 0; i < N; i++)
        psi[i] *= (__builtin_popcountll((uint64_t)i & mask) & 1) ? odd : even;
    return 0;
}

int orc_apply_op(int n, double complex *psi, const char *name,
                 const int *wires, int nw, int inverse, const double *params,
                 const double complex *matrix) {
    if (!strcmp(name, "Identity")) return 0;
    if (!strcmp(name, "MultiRZ") && nw > 6)
        return apply_multirz(n, psi, wires, nw, params ? params[0] : 0.0, inverse);
    if (nw > ORC_MAX_K) return -1;
    dc *M = (dc *)malloc(sizeof(dc) << (2 * nw));
    if (!M) return -2;
    int k = orc_gate_matrix(name, nw, params, inverse, M);
    if (k < 0) {
        if (!matrix) { free(M); return -1; }
        memcpy(M, matrix, sizeof(dc) << (2 * nw));
        if (inverse) dagger(M, 1 << nw);
        k = nw;
    }
    int rc = apply_dense(n, psi, wires, k, M);
    free(M);
    return rc;
}

int orc_apply_ops(int n, double complex *psi, int nops,
                  const char *const *names, const int *wires_flat,
                  const int *woff, const int *inverses, const double *params,
                  const double complex *const *matrices) {
    for (int i = 0; i < nops; i++) {
        int rc = orc_apply_op(n, psi, names[i], wires_flat + woff[i],
                              woff[i + 1] - woff[i], inverses ? inverses[i] : 0,
                              params + 3 * i, matrices ? matrices[i] : NULL);
        if (rc) return rc;
    }
    return 0;
}

int orc_apply_generator(int n, double complex *psi, const char *name,
                        const int *wires, int nw) {
    if (nw > ORC_MAX_K) return -1;
    dc *G = (dc *)malloc(sizeof(dc) << (2 * nw));
    double s;
    if (!G) return -2;
    int k = orc_generator(name, nw, G, &s);
    int rc = k < 0 ? -1 : apply_dense(n, psi, wires, k, G);
    free(G);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Measurements                                                             */
/* ------------------------------------------------------------------------ */

int orc_expval_matrix(int n, const double complex *psi, const int *wires,
                      int k, const double complex *M, double *re, double *im) {
    /* <psi|M|psi> = sum_i conj(psi_i) (M psi)_i, with M psi computed on a copy */
    if (check_wires(n, wires, k)) return -1;
    int64_t N = (int64_t)1 << n;
    dc *tmp = (dc *)malloc(sizeof(dc) * N);
    if (!tmp) return -2;
    memcpy(tmp, psi, sizeof(dc) * N);
    int rc = apply_dense(n, tmp, wires, k, M);
    double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si) schedule(static) if (N > 4096)
    for (int64_t i = 0; i < N; i++) {
        dc t = conj(psi[i]) * tmp[i];
        sr += creal(t);
        si += cimag(t);
    }
    free(tmp);
    *re = sr;
    *im = si;
    return rc;
}

int orc_expval_pauli_words(int n, const double complex *psi, int nwords,
                           const char *paulis, const int *wires_flat,
                           const int64_t *offsets,
                           const double complex *coeffs, double *per_word,
                           double *total) {
    /* P|i> = coef(i) |i ^ flip>,  coef(i) = i^{nY} (-1)^{popcount(i & zy)}
       <P> = sum_i conj(psi_{i^flip}) coef(i) psi_i
       [reference semantics: StateVectorCudaManaged.hpp:1071-1148] */
    int64_t N = (int64_t)1 << n;
    dc acc = 0;
    for (int w = 0; w < nwords; w++) {
        uint64_t flip = 0, zy = 0;
        int ny = 0;
        for (int64_t j = offsets[w]; j < offsets[w + 1]; j++) {
            int wire = wires_flat[j];
            if (wire < 0 || wire >= n) return -1;
            uint64_t bit = 1ULL << (n - 1 - wire);
            switch (paulis[j]) {
            case 'I': break;
            case 'X': flip |= bit; break;
            case 'Y': flip |= bit; zy |= bit; ny++; break;
            case 'Z': zy |= bit; break;
            default: return -1;
            }
        }
        dc ph = 1;
        for (int t = 0; t < (ny & 3); t++) ph *= I;
        double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si) schedule(static) if (N > 4096)
        for (int64_t i = 0; i < N; i++) {
            double sg = (__builtin_popcountll((uint64_t)i & zy) & 1) ? -1.0 : 1.0;
            dc t = conj(psi[(uint64_t)i ^ flip]) * psi[i] * sg;
            sr += creal(t);
            si += cimag(t);
        }
        dc e = ph * (sr + I * si);
        if (per_word) per_word[w] = creal(e);
        acc += creal(e) * coeffs[w];
    }
    *total = creal(acc);
    return 0;
}

int orc_spmv(int64_t nrows, const int64_t *indptr, const int64_t *indices,
             const double complex *data, const double complex *x,
             double complex *y) {
#pragma omp parallel for schedule(static) if (nrows > 4096)
    for (int64_t r = 0; r < nrows; r++) {
        dc acc = 0;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; k++) acc += data[k] * x[indices[k]];
        y[r] = acc;
    }
    return 0;
}

int orc_expval_csr(int n, const double complex *psi, const int64_t *indptr,
                   const int64_t *indices, const double complex *data,
                   double *re, double *im) {
    /* y = H psi, then <psi|y> [StateVectorCudaManaged.hpp:795-922] */
    int64_t N = (int64_t)1 << n;
    dc *y = (dc *)malloc(sizeof(dc) * N);
    if (!y) return -2;
    orc_spmv(N, indptr, indices, data, psi, y);
    double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si) schedule(static) if (N > 4096)
    for (int64_t i = 0; i < N; i++) {
        dc t = conj(psi[i]) * y[i];
        sr += creal(t);
        si += cimag(t);
    }
    free(y);
    *re = sr;
    *im = si;
    return 0;
}

int orc_probs(int n, const double complex *psi, const int *wires, int k,
              double *out) {
    if (k < 0 || k > n) return -1;
    for (int j = 0; j < k; j++)
        if (wires[j] < 0 || wires[j] >= n) return -1;
    int64_t N = (int64_t)1 << n, M = (int64_t)1 << k;
    memset(out, 0, sizeof(double) * M);
    for (int64_t i = 0; i < N; i++) {
        int64_t o = 0;
        for (int j = 0; j < k; j++) o |= (int64_t)((i >> (n - 1 - wires[j])) & 1) << j;
        double p = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
        out[o] += p;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Seeded sampler                                                           */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11.                                                 */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10_raw(const uint32_t ctr[4], const uint32_t key[2],
                           uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox4x32_10(uint64_t seed, uint64_t counter, uint32_t out[4]) {
    /* counter = (shot index lo, hi, 0, 0), key = (seed lo, seed hi) */
    uint32_t ctr[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), 0, 0};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10_raw(ctr, key, out);
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

#define ORC_FIXED_SCALE 1152921504606846976.0 /* 2^60 */

static int sample_from_cdf(int64_t N, const uint64_t *cdf, uint64_t shots,
                           uint64_t seed, uint64_t *out) {
    uint64_t total = cdf[N - 1];
    if (total == 0) return -1;
#pragma omp parallel for schedule(static) if (shots > 1024)
    for (int64_t s = 0; s < (int64_t)shots; s++) {
        uint32_t r[4];
        orc_philox4x32_10(seed, (uint64_t)s, r);
        uint64_t u53 = (((uint64_t)r[1] << 32) | r[0]) >> 11;
        uint64_t t = (uint64_t)(((unsigned __int128)u53 * total) >> 53);
        /* first index with cdf > t */
        int64_t lo = 0, hi = N - 1;
        while (lo < hi) {
            int64_t mid = lo + ((hi - lo) >> 1);
            if (cdf[mid] > t) hi = mid; else lo = mid + 1;
        }
        out[s] = (uint64_t)lo;
    }
    qsort(out, shots, sizeof(uint64_t), cmp_u64);
    return 0;
}

int orc_sample(int n, const double complex *psi, uint64_t shots,
               uint64_t seed, uint64_t *out) {
    int64_t N = (int64_t)1 << n;
    uint64_t *cdf = (uint64_t *)malloc(sizeof(uint64_t) * N);
    if (!cdf) return -2;
    uint64_t run = 0;
    for (int64_t i = 0; i < N; i++) {
        double re = creal(psi[i]), im = cimag(psi[i]);
        double p = re * re + im * im;
        run += (uint64_t)(p * ORC_FIXED_SCALE);
        cdf[i] = run;
    }
    int rc = sample_from_cdf(N, cdf, shots, seed, out);
    free(cdf);
    return rc;
}

int orc_sample_c64(int n, const float complex *psi, uint64_t shots,
                   uint64_t seed, uint64_t *out) {
    int64_t N = (int64_t)1 << n;
    uint64_t *cdf = (uint64_t *)malloc(sizeof(uint64_t) * N);
    if (!cdf) return -2;
    uint64_t run = 0;
    for (int64_t i = 0; i < N; i++) {
        double re = (double)crealf(psi[i]), im = (double)cimagf(psi[i]);
        double p = re * re + im * im;
        run += (uint64_t)(p * ORC_FIXED_SCALE);
        cdf[i] = run;
    }
    int rc = sample_from_cdf(N, cdf, shots, seed, out);
    free(cdf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Adjoint Jacobian [AdjointDiffGPU.hpp:499-596]                            */
/* ------------------------------------------------------------------------ */

static int is_state_prep(const char *name) {
    return !strcmp(name, "QubitStateVector") || !strcmp(name, "StatePrep") ||
           !strcmp(name, "BasisState");
}

int orc_adjoint_jacobian(int n, const double complex *psi, int nobs,
                         const double complex *const *bras, int nops,
                         const char *const *names, const int *wires_flat,
                         const int *woff, const int *inverses,
                         const double *params, const int *nparams,
                         const double complex *const *matrices,
                         const int64_t *trainable, int ntp, double *jac) {
    if (ntp <= 0) return -1;
    int64_t N = (int64_t)1 << n;
    size_t bytes = sizeof(dc) * (size_t)N;
    int num_param_ops = 0;
    for (int i = 0; i < nops; i++) num_param_ops += nparams[i] > 0;

    dc *lambda = (dc *)malloc(bytes), *mu = (dc *)malloc(bytes);
    dc **hl = (dc **)calloc((size_t)nobs, sizeof(dc *));
    int rc = 0;
    if (!lambda || !mu || !hl) { rc = -2; goto done; }
    memcpy(lambda, psi, bytes);
    for (int j = 0; j < nobs; j++) {
        hl[j] = (dc *)malloc(bytes);
        if (!hl[j]) { rc = -2; goto done; }
        memcpy(hl[j], bras[j], bytes);
    }
    for (int j = 0; j < nobs * ntp; j++) jac[j] = 0;

    int tp_pos = ntp - 1;              /* walks trainable[] backwards */
    int param_idx = num_param_ops - 1; /* index among parametrised ops */
    for (int op = nops - 1; op >= 0; op--) {
        if (nparams[op] > 1) { rc = -3; goto done; }
        if (is_state_prep(names[op])) continue;
        if (tp_pos < 0) break;
        const int *w = wires_flat + woff[op];
        int nw = woff[op + 1] - woff[op];
        const dc *mat = matrices ? matrices[op] : NULL;
        memcpy(mu, lambda, bytes);
        /* lambda <- U^dagger lambda */
        rc = orc_apply_op(n, lambda, names[op], w, nw, !inverses[op], params + 3 * op, mat);
        if (rc) goto done;
        if (nparams[op] > 0) {
            if (param_idx == trainable[tp_pos]) {
                double s;
                dc G[1 << 8];
                if (nw > 4 && strcmp(names[op], "MultiRZ")) { rc = -1; goto done; }
                if (!strcmp(names[op], "MultiRZ")) {
                    /* G = Z (x) ... (x) Z, diagonal by parity */
                    uint64_t mask = 0;
                    for (int j = 0; j < nw; j++) mask |= 1ULL << (n - 1 - w[j]);
                    for (int64_t i = 0; i < N; i++)
                        if (__builtin_popcountll((uint64_t)i & mask) & 1) mu[i] = -mu[i];
                    s = -0.5;
                } else {
                    int k = orc_generator(names[op], nw, G, &s);
                    if (k < 0) { rc = -1; goto done; }
                    rc = apply_dense(n, mu, w, k, G);
                    if (rc) goto done;
                }
                double sf = s * (inverses[op] ? -1.0 : 1.0);
                for (int j = 0; j < nobs; j++) {
                    double im = 0;
#pragma omp parallel for reduction(+ : im) schedule(static) if (N > 4096)
                    for (int64_t i = 0; i < N; i++) im += cimag(conj(hl[j][i]) * mu[i]);
                    jac[j * ntp + tp_pos] = -2.0 * sf * im;
                }
                tp_pos--;
            }
            param_idx--;
        }
        for (int j = 0; j < nobs; j++) {
            rc = orc_apply_op(n, hl[j], names[op], w, nw, !inverses[op], params + 3 * op, mat);
            if (rc) goto done;
        }
    }
done:
    if (hl)
        for (int j = 0; j < nobs; j++) free(hl[j]);
    free(hl);
    free(lambda);
    free(mu);
    return rc;
}
