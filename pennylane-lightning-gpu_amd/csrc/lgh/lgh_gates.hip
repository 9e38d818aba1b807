// Gate application: operation compiler (named gate -> kernel shape) and the
// gate kernels.
//
// Reference counterparts: StateVectorCudaManaged::applyOperation and the
// named gate methods (pennylane_lightning_gpu/src/simulator/
// StateVectorCudaManaged.hpp:198-560), applyParametricPauliGate (:1339-1386),
// applyDeviceMatrixGate (:1400-1472), the gate matrices of cuGates_host.hpp
// and the generators of GateGenerators.hpp:58-320.  The reference hands all
// of these to cuStateVec; here every gate is reduced to one of two kernel
// shapes (a dense 2^k x 2^k block on k target bits with optional controls,
// or a diagonal phase by control mask and parity mask) and applied by the
// HIP kernels below in one HBM pass.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lgh_fused.hpp"
#include "lgh_internal.hpp"

namespace lgh {

using C = std::complex<double>;
static const C kI(0, 1);

namespace {

struct GateInfo {
    int wires;    // fixed arity, 0 = variable
    int controls; // leading control wires
    int params;
};

GateInfo gate_info(int g) {
    switch (g) {
    case LGH_GATE_IDENTITY: return {0, 0, 0};
    case LGH_GATE_PAULIX: case LGH_GATE_PAULIY: case LGH_GATE_PAULIZ:
    case LGH_GATE_HADAMARD: case LGH_GATE_S: case LGH_GATE_T: return {1, 0, 0};
    case LGH_GATE_CNOT: case LGH_GATE_CY: case LGH_GATE_CZ: return {2, 1, 0};
    case LGH_GATE_SWAP: return {2, 0, 0};
    case LGH_GATE_CSWAP: return {3, 1, 0};
    case LGH_GATE_TOFFOLI: return {3, 2, 0};
    case LGH_GATE_PHASESHIFT: case LGH_GATE_RX: case LGH_GATE_RY: case LGH_GATE_RZ: return {1, 0, 1};
    case LGH_GATE_ROT: return {1, 0, 3};
    case LGH_GATE_CONTROLLEDPHASESHIFT: case LGH_GATE_CRX: case LGH_GATE_CRY: case LGH_GATE_CRZ:
        return {2, 1, 1};
    case LGH_GATE_CROT: return {2, 1, 3};
    case LGH_GATE_ISINGXX: case LGH_GATE_ISINGYY: case LGH_GATE_ISINGZZ:
    case LGH_GATE_SINGLEEXCITATION: case LGH_GATE_SINGLEEXCITATIONMINUS:
    case LGH_GATE_SINGLEEXCITATIONPLUS: return {2, 0, 1};
    case LGH_GATE_DOUBLEEXCITATION: case LGH_GATE_DOUBLEEXCITATIONMINUS:
    case LGH_GATE_DOUBLEEXCITATIONPLUS: return {4, 0, 1};
    case LGH_GATE_MULTIRZ: return {0, 0, 1};
    case LGH_GATE_MATRIX: return {0, -1, 0};
    default: throw Error(LGH_ERR_UNSUPPORTED, "unknown gate id " + std::to_string(g));
    }
}

using Mat = std::vector<C>;

Mat eye(int dim) {
    Mat m(size_t(dim) * dim, C(0));
    for (int i = 0; i < dim; ++i) m[size_t(i) * dim + i] = 1;
    return m;
}

Mat dagger_of(const Mat &m, int dim) {
    Mat d(m.size());
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) d[size_t(c) * dim + r] = std::conj(m[size_t(r) * dim + c]);
    return d;
}

Mat kron2(const Mat &a, const Mat &b, int da, int db) {
    Mat out(size_t(da * db) * (da * db));
    for (int r1 = 0; r1 < da; ++r1)
        for (int c1 = 0; c1 < da; ++c1)
            for (int r2 = 0; r2 < db; ++r2)
                for (int c2 = 0; c2 < db; ++c2)
                    out[size_t(r1 * db + r2) * (da * db) + c1 * db + c2] = a[r1 * da + c1] * b[r2 * db + c2];
    return out;
}

const Mat kX = {0, 1, 1, 0};
const Mat kY = {0, -kI, kI, 0};
const Mat kZ = {1, 0, 0, -1};

Mat rot(double phi, double theta, double omega) {
    // RZ(omega) RY(theta) RZ(phi)
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    return {std::exp(-kI * ((phi + omega) / 2)) * c, -std::exp(kI * ((phi - omega) / 2)) * s,
            std::exp(-kI * ((phi - omega) / 2)) * s, std::exp(kI * ((phi + omega) / 2)) * c};
}

// Target-block matrix of a named gate (controls excluded), forward direction.
Mat target_matrix(int g, const double *p, int ntargets) {
    const double h = p[0] / 2;
    const double c = std::cos(h), s = std::sin(h);
    switch (g) {
    case LGH_GATE_PAULIX: case LGH_GATE_CNOT: case LGH_GATE_TOFFOLI: return kX;
    case LGH_GATE_PAULIY: case LGH_GATE_CY: return kY;
    case LGH_GATE_PAULIZ: case LGH_GATE_CZ: return kZ;
    case LGH_GATE_HADAMARD: {
        const double r = 1.0 / std::sqrt(2.0);
        return {r, r, r, -r};
    }
    case LGH_GATE_S: return {1, 0, 0, kI};
    case LGH_GATE_T: return {1, 0, 0, std::exp(kI * (M_PI / 4))};
    case LGH_GATE_PHASESHIFT: case LGH_GATE_CONTROLLEDPHASESHIFT: return {1, 0, 0, std::exp(kI * p[0])};
    case LGH_GATE_RX: case LGH_GATE_CRX: return {c, -kI * s, -kI * s, c};
    case LGH_GATE_RY: case LGH_GATE_CRY: return {c, -s, s, c};
    case LGH_GATE_RZ: case LGH_GATE_CRZ: return {std::exp(-kI * h), 0, 0, std::exp(kI * h)};
    case LGH_GATE_ROT: case LGH_GATE_CROT: return rot(p[0], p[1], p[2]);
    case LGH_GATE_SWAP: case LGH_GATE_CSWAP: return {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
    case LGH_GATE_ISINGXX: {
        Mat m = eye(4);
        for (int i = 0; i < 4; ++i) {
            m[i * 4 + i] = c;
            m[i * 4 + 3 - i] = -kI * s;
        }
        return m;
    }
    case LGH_GATE_ISINGYY: {
        Mat yy = kron2(kY, kY, 2, 2), m = eye(4);
        for (int i = 0; i < 16; ++i) m[i] = m[i] * c - kI * s * yy[i];
        return m;
    }
    case LGH_GATE_ISINGZZ: {
        Mat m(16, C(0));
        m[0] = m[15] = std::exp(-kI * h);
        m[5] = m[10] = std::exp(kI * h);
        return m;
    }
    case LGH_GATE_SINGLEEXCITATION: case LGH_GATE_SINGLEEXCITATIONMINUS:
    case LGH_GATE_SINGLEEXCITATIONPLUS: {
        C d = g == LGH_GATE_SINGLEEXCITATION ? C(1) : std::exp((g == LGH_GATE_SINGLEEXCITATIONMINUS ? -kI : kI) * h);
        Mat m(16, C(0));
        m[0] = m[15] = d;
        m[5] = c; m[6] = -s; m[9] = s; m[10] = c;
        return m;
    }
    case LGH_GATE_DOUBLEEXCITATION: case LGH_GATE_DOUBLEEXCITATIONMINUS:
    case LGH_GATE_DOUBLEEXCITATIONPLUS: {
        C d = g == LGH_GATE_DOUBLEEXCITATION ? C(1) : std::exp((g == LGH_GATE_DOUBLEEXCITATIONMINUS ? -kI : kI) * h);
        Mat m(256, C(0));
        for (int i = 0; i < 16; ++i) m[i * 16 + i] = d;
        m[3 * 16 + 3] = c; m[3 * 16 + 12] = -s;
        m[12 * 16 + 3] = s; m[12 * 16 + 12] = c;
        return m;
    }
    case LGH_GATE_MULTIRZ: {
        const int dim = 1 << ntargets;
        Mat m(size_t(dim) * dim, C(0));
        for (int i = 0; i < dim; ++i) m[size_t(i) * dim + i] = std::exp((__builtin_popcount(i) & 1 ? kI : -kI) * h);
        return m;
    }
    case LGH_GATE_IDENTITY: return eye(1 << ntargets);
    default: throw Error(LGH_ERR_UNSUPPORTED, "no matrix for gate id " + std::to_string(g));
    }
}

int num_controls(const lgh_op &op) {
    GateInfo gi = gate_info(op.gate);
    return gi.controls < 0 ? op.num_ctrl : gi.controls;
}

uint64_t bit_of(int n, int wire) { return 1ull << (n - 1 - wire); }

} // namespace

void validate_op(const lgh_op &op, int n) {
    GateInfo gi = gate_info(op.gate);
    LGH_REQUIRE(op.num_wires >= 0, "negative wire count");
    LGH_REQUIRE(op.num_wires == 0 || op.wires, "null wire list");
    if (gi.wires > 0)
        LGH_REQUIRE(op.num_wires == gi.wires, "gate " + std::to_string(op.gate) + " expects " +
                                                   std::to_string(gi.wires) + " wires");
    if (op.gate == LGH_GATE_MULTIRZ) LGH_REQUIRE(op.num_wires >= 1, "MultiRZ needs at least one wire");
    if (op.gate == LGH_GATE_MATRIX) {
        LGH_REQUIRE(op.matrix, "matrix gate without a matrix");
        LGH_REQUIRE(op.num_ctrl >= 0 && op.num_ctrl < op.num_wires, "bad control count for matrix gate");
        LGH_REQUIRE(op.num_wires - op.num_ctrl <= kMaxLdsK, "matrix gates act on at most 13 target wires");
    }
    for (int i = 0; i < op.num_wires; ++i) {
        LGH_REQUIRE(op.wires[i] >= 0 && op.wires[i] < n, "wire index out of range");
        for (int j = 0; j < i; ++j) LGH_REQUIRE(op.wires[i] != op.wires[j], "repeated wire in one operation");
    }
}

std::vector<std::complex<double>> full_matrix(const lgh_op &op, bool dagger) {
    const int nc = num_controls(op);
    const int kt = op.num_wires - nc;
    const int dt = 1 << kt, dim = 1 << op.num_wires;
    Mat t;
    if (op.gate == LGH_GATE_MATRIX) {
        t.resize(size_t(dt) * dt);
        for (size_t i = 0; i < t.size(); ++i) t[i] = C(op.matrix[2 * i], op.matrix[2 * i + 1]);
    } else {
        t = target_matrix(op.gate, op.params, kt);
    }
    Mat m = eye(dim);
    const int off = dim - dt; // all controls set: bottom-right block
    for (int r = 0; r < dt; ++r)
        for (int c = 0; c < dt; ++c) m[size_t(off + r) * dim + off + c] = t[size_t(r) * dt + c];
    const bool inv = (op.inverse != 0) != dagger;
    return inv ? dagger_of(m, dim) : m;
}

KOp compile_op(const lgh_op &op, int n, bool dagger) {
    validate_op(op, n);
    KOp k;
    const bool inv = (op.inverse != 0) != dagger;
    const int nc = num_controls(op);
    const int kt = op.num_wires - nc;
    auto conj_if = [&](C z) { return inv ? std::conj(z) : z; };
    for (int i = 0; i < nc; ++i) {
        k.cbits.push_back(n - 1 - op.wires[i]);
        k.cmask |= bit_of(n, op.wires[i]);
    }
    const double h = op.params[0] / 2;
    switch (op.gate) {
    case LGH_GATE_IDENTITY: k.kind = KOp::NOP; return k;
    // diagonal gates: one pass over the amplitudes they change
    case LGH_GATE_PAULIZ: case LGH_GATE_S: case LGH_GATE_T: case LGH_GATE_PHASESHIFT:
    case LGH_GATE_CZ: case LGH_GATE_CONTROLLEDPHASESHIFT: {
        const int tw = op.wires[op.num_wires - 1];
        k.kind = KOp::DIAG;
        k.cbits.push_back(n - 1 - tw);
        k.cmask |= bit_of(n, tw);
        C ph = op.gate == LGH_GATE_PAULIZ || op.gate == LGH_GATE_CZ ? C(-1)
             : op.gate == LGH_GATE_S ? kI
             : op.gate == LGH_GATE_T ? std::exp(kI * (M_PI / 4))
                                     : std::exp(kI * op.params[0]);
        k.d_even = k.d_odd = conj_if(ph);
        return k;
    }
    case LGH_GATE_RZ: case LGH_GATE_CRZ: case LGH_GATE_ISINGZZ: case LGH_GATE_MULTIRZ: {
        k.kind = KOp::DIAG;
        for (int i = nc; i < op.num_wires; ++i) k.pmask |= bit_of(n, op.wires[i]);
        k.d_even = conj_if(std::exp(-kI * h));
        k.d_odd = conj_if(std::exp(kI * h));
        return k;
    }
    default: break;
    }
    k.kind = KOp::DENSE;
    k.k = kt;
    LGH_REQUIRE(kt >= 1 && kt <= kMaxLdsK, "dense gate target count out of range");
    for (int i = 0; i < kt; ++i) k.tbits[i] = n - 1 - op.wires[nc + i];
    const int dt = 1 << kt;
    Mat t;
    if (op.gate == LGH_GATE_MATRIX) {
        t.resize(size_t(dt) * dt);
        for (size_t i = 0; i < t.size(); ++i) t[i] = C(op.matrix[2 * i], op.matrix[2 * i + 1]);
    } else {
        t = target_matrix(op.gate, op.params, kt);
    }
    k.mat = inv ? dagger_of(t, dt) : t;
    return k;
}

KGen compile_generator(const lgh_op &op, int n, const KOp &shape) {
    (void)n;
    KGen g;
    g.present = true;
    g.scale = -0.5;
    switch (op.gate) {
    case LGH_GATE_RX: case LGH_GATE_CRX: g.mat = kX; break;
    case LGH_GATE_RY: case LGH_GATE_CRY: g.mat = kY; break;
    case LGH_GATE_RZ: case LGH_GATE_CRZ: case LGH_GATE_ISINGZZ: case LGH_GATE_MULTIRZ:
        g.g_even = 1;
        g.g_odd = -1;
        break;
    case LGH_GATE_PHASESHIFT: case LGH_GATE_CONTROLLEDPHASESHIFT:
        // P_1 (P_11): the DIAG shape already restricts to amplitudes with the bit(s) set
        g.g_even = g.g_odd = 1;
        g.scale = 1.0;
        break;
    case LGH_GATE_ISINGXX: g.mat = kron2(kX, kX, 2, 2); break;
    case LGH_GATE_ISINGYY: g.mat = kron2(kY, kY, 2, 2); break;
    case LGH_GATE_SINGLEEXCITATION: case LGH_GATE_SINGLEEXCITATIONMINUS:
    case LGH_GATE_SINGLEEXCITATIONPLUS: {
        C d = op.gate == LGH_GATE_SINGLEEXCITATION ? C(0) : C(op.gate == LGH_GATE_SINGLEEXCITATIONMINUS ? 1 : -1);
        g.mat.assign(16, C(0));
        g.mat[0] = g.mat[15] = d;
        g.mat[1 * 4 + 2] = -kI;
        g.mat[2 * 4 + 1] = kI;
        break;
    }
    case LGH_GATE_DOUBLEEXCITATION: case LGH_GATE_DOUBLEEXCITATIONMINUS:
    case LGH_GATE_DOUBLEEXCITATIONPLUS: {
        C d = op.gate == LGH_GATE_DOUBLEEXCITATION ? C(0) : C(op.gate == LGH_GATE_DOUBLEEXCITATIONMINUS ? 1 : -1);
        g.mat.assign(256, C(0));
        for (int i = 0; i < 16; ++i)
            if (i != 3 && i != 12) g.mat[i * 16 + i] = d;
        g.mat[3 * 16 + 12] = -kI;
        g.mat[12 * 16 + 3] = kI;
        break;
    }
    default:
        throw Error(LGH_ERR_UNSUPPORTED, "operation " + std::to_string(op.gate) +
                                             " is not supported by the adjoint differentiation method");
    }
    const bool dense_gen = !g.mat.empty();
    LGH_REQUIRE(dense_gen == (shape.kind == KOp::DENSE), "generator shape mismatch");
    return g;
}

} // namespace lgh

using namespace lgh;

// ---------------------------------------------------------------- kernels

namespace {

constexpr int kBlock = 256;

// 1-target dense gate with optional controls: one amplitude pair per thread,
// consecutive lanes on consecutive pair indices (for a target bit >= 6 each
// wavefront access is a contiguous 1 KiB for complex128).
template <class R>
__global__ __launch_bounds__(kBlock) void k_gate1_pair(cx<R> *__restrict__ sv, uint64_t ngroups, Fixed fx,
                                                       uint64_t cmask, uint64_t tbit, cx<R> m00, cx<R> m01,
                                                       cx<R> m10, cx<R> m11, uint64_t xcd_run) {
    const uint64_t g = xcd_block_runs(xcd_run) * kBlock + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t i0 = expand(g, fx) | cmask;
    const cx<R> a = ldnt(sv + i0), b = ldnt(sv + (i0 | tbit));
    stnt(sv + i0, cfma(m01, b, cmul(m00, a)));
    stnt(sv + (i0 | tbit), cfma(m11, b, cmul(m10, a)));
}

// 1-target dense gate without controls on target bit t < 6: the partner of
// amplitude i sits in lane (lane ^ 2^t) of the same wavefront, so each thread
// streams one amplitude and fetches its partner with a cross-lane shuffle.
template <class R>
__global__ __launch_bounds__(kBlock) void k_gate1_lane(cx<R> *__restrict__ sv, uint64_t len, int tmask, cx<R> m00,
                                                       cx<R> m01, cx<R> m10, cx<R> m11) {
    const uint64_t i = xcd_block() * kBlock + threadIdx.x;
    const bool hi = (threadIdx.x & tmask) != 0;
    const cx<R> cs = hi ? m11 : m00, cp = hi ? m10 : m01;
    const cx<R> v = i < len ? ldnt(sv + i) : cx<R>{R(0), R(0)};
    const cx<R> p = shfl_xor_cx(v, tmask);
    if (i < len) stnt(sv + i, cfma(cp, p, cmul(cs, v)));
}

// Workgroup order of the 1-target pair kernel (xcd_block_runs).  Measured on
// MI355X with 30-qubit complex128 states (one RX per target, in place): with
// each XCD sweeping its own eighth of the pairs, targets whose pair distance
// is byte-address bit 25, 29 or 33 stream at 5.7-5.8 TB/s; with all XCDs in
// one window (runs of 64 workgroups) they reach 6.3 TB/s, as do distances up
// to bit 16 (+2 %).  Other distances are 2-4 % faster with the eighths.
// Byte bits 24, 28 and 32 stay at 5.5-5.7 TB/s under every order tried
// (runs of 1-2048 workgroups, index swizzles, several pairs per thread,
// half-pairs across lanes).
uint64_t pair_xcd_run(int n, int tbit, size_t amp_bytes) {
    if (n < 27) return 0;
    int byte_bit = tbit;
    for (size_t b = amp_bytes; b > 1; b >>= 1) ++byte_bit;
    const bool one_window = (byte_bit >= 25 && byte_bit % 4 == 1) || byte_bit <= 16;
    return one_window ? 64 : 0;
}

template <int K> struct Offs {
    uint64_t o[1 << K];
};

// K-target dense gate (2 <= K <= 4) with the matrix staged in LDS and the
// 2^K amplitudes of a group held in registers.
template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_denseK(cx<R> *__restrict__ sv, uint64_t ngroups, Fixed fx,
                                                   uint64_t cmask, Offs<K> offs, const cx<R> *__restrict__ mat) {
    constexpr int D = 1 << K;
    __shared__ cx<R> sm[D * D];
    for (int t = threadIdx.x; t < D * D; t += kBlock) sm[t] = mat[t];
    __syncthreads();
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t base = expand(g, fx) | cmask;
    cx<R> v[D];
#pragma unroll
    for (int m = 0; m < D; ++m) v[m] = ldnt(sv + (base | offs.o[m]));
#pragma unroll
    for (int r = 0; r < D; ++r) {
        cx<R> acc{R(0), R(0)};
#pragma unroll
        for (int c = 0; c < D; ++c) acc = cfma(sm[r * D + c], v[c], acc);
        stnt(sv + (base | offs.o[r]), acc);
    }
}

// Large dense gates (5 <= K <= 13): one workgroup per group; the 2^K
// amplitudes are staged in LDS and the transposed matrix is streamed from
// global memory (coalesced across rows).
struct Tbits {
    int k;
    int b[kMaxLdsK];
};

__device__ __forceinline__ uint64_t deposit(uint32_t m, const Tbits &tb) {
    uint64_t o = 0;
    for (int j = 0; j < tb.k; ++j)
        if ((m >> (tb.k - 1 - j)) & 1u) o |= 1ull << tb.b[j];
    return o;
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_denseLDS(cx<R> *__restrict__ sv, Fixed fx, uint64_t cmask, Tbits tb,
                                                     const cx<R> *__restrict__ matT) {
    extern __shared__ unsigned char lds_raw[];
    cx<R> *x = reinterpret_cast<cx<R> *>(lds_raw);
    const uint32_t D = 1u << tb.k;
    const uint64_t base = expand(blockIdx.x, fx) | cmask;
    for (uint32_t m = threadIdx.x; m < D; m += kBlock) x[m] = sv[base | deposit(m, tb)];
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < D; r += kBlock) {
        cx<R> acc{R(0), R(0)};
        for (uint32_t c = 0; c < D; ++c) acc = cfma(matT[uint64_t(c) * D + r], x[c], acc);
        sv[base | deposit(r, tb)] = acc;
    }
}

// Diagonal phase: amplitudes with every cmask bit set are multiplied by
// d_even / d_odd according to the parity of (index & pmask).
template <class R>
__global__ __launch_bounds__(kBlock) void k_diag(cx<R> *__restrict__ sv, uint64_t count, Fixed fx, uint64_t cmask,
                                                 uint64_t pmask, cx<R> de, cx<R> dodd) {
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= count) return;
    const uint64_t i = expand(g, fx) | cmask;
    const bool odd = __popcll(i & pmask) & 1;
    stnt(sv + i, cmul(ldnt(sv + i), odd ? dodd : de));
}

template <class R> void launch_kop_t(lgh_sv *sv, const KOp &op) {
    StreamCtx *ctx = sv->ctx;
    hipStream_t st = ctx->stream;
    cx<R> *data = static_cast<cx<R> *>(sv->data);
    const int n = sv->n;
    if (op.kind == KOp::NOP) return;
    if (op.kind == KOp::DIAG) {
        Fixed fx = make_fixed(op.cbits);
        const uint64_t count = 1ull << (n - fx.cnt);
        const uint64_t blocks = (count + kBlock - 1) / kBlock;
        hipLaunchKernelGGL((k_diag<R>), dim3(unsigned(blocks)), dim3(kBlock), 0, st, data, count, fx, op.cmask,
                           op.pmask, to_cx<R>(op.d_even), to_cx<R>(op.d_odd));
        LGH_LAUNCH_CHECK();
        return;
    }
    std::vector<int> fixed_bits = op.cbits;
    for (int i = 0; i < op.k; ++i) fixed_bits.push_back(op.tbits[i]);
    Fixed fx = make_fixed(fixed_bits);
    const uint64_t ngroups = 1ull << (n - fx.cnt);
    const int D = 1 << op.k;
    if (op.k == 1) {
        const cx<R> m00 = to_cx<R>(op.mat[0]), m01 = to_cx<R>(op.mat[1]), m10 = to_cx<R>(op.mat[2]),
                    m11 = to_cx<R>(op.mat[3]);
        if (op.cmask == 0 && op.tbits[0] < 6 && n >= 6) {
            const uint64_t len = 1ull << n;
            hipLaunchKernelGGL((k_gate1_lane<R>), dim3(unsigned((len + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                               data, len, 1 << op.tbits[0], m00, m01, m10, m11);
        } else {
            const uint64_t blocks = (ngroups + kBlock - 1) / kBlock;
            hipLaunchKernelGGL((k_gate1_pair<R>), dim3(unsigned(blocks)), dim3(kBlock), 0, st, data, ngroups, fx,
                               op.cmask, 1ull << op.tbits[0], m00, m01, m10, m11,
                               pair_xcd_run(n, op.tbits[0], sizeof(cx<R>)));
        }
        LGH_LAUNCH_CHECK();
        return;
    }
    if (op.k <= kMaxRegK) {
        std::vector<cx<R>> m(size_t(D) * D);
        for (size_t i = 0; i < m.size(); ++i) m[i] = to_cx<R>(op.mat[i]);
        const cx<R> *dm = static_cast<const cx<R> *>(ctx->upload(m.data(), m.size() * sizeof(cx<R>)));
        const uint64_t blocks = (ngroups + kBlock - 1) / kBlock;
        auto fill = [&](auto &offs, int k) {
            for (int mi = 0; mi < (1 << k); ++mi) {
                uint64_t o = 0;
                for (int j = 0; j < k; ++j)
                    if ((mi >> (k - 1 - j)) & 1) o |= 1ull << op.tbits[j];
                offs.o[mi] = o;
            }
        };
        if (op.k == 2) {
            Offs<2> offs;
            fill(offs, 2);
            hipLaunchKernelGGL((k_denseK<R, 2>), dim3(unsigned(blocks)), dim3(kBlock), 0, st, data, ngroups, fx,
                               op.cmask, offs, dm);
        } else if (op.k == 3) {
            Offs<3> offs;
            fill(offs, 3);
            hipLaunchKernelGGL((k_denseK<R, 3>), dim3(unsigned(blocks)), dim3(kBlock), 0, st, data, ngroups, fx,
                               op.cmask, offs, dm);
        } else {
            Offs<4> offs;
            fill(offs, 4);
            hipLaunchKernelGGL((k_denseK<R, 4>), dim3(unsigned(blocks)), dim3(kBlock), 0, st, data, ngroups, fx,
                               op.cmask, offs, dm);
        }
        LGH_LAUNCH_CHECK();
        return;
    }
    // 5..13 targets: transposed matrix for coalesced row streaming
    std::vector<cx<R>> mt(size_t(D) * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) mt[size_t(c) * D + r] = to_cx<R>(op.mat[size_t(r) * D + c]);
    const size_t mbytes = mt.size() * sizeof(cx<R>);
    void *dm;
    if (mbytes <= (1u << 20)) {
        dm = ctx->upload(mt.data(), mbytes);
    } else {
        dm = ctx->workspace(mbytes);
        LGH_HIP(hipMemcpyAsync(dm, mt.data(), mbytes, hipMemcpyHostToDevice, st));
        ctx->sync(); // mt is a host temporary
    }
    Tbits tb{};
    tb.k = op.k;
    for (int j = 0; j < op.k; ++j) tb.b[j] = op.tbits[j];
    const size_t lds = size_t(D) * sizeof(cx<R>);
    LGH_REQUIRE(lds <= 160 * 1024, "matrix gate too large for one workgroup");
    hipLaunchKernelGGL((k_denseLDS<R>), dim3(unsigned(ngroups)), dim3(kBlock), lds, st, data, fx, op.cmask, tb,
                       static_cast<const cx<R> *>(dm));
    LGH_LAUNCH_CHECK();
}

} // namespace

namespace lgh {
void launch_kop(lgh_sv *sv, const KOp &op) {
    if (sv->dtype == LGH_C128)
        launch_kop_t<double>(sv, op);
    else
        launch_kop_t<float>(sv, op);
}
} // namespace lgh

// ---------------------------------------------------------------- C ABI

extern "C" {

lgh_status lgh_apply(lgh_sv *sv, const lgh_op *ops, int64_t num_ops) {
    return guard([&] {
        Range rg("lgh_apply");
        LGH_REQUIRE(sv, "null state");
        LGH_REQUIRE(num_ops == 0 || ops, "null op list");
        // validate everything before launching anything
        for (int64_t i = 0; i < num_ops; ++i) validate_op(ops[i], sv->n);
        DeviceGuard g(sv->device);
        if (num_ops >= 2 && fusion_enabled()) {
            std::vector<KOp> kops;
            kops.reserve(size_t(num_ops));
            for (int64_t i = 0; i < num_ops; ++i) kops.push_back(compile_op(ops[i], sv->n, false));
            fused_apply(sv, kops);
            return;
        }
        for (int64_t i = 0; i < num_ops; ++i) launch_kop(sv, compile_op(ops[i], sv->n, false));
    });
}

lgh_status lgh_apply_generator(lgh_sv *sv, const lgh_op *op, double *scale) {
    return guard([&] {
        LGH_REQUIRE(sv && op, "null pointer");
        validate_op(*op, sv->n);
        KOp shape = compile_op(*op, sv->n, false);
        KGen gen = compile_generator(*op, sv->n, shape);
        if (scale) *scale = gen.scale;
        DeviceGuard g(sv->device);
        KOp k;
        if (op->gate == LGH_GATE_MULTIRZ || op->gate == LGH_GATE_ISINGZZ || op->gate == LGH_GATE_RZ) {
            // Z...Z: diagonal by parity over all op wires
            k.kind = KOp::DIAG;
            k.pmask = shape.pmask;
            k.d_even = 1;
            k.d_odd = -1;
        } else {
            // generator as a dense operator on all op wires (controls included)
            const int nw = op->num_wires;
            const int dim = 1 << nw;
            std::vector<std::complex<double>> full(size_t(dim) * dim, 0);
            if (shape.kind == KOp::DENSE) {
                const int dt = 1 << shape.k, off = dim - dt;
                for (int r = 0; r < dt; ++r)
                    for (int c = 0; c < dt; ++c) full[size_t(off + r) * dim + off + c] = gen.mat[size_t(r) * dt + c];
            } else { // diagonal generator: eigenvalue by control mask and parity
                for (int m = 0; m < dim; ++m) {
                    uint64_t idx = 0;
                    for (int i = 0; i < nw; ++i)
                        if ((m >> (nw - 1 - i)) & 1) idx |= 1ull << (sv->n - 1 - op->wires[i]);
                    double v = 0;
                    if ((idx & shape.cmask) == shape.cmask)
                        v = (__builtin_popcountll(idx & shape.pmask) & 1) ? gen.g_odd : gen.g_even;
                    full[size_t(m) * dim + m] = v;
                }
            }
            k.kind = KOp::DENSE;
            k.k = nw;
            for (int i = 0; i < nw; ++i) k.tbits[i] = sv->n - 1 - op->wires[i];
            k.mat = full;
        }
        launch_kop(sv, k);
    });
}

} // extern "C"

extern "C" lgh_status lgh_gate_matrix(const lgh_op *op, double *out) {
    return guard([&] {
        LGH_REQUIRE(op && out, "null pointer");
        int max_wire = 0;
        for (int i = 0; i < op->num_wires; ++i) max_wire = std::max(max_wire, op->wires ? op->wires[i] + 1 : 0);
        validate_op(*op, std::max(max_wire, op->num_wires));
        LGH_REQUIRE(op->num_wires <= 10, "matrix too large");
        const auto m = full_matrix(*op, false);
        for (size_t i = 0; i < m.size(); ++i) {
            out[2 * i] = m[i].real();
            out[2 * i + 1] = m[i].imag();
        }
    });
}
