// Adjoint Jacobian (arXiv:2009.02823) with the generator product and the
// overlap fused into the reverse step.
//
// Reference: AdjointJacobianGPU::adjointJacobian
// (pennylane_lightning_gpu/src/algorithms/AdjointDiffGPU.hpp:499-596), which
// per trainable op copies lambda into mu, applies U^dagger to lambda, applies
// the generator to mu, takes one cuBLAS dot per observable followed by a
// blocking device-to-host copy, and applies U^dagger to every observable
// state.  Here one kernel pass per trainable op reads lambda and every
// H|lambda_j>, accumulates Im<H lambda_j | G lambda> without materialising
// mu, and writes U^dagger lambda and U^dagger H|lambda_j>; the Jacobian stays
// on the device until one copy at the end.
//
//   jac[j][t] = -2 * s * (inverse ? -1 : 1) * Im<H lambda_j | G lambda>
//   (AdjointDiffGPU.hpp:132-162, :578-582; s from :96-114)
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "lgh_fused.hpp"
#include "lgh_internal.hpp"

using namespace lgh;

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kBras = 8; // observable states per launch

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = lane < kWaves ? sh[k][lane] : 0.0;
            v[k] = wave_sum(s);
        }
    }
}

template <class R> __device__ __forceinline__ cx<double> up(cx<R> a) { return cx<double>{double(a.re), double(a.im)}; }

// Im(conj(b) * q)
template <class R> __device__ __forceinline__ double im_dot(cx<R> b, cx<double> q) {
    return double(b.re) * q.im - double(b.im) * q.re;
}

// Dense generator step for one 1-target gate with optional controls: one
// amplitude pair per thread.  U^dagger in the state precision, G in double.
template <class R, int NB>
__global__ __launch_bounds__(kBlock) void k_adj_pair1(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                      int upd_psi, uint64_t ngroups, Fixed fx, uint64_t cmask,
                                                      uint64_t tbit, cx<R> u00, cx<R> u01, cx<R> u10, cx<R> u11,
                                                      int with_gen, cx<double> g00, cx<double> g01, cx<double> g10,
                                                      cx<double> g11, double *__restrict__ partials) {
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = 0.0;
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g < ngroups) {
        const uint64_t i0 = expand(g, fx) | cmask, i1 = i0 | tbit;
        const cx<R> a0 = ldnt(psi + i0), a1 = ldnt(psi + i1);
        const cx<double> q0 = cfma(g01, up(a1), cmul(g00, up(a0)));
        const cx<double> q1 = cfma(g11, up(a1), cmul(g10, up(a0)));
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                cx<R> *bj = bras[j];
                const cx<R> b0 = ldnt(bj + i0), b1 = ldnt(bj + i1);
                acc[j] += im_dot(b0, q0) + im_dot(b1, q1);
                stnt(bj + i0, cfma(u01, b1, cmul(u00, b0)));
                stnt(bj + i1, cfma(u11, b1, cmul(u10, b0)));
            }
        }
        if (upd_psi) {
            stnt(psi + i0, cfma(u01, a1, cmul(u00, a0)));
            stnt(psi + i1, cfma(u11, a1, cmul(u10, a0)));
        }
    }
    if (!with_gen) return;
    block_sum<NB>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

// Same step for an uncontrolled target bit t < 6: each thread streams one
// amplitude of every state and takes its partner from lane ^ 2^t.
template <class R, int NB>
__global__ __launch_bounds__(kBlock) void k_adj_lane1(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                      int upd_psi, uint64_t len, int tmask, cx<R> u00, cx<R> u01,
                                                      cx<R> u10, cx<R> u11, int with_gen, cx<double> g00,
                                                      cx<double> g01, cx<double> g10, cx<double> g11,
                                                      double *__restrict__ partials) {
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = 0.0;
    const uint64_t i = xcd_block() * kBlock + threadIdx.x;
    const bool ok = i < len;
    const bool hi = (threadIdx.x & tmask) != 0;
    const cx<R> us = hi ? u11 : u00, up_ = hi ? u10 : u01;
    const cx<double> gs = hi ? g11 : g00, gp = hi ? g10 : g01;
    const cx<R> a = ok ? ldnt(psi + i) : cx<R>{R(0), R(0)};
    const cx<R> ap = shfl_xor_cx(a, tmask);
    const cx<double> q = cfma(gp, up(ap), cmul(gs, up(a)));
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < nb) {
            cx<R> *bj = bras[j];
            const cx<R> b = ok ? ldnt(bj + i) : cx<R>{R(0), R(0)};
            const cx<R> bp = shfl_xor_cx(b, tmask);
            acc[j] += im_dot(b, q);
            if (ok) stnt(bj + i, cfma(up_, bp, cmul(us, b)));
        }
    }
    if (upd_psi && ok) stnt(psi + i, cfma(up_, ap, cmul(us, a)));
    if (!with_gen) return;
    block_sum<NB>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

// K-target dense step (K = 2, 4): matrices staged in LDS, one group per thread.
template <class R, int K, int NB>
__global__ __launch_bounds__(kBlock) void k_adj_dense(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                      int upd_psi, uint64_t ngroups, Fixed fx, uint64_t cmask,
                                                      const uint64_t *__restrict__ offs_g,
                                                      const cx<R> *__restrict__ udag_g,
                                                      const cx<double> *__restrict__ gen_g,
                                                      double *__restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ cx<R> U[D * D];
    __shared__ cx<double> G[D * D];
    __shared__ uint64_t O[D];
    for (int t = threadIdx.x; t < D * D; t += kBlock) {
        U[t] = udag_g[t];
        if (gen_g) G[t] = gen_g[t];
    }
    for (int t = threadIdx.x; t < D; t += kBlock) O[t] = offs_g[t];
    __syncthreads();
    const bool with_gen = gen_g != nullptr;
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = 0.0;
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g < ngroups) {
        const uint64_t base = expand(g, fx) | cmask;
        cx<R> a[D];
#pragma unroll
        for (int m = 0; m < D; ++m) a[m] = ldnt(psi + (base | O[m]));
        cx<double> q[D];
        if (with_gen) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                cx<double> s{0.0, 0.0};
#pragma unroll
                for (int c = 0; c < D; ++c) s = cfma(G[r * D + c], up(a[c]), s);
                q[r] = s;
            }
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                cx<R> *bj = bras[j];
                cx<R> b[D];
#pragma unroll
                for (int m = 0; m < D; ++m) b[m] = ldnt(bj + (base | O[m]));
                if (with_gen) {
#pragma unroll
                    for (int m = 0; m < D; ++m) acc[j] += im_dot(b[m], q[m]);
                }
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    cx<R> s{R(0), R(0)};
#pragma unroll
                    for (int c = 0; c < D; ++c) s = cfma(U[r * D + c], b[c], s);
                    stnt(bj + (base | O[r]), s);
                }
            }
        }
        if (upd_psi) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                cx<R> s{R(0), R(0)};
#pragma unroll
                for (int c = 0; c < D; ++c) s = cfma(U[r * D + c], a[c], s);
                stnt(psi + (base | O[r]), s);
            }
        }
    }
    if (!with_gen) return;
    block_sum<NB>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

template <class R, int NB>
__global__ __launch_bounds__(kBlock) void k_adj_diag(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                     int upd_psi, uint64_t count, Fixed fx, uint64_t cmask,
                                                     uint64_t pmask, cx<R> ue, cx<R> uo, int with_gen, double ge,
                                                     double go, double *__restrict__ partials) {
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = 0.0;
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g < count) {
        const uint64_t i = expand(g, fx) | cmask;
        const bool odd = __popcll(i & pmask) & 1;
        const cx<R> a = ldnt(psi + i);
        const cx<R> u = odd ? uo : ue;
        const double gv = odd ? go : ge;
        const cx<double> q{gv * double(a.re), gv * double(a.im)};
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                const cx<R> b = ldnt(bras[j] + i);
                acc[j] += im_dot(b, q);
                stnt(bras[j] + i, cmul(u, b));
            }
        }
        if (upd_psi) stnt(psi + i, cmul(u, a));
    }
    if (!with_gen) return;
    block_sum<NB>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

// Row sums of the per-workgroup partials in two deterministic stages:
// k_row_chunks sums fixed contiguous slices of each row (grid: slices x rows),
// k_jac_store sums the slice results of row j and scales them into the
// Jacobian column.
constexpr int kSlices = 256;

__global__ __launch_bounds__(kBlock) void k_row_chunks(const double *__restrict__ partials, int nparts,
                                                       double *__restrict__ slices) {
    const int row = blockIdx.y;
    const int per = (nparts + kSlices - 1) / kSlices;
    const int lo = blockIdx.x * per, hi = min(lo + per, nparts);
    const double *p = partials + size_t(row) * nparts;
    double v[1] = {0.0};
    for (int b = lo + threadIdx.x; b < hi; b += kBlock) v[0] += p[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) slices[size_t(row) * kSlices + blockIdx.x] = v[0];
}

__global__ __launch_bounds__(kBlock) void k_jac_store(const double *__restrict__ slices, double *__restrict__ jac,
                                                      int64_t ntp, int64_t col, double factor) {
    double v[1] = {threadIdx.x < kSlices ? slices[size_t(blockIdx.x) * kSlices + threadIdx.x] : 0.0};
    block_sum<1>(v);
    if (threadIdx.x == 0) jac[blockIdx.x * ntp + col] = factor * v[0];
}

bool is_prep(const lgh_op &op) { return op.gate < 0; }

// Launch geometry of one fused step: workgroups of kBlock threads, one
// amplitude (lane kernel), pair or group per thread.
struct StepShape {
    enum Kind { LANE1, PAIR1, DENSE, DIAG } kind;
    Fixed fx;
    uint64_t count;
    unsigned blocks;
};

StepShape step_shape(int n, const KOp &udag) {
    StepShape s{};
    if (udag.kind == KOp::DIAG) {
        s.kind = StepShape::DIAG;
        s.fx = make_fixed(udag.cbits);
    } else {
        std::vector<int> fixed_bits = udag.cbits;
        for (int i = 0; i < udag.k; ++i) fixed_bits.push_back(udag.tbits[i]);
        s.fx = make_fixed(fixed_bits);
        if (udag.k == 1)
            s.kind = (udag.cmask == 0 && udag.tbits[0] < 6 && n >= 6) ? StepShape::LANE1 : StepShape::PAIR1;
        else
            s.kind = StepShape::DENSE;
    }
    s.count = s.kind == StepShape::LANE1 ? (1ull << n) : (1ull << (n - s.fx.cnt));
    s.blocks = unsigned((s.count + kBlock - 1) / kBlock);
    return s;
}

// Calls f(std::integral_constant<int, NB>) with the smallest NB in {1, 2, 4, 8} >= nb.
template <class F> void with_nb(int nb, F &&f) {
    if (nb <= 1)
        f(std::integral_constant<int, 1>{});
    else if (nb <= 2)
        f(std::integral_constant<int, 2>{});
    else if (nb <= 4)
        f(std::integral_constant<int, 4>{});
    else
        f(std::integral_constant<int, 8>{});
}

template <class R>
void fused_step(lgh_sv *lam, cx<R> **bras_dev, int nobs, const KOp &udag, const KGen *gen, double *partials) {
    StreamCtx *ctx = lam->ctx;
    hipStream_t st = ctx->stream;
    cx<R> *psi = static_cast<cx<R> *>(lam->data);
    const int nbatches = std::max(1, (nobs + kBras - 1) / kBras);
    const StepShape sh = step_shape(lam->n, udag);
    const int wg = gen ? 1 : 0;
    auto batch = [&](int b, auto launch) {
        const int nb = std::max(0, std::min(kBras, nobs - b * kBras));
        launch(bras_dev + b * kBras, nb, int(b == nbatches - 1), partials + size_t(b) * kBras * sh.blocks);
        LGH_LAUNCH_CHECK();
    };
    if (sh.kind == StepShape::DIAG) {
        const double ge = gen ? gen->g_even : 0.0, go = gen ? gen->g_odd : 0.0;
        for (int b = 0; b < nbatches; ++b)
            batch(b, [&](cx<R> **bp, int nb, int upd, double *pb) {
                with_nb(nb, [&](auto NBc) {
                    hipLaunchKernelGGL((k_adj_diag<R, decltype(NBc)::value>), dim3(sh.blocks), dim3(kBlock), 0, st,
                                       psi, bp, nb, upd, sh.count, sh.fx, udag.cmask, udag.pmask,
                                       to_cx<R>(udag.d_even), to_cx<R>(udag.d_odd), wg, ge, go, pb);
                });
            });
        return;
    }
    if (sh.kind != StepShape::DENSE) {
        const cx<R> u00 = to_cx<R>(udag.mat[0]), u01 = to_cx<R>(udag.mat[1]), u10 = to_cx<R>(udag.mat[2]),
                    u11 = to_cx<R>(udag.mat[3]);
        cx<double> gm[4] = {};
        if (gen)
            for (int i = 0; i < 4; ++i) gm[i] = to_cx<double>(gen->mat[size_t(i)]);
        for (int b = 0; b < nbatches; ++b)
            batch(b, [&](cx<R> **bp, int nb, int upd, double *pb) {
                with_nb(nb, [&](auto NBc) {
                    constexpr int NB = decltype(NBc)::value;
                    if (sh.kind == StepShape::LANE1)
                        hipLaunchKernelGGL((k_adj_lane1<R, NB>), dim3(sh.blocks), dim3(kBlock), 0, st, psi, bp, nb,
                                           upd, sh.count, 1 << udag.tbits[0], u00, u01, u10, u11, wg, gm[0], gm[1],
                                           gm[2], gm[3], pb);
                    else
                        hipLaunchKernelGGL((k_adj_pair1<R, NB>), dim3(sh.blocks), dim3(kBlock), 0, st, psi, bp, nb,
                                           upd, sh.count, sh.fx, udag.cmask, 1ull << udag.tbits[0], u00, u01, u10,
                                           u11, wg, gm[0], gm[1], gm[2], gm[3], pb);
                });
            });
        return;
    }
    // DENSE with K in {2, 4}: one upload of offsets | U^dagger | G
    const int D = 1 << udag.k;
    const size_t b_off = sizeof(uint64_t) * D, b_u = sizeof(cx<R>) * D * D, b_g = sizeof(cx<double>) * D * D;
    const size_t o_u = (b_off + 15) & ~size_t(15), o_g = (o_u + b_u + 15) & ~size_t(15);
    std::vector<unsigned char> blob(o_g + b_g);
    uint64_t *offs = reinterpret_cast<uint64_t *>(blob.data());
    for (int m = 0; m < D; ++m) {
        uint64_t o = 0;
        for (int j = 0; j < udag.k; ++j)
            if ((m >> (udag.k - 1 - j)) & 1) o |= 1ull << udag.tbits[j];
        offs[m] = o;
    }
    cx<R> *u = reinterpret_cast<cx<R> *>(blob.data() + o_u);
    cx<double> *gmat = reinterpret_cast<cx<double> *>(blob.data() + o_g);
    for (int i = 0; i < D * D; ++i) {
        u[i] = to_cx<R>(udag.mat[size_t(i)]);
        gmat[i] = gen ? to_cx<double>(gen->mat[size_t(i)]) : cx<double>{0.0, 0.0};
    }
    unsigned char *dblob = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
    const uint64_t *d_off = reinterpret_cast<const uint64_t *>(dblob);
    const cx<R> *d_u = reinterpret_cast<const cx<R> *>(dblob + o_u);
    const cx<double> *d_g = gen ? reinterpret_cast<const cx<double> *>(dblob + o_g) : nullptr;
    for (int b = 0; b < nbatches; ++b)
        batch(b, [&](cx<R> **bp, int nb, int upd, double *pb) {
            if (udag.k != 2 && udag.k != 4)
                throw Error(LGH_ERR_UNSUPPORTED, "no fused adjoint kernel for a " + std::to_string(udag.k) +
                                                     "-target generator");
            with_nb(nb, [&](auto NBc) {
                constexpr int NB = decltype(NBc)::value;
                if (udag.k == 2)
                    hipLaunchKernelGGL((k_adj_dense<R, 2, NB>), dim3(sh.blocks), dim3(kBlock), 0, st, psi, bp, nb,
                                       upd, sh.count, sh.fx, udag.cmask, d_off, d_u, d_g, pb);
                else
                    hipLaunchKernelGGL((k_adj_dense<R, 4, NB>), dim3(sh.blocks), dim3(kBlock), 0, st, psi, bp, nb,
                                       upd, sh.count, sh.fx, udag.cmask, d_off, d_u, d_g, pb);
            });
        });
}

// dst = Z-word(zmask) src: the sign (-1)^popcount(i & zmask) per amplitude
template <class R>
__global__ __launch_bounds__(kBlock) void k_zsign(cx<R> *__restrict__ dst, const cx<R> *__restrict__ src, uint64_t len,
                                                  uint64_t zmask) {
    const uint64_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= len) return;
    const cx<R> a = ldnt(src + i);
    stnt(dst + i, (__popcll(i & zmask) & 1) ? cx<R>{-a.re, -a.im} : a);
}

template <class R> void zsign(lgh_sv *dst, const void *src, uint64_t zmask, hipStream_t stream) {
    const uint64_t len = dst->length();
    hipLaunchKernelGGL(k_zsign<R>, dim3(unsigned((len + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       static_cast<cx<R> *>(dst->data), static_cast<const cx<R> *>(src), len, zmask);
    LGH_LAUNCH_CHECK();
}

// zmasks (nullable): observable j is the Z word zmasks[j] and bras[j] is
// scratch that the sweep fills with Z_j psi (inside its first reverse pass
// where that pass is fused).
template <class R>
void adjoint_t(const lgh_sv *psi_in, lgh_sv *const *bras, int nobs, const lgh_op *ops, int64_t nops,
               const int64_t *tp, int64_t ntp, bool apply_ops, double *jac_host, const uint64_t *zmasks) {
    const int n = psi_in->n;
    StreamCtx *ctx = psi_in->ctx;
    for (int j = 0; j < nobs; ++j)
        if (bras[j]->ctx != ctx) bras[j]->ctx->sync();
    // lambda: working copy of psi
    lgh_sv lam = *psi_in;
    lam.data = pool_alloc(psi_in->device, psi_in->bytes);
    struct Free {
        lgh_sv *s;
        ~Free() {
            (void)hipStreamSynchronize(s->ctx->stream);
            pool_free(s->device, s->data, s->bytes);
        }
    } free_lam{&lam};
    // The fused sweep reads lambda's starting values from psi itself when no
    // forward application is asked for; the other forms work on a copy.
    const bool fused_path = fusion_enabled() && fused_partials_bytes(n, psi_in->device) > 0 && nobs > 0;
    const bool from_psi = fused_path && !apply_ops;
    if (!from_psi)
        LGH_HIP(hipMemcpyAsync(lam.data, psi_in->data, psi_in->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (apply_ops)
        for (int64_t i = 0; i < nops; ++i)
            if (!is_prep(ops[i])) launch_kop(&lam, compile_op(ops[i], n, false));
    // Z-word observable states: formed here unless the fused sweep forms them
    if (zmasks && !fused_path)
        for (int j = 0; j < nobs; ++j) zsign<R>(bras[j], lam.data, zmasks[j], ctx->stream);

    // workspace: [jac (nobs*ntp) | bra pointers | partials]
    // the widest step shape has one thread per amplitude
    const unsigned max_blocks = unsigned(((1ull << n) + kBlock - 1) / kBlock);
    const int nbatches = std::max(1, (nobs + kBras - 1) / kBras);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_jac = al(sizeof(double) * size_t(std::max<int64_t>(1, nobs * ntp)));
    const size_t b_ptr = al(sizeof(void *) * size_t(nbatches * kBras));
    const size_t b_part = al(sizeof(double) * size_t(nbatches) * kBras * max_blocks);
    const size_t b_slices = al(sizeof(double) * kBras * kSlices);
    const size_t b_fused = al(fused_partials_bytes(n, psi_in->device));
    char *ws = static_cast<char *>(ctx->workspace(b_jac + b_ptr + b_part + b_slices + b_fused));
    double *slices = reinterpret_cast<double *>(ws + b_jac + b_ptr + b_part);
    double *fpartials = reinterpret_cast<double *>(ws + b_jac + b_ptr + b_part + b_slices);
    double *jac_dev = reinterpret_cast<double *>(ws);
    cx<R> **bras_dev = reinterpret_cast<cx<R> **>(ws + b_jac);
    double *partials = reinterpret_cast<double *>(ws + b_jac + b_ptr);
    std::vector<void *> hp(size_t(nbatches * kBras), nullptr);
    for (int j = 0; j < nobs; ++j) hp[size_t(j)] = bras[j]->data;
    LGH_HIP(hipMemcpyAsync(bras_dev, hp.data(), sizeof(void *) * hp.size(), hipMemcpyHostToDevice, ctx->stream));
    LGH_HIP(hipMemsetAsync(jac_dev, 0, sizeof(double) * size_t(std::max<int64_t>(1, nobs * ntp)), ctx->stream));
    ctx->sync(); // hp is a host temporary

    if (fused_path) {
        // Fused reverse sweep, one observable state at a time on a fresh copy
        // of lambda.  Ops before the earliest trainable one are never visited.
        std::vector<KOp> udag;
        std::vector<KGen> gens;
        std::vector<int> slot;
        std::vector<double> fac;
        int64_t pidx = -1, k = 0;
        bool started = false;
        for (int64_t i = 0; i < nops; ++i) {
            const lgh_op &o = ops[i];
            if (is_prep(o)) continue;
            int col = -1;
            if (o.num_params > 0) {
                ++pidx;
                if (k < ntp && pidx == tp[k]) col = int(k++);
            }
            if (col < 0 && !started) continue;
            started = true;
            udag.push_back(compile_op(o, n, true));
            if (col >= 0) {
                gens.push_back(compile_generator(o, n, udag.back()));
                fac.push_back(-2.0 * gens.back().scale * (o.inverse ? -1.0 : 1.0));
            } else {
                gens.emplace_back();
                fac.push_back(0.0);
            }
            slot.push_back(col);
        }
        lgh_sv lam0 = lam; // forward state every observable sweep starts from (after apply_ops)
        lam0.data = nullptr;
        if (nobs > 1 && !from_psi) {
            lam0.data = pool_alloc(psi_in->device, psi_in->bytes);
            LGH_HIP(hipMemcpyAsync(lam0.data, lam.data, lam.bytes, hipMemcpyDeviceToDevice, ctx->stream));
        }
        struct Free0 {
            lgh_sv *s;
            ~Free0() {
                if (!s->data) return;
                (void)hipStreamSynchronize(s->ctx->stream);
                pool_free(s->device, s->data, s->bytes);
            }
        } free_lam0{&lam0};
        for (int j = 0; j < nobs; ++j) {
            if (j > 0 && !from_psi)
                LGH_HIP(hipMemcpyAsync(lam.data, lam0.data, lam.bytes, hipMemcpyDeviceToDevice, ctx->stream));
            double *jrow = jac_dev + size_t(j) * ntp;
            auto standalone = [&](const KOp &u, const KGen *g, int64_t col, double f) {
                const bool fusedk = u.kind == KOp::DIAG || (u.kind == KOp::DENSE && (u.k == 1 || u.k == 2 || u.k == 4));
                if (!fusedk) {
                    launch_kop(&lam, u);
                    launch_kop(bras[j], u);
                    return;
                }
                fused_step<R>(&lam, bras_dev + j, 1, u, g, partials);
                if (g) {
                    const unsigned blocks = step_shape(n, u).blocks;
                    hipLaunchKernelGGL(k_row_chunks, dim3(kSlices, 1), dim3(kBlock), 0, ctx->stream, partials,
                                       int(blocks), slices);
                    hipLaunchKernelGGL(k_jac_store, dim3(1), dim3(kBlock), 0, ctx->stream, slices, jrow, ntp, col, f);
                    LGH_LAUNCH_CHECK();
                }
            };
            if (zmasks) {
                // the sweep's starting lambda: psi itself, or lam's copy of it
                const void *start = from_psi ? psi_in->data : lam.data;
                BraFromZ bz{zmasks[j], [&, j, start] { zsign<R>(bras[j], start, zmasks[j], ctx->stream); }};
                fused_adjoint_one(&lam, from_psi ? psi_in->data : nullptr, bras[j], udag, gens, slot, fac, jrow,
                                  fpartials, standalone, &bz);
            } else {
                fused_adjoint_one(&lam, from_psi ? psi_in->data : nullptr, bras[j], udag, gens, slot, fac, jrow,
                                  fpartials, standalone);
            }
        }
        if (nobs * ntp > 0)
            LGH_HIP(hipMemcpyAsync(jac_host, jac_dev, sizeof(double) * size_t(nobs * ntp), hipMemcpyDeviceToHost,
                                   ctx->stream));
        ctx->sync();
        return;
    }

    int64_t num_param_ops = 0;
    for (int64_t i = 0; i < nops; ++i)
        if (!is_prep(ops[i]) && ops[i].num_params > 0) ++num_param_ops;
    int64_t tp_pos = ntp - 1;
    int64_t param_idx = num_param_ops - 1;
    for (int64_t op = nops - 1; op >= 0; --op) {
        const lgh_op &o = ops[op];
        if (is_prep(o)) continue;
        if (tp_pos < 0) break;
        const KOp udag = compile_op(o, n, true);
        const bool trainable = o.num_params > 0 && param_idx == tp[tp_pos];
        if (trainable && nobs > 0) {
            const KGen gen = compile_generator(o, n, udag);
            if (udag.kind == KOp::DENSE && udag.k != 1 && udag.k != 2 && udag.k != 4)
                throw Error(LGH_ERR_UNSUPPORTED, "no fused adjoint kernel for this operation");
            fused_step<R>(&lam, bras_dev, nobs, udag, &gen, partials);
            const unsigned blocks = step_shape(n, udag).blocks;
            const double factor = -2.0 * gen.scale * (o.inverse ? -1.0 : 1.0);
            for (int b = 0; b < nbatches; ++b) {
                const int nb = std::min(kBras, nobs - b * kBras);
                if (nb <= 0) break;
                hipLaunchKernelGGL(k_row_chunks, dim3(kSlices, nb), dim3(kBlock), 0, ctx->stream,
                                   partials + size_t(b) * kBras * blocks, int(blocks), slices);
                hipLaunchKernelGGL(k_jac_store, dim3(nb), dim3(kBlock), 0, ctx->stream, slices,
                                   jac_dev + size_t(b) * kBras * ntp, ntp, tp_pos, factor);
                LGH_LAUNCH_CHECK();
            }
        } else if (udag.kind == KOp::DIAG ||
                   (udag.kind == KOp::DENSE && (udag.k == 1 || udag.k == 2 || udag.k == 4))) {
            fused_step<R>(&lam, bras_dev, nobs, udag, nullptr, partials); // U^dagger on every state, one pass
        } else if (udag.kind != KOp::NOP) {
            launch_kop(&lam, udag);
            for (int j = 0; j < nobs; ++j) launch_kop(bras[j], udag);
        }
        if (trainable) --tp_pos;
        if (o.num_params > 0) --param_idx;
    }
    if (nobs * ntp > 0) {
        LGH_HIP(hipMemcpyAsync(jac_host, jac_dev, sizeof(double) * size_t(nobs * ntp), hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    ctx->sync();
}

} // namespace

extern "C" {

static void adjoint_entry(const lgh_sv *psi, lgh_sv *const *bras, int32_t num_obs, const lgh_op *ops,
                      int64_t num_ops, const int64_t *trainable, int64_t num_trainable, int32_t apply_operations,
                      double *jac, const uint64_t *zmasks) {
    LGH_REQUIRE(psi, "null state");
    LGH_REQUIRE(num_trainable > 0, "No trainable parameters provided.");
    LGH_REQUIRE(trainable && (num_obs == 0 || (bras && jac)) && (num_ops == 0 || ops), "null pointer");
    LGH_REQUIRE(num_obs >= 0, "negative observable count");
    for (int32_t j = 0; j < num_obs; ++j) {
        LGH_REQUIRE(bras[j], "null observable state");
        LGH_REQUIRE(bras[j]->n == psi->n && bras[j]->dtype == psi->dtype && bras[j]->device == psi->device,
                    "observable state does not match psi");
    }
    for (int64_t i = 1; i < num_trainable; ++i)
        LGH_REQUIRE(trainable[i] > trainable[i - 1], "trainable parameters must be sorted and unique");
    for (int64_t i = 0; i < num_ops; ++i) {
        if (ops[i].gate < 0) continue; // state preparation: skipped by the reverse sweep
        LGH_REQUIRE(ops[i].num_params <= 1,
                    "The operation is not supported using the adjoint differentiation method");
        validate_op(ops[i], psi->n);
    }
    if (zmasks)
        for (int32_t j = 0; j < num_obs; ++j)
            LGH_REQUIRE((zmasks[j] >> psi->n) == 0, "Z-word mask beyond the state's qubits");
    DeviceGuard g(psi->device);
    if (psi->dtype == LGH_C128)
        adjoint_t<double>(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations != 0, jac,
                          zmasks);
    else
        adjoint_t<float>(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations != 0, jac,
                         zmasks);
}

lgh_status lgh_adjoint_jacobian(const lgh_sv *psi, lgh_sv *const *bras, int32_t num_obs, const lgh_op *ops,
                                int64_t num_ops, const int64_t *trainable, int64_t num_trainable,
                                int32_t apply_operations, double *jac) {
    return guard([&] {
        Range rg("lgh_adjoint_jacobian");
        adjoint_entry(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations, jac, nullptr);
    });
}

lgh_status lgh_adjoint_jacobian_zwords(const lgh_sv *psi, lgh_sv *const *bras, const uint64_t *zmasks,
                                       int32_t num_obs, const lgh_op *ops, int64_t num_ops, const int64_t *trainable,
                                       int64_t num_trainable, int32_t apply_operations, double *jac) {
    return guard([&] {
        Range rg("lgh_adjoint_jacobian_zwords");
        LGH_REQUIRE(num_obs == 0 || zmasks, "null pointer");
        adjoint_entry(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations, jac, zmasks);
    });
}

} // extern "C"
