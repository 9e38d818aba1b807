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

template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_adj_dense(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                      int upd_psi, uint64_t ngroups, uint64_t chunk, Fixed fx,
                                                      uint64_t cmask, const uint64_t *__restrict__ offs_g,
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
    double acc[kBras];
#pragma unroll
    for (int j = 0; j < kBras; ++j) acc[j] = 0.0;
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, ngroups);
    for (uint64_t g = lo + threadIdx.x; g < hi; g += kBlock) {
        const uint64_t base = expand(g, fx) | cmask;
        cx<R> a[D];
#pragma unroll
        for (int m = 0; m < D; ++m) a[m] = psi[base | O[m]];
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
        for (int j = 0; j < kBras; ++j) {
            if (j < nb) {
                cx<R> *bj = bras[j];
                cx<R> b[D];
#pragma unroll
                for (int m = 0; m < D; ++m) b[m] = bj[base | O[m]];
                if (with_gen) {
#pragma unroll
                    for (int m = 0; m < D; ++m) acc[j] += im_dot(b[m], q[m]);
                }
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    cx<R> s{R(0), R(0)};
#pragma unroll
                    for (int c = 0; c < D; ++c) s = cfma(U[r * D + c], b[c], s);
                    bj[base | O[r]] = s;
                }
            }
        }
        if (upd_psi) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                cx<R> s{R(0), R(0)};
#pragma unroll
                for (int c = 0; c < D; ++c) s = cfma(U[r * D + c], a[c], s);
                psi[base | O[r]] = s;
            }
        }
    }
    if (!with_gen) return;
    block_sum<kBras>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_adj_diag(cx<R> *__restrict__ psi, cx<R> *const *__restrict__ bras, int nb,
                                                     int upd_psi, uint64_t count, uint64_t chunk, Fixed fx,
                                                     uint64_t cmask, uint64_t pmask, cx<R> ue, cx<R> uo, int with_gen,
                                                     double ge, double go, double *__restrict__ partials) {
    double acc[kBras];
#pragma unroll
    for (int j = 0; j < kBras; ++j) acc[j] = 0.0;
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, count);
    for (uint64_t g = lo + threadIdx.x; g < hi; g += kBlock) {
        const uint64_t i = expand(g, fx) | cmask;
        const bool odd = __popcll(i & pmask) & 1;
        const cx<R> a = psi[i];
        const cx<R> u = odd ? uo : ue;
        const double gv = odd ? go : ge;
        const cx<double> q{gv * double(a.re), gv * double(a.im)};
#pragma unroll
        for (int j = 0; j < kBras; ++j) {
            if (j < nb) {
                const cx<R> b = bras[j][i];
                acc[j] += im_dot(b, q);
                bras[j][i] = cmul(u, b);
            }
        }
        if (upd_psi) psi[i] = cmul(u, a);
    }
    if (!with_gen) return;
    block_sum<kBras>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; ++j) partials[size_t(j) * gridDim.x + blockIdx.x] = acc[j];
}

// jac[j * ntp + col] = factor * sum_b partials[j][b]; one workgroup per j.
__global__ __launch_bounds__(kBlock) void k_jac_store(const double *__restrict__ partials, int nparts,
                                                      double *__restrict__ jac, int64_t ntp, int64_t col,
                                                      double factor) {
    const double *p = partials + size_t(blockIdx.x) * nparts;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nparts; b += kBlock) v[0] += p[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) jac[blockIdx.x * ntp + col] = factor * v[0];
}

struct Chunks {
    uint64_t chunk;
    unsigned blocks;
};
Chunks chunks_for(uint64_t count, int device) {
    const uint64_t max_blocks = uint64_t(std::max(1, sm_count(device))) * 8;
    uint64_t blocks = (count + kBlock * 4 - 1) / (kBlock * 4);
    blocks = std::max<uint64_t>(1, std::min(blocks, max_blocks));
    const uint64_t chunk = (count + blocks - 1) / blocks;
    blocks = (count + chunk - 1) / chunk;
    return {chunk, unsigned(std::max<uint64_t>(1, blocks))};
}

bool is_prep(const lgh_op &op) { return op.gate < 0; }

template <class R>
void fused_step(lgh_sv *lam, cx<R> **bras_dev, int nobs, const KOp &udag, const KGen &gen, double *partials,
                unsigned max_blocks_hint) {
    (void)max_blocks_hint;
    StreamCtx *ctx = lam->ctx;
    hipStream_t st = ctx->stream;
    const int n = lam->n;
    cx<R> *psi = static_cast<cx<R> *>(lam->data);
    const int nbatches = std::max(1, (nobs + kBras - 1) / kBras);
    if (udag.kind == KOp::DIAG) {
        Fixed fx = make_fixed(udag.cbits);
        const uint64_t count = 1ull << (n - fx.cnt);
        Chunks ch = chunks_for(count, lam->device);
        for (int b = 0; b < nbatches; ++b) {
            const int nb = std::min(kBras, nobs - b * kBras);
            hipLaunchKernelGGL(k_adj_diag<R>, dim3(ch.blocks), dim3(kBlock), 0, st, psi, bras_dev + b * kBras,
                               std::max(nb, 0), int(b == nbatches - 1), count, ch.chunk, fx, udag.cmask, udag.pmask,
                               to_cx<R>(udag.d_even), to_cx<R>(udag.d_odd), 1, gen.g_even, gen.g_odd,
                               partials + size_t(b) * kBras * ch.blocks);
            LGH_LAUNCH_CHECK();
        }
        return;
    }
    // DENSE with K in {1, 2, 4}
    std::vector<int> fixed_bits = udag.cbits;
    for (int i = 0; i < udag.k; ++i) fixed_bits.push_back(udag.tbits[i]);
    Fixed fx = make_fixed(fixed_bits);
    const uint64_t ngroups = 1ull << (n - fx.cnt);
    Chunks ch = chunks_for(ngroups, lam->device);
    const int D = 1 << udag.k;
    // one upload: offsets | U^dagger (state precision) | G (double)
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
    cx<double> *gm = reinterpret_cast<cx<double> *>(blob.data() + o_g);
    for (int i = 0; i < D * D; ++i) {
        u[i] = to_cx<R>(udag.mat[size_t(i)]);
        gm[i] = to_cx<double>(gen.mat[size_t(i)]);
    }
    unsigned char *dblob = static_cast<unsigned char *>(ctx->upload(blob.data(), blob.size()));
    const uint64_t *d_off = reinterpret_cast<const uint64_t *>(dblob);
    const cx<R> *d_u = reinterpret_cast<const cx<R> *>(dblob + o_u);
    const cx<double> *d_g = reinterpret_cast<const cx<double> *>(dblob + o_g);
    for (int b = 0; b < nbatches; ++b) {
        const int nb = std::max(0, std::min(kBras, nobs - b * kBras));
        double *pb = partials + size_t(b) * kBras * ch.blocks;
        const int upd = int(b == nbatches - 1);
        switch (udag.k) {
        case 1:
            hipLaunchKernelGGL((k_adj_dense<R, 1>), dim3(ch.blocks), dim3(kBlock), 0, st, psi, bras_dev + b * kBras, nb,
                               upd, ngroups, ch.chunk, fx, udag.cmask, d_off, d_u, d_g, pb);
            break;
        case 2:
            hipLaunchKernelGGL((k_adj_dense<R, 2>), dim3(ch.blocks), dim3(kBlock), 0, st, psi, bras_dev + b * kBras, nb,
                               upd, ngroups, ch.chunk, fx, udag.cmask, d_off, d_u, d_g, pb);
            break;
        case 4:
            hipLaunchKernelGGL((k_adj_dense<R, 4>), dim3(ch.blocks), dim3(kBlock), 0, st, psi, bras_dev + b * kBras, nb,
                               upd, ngroups, ch.chunk, fx, udag.cmask, d_off, d_u, d_g, pb);
            break;
        default:
            throw Error(LGH_ERR_UNSUPPORTED, "no fused adjoint kernel for a " + std::to_string(udag.k) +
                                                 "-target generator");
        }
        LGH_LAUNCH_CHECK();
    }
}

// Partials of a fused step live in rows of kBras * blocks per batch; this
// returns the number of workgroups the step uses so the store kernel knows
// the row length.
unsigned step_blocks(const lgh_sv *lam, const KOp &udag) {
    std::vector<int> fixed_bits = udag.cbits;
    if (udag.kind == KOp::DENSE)
        for (int i = 0; i < udag.k; ++i) fixed_bits.push_back(udag.tbits[i]);
    const uint64_t count = 1ull << (lam->n - int(fixed_bits.size()));
    return chunks_for(count, lam->device).blocks;
}

template <class R>
void adjoint_t(const lgh_sv *psi_in, lgh_sv *const *bras, int nobs, const lgh_op *ops, int64_t nops,
               const int64_t *tp, int64_t ntp, bool apply_ops, double *jac_host) {
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
    LGH_HIP(hipMemcpyAsync(lam.data, psi_in->data, psi_in->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (apply_ops)
        for (int64_t i = 0; i < nops; ++i)
            if (!is_prep(ops[i])) launch_kop(&lam, compile_op(ops[i], n, false));

    // workspace: [jac (nobs*ntp) | bra pointers | partials]
    const unsigned max_blocks = unsigned(std::max(1, sm_count(psi_in->device)) * 8);
    const int nbatches = std::max(1, (nobs + kBras - 1) / kBras);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_jac = al(sizeof(double) * size_t(std::max<int64_t>(1, nobs * ntp)));
    const size_t b_ptr = al(sizeof(void *) * size_t(nbatches * kBras));
    const size_t b_part = al(sizeof(double) * size_t(nbatches) * kBras * max_blocks);
    char *ws = static_cast<char *>(ctx->workspace(b_jac + b_ptr + b_part));
    double *jac_dev = reinterpret_cast<double *>(ws);
    cx<R> **bras_dev = reinterpret_cast<cx<R> **>(ws + b_jac);
    double *partials = reinterpret_cast<double *>(ws + b_jac + b_ptr);
    std::vector<void *> hp(size_t(nbatches * kBras), nullptr);
    for (int j = 0; j < nobs; ++j) hp[size_t(j)] = bras[j]->data;
    LGH_HIP(hipMemcpyAsync(bras_dev, hp.data(), sizeof(void *) * hp.size(), hipMemcpyHostToDevice, ctx->stream));
    LGH_HIP(hipMemsetAsync(jac_dev, 0, sizeof(double) * size_t(std::max<int64_t>(1, nobs * ntp)), ctx->stream));
    ctx->sync(); // hp is a host temporary

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
            fused_step<R>(&lam, bras_dev, nobs, udag, gen, partials, max_blocks);
            const unsigned blocks = step_blocks(&lam, udag);
            const double factor = -2.0 * gen.scale * (o.inverse ? -1.0 : 1.0);
            for (int b = 0; b < nbatches; ++b) {
                const int nb = std::min(kBras, nobs - b * kBras);
                if (nb <= 0) break;
                hipLaunchKernelGGL(k_jac_store, dim3(nb), dim3(kBlock), 0, ctx->stream,
                                   partials + size_t(b) * kBras * blocks, int(blocks),
                                   jac_dev + size_t(b) * kBras * ntp, ntp, tp_pos, factor);
                LGH_LAUNCH_CHECK();
            }
        } else {
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

lgh_status lgh_adjoint_jacobian(const lgh_sv *psi, lgh_sv *const *bras, int32_t num_obs, const lgh_op *ops,
                                int64_t num_ops, const int64_t *trainable, int64_t num_trainable,
                                int32_t apply_operations, double *jac) {
    return guard([&] {
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
        DeviceGuard g(psi->device);
        if (psi->dtype == LGH_C128)
            adjoint_t<double>(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations != 0, jac);
        else
            adjoint_t<float>(psi, bras, num_obs, ops, num_ops, trainable, num_trainable, apply_operations != 0, jac);
    });
}

} // extern "C"
