// Measurements and BLAS-like helpers: inner products, axpy/scale, dense and
// Pauli-word expectation values, resident CSR SpMV (+ fused dot),
// probabilities and the seeded sampler.
//
// Reference counterparts (pennylane_lightning_gpu/src/...):
//   innerProdC_CUDA / scaleAndAddC_CUDA / scaleC_CUDA   util/cuda_helpers.hpp:456-560
//   getExpectationValueDeviceMatrix                    simulator/StateVectorCudaManaged.hpp:1649-1713
//   getExpectationValuePauliWords                      simulator/StateVectorCudaManaged.hpp:1071-1148
//   getExpectationValueOnSparseSpMV                    simulator/StateVectorCudaManaged.hpp:795-922
//   SparseHamiltonianGPU::applyInPlace                 algorithms/ObservablesGPU.hpp:469-587
//   probability                                        simulator/StateVectorCudaManaged.hpp:931-970
//   generate_samples                                   simulator/StateVectorCudaManaged.hpp:982-1061
//
// Every reduction is deterministic for a given problem size: each workgroup
// reduces a fixed contiguous chunk (wavefront shuffles, then an LDS tree) to
// one partial, and a second kernel sums the partials in a fixed order.
#include <algorithm>
#include <cstring>
#include <map>

#include <rocprim/device/device_radix_sort.hpp>

#include "lgh_internal.hpp"

namespace lgh {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-wide sums of NV values; the results are valid in thread 0.
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
            s = wave_sum(s);
            v[k] = s;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_final_reduce(const double *__restrict__ partials, int nparts,
                                                         double *__restrict__ out) {
    // one workgroup per value
    const double *p = partials + size_t(blockIdx.x) * nparts;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nparts; b += kBlock) v[0] += p[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// chunked partitioning of [0, count) over a bounded number of workgroups
struct Chunks {
    uint64_t chunk;
    unsigned blocks;
};
Chunks chunks_for(uint64_t count, int device, uint64_t per_thread = 4) {
    const uint64_t max_blocks = uint64_t(std::max(1, sm_count(device))) * 8;
    uint64_t blocks = (count + kBlock * per_thread - 1) / (kBlock * per_thread);
    blocks = std::max<uint64_t>(1, std::min(blocks, max_blocks));
    uint64_t chunk = (count + blocks - 1) / blocks;
    blocks = (count + chunk - 1) / chunk;
    return {chunk, unsigned(std::max<uint64_t>(1, blocks))};
}

} // namespace

void launch_reduce(StreamCtx *ctx, const double *partials, int nparts, int nvals, double *out_dev) {
    hipLaunchKernelGGL(k_final_reduce, dim3(nvals), dim3(kBlock), 0, ctx->stream, partials, nparts, out_dev);
    LGH_LAUNCH_CHECK();
}

// Reduction workspace layout: [partials (nvals * nparts) | results (nvals)].
struct RedBuf {
    double *partials;
    double *result;
};
static RedBuf red_buf(StreamCtx *ctx, int nvals, int nparts, size_t extra_offset = 0) {
    char *ws = static_cast<char *>(ctx->workspace(extra_offset + sizeof(double) * size_t(nvals) * (nparts + 1) + 256));
    RedBuf r;
    r.partials = reinterpret_cast<double *>(ws + extra_offset);
    r.result = r.partials + size_t(nvals) * nparts;
    return r;
}

static void fetch(StreamCtx *ctx, const double *dev, int nvals, double *out) {
    double *h = static_cast<double *>(ctx->host_buffer(sizeof(double) * nvals));
    LGH_HIP(hipMemcpyAsync(h, dev, sizeof(double) * nvals, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::memcpy(out, h, sizeof(double) * nvals);
}

} // namespace lgh

using namespace lgh;

namespace {

// ---------------------------------------------------------------- BLAS-like

template <class R>
__global__ __launch_bounds__(kBlock) void k_inner(const cx<R> *__restrict__ a, const cx<R> *__restrict__ b,
                                                  uint64_t len, uint64_t chunk, double *__restrict__ partials) {
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, len);
    double v[2] = {0.0, 0.0};
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const cx<R> x = a[i], y = b[i];
        v[0] += double(x.re) * y.re + double(x.im) * y.im;
        v[1] += double(x.re) * y.im - double(x.im) * y.re;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[gridDim.x + blockIdx.x] = v[1];
    }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_axpy(cx<R> *__restrict__ y, const cx<R> *__restrict__ x, uint64_t len,
                                                 cx<R> a) {
    const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < len) y[i] = cfma(a, x[i], y[i]);
}

// dst = P src for the Pauli word P with X/Y bits xmask, Y/Z bits zmask and
// ny Y factors: (P src)[i] = i^ny (-1)^popcount(j & zmask) src[j], j = i ^ xmask.
// One read and one write per amplitude; the reads of a wavefront stay in the
// same cache lines as its writes when xmask has only low bits, and are whole
// lines otherwise.
template <class R>
__global__ __launch_bounds__(kBlock) void k_pauli_from(cx<R> *__restrict__ dst, const cx<R> *__restrict__ src,
                                                       uint64_t len, uint64_t xmask, uint64_t zmask, int ny) {
    const uint64_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= len) return;
    const uint64_t j = i ^ xmask;
    cx<R> v = ldnt(src + j);
    if (ny & 2) v = cx<R>{-v.re, -v.im};       // i^2 = -1
    if (ny & 1) v = cx<R>{-v.im, v.re};        // times i
    if (__popcll(j & zmask) & 1) v = cx<R>{-v.re, -v.im};
    stnt(dst + i, v);
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_scale(cx<R> *__restrict__ x, uint64_t len, cx<R> a) {
    const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < len) x[i] = cmul(a, x[i]);
}

// ---------------------------------------------------------------- dense expval

template <int K> struct Offs {
    uint64_t o[1 << K];
};

template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_expval_reg(const cx<R> *__restrict__ sv, uint64_t ngroups, uint64_t chunk,
                                                       Fixed fx, Offs<K> offs, const cx<double> *__restrict__ mat,
                                                       double *__restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ cx<double> sm[D * D];
    for (int t = threadIdx.x; t < D * D; t += kBlock) sm[t] = mat[t];
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, ngroups);
    double acc[2] = {0.0, 0.0};
    for (uint64_t g = lo + threadIdx.x; g < hi; g += kBlock) {
        const uint64_t base = expand(g, fx);
        cx<double> v[D];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            cx<R> t = sv[base | offs.o[m]];
            v[m] = cx<double>{double(t.re), double(t.im)};
        }
#pragma unroll
        for (int r = 0; r < D; ++r) {
            cx<double> y{0.0, 0.0};
#pragma unroll
            for (int c = 0; c < D; ++c) y = cfma(sm[r * D + c], v[c], y);
            acc[0] += v[r].re * y.re + v[r].im * y.im;
            acc[1] += v[r].re * y.im - v[r].im * y.re;
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
        partials[gridDim.x + blockIdx.x] = acc[1];
    }
}

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

// 5..13 targets: one workgroup per group, amplitudes in LDS, transposed
// matrix streamed from global memory.
template <class R>
__global__ __launch_bounds__(kBlock) void k_expval_lds(const cx<R> *__restrict__ sv, Fixed fx, Tbits tb,
                                                       const cx<double> *__restrict__ matT,
                                                       double *__restrict__ partials) {
    extern __shared__ unsigned char lds_raw[];
    cx<double> *x = reinterpret_cast<cx<double> *>(lds_raw);
    const uint32_t D = 1u << tb.k;
    const uint64_t base = expand(blockIdx.x, fx);
    for (uint32_t m = threadIdx.x; m < D; m += kBlock) {
        cx<R> t = sv[base | deposit(m, tb)];
        x[m] = cx<double>{double(t.re), double(t.im)};
    }
    __syncthreads();
    double acc[2] = {0.0, 0.0};
    for (uint32_t r = threadIdx.x; r < D; r += kBlock) {
        cx<double> y{0.0, 0.0};
        for (uint32_t c = 0; c < D; ++c) y = cfma(matT[uint64_t(c) * D + r], x[c], y);
        acc[0] += x[r].re * y.re + x[r].im * y.im;
        acc[1] += x[r].re * y.im - x[r].im * y.re;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
        partials[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ---------------------------------------------------------------- Pauli words
//
// <P> for a word with X/Y bits f (flip mask), Z/Y bits zy and nY factors Y:
// P|i> = i^nY (-1)^popcount(i & zy) |i ^ f>, so
//   <P> = Re( i^nY sum_i (-1)^popcount(i & zy) conj(psi_{i^f}) psi_i ),
// real for a Hermitian word.  Kernels accumulate the term's real and
// imaginary parts and keep the one i^nY turns into the real part (`phase` =
// nY mod 4: +re, -im, -re, +im).
//
// Words are grouped into passes by tiles: a pass holds a set of index bits
// (the three lowest, so that loads cover 128-byte runs, and up to eight more
// chosen per pass); a workgroup stages each tile of 2^11 amplitudes (all
// combinations of the pass's bits for one value of the others) in LDS, and
// every word of the pass whose flip bits are tile bits finds each partner
// i ^ f in the tile.  One read of the state then serves every word of the
// pass, whatever their flip masks.  Words with more flip bits than a tile
// holds take the pair kernel below, one pass per flip mask.

constexpr int kPTileBits = 11;       // amplitudes per tile: 2^11
constexpr int kPFixedBits = 3;       // lowest index bits in every tile
constexpr int kPassWords = 32;       // words per tile launch
constexpr int kMaxWordsPerPass = 8;  // words per pair-kernel launch

// a pass's words as kernel arguments
struct PWords {
    int count;
    uint32_t flip[kPassWords];  // flip mask in tile-local bits
    uint32_t zyl[kPassWords];   // Z/Y mask in tile-local bits
    uint32_t phase[kPassWords]; // nY mod 4
    uint64_t zy[kPassWords];    // Z/Y mask over global index bits (only those outside the tile are read)
};
// tile-local bit r is global index bit bit[r]
struct PTile {
    int g;
    int bit[kPTileBits];
};

// Sign flip of a double by an integer xor of its high word (bit 31 of
// `neg`): one instruction instead of a multiply by -1.
__device__ __forceinline__ double flip_sign(double x, uint32_t neg) {
    const long long b = __double_as_longlong(x) ^ (static_cast<long long>(neg & 0x80000000u) << 32);
    return __longlong_as_double(b);
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_pauli_tile(const cx<R> *__restrict__ sv, uint64_t ntiles,
                                                       uint64_t tiles_per_block, Fixed tfx, PTile tm, PWords pw,
                                                       double *__restrict__ partials) {
    constexpr int kAmps = 1 << kPTileBits;
    constexpr int kPer = kAmps / kBlock;
    __shared__ cx<R> tile[kAmps];
    const int tid = threadIdx.x;
    const int tsz = 1 << tm.g;
    uint64_t off[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int j = tid + k * kBlock;
        uint64_t o = 0;
#pragma unroll
        for (int r = 0; r < kPTileBits; ++r)
            if (r < tm.g && ((j >> r) & 1)) o |= 1ull << tm.bit[r];
        off[k] = o;
    }
    double acc[kPassWords];
#pragma unroll
    for (int q = 0; q < kPassWords; ++q) acc[q] = 0.0;
    const uint64_t t0 = uint64_t(blockIdx.x) * tiles_per_block, t1 = min(t0 + tiles_per_block, ntiles);
    for (uint64_t t = t0; t < t1; ++t) {
        const uint64_t base = expand(t, tfx);
        cx<R> v[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int j = tid + k * kBlock;
            v[k] = cx<R>{R(0), R(0)};
            if (j < tsz) {
                v[k] = ldnt(sv + (base | off[k]));
                tile[j] = v[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPassWords; ++q) {
            if (q >= pw.count) break;
            const uint32_t f = pw.flip[q], zl = pw.zyl[q], ph = pw.phase[q];
            // the word's sign outside the tile is uniform over the tile
            const uint32_t pbase = uint32_t(__popcll(base & pw.zy[q]) & 1) << 31;
            // only the component i^nY turns into the real part is summed
            double sum = 0.0;
            if ((ph & 1u) == 0) {
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const int j = tid + k * kBlock;
                    if (j < tsz) {
                        const cx<R> p = tile[j ^ int(f)];
                        const double x = double(p.re) * v[k].re + double(p.im) * v[k].im;
                        sum += flip_sign(x, (uint32_t(__popc(uint32_t(j) & zl)) << 31) ^ pbase);
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < kPer; ++k) {
                    const int j = tid + k * kBlock;
                    if (j < tsz) {
                        const cx<R> p = tile[j ^ int(f)];
                        const double x = double(p.re) * v[k].im - double(p.im) * v[k].re;
                        sum += flip_sign(x, (uint32_t(__popc(uint32_t(j) & zl)) << 31) ^ pbase);
                    }
                }
            }
            // i^nY: +re, -im, -re, +im
            acc[q] += (ph == 1u || ph == 2u) ? -sum : sum;
        }
        __syncthreads(); // the next tile overwrites the LDS copy
    }
    block_sum<kPassWords>(acc);
    if (tid == 0)
        for (int q = 0; q < pw.count; ++q) partials[size_t(q) * gridDim.x + blockIdx.x] = acc[q];
}

struct WordSet {
    int count;
    uint64_t zy[kMaxWordsPerPass];
    uint32_t phase[kMaxWordsPerPass];
};

// <P_w> for words that share one flip mask f.  With f != 0 each pair
// (i, i^f) is read once: its two terms s_i conj(psi_j) psi_i and
// s_j conj(psi_i) psi_j are added together.
template <class R>
__global__ __launch_bounds__(kBlock) void k_pauli(const cx<R> *__restrict__ sv, uint64_t count, uint64_t chunk,
                                                  int pivot, uint64_t flip, WordSet ws,
                                                  double *__restrict__ partials) {
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, count);
    double acc[2 * kMaxWordsPerPass];
#pragma unroll
    for (int k = 0; k < 2 * kMaxWordsPerPass; ++k) acc[k] = 0.0;
    for (uint64_t g = lo + threadIdx.x; g < hi; g += kBlock) {
        if (pivot < 0) {
            const cx<R> a = sv[g];
            const double p = double(a.re) * a.re + double(a.im) * a.im;
#pragma unroll
            for (int w = 0; w < kMaxWordsPerPass; ++w)
                if (w < ws.count) acc[2 * w] += (__popcll(g & ws.zy[w]) & 1) ? -p : p;
        } else {
            const uint64_t lo_bits = g & ((1ull << pivot) - 1ull);
            const uint64_t i = ((g >> pivot) << (pivot + 1)) | lo_bits;
            const uint64_t j = i ^ flip;
            const cx<R> a = sv[i], b = sv[j];
            // conj(b) a and conj(a) b = conj(conj(b) a)
            const double re = double(b.re) * a.re + double(b.im) * a.im;
            const double im = double(b.re) * a.im - double(b.im) * a.re;
#pragma unroll
            for (int w = 0; w < kMaxWordsPerPass; ++w) {
                if (w < ws.count) {
                    const double si = (__popcll(i & ws.zy[w]) & 1) ? -1.0 : 1.0;
                    const double sj = (__popcll(j & ws.zy[w]) & 1) ? -1.0 : 1.0;
                    acc[2 * w] += (si + sj) * re;
                    acc[2 * w + 1] += (si - sj) * im;
                }
            }
        }
    }
    double val[kMaxWordsPerPass];
#pragma unroll
    for (int w = 0; w < kMaxWordsPerPass; ++w) {
        const uint32_t ph = w < ws.count ? ws.phase[w] : 0u;
        const double re = acc[2 * w], im = acc[2 * w + 1];
        val[w] = ph == 0 ? re : ph == 1 ? -im : ph == 2 ? -re : im;
    }
    block_sum<kMaxWordsPerPass>(val);
    if (threadIdx.x == 0)
        for (int w = 0; w < ws.count; ++w) partials[size_t(w) * gridDim.x + blockIdx.x] = val[w];
}

// ---------------------------------------------------------------- CSR

template <class I, class R, int L>
__global__ __launch_bounds__(kBlock) void k_csr(const I *__restrict__ indptr, const I *__restrict__ indices,
                                                const cx<R> *__restrict__ vals, const cx<R> *__restrict__ x,
                                                cx<R> *__restrict__ y, int64_t rows, double *__restrict__ partials) {
    // L lanes per row; rows assigned round-robin over all sub-groups
    const int lane = threadIdx.x & (L - 1);
    const int64_t sub = (int64_t(blockIdx.x) * kBlock + threadIdx.x) / L;
    const int64_t nsub = int64_t(gridDim.x) * kBlock / L;
    double acc[2] = {0.0, 0.0};
    for (int64_t r = sub; r < rows; r += nsub) {
        const int64_t b = indptr[r], e = indptr[r + 1];
        double yr = 0.0, yi = 0.0;
        for (int64_t k = b + lane; k < e; k += L) {
            const cx<R> v = vals[k];
            const cx<R> xv = x[indices[k]];
            yr += double(v.re) * xv.re - double(v.im) * xv.im;
            yi += double(v.re) * xv.im + double(v.im) * xv.re;
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            yr += __shfl_xor(yr, off, L);
            yi += __shfl_xor(yi, off, L);
        }
        if (lane == 0) {
            if (y) {
                y[r] = cx<R>{R(yr), R(yi)};
            } else {
                const cx<R> p = x[r];
                acc[0] += double(p.re) * yr + double(p.im) * yi;
                acc[1] += double(p.re) * yi - double(p.im) * yr;
            }
        }
    }
    if (y) return;
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc[0];
        partials[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ---------------------------------------------------------------- probabilities

struct WireBits {
    int k;
    int b[32]; // state-index bit of wires[j]; output bit j
};

__device__ __forceinline__ uint32_t gather_bits(uint64_t i, const WireBits &wb) {
    uint32_t o = 0;
    for (int j = 0; j < wb.k; ++j) o |= uint32_t((i >> wb.b[j]) & 1ull) << j;
    return o;
}

// k <= 10 output bins: LDS histogram per workgroup, then a per-bin sum over
// workgroups.
template <class R>
__global__ __launch_bounds__(kBlock) void k_probs_hist(const cx<R> *__restrict__ sv, uint64_t len, uint64_t chunk,
                                                       WireBits wb, double *__restrict__ partials) {
    __shared__ double hist[1024];
    const int nb = 1 << wb.k;
    for (int t = threadIdx.x; t < nb; t += kBlock) hist[t] = 0.0;
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * chunk, hi = min(lo + chunk, len);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const cx<R> a = sv[i];
        atomicAdd(&hist[gather_bits(i, wb)], double(a.re) * a.re + double(a.im) * a.im);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nb; t += kBlock) partials[size_t(t) * gridDim.x + blockIdx.x] = hist[t];
}

// k > 10: one thread per output bin, sequential sum over the other bits.
struct Scatter {
    int nrest;
    int rest[64]; // state-index bits not in the wire set, ascending
};

template <class R>
__global__ __launch_bounds__(kBlock) void k_probs_bins(const cx<R> *__restrict__ sv, WireBits wb, Scatter sc,
                                                       double *__restrict__ out) {
    const uint64_t bin = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (bin >= (1ull << wb.k)) return;
    uint64_t base = 0;
    for (int j = 0; j < wb.k; ++j) base |= ((bin >> j) & 1ull) << wb.b[j];
    const uint64_t m = 1ull << sc.nrest;
    double s = 0.0;
    for (uint64_t c = 0; c < m; ++c) {
        uint64_t idx = base;
        for (int q = 0; q < sc.nrest; ++q) idx |= ((c >> q) & 1ull) << sc.rest[q];
        const cx<R> a = sv[idx];
        s += double(a.re) * a.re + double(a.im) * a.im;
    }
    out[bin] = s;
}

// ---------------------------------------------------------------- sampler

// Philox4x32-10 (Salmon et al., SC'11); counter = (shot lo, shot hi, 0, 0),
// key = (seed lo, seed hi).
__device__ __forceinline__ void philox(uint64_t seed, uint64_t ctr, uint32_t out[4]) {
    uint32_t c0 = uint32_t(ctr), c1 = uint32_t(ctr >> 32), c2 = 0, c3 = 0;
    uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1, n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

constexpr double kFixedScale = 1152921504606846976.0; // 2^60

// |a|^2 in 2^-60 fixed point with explicit roundings (no contraction), so
// that the value is the same bit pattern the oracle computes.
template <class R> __device__ __forceinline__ uint64_t fixed_prob(cx<R> a) {
    const double re = double(a.re), im = double(a.im);
    const double p = __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
    return uint64_t(__dmul_rn(p, kFixedScale));
}

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <class R>
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(const cx<R> *__restrict__ sv, uint64_t len,
                                                           uint64_t *__restrict__ tile_sums) {
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    uint64_t s = 0;
    for (int e = 0; e < kScanItems; ++e) {
        const uint64_t i = base + uint64_t(e) * kBlock + threadIdx.x;
        if (i < len) s += fixed_prob(sv[i]);
    }
    __shared__ uint64_t sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = sh[0];
}

// exclusive scan of the tile sums in one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_tile_offsets(uint64_t *__restrict__ tile_sums, uint64_t ntiles,
                                                            uint64_t *__restrict__ total) {
    __shared__ uint64_t sh[1024];
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t lo = threadIdx.x * per, hi = min(lo + per, ntiles);
    uint64_t s = 0;
    for (uint64_t i = lo; i < hi; ++i) s += tile_sums[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 thread sums
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = threadIdx.x >= unsigned(off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = sh[threadIdx.x] - s; // exclusive prefix of this thread's segment
    for (uint64_t i = lo; i < hi; ++i) {
        const uint64_t t = tile_sums[i];
        tile_sums[i] = run;
        run += t;
    }
    if (threadIdx.x == 1023) *total = sh[1023];
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_scan_cdf(const cx<R> *__restrict__ sv, uint64_t len,
                                                     const uint64_t *__restrict__ tile_offsets,
                                                     uint64_t *__restrict__ cdf) {
    // each thread owns kScanItems consecutive elements of the tile
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
    uint64_t q[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
        const uint64_t i = base + e;
        q[e] = i < len ? fixed_prob(sv[i]) : 0;
        s += q[e];
    }
    __shared__ uint64_t sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint64_t v = threadIdx.x >= unsigned(off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = tile_offsets[blockIdx.x] + sh[threadIdx.x] - s;
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
        const uint64_t i = base + e;
        run += q[e];
        if (i < len) cdf[i] = run;
    }
}

__global__ __launch_bounds__(kBlock) void k_sample_search(const uint64_t *__restrict__ cdf, uint64_t len,
                                                          const uint64_t *__restrict__ total_ptr, uint64_t shots,
                                                          uint64_t seed, uint64_t *__restrict__ out) {
    const uint64_t s = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (s >= shots) return;
    const uint64_t total = *total_ptr;
    uint32_t r[4];
    philox(seed, s, r);
    const uint64_t u53 = ((uint64_t(r[1]) << 32) | r[0]) >> 11;
    // t = floor(u53 * total / 2^53) with a 128-bit product
    const uint64_t hi = __umul64hi(u53, total), lo = u53 * total;
    const uint64_t t = (hi << 11) | (lo >> 53);
    uint64_t a = 0, b = len - 1;
    while (a < b) {
        const uint64_t mid = a + ((b - a) >> 1);
        if (cdf[mid] > t)
            b = mid;
        else
            a = mid + 1;
    }
    out[s] = a;
}

__global__ __launch_bounds__(kBlock) void k_expand_bits(const uint64_t *__restrict__ idx, uint64_t shots, int n,
                                                        uint64_t *__restrict__ bits) {
    const uint64_t t = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= shots * uint64_t(n)) return;
    const uint64_t s = t / n;
    const int col = int(t - s * n);
    bits[t] = (idx[s] >> (n - 1 - col)) & 1ull;
}

// ---------------------------------------------------------------- host helpers

template <class R> const cx<R> *cdata(const lgh_sv *sv) { return static_cast<const cx<R> *>(sv->data); }
template <class R> cx<R> *mdata(lgh_sv *sv) { return static_cast<cx<R> *>(sv->data); }

template <class R> void inner_t(const lgh_sv *a, const lgh_sv *b, double *re, double *im) {
    StreamCtx *ctx = a->ctx;
    Chunks ch = chunks_for(a->length(), a->device);
    RedBuf rb = red_buf(ctx, 2, ch.blocks);
    hipLaunchKernelGGL(k_inner<R>, dim3(ch.blocks), dim3(kBlock), 0, ctx->stream, cdata<R>(a), cdata<R>(b),
                       a->length(), ch.chunk, rb.partials);
    LGH_LAUNCH_CHECK();
    launch_reduce(ctx, rb.partials, ch.blocks, 2, rb.result);
    double out[2];
    fetch(ctx, rb.result, 2, out);
    *re = out[0];
    *im = out[1];
}

template <class R> void expval_t(const lgh_sv *sv, const int32_t *wires, int k, const double *matrix, double *re,
                                 double *im) {
    StreamCtx *ctx = sv->ctx;
    const int n = sv->n;
    const int D = 1 << k;
    std::vector<int> tb(k);
    for (int j = 0; j < k; ++j) tb[j] = n - 1 - wires[j];
    Fixed fx = make_fixed(tb);
    const uint64_t ngroups = 1ull << (n - k);
    std::vector<cx<double>> m(size_t(D) * D);
    if (k <= kMaxRegK) {
        for (size_t i = 0; i < m.size(); ++i) m[i] = cx<double>{matrix[2 * i], matrix[2 * i + 1]};
        const cx<double> *dm = static_cast<const cx<double> *>(ctx->upload(m.data(), m.size() * sizeof(m[0])));
        Chunks ch = chunks_for(ngroups, sv->device, 2);
        RedBuf rb = red_buf(ctx, 2, ch.blocks);
        auto run = [&](auto offs, auto kern) {
            for (int mi = 0; mi < D; ++mi) {
                uint64_t o = 0;
                for (int j = 0; j < k; ++j)
                    if ((mi >> (k - 1 - j)) & 1) o |= 1ull << tb[j];
                offs.o[mi] = o;
            }
            hipLaunchKernelGGL(kern, dim3(ch.blocks), dim3(kBlock), 0, ctx->stream, cdata<R>(sv), ngroups, ch.chunk,
                               fx, offs, dm, rb.partials);
        };
        if (k == 1) run(Offs<1>{}, k_expval_reg<R, 1>);
        else if (k == 2) run(Offs<2>{}, k_expval_reg<R, 2>);
        else if (k == 3) run(Offs<3>{}, k_expval_reg<R, 3>);
        else run(Offs<4>{}, k_expval_reg<R, 4>);
        LGH_LAUNCH_CHECK();
        launch_reduce(ctx, rb.partials, ch.blocks, 2, rb.result);
        double out[2];
        fetch(ctx, rb.result, 2, out);
        *re = out[0];
        *im = out[1];
        return;
    }
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c)
            m[size_t(c) * D + r] = cx<double>{matrix[2 * (size_t(r) * D + c)], matrix[2 * (size_t(r) * D + c) + 1]};
    const size_t mbytes = m.size() * sizeof(m[0]);
    const size_t moff = (mbytes + 255) & ~size_t(255);
    RedBuf rb = red_buf(ctx, 2, int(ngroups), moff);
    void *dm = ctx->workspace(moff); // matrix at the start of the workspace
    LGH_HIP(hipMemcpyAsync(dm, m.data(), mbytes, hipMemcpyHostToDevice, ctx->stream));
    Tbits t{};
    t.k = k;
    for (int j = 0; j < k; ++j) t.b[j] = tb[j];
    const size_t lds = size_t(D) * sizeof(cx<double>);
    LGH_REQUIRE(lds <= 160 * 1024, "observable matrix too large");
    hipLaunchKernelGGL(k_expval_lds<R>, dim3(unsigned(ngroups)), dim3(kBlock), lds, ctx->stream, cdata<R>(sv), fx, t,
                       static_cast<const cx<double> *>(dm), rb.partials);
    LGH_LAUNCH_CHECK();
    launch_reduce(ctx, rb.partials, int(ngroups), 2, rb.result);
    double out[2];
    fetch(ctx, rb.result, 2, out); // synchronises: m may go out of scope
    *re = out[0];
    *im = out[1];
}

// sum_w Re(c_w <P_w>) with <P_w> real, in word order; per_word (nullable) gets <P_w>
double weighted_total(int64_t nwords, const double *coeffs, const double *vals, double *per_word) {
    double tot = 0.0;
    for (int64_t w = 0; w < nwords; ++w) {
        tot += coeffs[2 * w] * vals[w];
        if (per_word) per_word[w] = vals[w];
    }
    return tot;
}

// Parses word w into its flip mask, Z/Y mask and number of Y factors.
void parse_word(int n, const char *paulis, const int32_t *wires, int64_t b, int64_t e, uint64_t &flip,
                uint64_t &zy, int32_t &ny) {
    flip = zy = 0;
    ny = 0;
    for (int64_t j = b; j < e; ++j) {
        const int wire = wires[j];
        LGH_REQUIRE(wire >= 0 && wire < n, "Pauli word wire out of range");
        const uint64_t bit = 1ull << (n - 1 - wire);
        switch (paulis[j]) {
        case 'I': break;
        case 'X': flip |= bit; break;
        case 'Y': flip |= bit; zy |= bit; ny++; break;
        case 'Z': zy |= bit; break;
        default: LGH_REQUIRE(false, "unknown Pauli factor (words use I, X, Y, Z)");
        }
    }
}

// Tile geometry of the Pauli passes on an n-qubit state.
struct PGeo {
    int g;          // tile bits
    int fixed;      // lowest index bits in every tile
    int chosen_max; // bits chosen per pass
};
PGeo pauli_geo(int n) {
    PGeo p;
    p.g = std::min(kPTileBits, n);
    p.fixed = std::min(kPFixedBits, n);
    p.chosen_max = p.g - p.fixed;
    return p;
}

// Groups words into tile passes without allocating: pass_of[w] gets the pass
// of word w (-1: its flip bits do not fit a tile), pass_mask[p] the chosen
// bits of pass p; returns the number of passes.  Each pass is filled first
// with the uncovered words of most flip bits that still fit the pass's
// budget of chosen bits (first fit, decreasing), then takes every
// uncovered word whose flip bits lie in the tile.
int64_t plan_pauli_tiles(int n, int64_t nwords, const uint64_t *flip, int32_t *pass_of, uint64_t *pass_mask) {
    const PGeo geo = pauli_geo(n);
    const uint64_t fixed_mask = (1ull << geo.fixed) - 1ull;
    int64_t left = 0;
    for (int64_t w = 0; w < nwords; ++w) {
        const int c = __builtin_popcountll(flip[w] & ~fixed_mask);
        pass_of[w] = c <= geo.chosen_max ? -2 : -1; // -2: still to place
        left += pass_of[w] == -2;
    }
    int64_t np = 0;
    while (left > 0) {
        uint64_t chosen = 0;
        for (int c = geo.chosen_max; c >= 0; --c)
            for (int64_t w = 0; w < nwords; ++w) {
                if (pass_of[w] != -2) continue;
                const uint64_t rest = flip[w] & ~fixed_mask;
                if (__builtin_popcountll(rest) == c && __builtin_popcountll(chosen | rest) <= geo.chosen_max)
                    chosen |= rest;
            }
        for (int b = geo.fixed; b < n && __builtin_popcountll(chosen) < geo.chosen_max; ++b) chosen |= 1ull << b;
        for (int64_t w = 0; w < nwords; ++w)
            if (pass_of[w] == -2 && ((flip[w] & ~fixed_mask) & ~chosen) == 0) {
                pass_of[w] = int32_t(np);
                --left;
            }
        pass_mask[np++] = chosen;
    }
    return np;
}

struct PauliBufs {
    double *partials; // device, partials_cap doubles
    size_t partials_cap;
    double *results;  // device, one double per word
};

// Device doubles of reduction partials the Pauli passes of an n-qubit state need.
size_t pauli_partials_doubles(int n, int device) {
    const uint64_t len = 1ull << n;
    const unsigned pair_blocks = chunks_for(len, device).blocks;
    const unsigned tile_blocks = unsigned(std::max(1, sm_count(device))) * 4;
    return std::max(size_t(kMaxWordsPerPass) * pair_blocks, size_t(kPassWords) * tile_blocks);
}

// Launches every pass of the words (tile passes first, in pass order, then
// pair passes in ascending flip mask) and their reductions into
// bufs.results; slot_word[s] is the word whose value lands in result slot s.
// Needs no allocation: the caller provides every array (nwords entries each).
template <class R>
void pauli_run(const lgh_sv *sv, int64_t nwords, const uint64_t *flip, const uint64_t *zy, const int32_t *ny,
               int32_t *pass_of, uint64_t *pass_mask, bool *done, int32_t *slot_word, const PauliBufs &bufs) {
    StreamCtx *ctx = sv->ctx;
    const int n = sv->n;
    const uint64_t len = sv->length();
    const PGeo geo = pauli_geo(n);
    const int64_t npasses = plan_pauli_tiles(n, nwords, flip, pass_of, pass_mask);
    int64_t nslots = 0;
    // tile passes
    const uint64_t ntiles = len >> geo.g;
    const unsigned max_blocks = unsigned(std::max(1, sm_count(sv->device))) * 4;
    const uint64_t per_block = (ntiles + max_blocks - 1) / max_blocks;
    const unsigned tblocks = unsigned((ntiles + per_block - 1) / per_block);
    LGH_REQUIRE(size_t(kPassWords) * tblocks <= bufs.partials_cap, "Pauli partials buffer too small");
    for (int64_t p = 0; p < npasses; ++p) {
        PTile tm{};
        tm.g = geo.g;
        int r = 0;
        for (int b = 0; b < geo.fixed; ++b) tm.bit[r++] = b;
        for (int b = geo.fixed; b < n; ++b)
            if ((pass_mask[p] >> b) & 1) tm.bit[r++] = b;
        LGH_REQUIRE(r == geo.g, "Pauli tile: bad bit count");
        Fixed tfx{};
        tfx.cnt = geo.g;
        for (int q = 0; q < geo.g; ++q) tfx.pos[q] = tm.bit[q]; // ascending
        PWords pw{};
        auto launch = [&] {
            hipLaunchKernelGGL(k_pauli_tile<R>, dim3(tblocks), dim3(kBlock), 0, ctx->stream, cdata<R>(sv), ntiles,
                               per_block, tfx, tm, pw, bufs.partials);
            LGH_LAUNCH_CHECK();
            launch_reduce(ctx, bufs.partials, int(tblocks), pw.count, bufs.results + (nslots - pw.count));
            pw.count = 0;
        };
        for (int64_t w = 0; w < nwords; ++w) {
            if (pass_of[w] != p) continue;
            uint32_t fl = 0, zl = 0;
            uint64_t tile_mask = 0;
            for (int q = 0; q < geo.g; ++q) {
                if ((flip[w] >> tm.bit[q]) & 1) fl |= 1u << q;
                if ((zy[w] >> tm.bit[q]) & 1) zl |= 1u << q;
                tile_mask |= 1ull << tm.bit[q];
            }
            pw.flip[pw.count] = fl;
            pw.zyl[pw.count] = zl;
            pw.zy[pw.count] = zy[w] & ~tile_mask;
            pw.phase[pw.count] = uint32_t(ny[w] & 3);
            ++pw.count;
            slot_word[nslots++] = int32_t(w);
            if (pw.count == kPassWords) launch();
        }
        if (pw.count > 0) launch();
    }
    // words wider than a tile: one pair pass per flip mask, ascending
    for (int64_t w = 0; w < nwords; ++w) done[w] = pass_of[w] >= 0;
    for (;;) {
        bool any = false;
        uint64_t f = 0;
        for (int64_t w = 0; w < nwords; ++w)
            if (!done[w] && (!any || flip[w] < f)) {
                f = flip[w];
                any = true;
            }
        if (!any) break;
        const int pivot = f ? __builtin_ctzll(f) : -1;
        const uint64_t count = f ? len / 2 : len;
        const Chunks ch = chunks_for(count, sv->device);
        LGH_REQUIRE(size_t(kMaxWordsPerPass) * ch.blocks <= bufs.partials_cap, "Pauli partials buffer too small");
        WordSet ws{};
        auto launch = [&] {
            hipLaunchKernelGGL(k_pauli<R>, dim3(ch.blocks), dim3(kBlock), 0, ctx->stream, cdata<R>(sv), count,
                               ch.chunk, pivot, f, ws, bufs.partials);
            LGH_LAUNCH_CHECK();
            launch_reduce(ctx, bufs.partials, ch.blocks, ws.count, bufs.results + (nslots - ws.count));
            ws.count = 0;
        };
        for (int64_t w = 0; w < nwords; ++w) {
            if (done[w] || flip[w] != f) continue;
            done[w] = true;
            ws.zy[ws.count] = zy[w];
            ws.phase[ws.count] = uint32_t(ny[w] & 3);
            ++ws.count;
            slot_word[nslots++] = int32_t(w);
            if (ws.count == kMaxWordsPerPass) launch();
        }
        if (ws.count > 0) launch();
    }
    LGH_REQUIRE(nslots == nwords, "Pauli words: every word gets one result slot");
}

template <class R>
void pauli_t(const lgh_sv *sv, int64_t nwords, const char *paulis, const int32_t *wires, const int64_t *offsets,
             const double *coeffs, double *per_word, double *result) {
    StreamCtx *ctx = sv->ctx;
    const size_t nw = size_t(nwords);
    std::vector<uint64_t> flip(nw), zy(nw), pass_mask(std::max<size_t>(1, nw));
    std::vector<int32_t> ny(nw), pass_of(nw), slot_word(nw);
    std::unique_ptr<bool[]> done(new bool[std::max<size_t>(1, nw)]);
    for (int64_t w = 0; w < nwords; ++w)
        parse_word(sv->n, paulis, wires, offsets[w], offsets[w + 1], flip[size_t(w)], zy[size_t(w)], ny[size_t(w)]);
    const size_t pcap = pauli_partials_doubles(sv->n, sv->device);
    char *ws_base = static_cast<char *>(ctx->workspace(sizeof(double) * (pcap + nw) + 256));
    PauliBufs bufs{reinterpret_cast<double *>(ws_base), pcap, reinterpret_cast<double *>(ws_base) + pcap};
    if (nwords > 0)
        pauli_run<R>(sv, nwords, flip.data(), zy.data(), ny.data(), pass_of.data(), pass_mask.data(), done.get(),
                     slot_word.data(), bufs);
    std::vector<double> out(std::max<size_t>(1, nw)), vals(std::max<size_t>(1, nw));
    if (nwords > 0) fetch(ctx, bufs.results, int(nwords), out.data());
    for (size_t q = 0; q < nw; ++q) vals[size_t(slot_word[q])] = out[q];
    *result = weighted_total(nwords, coeffs, vals.data(), per_word);
}

// ---------------------------------------------------------------- Pauli words without allocation
//
// lgh_expval_pauli_words_noalloc runs the passes of pauli_t (same kernels,
// same grouping, same chunking, so the same results bit for bit) with its
// reduction partials and word results in `scratch`, a device array of fixed
// size that every call on one device shares.  The host side keeps its
// bookkeeping in fixed-size arrays on the stack.  Nothing is allocated after
// the first call on a device, which looks up the array's device address (and
// the device's CU count used to size the passes).
//
// Layout of `scratch` for one call: reduction partials of one launch |
// one result per word.  Every entry a call reads was written earlier in the
// same call (each launch writes all its partials, each reduction its
// words' results), so nothing needs clearing; two calls that overlap on one
// device race on the array.

constexpr size_t kNoAllocScratchBytes = size_t(512) << 10;
constexpr int64_t kNoAllocMaxWords = 1024;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up once
// per device.  The array's host-side name is not a device address.
void *scratch_on(int device) {
    LGH_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        LGH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        LGH_REQUIRE(p != nullptr, "scratch array has no device address");
        (void)sm_count(device); // cached from here on: later calls only read it
        g_scratch_dev[device] = p;
    }
    return g_scratch_dev[device];
}

template <class R>
void pauli_noalloc_t(const lgh_sv *sv, int64_t nwords, const char *paulis, const int32_t *wires,
                     const int64_t *offsets, const double *coeffs, double *per_word, double *result) {
    LGH_REQUIRE(nwords <= kNoAllocMaxWords, "lgh_expval_pauli_words_noalloc takes at most " +
                                                std::to_string(kNoAllocMaxWords) + " words (got " +
                                                std::to_string(nwords) + ")");
    StreamCtx *ctx = sv->ctx;
    char *base = static_cast<char *>(scratch_on(sv->device));
    uint64_t flip[kNoAllocMaxWords], zy[kNoAllocMaxWords], pass_mask[kNoAllocMaxWords];
    int32_t ny[kNoAllocMaxWords], pass_of[kNoAllocMaxWords], slot_word[kNoAllocMaxWords];
    bool done[kNoAllocMaxWords];
    for (int64_t w = 0; w < nwords; ++w) parse_word(sv->n, paulis, wires, offsets[w], offsets[w + 1], flip[w], zy[w], ny[w]);
    const size_t pcap = pauli_partials_doubles(sv->n, sv->device);
    const size_t need = sizeof(double) * (pcap + size_t(nwords));
    LGH_REQUIRE(need <= kNoAllocScratchBytes, "lgh_expval_pauli_words_noalloc: " + std::to_string(need) +
                                                  " bytes of scratch needed, " +
                                                  std::to_string(kNoAllocScratchBytes) + " available");
    PauliBufs bufs{reinterpret_cast<double *>(base), pcap, reinterpret_cast<double *>(base) + pcap};
    if (nwords > 0) pauli_run<R>(sv, nwords, flip, zy, ny, pass_of, pass_mask, done, slot_word, bufs);
    double out[kNoAllocMaxWords];
    double vals[kNoAllocMaxWords];
    if (nwords > 0) {
        LGH_HIP(hipMemcpyAsync(out, bufs.results, sizeof(double) * size_t(nwords), hipMemcpyDeviceToHost,
                               ctx->stream));
        ctx->sync();
    }
    for (int64_t q = 0; q < nwords; ++q) vals[slot_word[q]] = out[q];
    *result = weighted_total(nwords, coeffs, vals, per_word);
}

template <class I, class R>
void csr_launch(const lgh_csr *csr, const cx<R> *x, cx<R> *y, StreamCtx *ctx, double *partials, unsigned blocks) {
    const I *ip = static_cast<const I *>(csr->indptr);
    const I *ix = static_cast<const I *>(csr->indices);
    const cx<R> *vv = static_cast<const cx<R> *>(csr->values);
    const double avg = csr->avg_row;
    if (avg <= 6)
        hipLaunchKernelGGL((k_csr<I, R, 4>), dim3(blocks), dim3(kBlock), 0, ctx->stream, ip, ix, vv, x, y, csr->rows,
                           partials);
    else if (avg <= 12)
        hipLaunchKernelGGL((k_csr<I, R, 8>), dim3(blocks), dim3(kBlock), 0, ctx->stream, ip, ix, vv, x, y, csr->rows,
                           partials);
    else if (avg <= 24)
        hipLaunchKernelGGL((k_csr<I, R, 16>), dim3(blocks), dim3(kBlock), 0, ctx->stream, ip, ix, vv, x, y,
                           csr->rows, partials);
    else
        hipLaunchKernelGGL((k_csr<I, R, 32>), dim3(blocks), dim3(kBlock), 0, ctx->stream, ip, ix, vv, x, y,
                           csr->rows, partials);
    LGH_LAUNCH_CHECK();
}

template <class R> void csr_run(const lgh_csr *csr, const lgh_sv *x, lgh_sv *y, double *re, double *im) {
    StreamCtx *ctx = x->ctx;
    const unsigned blocks = unsigned(std::max<int64_t>(
        1, std::min<int64_t>((csr->rows * 16 + kBlock - 1) / kBlock, int64_t(sm_count(x->device)) * 8)));
    if (y) {
        if (csr->idx64)
            csr_launch<int64_t, R>(csr, cdata<R>(x), mdata<R>(y), ctx, nullptr, blocks);
        else
            csr_launch<int32_t, R>(csr, cdata<R>(x), mdata<R>(y), ctx, nullptr, blocks);
        return;
    }
    RedBuf rb = red_buf(ctx, 2, int(blocks));
    if (csr->idx64)
        csr_launch<int64_t, R>(csr, cdata<R>(x), nullptr, ctx, rb.partials, blocks);
    else
        csr_launch<int32_t, R>(csr, cdata<R>(x), nullptr, ctx, rb.partials, blocks);
    launch_reduce(ctx, rb.partials, int(blocks), 2, rb.result);
    double out[2];
    fetch(ctx, rb.result, 2, out);
    *re = out[0];
    *im = out[1];
}

template <class R> void probs_t(const lgh_sv *sv, const int32_t *wires, int k, double *out) {
    StreamCtx *ctx = sv->ctx;
    const int n = sv->n;
    WireBits wb{};
    wb.k = k;
    std::vector<bool> used(size_t(n), false);
    for (int j = 0; j < k; ++j) {
        wb.b[j] = n - 1 - wires[j];
        used[size_t(wb.b[j])] = true;
    }
    const int nb = 1 << k;
    if (k <= 10) {
        Chunks ch = chunks_for(sv->length(), sv->device, 8);
        RedBuf rb = red_buf(ctx, nb, ch.blocks);
        hipLaunchKernelGGL(k_probs_hist<R>, dim3(ch.blocks), dim3(kBlock), 0, ctx->stream, cdata<R>(sv),
                           sv->length(), ch.chunk, wb, rb.partials);
        LGH_LAUNCH_CHECK();
        launch_reduce(ctx, rb.partials, ch.blocks, nb, rb.result);
        fetch(ctx, rb.result, nb, out);
        return;
    }
    Scatter sc{};
    for (int b = 0; b < n; ++b)
        if (!used[size_t(b)]) sc.rest[sc.nrest++] = b;
    double *dout = static_cast<double *>(ctx->workspace(sizeof(double) * size_t(nb)));
    hipLaunchKernelGGL(k_probs_bins<R>, dim3(unsigned((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       cdata<R>(sv), wb, sc, dout);
    LGH_LAUNCH_CHECK();
    LGH_HIP(hipMemcpyAsync(out, dout, sizeof(double) * size_t(nb), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
}

// Sampler workspace: [cdf (len) | tile sums | total | indices (shots) | sorted
// (shots) | radix-sort temporary].  Returns the device pointer of the
// sorted indices.
template <class R> uint64_t *sample_t(const lgh_sv *sv, uint64_t shots, uint64_t seed) {
    StreamCtx *ctx = sv->ctx;
    const uint64_t len = sv->length();
    const uint64_t ntiles = (len + kScanTile - 1) / kScanTile;
    size_t temp_bytes = 0;
    LGH_HIP(rocprim::radix_sort_keys(nullptr, temp_bytes, static_cast<uint64_t *>(nullptr),
                                     static_cast<uint64_t *>(nullptr), size_t(shots), 0u, unsigned(sv->n),
                                     ctx->stream));
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_tiles = al(len * 8), o_total = o_tiles + al(ntiles * 8), o_idx = o_total + 256,
                 o_sorted = o_idx + al(shots * 8), o_temp = o_sorted + al(shots * 8);
    char *ws = static_cast<char *>(ctx->workspace(o_temp + temp_bytes + 256));
    uint64_t *cdf = reinterpret_cast<uint64_t *>(ws);
    uint64_t *tiles = reinterpret_cast<uint64_t *>(ws + o_tiles);
    uint64_t *total = reinterpret_cast<uint64_t *>(ws + o_total);
    uint64_t *idx = reinterpret_cast<uint64_t *>(ws + o_idx);
    uint64_t *sorted = reinterpret_cast<uint64_t *>(ws + o_sorted);
    hipLaunchKernelGGL(k_scan_tile_sums<R>, dim3(unsigned(ntiles)), dim3(kBlock), 0, ctx->stream, cdata<R>(sv), len,
                       tiles);
    hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, ctx->stream, tiles, ntiles, total);
    hipLaunchKernelGGL(k_scan_cdf<R>, dim3(unsigned(ntiles)), dim3(kBlock), 0, ctx->stream, cdata<R>(sv), len, tiles,
                       cdf);
    LGH_LAUNCH_CHECK();
    // a state with no probability mass cannot be sampled
    uint64_t htotal = 0;
    LGH_HIP(hipMemcpyAsync(&htotal, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    LGH_REQUIRE(htotal > 0, "cannot sample from a zero state");
    hipLaunchKernelGGL(k_sample_search, dim3(unsigned((shots + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       cdf, len, total, shots, seed, idx);
    LGH_LAUNCH_CHECK();
    LGH_HIP(rocprim::radix_sort_keys(ws + o_temp, temp_bytes, idx, sorted, size_t(shots), 0u, unsigned(sv->n),
                                     ctx->stream));
    return sorted;
}

template <class F> void by_dtype(const lgh_sv *sv, F &&f) {
    if (sv->dtype == LGH_C128)
        f(double{});
    else
        f(float{});
}

} // namespace

extern "C" {

lgh_status lgh_inner(const lgh_sv *a, const lgh_sv *b, double *re, double *im) {
    return guard([&] {
        LGH_REQUIRE(a && b && re && im, "null pointer");
        LGH_REQUIRE(a->n == b->n && a->dtype == b->dtype, "states differ in size or precision");
        DeviceGuard g(a->device);
        if (b->ctx != a->ctx) b->ctx->sync();
        by_dtype(a, [&](auto r) { inner_t<decltype(r)>(a, b, re, im); });
    });
}

lgh_status lgh_axpy(lgh_sv *y, double are, double aim, const lgh_sv *x) {
    return guard([&] {
        LGH_REQUIRE(x && y, "null state");
        LGH_REQUIRE(x->n == y->n && x->dtype == y->dtype, "states differ in size or precision");
        DeviceGuard g(y->device);
        if (x->ctx != y->ctx) x->ctx->sync();
        const unsigned blocks = unsigned((y->length() + kBlock - 1) / kBlock);
        by_dtype(y, [&](auto r) {
            using R = decltype(r);
            hipLaunchKernelGGL(k_axpy<R>, dim3(blocks), dim3(kBlock), 0, y->ctx->stream, mdata<R>(y), cdata<R>(x),
                               y->length(), cx<R>{R(are), R(aim)});
        });
        LGH_LAUNCH_CHECK();
    });
}

lgh_status lgh_sv_pauli_word_from(lgh_sv *dst, const lgh_sv *src, const char *paulis, const int32_t *wires,
                                  int32_t num_wires) {
    return guard([&] {
        LGH_REQUIRE(dst && src, "null state");
        LGH_REQUIRE(dst != src && dst->data != src->data, "source and destination must differ");
        LGH_REQUIRE(dst->n == src->n && dst->dtype == src->dtype, "states differ in size or precision");
        LGH_REQUIRE(num_wires >= 0 && (num_wires == 0 || (paulis && wires)), "null Pauli word");
        const int n = dst->n;
        uint64_t xmask = 0, zmask = 0, seen = 0;
        int ny = 0;
        for (int32_t k = 0; k < num_wires; ++k) {
            LGH_REQUIRE(wires[k] >= 0 && wires[k] < n, "wire out of range");
            const uint64_t bit = 1ull << (n - 1 - wires[k]);
            LGH_REQUIRE(!(seen & bit), "repeated wire in a Pauli word");
            seen |= bit;
            switch (paulis[k]) {
            case 'I': break;
            case 'X': xmask |= bit; break;
            case 'Y': xmask |= bit; zmask |= bit; ++ny; break;
            case 'Z': zmask |= bit; break;
            default: throw Error(LGH_ERR_INVALID_ARGUMENT, "Pauli word letters are I, X, Y, Z");
            }
        }
        DeviceGuard g(dst->device);
        if (src->ctx != dst->ctx) src->ctx->sync();
        const uint64_t len = dst->length();
        unsigned blocks = unsigned((len + kBlock - 1) / kBlock);
        if (blocks >= 8) blocks = (blocks + 7) / 8 * 8; // whole XCD rounds for the block order
        by_dtype(dst, [&](auto r) {
            using R = decltype(r);
            hipLaunchKernelGGL(k_pauli_from<R>, dim3(blocks), dim3(kBlock), 0, dst->ctx->stream, mdata<R>(dst),
                               cdata<R>(src), len, xmask, zmask, ny & 3);
        });
        LGH_LAUNCH_CHECK();
    });
}

lgh_status lgh_scale(lgh_sv *x, double are, double aim) {
    return guard([&] {
        LGH_REQUIRE(x, "null state");
        DeviceGuard g(x->device);
        const unsigned blocks = unsigned((x->length() + kBlock - 1) / kBlock);
        by_dtype(x, [&](auto r) {
            using R = decltype(r);
            hipLaunchKernelGGL(k_scale<R>, dim3(blocks), dim3(kBlock), 0, x->ctx->stream, mdata<R>(x), x->length(),
                               cx<R>{R(are), R(aim)});
        });
        LGH_LAUNCH_CHECK();
    });
}

lgh_status lgh_expval_matrix(const lgh_sv *sv, const int32_t *wires, int32_t num_wires, const double *matrix,
                             double *re, double *im) {
    return guard([&] {
        Range rg("lgh_expval_matrix");
        LGH_REQUIRE(sv && wires && matrix && re && im, "null pointer");
        LGH_REQUIRE(num_wires >= 1 && num_wires <= kMaxLdsK && num_wires <= sv->n,
                    "observable must act on 1..13 wires of the state");
        for (int i = 0; i < num_wires; ++i) {
            LGH_REQUIRE(wires[i] >= 0 && wires[i] < sv->n, "wire index out of range");
            for (int j = 0; j < i; ++j) LGH_REQUIRE(wires[i] != wires[j], "repeated wire");
        }
        DeviceGuard g(sv->device);
        by_dtype(sv, [&](auto r) { expval_t<decltype(r)>(sv, wires, num_wires, matrix, re, im); });
    });
}

lgh_status lgh_expval_pauli_words(const lgh_sv *sv, int64_t num_words, const char *paulis, const int32_t *wires,
                                  const int64_t *offsets, const double *coeffs, double *per_word, double *result) {
    return guard([&] {
        Range rg("lgh_expval_pauli_words");
        LGH_REQUIRE(sv && result, "null pointer");
        LGH_REQUIRE(num_words >= 0, "negative word count");
        LGH_REQUIRE(num_words == 0 || (paulis && offsets && coeffs), "null pointer");
        for (int64_t w = 0; w < num_words; ++w)
            LGH_REQUIRE(offsets[w + 1] >= offsets[w] && (offsets[w + 1] == offsets[w] || wires),
                        "bad Pauli word offsets");
        DeviceGuard g(sv->device);
        by_dtype(sv, [&](auto r) {
            pauli_t<decltype(r)>(sv, num_words, paulis, wires, offsets, coeffs, per_word, result);
        });
    });
}

lgh_status lgh_pauli_plan(int32_t num_qubits, int64_t num_words, const char *paulis, const int32_t *wires,
                          const int64_t *offsets, int32_t *tile_passes, int32_t *tile_launches,
                          int32_t *pair_launches) {
    return guard([&] {
        LGH_REQUIRE(num_qubits >= 1 && num_qubits <= 36, "num_qubits must be in [1, 36]");
        LGH_REQUIRE(num_words >= 0 && (num_words == 0 || (paulis && offsets)), "bad word list");
        const size_t nw = size_t(num_words);
        std::vector<uint64_t> flip(nw), zy(nw), mask(std::max<size_t>(1, nw));
        std::vector<int32_t> ny(nw), pass_of(nw);
        for (int64_t w = 0; w < num_words; ++w) {
            LGH_REQUIRE(offsets[w + 1] >= offsets[w] && (offsets[w + 1] == offsets[w] || wires), "bad Pauli word offsets");
            parse_word(num_qubits, paulis, wires, offsets[w], offsets[w + 1], flip[size_t(w)], zy[size_t(w)],
                       ny[size_t(w)]);
        }
        const int64_t np = num_words ? plan_pauli_tiles(num_qubits, num_words, flip.data(), pass_of.data(), mask.data()) : 0;
        std::vector<int64_t> per(size_t(std::max<int64_t>(1, np)), 0);
        std::map<uint64_t, int64_t> pair;
        for (size_t w = 0; w < nw; ++w) {
            if (pass_of[w] >= 0)
                ++per[size_t(pass_of[w])];
            else
                ++pair[flip[w]];
        }
        int64_t tl = 0, pl = 0;
        for (int64_t p = 0; p < np; ++p) tl += (per[size_t(p)] + kPassWords - 1) / kPassWords;
        for (auto &kv : pair) pl += (kv.second + kMaxWordsPerPass - 1) / kMaxWordsPerPass;
        if (tile_passes) *tile_passes = int32_t(np);
        if (tile_launches) *tile_launches = int32_t(tl);
        if (pair_launches) *pair_launches = int32_t(pl);
    });
}

lgh_status lgh_expval_pauli_words_noalloc(const lgh_sv *sv, int64_t num_words, const char *paulis,
                                          const int32_t *wires, const int64_t *offsets, const double *coeffs,
                                          double *per_word, double *result) {
    return guard([&] {
        LGH_REQUIRE(sv && result, "null pointer");
        LGH_REQUIRE(num_words >= 0, "negative word count");
        LGH_REQUIRE(num_words == 0 || (paulis && offsets && coeffs), "null pointer");
        for (int64_t w = 0; w < num_words; ++w)
            LGH_REQUIRE(offsets[w + 1] >= offsets[w] && (offsets[w + 1] == offsets[w] || wires),
                        "bad Pauli word offsets");
        DeviceGuard g(sv->device);
        by_dtype(sv, [&](auto r) {
            pauli_noalloc_t<decltype(r)>(sv, num_words, paulis, wires, offsets, coeffs, per_word, result);
        });
    });
}

lgh_status lgh_csr_create(lgh_csr **out, int device, lgh_dtype dtype, int64_t num_rows, const int64_t *indptr,
                          const int64_t *indices, const double *values, int64_t nnz) {
    return guard([&] {
        LGH_REQUIRE(out && indptr && (nnz == 0 || (indices && values)), "null pointer");
        LGH_REQUIRE(num_rows >= 1, "CSR matrix needs at least one row");
        LGH_REQUIRE(indptr[0] == 0 && indptr[num_rows] == nnz, "CSR offsets do not match nnz");
        for (int64_t r = 0; r < num_rows; ++r) LGH_REQUIRE(indptr[r + 1] >= indptr[r], "CSR offsets must not decrease");
        for (int64_t k = 0; k < nnz; ++k)
            LGH_REQUIRE(indices[k] >= 0 && indices[k] < num_rows, "CSR column index out of range");
        auto c = std::make_unique<lgh_csr>();
        c->device = device;
        c->dtype = dtype;
        c->rows = num_rows;
        c->nnz = nnz;
        c->avg_row = double(nnz) / double(num_rows);
        c->idx64 = dtype == LGH_C128 || num_rows >= (int64_t(1) << 31) || nnz >= (int64_t(1) << 31);
        DeviceGuard g(device);
        const size_t isz = c->idx64 ? 8 : 4;
        c->b_ptr = isz * size_t(num_rows + 1);
        c->b_idx = isz * size_t(std::max<int64_t>(nnz, 1));
        c->b_val = (dtype == LGH_C128 ? 16 : 8) * size_t(std::max<int64_t>(nnz, 1));
        LGH_HIP(hipMalloc(&c->indptr, c->b_ptr));
        LGH_HIP(hipMalloc(&c->indices, c->b_idx));
        LGH_HIP(hipMalloc(&c->values, c->b_val));
        if (c->idx64) {
            LGH_HIP(hipMemcpy(c->indptr, indptr, c->b_ptr, hipMemcpyHostToDevice));
            if (nnz) LGH_HIP(hipMemcpy(c->indices, indices, 8 * size_t(nnz), hipMemcpyHostToDevice));
        } else {
            std::vector<int32_t> p(indptr, indptr + num_rows + 1), ix(indices, indices + nnz);
            LGH_HIP(hipMemcpy(c->indptr, p.data(), c->b_ptr, hipMemcpyHostToDevice));
            if (nnz) LGH_HIP(hipMemcpy(c->indices, ix.data(), 4 * size_t(nnz), hipMemcpyHostToDevice));
        }
        if (dtype == LGH_C128) {
            if (nnz) LGH_HIP(hipMemcpy(c->values, values, 16 * size_t(nnz), hipMemcpyHostToDevice));
        } else {
            std::vector<float> v(2 * size_t(nnz));
            for (size_t i = 0; i < v.size(); ++i) v[i] = float(values[i]);
            if (nnz) LGH_HIP(hipMemcpy(c->values, v.data(), 8 * size_t(nnz), hipMemcpyHostToDevice));
        }
        *out = c.release();
    });
}

lgh_status lgh_csr_destroy(lgh_csr *csr) {
    return guard([&] {
        if (!csr) return;
        DeviceGuard g(csr->device);
        (void)hipDeviceSynchronize();
        (void)hipFree(csr->indptr);
        (void)hipFree(csr->indices);
        (void)hipFree(csr->values);
        delete csr;
    });
}

lgh_status lgh_expval_csr(const lgh_sv *sv, const lgh_csr *csr, double *re, double *im) {
    return guard([&] {
        Range rg("lgh_expval_csr");
        LGH_REQUIRE(sv && csr && re && im, "null pointer");
        LGH_REQUIRE(csr->rows == int64_t(sv->length()), "CSR dimension does not match the state");
        LGH_REQUIRE(csr->dtype == sv->dtype && csr->device == sv->device, "CSR precision/device mismatch");
        DeviceGuard g(sv->device);
        by_dtype(sv, [&](auto r) { csr_run<decltype(r)>(csr, sv, nullptr, re, im); });
    });
}

lgh_status lgh_apply_csr(lgh_sv *out, const lgh_csr *csr, const lgh_sv *in) {
    return guard([&] {
        LGH_REQUIRE(out && csr && in, "null pointer");
        LGH_REQUIRE(out != in, "in-place CSR product is not supported");
        LGH_REQUIRE(csr->rows == int64_t(in->length()) && out->n == in->n && out->dtype == in->dtype,
                    "CSR dimension does not match the state");
        LGH_REQUIRE(csr->dtype == in->dtype && csr->device == in->device, "CSR precision/device mismatch");
        DeviceGuard g(in->device);
        if (in->ctx != out->ctx) in->ctx->sync();
        by_dtype(in, [&](auto r) { csr_run<decltype(r)>(csr, in, out, nullptr, nullptr); });
    });
}

lgh_status lgh_probs(const lgh_sv *sv, const int32_t *wires, int32_t num_wires, double *out) {
    return guard([&] {
        Range rg("lgh_probs");
        LGH_REQUIRE(sv && out && (num_wires == 0 || wires), "null pointer");
        LGH_REQUIRE(num_wires >= 0 && num_wires <= sv->n && num_wires <= 30, "bad wire count");
        for (int i = 0; i < num_wires; ++i) {
            LGH_REQUIRE(wires[i] >= 0 && wires[i] < sv->n, "wire index out of range");
            for (int j = 0; j < i; ++j) LGH_REQUIRE(wires[i] != wires[j], "repeated wire");
        }
        DeviceGuard g(sv->device);
        by_dtype(sv, [&](auto r) { probs_t<decltype(r)>(sv, wires, num_wires, out); });
    });
}

lgh_status lgh_sample(const lgh_sv *sv, uint64_t shots, uint64_t seed, uint64_t *indices) {
    return guard([&] {
        Range rg("lgh_sample");
        LGH_REQUIRE(sv && (shots == 0 || indices), "null pointer");
        LGH_REQUIRE(shots < (1ull << 32), "too many shots in one call");
        if (shots == 0) return;
        DeviceGuard g(sv->device);
        uint64_t *d = nullptr;
        by_dtype(sv, [&](auto r) { d = sample_t<decltype(r)>(sv, shots, seed); });
        LGH_HIP(hipMemcpyAsync(indices, d, 8 * shots, hipMemcpyDeviceToHost, sv->ctx->stream));
        sv->ctx->sync();
    });
}

lgh_status lgh_sample_bits(const lgh_sv *sv, uint64_t shots, uint64_t seed, uint64_t *bits) {
    return guard([&] {
        LGH_REQUIRE(sv && (shots == 0 || bits), "null pointer");
        LGH_REQUIRE(shots < (1ull << 32), "too many shots in one call");
        if (shots == 0) return;
        DeviceGuard g(sv->device);
        uint64_t *d = nullptr;
        by_dtype(sv, [&](auto r) { d = sample_t<decltype(r)>(sv, shots, seed); });
        const uint64_t total = shots * uint64_t(sv->n);
        // the expanded bits go to a separate buffer (the workspace holds the indices)
        uint64_t *dbits = static_cast<uint64_t *>(pool_alloc(sv->device, 8 * total));
        hipLaunchKernelGGL(k_expand_bits, dim3(unsigned((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           sv->ctx->stream, d, shots, sv->n, dbits);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(bits, dbits, 8 * total, hipMemcpyDeviceToHost, sv->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(sv->ctx->stream);
        pool_free(sv->device, dbits, 8 * total);
        LGH_HIP(e);
    });
}

} // extern "C"
