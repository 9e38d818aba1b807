// Internal declarations shared by the HIP translation units of liblgh.so.
// Nothing here is part of the C ABI (include/lgh.h).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "lgh.h"

namespace lgh {

// ---------------------------------------------------------------- errors

struct Error : std::runtime_error {
    lgh_status code;
    Error(lgh_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define LGH_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t lgh_e_ = (expr);                                                     \
        if (lgh_e_ != hipSuccess)                                                       \
            throw ::lgh::Error(lgh_e_ == hipErrorOutOfMemory ? LGH_ERR_OUT_OF_MEMORY    \
                                                             : LGH_ERR_DEVICE,          \
                               std::string(#expr) + ": " + hipGetErrorString(lgh_e_)); \
    } while (0)

#define LGH_REQUIRE(cond, msg)                                                  \
    do {                                                                        \
        if (!(cond)) throw ::lgh::Error(LGH_ERR_INVALID_ARGUMENT, (msg));       \
    } while (0)

#define LGH_LAUNCH_CHECK() LGH_HIP(hipGetLastError())

template <class F> lgh_status guard(F &&f) {
    try {
        f();
        return LGH_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return LGH_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return LGH_ERR_INVALID_ARGUMENT;
    }
}

// ---------------------------------------------------------------- device types

// Interleaved complex amplitude; 16-byte aligned for complex128 so that one
// amplitude is one global_load_dwordx4.
template <class R> struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <class R> __host__ __device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <class R> __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
    return cx<R>{a.re + b.re, a.im + b.im};
}
template <class R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * b + c
template <class R> __device__ __forceinline__ cx<R> cfma(cx<R> a, cx<R> b, cx<R> c) {
    return cx<R>{c.re + a.re * b.re - a.im * b.im, c.im + a.re * b.im + a.im * b.re};
}

template <class R> cx<R> to_cx(std::complex<double> z) {
    return cx<R>{static_cast<R>(z.real()), static_cast<R>(z.imag())};
}

// Streaming accesses of whole amplitudes (one 16 B / 8 B vector access) with
// the nontemporal hint: every state pass touches each amplitude once, so
// keeping it out of the caches leaves them to the data that is re-used.
template <class R> using vec2_t = R __attribute__((ext_vector_type(2)));

// State vectors live in global memory: the accesses name that address space,
// so that they compile to global (not flat) instructions even where the
// pointer went through an opaque register copy.  A flat access also counts
// as an LDS / scalar-memory access, and every wait for a descriptor load
// would then wait for the tile's stores as well.
template <class R> using gvec2_t = __attribute__((address_space(1))) vec2_t<R>;

template <class R> __device__ __forceinline__ cx<R> ldnt(const cx<R> *p) {
    const vec2_t<R> v = __builtin_nontemporal_load((const gvec2_t<R> *)(p));
    return cx<R>{v.x, v.y};
}
template <class R> __device__ __forceinline__ void stnt(cx<R> *p, cx<R> a) {
    vec2_t<R> v;
    v.x = a.re;
    v.y = a.im;
    __builtin_nontemporal_store(v, (gvec2_t<R> *)(p));
}

// XCD-aware workgroup order: the hardware hands consecutive workgroups to the
// 8 XCDs round-robin; remapping so that XCD k processes the k-th contiguous
// eighth of the index space keeps each XCD's memory stream sequential.
__device__ __forceinline__ uint64_t xcd_block() {
    const uint64_t g = gridDim.x;
    if (g % 8 != 0) return blockIdx.x;
    return (blockIdx.x % 8) * (g / 8) + blockIdx.x / 8;
}

// XCD-interleaved order: the 8 XCDs take consecutive runs of `run`
// workgroups in turn (XCD k takes runs k, k + 8, k + 16, ...), so that all
// XCDs stream through one window of the index space at a time instead of
// eight windows an eighth apart.  Falls back to xcd_block() unless `run`
// divides gridDim.x / 8.
__device__ __forceinline__ uint64_t xcd_block_runs(uint64_t run) {
    const uint64_t g = gridDim.x;
    if (run == 0 || g % (8 * run) != 0) return xcd_block();
    const uint64_t k = blockIdx.x % 8, j = blockIdx.x / 8;
    return ((j / run) * 8 + k) * run + (j % run);
}

template <class R> __device__ __forceinline__ cx<R> shfl_xor_cx(cx<R> a, int mask) {
    return cx<R>{__shfl_xor(a.re, mask, 64), __shfl_xor(a.im, mask, 64)};
}

// Value of lane (lane ^ 2^L) without going through LDS: DPP moves inside
// 16-lane rows for L <= 3 -- quad_perm [1,0,3,2] (xor 1), quad_perm
// [2,3,0,1] (xor 2), row_half_mirror (xor 7 in each 8 lanes) followed by
// quad_perm [3,2,1,0] (xor 3), together xor 4, and row_ror:8 (xor 8) -- and
// the gfx950 permlane swaps for L = 4, 5 (odd 16- or 32-lane rows of the first
// operand trade places with even rows of the second).  Every lane reads a
// lane of its own row, so bound_ctrl (0 for out-of-row reads) never applies;
// it lets the move skip initialising its destination.
template <int L> __device__ __forceinline__ int xor_lane_b32(int v, int lane) {
    static_assert(L >= 0 && L < 6, "lane bit");
    if constexpr (L == 0) {
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true); // quad_perm [1,0,3,2]
    } else if constexpr (L == 1) {
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true); // quad_perm [2,3,0,1]
    } else if constexpr (L == 2) {
        const int t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true); // row_half_mirror
        return __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, true);         // quad_perm [3,2,1,0]
    } else if constexpr (L == 3) {
        return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, true); // row_ror:8
    } else if constexpr (L == 4) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return int((lane & 16) ? p[0] : p[1]);
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return int((lane & 32) ? p[0] : p[1]);
    }
}
template <int L> __device__ __forceinline__ float xor_lane(float v, int lane) {
    return __int_as_float(xor_lane_b32<L>(__float_as_int(v), lane));
}
template <int L> __device__ __forceinline__ double xor_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane_b32<L>(int(b & 0xffffffffll), lane);
    const int hi = xor_lane_b32<L>(int(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int L, class R> __device__ __forceinline__ cx<R> xor_lane_cx(cx<R> a, int lane) {
    return cx<R>{xor_lane<L>(a.re, lane), xor_lane<L>(a.im, lane)};
}

// Row swap for lane bit L = 4 or 5 between two registers x, y: afterwards
// every lane holds in (x, y) the two amplitudes that differ only in lane
// bit L (x: bit clear, y: bit set), taken from x in the lanes with the bit
// clear and from y in the others.  Applying it again restores the layout.
template <int L> __device__ __forceinline__ void row_swap_b32(int &x, int &y) {
    static_assert(L == 4 || L == 5, "row swaps exist for lane bits 4 and 5");
    if constexpr (L == 4) {
        const auto p = __builtin_amdgcn_permlane16_swap(x, y, false, false);
        x = int(p[0]);
        y = int(p[1]);
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(x, y, false, false);
        x = int(p[0]);
        y = int(p[1]);
    }
}
template <int L> __device__ __forceinline__ void row_swap(float &x, float &y) {
    int a = __float_as_int(x), b = __float_as_int(y);
    row_swap_b32<L>(a, b);
    x = __int_as_float(a);
    y = __int_as_float(b);
}
template <int L> __device__ __forceinline__ void row_swap(double &x, double &y) {
    const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
    int xl = int(bx & 0xffffffffll), xh = int(bx >> 32), yl = int(by & 0xffffffffll), yh = int(by >> 32);
    row_swap_b32<L>(xl, yl);
    row_swap_b32<L>(xh, yh);
    x = __longlong_as_double((static_cast<long long>(xh) << 32) | static_cast<unsigned>(xl));
    y = __longlong_as_double((static_cast<long long>(yh) << 32) | static_cast<unsigned>(yl));
}
template <int L, class R> __device__ __forceinline__ void row_swap_cx(cx<R> &x, cx<R> &y) {
    row_swap<L>(x.re, y.re);
    row_swap<L>(x.im, y.im);
}

constexpr int kMaxFixed = 16; // target + control bits that a kernel enumerates around
constexpr int kMaxRegK = 4;   // dense gates held in registers up to 4 targets
constexpr int kMaxLdsK = 13;  // dense gates staged through LDS up to 13 targets

// Bit positions (ascending) where a zero is inserted into a group index.
struct Fixed {
    int cnt;
    int pos[kMaxFixed];
};

__device__ __forceinline__ uint64_t expand(uint64_t g, const Fixed &f) {
#pragma unroll 4
    for (int k = 0; k < f.cnt; ++k) {
        const int p = f.pos[k];
        const uint64_t lo = g & ((1ull << p) - 1ull);
        g = ((g >> p) << (p + 1)) | lo;
    }
    return g;
}

Fixed make_fixed(std::vector<int> bits);

// ---------------------------------------------------------------- runtime objects

// Per (device, stream) context: the stream, a grow-only device workspace,
// a pinned host buffer for results and a pinned upload ring for small
// operands (gate matrices, pointer tables).
struct StreamCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    int refs = 0;

    void *ws = nullptr;   // device workspace
    size_t ws_bytes = 0;
    void *host = nullptr; // pinned host staging for results
    size_t host_bytes = 0;

    void *up_host = nullptr; // pinned upload ring (host side)
    void *up_dev = nullptr;  // device side of the ring
    size_t up_bytes = 0;
    size_t up_off = 0;

    void *workspace(size_t bytes);
    void *host_buffer(size_t bytes);
    // Copies `bytes` from host memory into the device ring (stream ordered)
    // and returns the device address.  Valid until the ring wraps, which
    // synchronises the stream first.
    void *upload(const void *src, size_t bytes);
    void sync();
};

StreamCtx *acquire_ctx(int device, hipStream_t stream);
void release_ctx(StreamCtx *ctx);

// Caching device allocator (exact-size free lists per device).
void *pool_alloc(int device, size_t bytes);
void pool_free(int device, void *ptr, size_t bytes);

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev);
    ~DeviceGuard();
};

int sm_count(int device);

// roctx ranges around the library's entry points (lgh_apply, the adjoint
// sweep, the measurements, batched runs), for timelines recorded with
// `rocprofv3 --marker-trace`.  Off unless the environment sets LGH_ROCTX=1,
// so that an untraced run pays one branch per call.
bool roctx_enabled();
void roctx_push(const char *name);
void roctx_pop();
struct Range {
    bool on;
    explicit Range(const char *name) : on(roctx_enabled()) {
        if (on) roctx_push(name);
    }
    ~Range() {
        if (on) roctx_pop();
    }
};

// Optional launch timing (lgh_profile_enable): HIP events recorded on the
// launch stream around each launch of the hot kernels, summed per kind by
// lgh_profile_read.
enum ProfKind { kProfFusedPass = 0, kProfFusedAdj = 1, kProfBatch = 2, kProfKinds = 3 };
struct ProfScope {
    int kind;
    hipStream_t stream;
    hipEvent_t start = nullptr;
    ProfScope(int kind, hipStream_t stream);
    ~ProfScope();
};

} // namespace lgh

struct lgh_sv {
    int n = 0;
    lgh_dtype dtype = LGH_C128;
    int device = 0;
    lgh::StreamCtx *ctx = nullptr;
    void *data = nullptr;
    size_t bytes = 0;
    uint64_t length() const { return 1ull << n; }
    size_t elem() const { return dtype == LGH_C128 ? 16 : 8; }
};

struct lgh_csr {
    int device = 0;
    lgh_dtype dtype = LGH_C128;
    int64_t rows = 0, nnz = 0;
    bool idx64 = true;
    void *indptr = nullptr;  // int64 or int32, rows + 1
    void *indices = nullptr; // int64 or int32, nnz
    void *values = nullptr;  // cx<R>, nnz
    size_t b_ptr = 0, b_idx = 0, b_val = 0;
    double avg_row = 0;
};

namespace lgh {

// ---------------------------------------------------------------- compiled operations

// A gate reduced to one of two kernel shapes.
//  DENSE: a 2^k x 2^k matrix on target bits tbits[0..k) (tbits[0] is the most
//         significant bit of the matrix index) applied where all cmask bits are 1.
//  DIAG:  amplitudes with all cmask bits set are multiplied by d_even or d_odd
//         according to the parity of (index & pmask).
struct KOp {
    enum Kind { NOP, DENSE, DIAG } kind = NOP;
    int k = 0;
    int tbits[kMaxLdsK] = {};
    uint64_t cmask = 0;
    std::vector<int> cbits;
    std::vector<std::complex<double>> mat;
    uint64_t pmask = 0;
    std::complex<double> d_even{1, 0}, d_odd{1, 0};
};

// Generator data aligned with a KOp of the same op: for DENSE a 2^k x 2^k
// matrix on the same targets (restricted to the control subspace), for DIAG
// real eigenvalues by parity.
struct KGen {
    bool present = false;
    std::vector<std::complex<double>> mat;
    double g_even = 0, g_odd = 0;
    double scale = 0;
};

void validate_op(const lgh_op &op, int n);
// U (or U^dagger when `dagger`) of op as a kernel op.
KOp compile_op(const lgh_op &op, int n, bool dagger);
// Generator of op matched to compile_op's kernel shape.
KGen compile_generator(const lgh_op &op, int n, const KOp &shape);
// Full matrix over all op wires (controls included) -- host utility.
std::vector<std::complex<double>> full_matrix(const lgh_op &op, bool dagger);

// Launch a compiled op on one state.
void launch_kop(lgh_sv *sv, const KOp &op);

// Device-side reduction helpers (implemented in lgh_measure.hip).
// Sums `count` partial values (stride 1) into out_dev[0] with a fixed order.
void launch_reduce(StreamCtx *ctx, const double *partials, int nparts, int nvals, double *out_dev);

} // namespace lgh
