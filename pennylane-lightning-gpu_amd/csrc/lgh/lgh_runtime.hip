// Runtime pieces of liblgh.so: errors, devices, streams, the caching
// allocator, state-vector lifecycle, initialisation kernels and copies.
//
// Reference counterparts: DataBuffer / DevTag / DevicePool
// (pennylane_lightning_gpu/src/util/DataBuffer.hpp:16-224, DevTag.hpp:14-56,
// DevicePool.hpp:15-128), StateVectorCudaBase (StateVectorCudaBase.hpp:54-313)
// and initSV.cu:64-117.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <unordered_map>

#include <cstdlib>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "lgh_internal.hpp"

namespace lgh {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

DeviceGuard::DeviceGuard(int dev) {
    LGH_HIP(hipGetDevice(&prev));
    if (prev != dev) LGH_HIP(hipSetDevice(dev));
}
DeviceGuard::~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
}

bool roctx_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("LGH_ROCTX");
        return e && e[0] == '1';
    }();
    return on;
}
void roctx_push(const char *name) { roctxRangePush(name); }
void roctx_pop() { roctxRangePop(); }

int sm_count(int device) {
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(device);
    if (it != cache.end()) return it->second;
    int v = 0;
    LGH_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device));
    cache[device] = v;
    return v;
}

namespace {
std::atomic<int> g_prof{0};
struct ProfRec {
    int kind;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof_recs;
} // namespace

ProfScope::ProfScope(int k, hipStream_t s) : kind(k), stream(s) {
    if (!g_prof.load()) return;
    if (hipEventCreate(&start) != hipSuccess) {
        start = nullptr;
        return;
    }
    if (hipEventRecord(start, stream) != hipSuccess) {
        (void)hipEventDestroy(start);
        start = nullptr;
    }
}

ProfScope::~ProfScope() {
    if (!start) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream) != hipSuccess) {
        if (stop) (void)hipEventDestroy(stop);
        (void)hipEventDestroy(start);
        return;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_recs.push_back(ProfRec{kind, start, stop});
}

Fixed make_fixed(std::vector<int> bits) {
    LGH_REQUIRE(static_cast<int>(bits.size()) <= kMaxFixed, "too many target/control wires for one operation");
    std::sort(bits.begin(), bits.end());
    Fixed f{};
    f.cnt = static_cast<int>(bits.size());
    for (int i = 0; i < f.cnt; ++i) f.pos[i] = bits[i];
    return f;
}

// ---------------------------------------------------------------- caching allocator

namespace {
struct Pool {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void *>> free_lists;
    size_t cached_bytes = 0;
};
Pool &pool() {
    static Pool *p = new Pool();
    return *p;
}
void trim_locked(Pool &p, int device) {
    for (auto it = p.free_lists.begin(); it != p.free_lists.end();) {
        if (device < 0 || it->first.first == device) {
            DeviceGuard g(it->first.first);
            for (void *ptr : it->second) {
                (void)hipFree(ptr);
                p.cached_bytes -= it->first.second;
            }
            it = p.free_lists.erase(it);
        } else {
            ++it;
        }
    }
}
} // namespace

// Requests are served from size classes so that buffers of similar sizes
// (batches of varying circuit counts, descriptor blocks) are reused instead
// of accumulating one free list per exact size: up to 64 KiB the next power
// of two, above that the next multiple of an eighth of the power of two
// below (at most 12.5 % over the request).
size_t size_class(size_t bytes) {
    if (bytes <= (size_t(64) << 10)) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    int hb = 63 - __builtin_clzll(uint64_t(bytes));
    const size_t step = size_t(1) << (hb - 3);
    return (bytes + step - 1) / step * step;
}

void *pool_alloc(int device, size_t bytes) {
    bytes = size_class(bytes);
    Pool &p = pool();
    {
        std::lock_guard<std::mutex> lk(p.mu);
        auto it = p.free_lists.find({device, bytes});
        if (it != p.free_lists.end() && !it->second.empty()) {
            void *ptr = it->second.back();
            it->second.pop_back();
            p.cached_bytes -= bytes;
            return ptr;
        }
    }
    DeviceGuard g(device);
    void *ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(p.mu);
        trim_locked(p, device);
        e = hipMalloc(&ptr, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? LGH_ERR_OUT_OF_MEMORY : LGH_ERR_DEVICE,
                    "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    return ptr;
}

void pool_free(int device, void *ptr, size_t bytes) {
    if (!ptr) return;
    bytes = size_class(bytes);
    Pool &p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    p.free_lists[{device, bytes}].push_back(ptr);
    p.cached_bytes += bytes;
}

// ---------------------------------------------------------------- stream contexts

namespace {
struct CtxRegistry {
    std::mutex mu;
    std::map<std::pair<int, hipStream_t>, StreamCtx *> by_key;
    std::map<int, hipStream_t> lib_stream;
};
CtxRegistry &registry() {
    static CtxRegistry *r = new CtxRegistry();
    return *r;
}
} // namespace

StreamCtx *acquire_ctx(int device, hipStream_t stream) {
    int count = 0;
    LGH_HIP(hipGetDeviceCount(&count));
    LGH_REQUIRE(device >= 0 && device < count, "device index out of range");
    CtxRegistry &r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    bool owned = false;
    if (!stream) {
        auto it = r.lib_stream.find(device);
        if (it == r.lib_stream.end()) {
            DeviceGuard g(device);
            hipStream_t s;
            LGH_HIP(hipStreamCreate(&s)); // default priority, blocking w.r.t. the null stream
            r.lib_stream[device] = s;
            stream = s;
        } else {
            stream = it->second;
        }
        owned = true;
    }
    auto key = std::make_pair(device, stream);
    auto it = r.by_key.find(key);
    StreamCtx *c;
    if (it == r.by_key.end()) {
        c = new StreamCtx();
        c->device = device;
        c->stream = stream;
        c->owns_stream = owned;
        r.by_key[key] = c;
    } else {
        c = it->second;
    }
    c->refs++;
    return c;
}

static void free_ctx_buffers(StreamCtx *c) {
    DeviceGuard g(c->device);
    if (c->ws) (void)hipFree(c->ws);
    if (c->host) (void)hipHostFree(c->host);
    if (c->up_host) (void)hipHostFree(c->up_host);
    if (c->up_dev) (void)hipFree(c->up_dev);
    c->ws = c->host = c->up_host = c->up_dev = nullptr;
    c->ws_bytes = c->host_bytes = c->up_bytes = c->up_off = 0;
}

void release_ctx(StreamCtx *c) {
    if (!c) return;
    CtxRegistry &r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    if (--c->refs > 0 || c->owns_stream) return;
    // a caller-owned stream: drop our buffers, keep the stream
    (void)hipStreamSynchronize(c->stream);
    free_ctx_buffers(c);
    r.by_key.erase({c->device, c->stream});
    delete c;
}

void StreamCtx::sync() { LGH_HIP(hipStreamSynchronize(stream)); }

void *StreamCtx::workspace(size_t bytes) {
    if (bytes <= ws_bytes) return ws;
    DeviceGuard g(device);
    sync();
    if (ws) LGH_HIP(hipFree(ws));
    ws = nullptr;
    size_t nb = std::max<size_t>(bytes, std::max<size_t>(ws_bytes * 3 / 2, 1 << 20));
    LGH_HIP(hipMalloc(&ws, nb));
    ws_bytes = nb;
    return ws;
}

void *StreamCtx::host_buffer(size_t bytes) {
    if (bytes <= host_bytes) return host;
    DeviceGuard g(device);
    sync();
    if (host) LGH_HIP(hipHostFree(host));
    host = nullptr;
    size_t nb = std::max<size_t>(bytes, 1 << 16);
    LGH_HIP(hipHostMalloc(&host, nb, hipHostMallocDefault));
    host_bytes = nb;
    return host;
}

void *StreamCtx::upload(const void *src, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    DeviceGuard g(device);
    if (need > up_bytes) {
        sync();
        if (up_host) LGH_HIP(hipHostFree(up_host));
        if (up_dev) LGH_HIP(hipFree(up_dev));
        up_host = up_dev = nullptr;
        size_t nb = std::max<size_t>(need, 4 << 20);
        LGH_HIP(hipHostMalloc(&up_host, nb, hipHostMallocDefault));
        LGH_HIP(hipMalloc(&up_dev, nb));
        up_bytes = nb;
        up_off = 0;
    }
    if (up_off + need > up_bytes) {
        sync(); // every earlier copy out of the ring has completed
        up_off = 0;
    }
    char *h = static_cast<char *>(up_host) + up_off;
    char *d = static_cast<char *>(up_dev) + up_off;
    std::memcpy(h, src, bytes);
    LGH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    up_off += need;
    return d;
}

} // namespace lgh

using namespace lgh;

// ---------------------------------------------------------------- kernels

namespace {

// One write pass: zero everywhere, `value` at `index` (initSV.cu:107-117 and
// DataBuffer::zeroInit fused).
template <class R>
__global__ __launch_bounds__(256) void k_fill_basis(cx<R> *__restrict__ sv, uint64_t len, uint64_t index,
                                                    cx<R> value) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x * 4 + threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e, i += blockDim.x) {
        if (i < len) sv[i] = (i == index) ? value : cx<R>{R(0), R(0)};
    }
}

// sv[idx[i]] = val[i] (initSV.cu:64-71)
template <class R>
__global__ __launch_bounds__(256) void k_scatter(cx<R> *__restrict__ sv, uint64_t len,
                                                 const int64_t *__restrict__ idx,
                                                 const cx<R> *__restrict__ val, uint64_t count) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        int64_t t = idx[i];
        if (t >= 0 && uint64_t(t) < len) sv[t] = val[i];
    }
}

template <class R> void fill_basis(lgh_sv *sv, uint64_t index, double re, double im) {
    const uint64_t len = sv->length();
    const uint64_t per_block = 256 * 4;
    const uint64_t blocks = (len + per_block - 1) / per_block;
    hipLaunchKernelGGL(k_fill_basis<R>, dim3(unsigned(blocks)), dim3(256), 0, sv->ctx->stream,
                       static_cast<cx<R> *>(sv->data), len, index, cx<R>{R(re), R(im)});
    LGH_LAUNCH_CHECK();
}

} // namespace

// ---------------------------------------------------------------- C ABI

extern "C" {

const char *lgh_last_error(void) { return g_last_error.c_str(); }
int lgh_abi_version(void) { return LGH_ABI_VERSION; }

lgh_status lgh_device_count(int *count) {
    return guard([&] {
        LGH_REQUIRE(count, "null pointer");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c = 0;
        }
        *count = c;
    });
}

lgh_status lgh_device_arch(int device, char *name, size_t len) {
    return guard([&] {
        LGH_REQUIRE(name && len > 0, "null pointer");
        hipDeviceProp_t prop;
        LGH_HIP(hipGetDeviceProperties(&prop, device));
        std::string arch = prop.gcnArchName;
        auto colon = arch.find(':');
        if (colon != std::string::npos) arch = arch.substr(0, colon);
        std::strncpy(name, arch.c_str(), len - 1);
        name[len - 1] = 0;
    });
}

lgh_status lgh_device_supported(int device, int *supported) {
    return guard([&] {
        LGH_REQUIRE(supported, "null pointer");
        char buf[64];
        *supported = 0;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
            (void)hipGetLastError();
            return;
        }
        if (lgh_device_arch(device, buf, sizeof buf) == LGH_OK) *supported = std::strcmp(buf, "gfx950") == 0;
    });
}

lgh_status lgh_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
    return guard([&] {
        DeviceGuard g(device);
        size_t f = 0, t = 0;
        LGH_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

lgh_status lgh_device_synchronize(int device) {
    return guard([&] {
        DeviceGuard g(device);
        LGH_HIP(hipDeviceSynchronize());
    });
}

lgh_status lgh_pool_trim(int device) {
    return guard([&] {
        Pool &p = pool();
        std::lock_guard<std::mutex> lk(p.mu);
        trim_locked(p, device);
    });
}

lgh_status lgh_device_reset(void) {
    return guard([&] {
        {
            Pool &p = pool();
            std::lock_guard<std::mutex> lk(p.mu);
            trim_locked(p, -1);
        }
        CtxRegistry &r = registry();
        std::lock_guard<std::mutex> lk(r.mu);
        for (auto &kv : r.by_key) free_ctx_buffers(kv.second);
        for (auto &kv : r.lib_stream) {
            DeviceGuard g(kv.first);
            (void)hipStreamDestroy(kv.second);
        }
        for (auto &kv : r.by_key) delete kv.second;
        r.by_key.clear();
        r.lib_stream.clear();
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
        for (int d = 0; d < count; ++d) {
            DeviceGuard g(d);
            LGH_HIP(hipDeviceReset());
        }
    });
}

lgh_status lgh_stream_create(int device, void **stream) {
    return guard([&] {
        LGH_REQUIRE(stream, "null pointer");
        DeviceGuard g(device);
        hipStream_t s;
        LGH_HIP(hipStreamCreate(&s)); // default priority
        *stream = s;
    });
}

lgh_status lgh_stream_destroy(void *stream) {
    return guard([&] {
        LGH_REQUIRE(stream, "null stream");
        LGH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
        LGH_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    });
}

lgh_status lgh_stream_synchronize(void *stream) {
    return guard([&] { LGH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream))); });
}

lgh_status lgh_sv_create(lgh_sv **out, int num_qubits, lgh_dtype dtype, int device, void *stream) {
    return guard([&] {
        LGH_REQUIRE(out, "null pointer");
        *out = nullptr;
        LGH_REQUIRE(num_qubits >= 1 && num_qubits <= 36, "num_qubits must be in [1, 36]");
        LGH_REQUIRE(dtype == LGH_C64 || dtype == LGH_C128, "unknown dtype");
        auto sv = std::make_unique<lgh_sv>();
        sv->n = num_qubits;
        sv->dtype = dtype;
        sv->device = device;
        sv->ctx = acquire_ctx(device, static_cast<hipStream_t>(stream));
        sv->bytes = sv->elem() << num_qubits;
        try {
            sv->data = pool_alloc(device, sv->bytes);
        } catch (...) {
            release_ctx(sv->ctx);
            throw;
        }
        *out = sv.release();
    });
}

lgh_status lgh_sv_destroy(lgh_sv *sv) {
    return guard([&] {
        if (!sv) return;
        // pending work on the buffer must finish before it is reused
        (void)hipStreamSynchronize(sv->ctx->stream);
        pool_free(sv->device, sv->data, sv->bytes);
        release_ctx(sv->ctx);
        delete sv;
    });
}

lgh_status lgh_sv_info(const lgh_sv *sv, int *num_qubits, lgh_dtype *dtype, int *device, void **stream,
                       void **device_data) {
    return guard([&] {
        LGH_REQUIRE(sv, "null state");
        if (num_qubits) *num_qubits = sv->n;
        if (dtype) *dtype = sv->dtype;
        if (device) *device = sv->device;
        if (stream) *stream = sv->ctx->stream;
        if (device_data) *device_data = sv->data;
    });
}

lgh_status lgh_sv_reset(lgh_sv *sv) { return lgh_sv_set_basis(sv, 0, 1.0, 0.0); }

lgh_status lgh_sv_set_basis(lgh_sv *sv, uint64_t index, double re, double im) {
    return guard([&] {
        LGH_REQUIRE(sv, "null state");
        LGH_REQUIRE(index < sv->length(), "basis index out of range");
        DeviceGuard g(sv->device);
        if (sv->dtype == LGH_C128)
            fill_basis<double>(sv, index, re, im);
        else
            fill_basis<float>(sv, index, re, im);
    });
}

lgh_status lgh_sv_set_indexed(lgh_sv *sv, const int64_t *indices, const void *values, uint64_t count) {
    return guard([&] {
        LGH_REQUIRE(sv, "null state");
        LGH_REQUIRE(count == 0 || (indices && values), "null pointer");
        for (uint64_t i = 0; i < count; ++i)
            LGH_REQUIRE(indices[i] >= 0 && uint64_t(indices[i]) < sv->length(), "state index out of range");
        DeviceGuard g(sv->device);
        StreamCtx *c = sv->ctx;
        const uint64_t len = sv->length();
        const size_t vb = count * sv->elem(), ib = count * sizeof(int64_t);
        if (sv->dtype == LGH_C128)
            fill_basis<double>(sv, len, 0, 0); // index == len: all zeros
        else
            fill_basis<float>(sv, len, 0, 0);
        if (count == 0) return;
        char *ws = static_cast<char *>(c->workspace(ib + vb + 256));
        LGH_HIP(hipMemcpyAsync(ws, indices, ib, hipMemcpyHostToDevice, c->stream));
        LGH_HIP(hipMemcpyAsync(ws + ((ib + 255) & ~size_t(255)), values, vb, hipMemcpyHostToDevice, c->stream));
        const unsigned blocks = unsigned((count + 255) / 256);
        const int64_t *d_idx = reinterpret_cast<const int64_t *>(ws);
        const void *d_val = ws + ((ib + 255) & ~size_t(255));
        if (sv->dtype == LGH_C128)
            hipLaunchKernelGGL(k_scatter<double>, dim3(blocks), dim3(256), 0, c->stream,
                               static_cast<cx<double> *>(sv->data), len, d_idx,
                               static_cast<const cx<double> *>(d_val), count);
        else
            hipLaunchKernelGGL(k_scatter<float>, dim3(blocks), dim3(256), 0, c->stream,
                               static_cast<cx<float> *>(sv->data), len, d_idx,
                               static_cast<const cx<float> *>(d_val), count);
        LGH_LAUNCH_CHECK();
        c->sync(); // host buffers are borrowed for the duration of the call
    });
}

lgh_status lgh_sv_copy_h2d(lgh_sv *sv, const void *host, uint64_t count) {
    return guard([&] {
        LGH_REQUIRE(sv && host, "null pointer");
        LGH_REQUIRE(count == sv->length(), "host buffer length does not match the state");
        DeviceGuard g(sv->device);
        LGH_HIP(hipMemcpyAsync(sv->data, host, sv->bytes, hipMemcpyHostToDevice, sv->ctx->stream));
        sv->ctx->sync();
    });
}

lgh_status lgh_sv_copy_d2h(const lgh_sv *sv, void *host, uint64_t count) {
    return guard([&] {
        LGH_REQUIRE(sv && host, "null pointer");
        LGH_REQUIRE(count <= sv->length(), "host buffer longer than the state");
        DeviceGuard g(sv->device);
        LGH_HIP(hipMemcpyAsync(host, sv->data, count * sv->elem(), hipMemcpyDeviceToHost, sv->ctx->stream));
        sv->ctx->sync();
    });
}

lgh_status lgh_sv_copy_d2d(lgh_sv *dst, const lgh_sv *src) {
    return guard([&] {
        LGH_REQUIRE(dst && src, "null state");
        LGH_REQUIRE(dst->n == src->n && dst->dtype == src->dtype, "states differ in size or precision");
        if (dst == src) return;
        DeviceGuard g(dst->device);
        if (src->ctx != dst->ctx) src->ctx->sync();
        LGH_HIP(hipMemcpyAsync(dst->data, src->data, dst->bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
    });
}

lgh_status lgh_sv_synchronize(const lgh_sv *sv) {
    return guard([&] {
        LGH_REQUIRE(sv, "null state");
        sv->ctx->sync();
    });
}

} // extern "C"

extern "C" lgh_status lgh_profile_enable(int enabled) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        for (ProfRec &r : g_prof_recs) {
            (void)hipEventSynchronize(r.b);
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        g_prof_recs.clear();
        g_prof.store(enabled ? 1 : 0);
    });
}

extern "C" lgh_status lgh_profile_read(int kind, int64_t *launches, double *total_ms) {
    return guard([&] {
        LGH_REQUIRE(kind >= 0 && kind < kProfKinds, "unknown profile kind");
        LGH_REQUIRE(launches && total_ms, "null pointer");
        std::lock_guard<std::mutex> lk(g_prof_mu);
        int64_t n = 0;
        double t = 0.0;
        for (ProfRec &r : g_prof_recs) {
            if (r.kind != kind) continue;
            LGH_HIP(hipEventSynchronize(r.b));
            float ms = 0.f;
            LGH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            ++n;
            t += ms;
        }
        *launches = n;
        *total_ms = t;
    });
}
