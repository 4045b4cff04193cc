// bs_capi.hip -- the C ABI of libbs_mi355.so (declared in include/bs_mi355.h).
//
// Argument checks, pointer validation and error mapping sit here; the kernels
// are in bs_kernels.hip and the serving engine in bs_serve.hip.  Nothing in
// this library prices on the CPU: every entry point that computes launches a
// HIP kernel, and fails with BS_ERR_NODEVICE / BS_ERR_HIP when it cannot.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "bs_internal.hpp"

namespace bs {

Config& config() {
    static Config c;
    return c;
}

static thread_local std::string t_last_error;

int set_error(int status, const std::string& msg) noexcept {
    try {
        t_last_error = msg;
    } catch (...) {
        t_last_error.clear();
    }
    return status;
}

int guard_failed(int status, const char* what, const char* detail) noexcept {
    try {
        t_last_error = std::string(what) + ": " + detail;
    } catch (...) {
        t_last_error.clear();   // no room even for the message
    }
    return status;
}

int hip_error(hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" +
                    hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return set_error(BS_ERR_NOMEM, m);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return set_error(BS_ERR_NODEVICE, m);
    return set_error(BS_ERR_HIP, m);
}

int check_device_range(const void* p, size_t bytes, const char* name) {
    if (p == nullptr) return set_error(BS_ERR_INVALID, std::string(name) + " is NULL");
    if (!config().check_pointers) return BS_OK;
    hipPointerAttribute_t attr;
    memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BS_ERR_INVALID, std::string(name) + " is not memory the GPU can access");
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged &&
        !(attr.type == hipMemoryTypeHost && attr.devicePointer != nullptr))
        return set_error(BS_ERR_INVALID,
                         std::string(name) + " is not device, managed or pinned memory");
    if (attr.type == hipMemoryTypeDevice) {
        hipDeviceptr_t base = nullptr;
        size_t size = 0;
        e = hipMemGetAddressRange(&base, &size, const_cast<void*>(p));
        if (e == hipSuccess) {
            const char* b = static_cast<const char*>(base);
            const char* q = static_cast<const char*>(p);
            if (q + bytes > b + size)
                return set_error(BS_ERR_INVALID,
                                 std::string(name) + " allocation is smaller than n elements");
        } else {
            (void)hipGetLastError();
        }
    }
    return BS_OK;
}

}  // namespace bs

using namespace bs;

namespace {

hipStream_t as_stream(bs_stream_t s) { return static_cast<hipStream_t>(s); }

int check_n(int64_t n) {
    if (n < 0) return set_error(BS_ERR_INVALID, "n < 0");
    return BS_OK;
}

template <class T>
int check_soa(const T* s, const T* k, const T* r, const T* v, const T* t, int64_t n) {
    const size_t b = (size_t)n * sizeof(T);
    int rc;
    if ((rc = check_device_range(s, b, "s")) || (rc = check_device_range(k, b, "k")) ||
        (rc = check_device_range(r, b, "r")) || (rc = check_device_range(v, b, "v")) ||
        (rc = check_device_range(t, b, "t")))
        return rc;
    return BS_OK;
}

int done(hipError_t e, const char* what) { return e == hipSuccess ? BS_OK : hip_error(e, what); }

// Per-thread pinned word used by the scalar convenience forms.
float* scalar_slot() {
    static thread_local float* slot = nullptr;
    if (slot == nullptr) {
        if (hipHostMalloc((void**)&slot, sizeof(float) * 4, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            slot = nullptr;
        }
    }
    return slot;
}

}  // namespace

// ---- repeated passes ---------------------------------------------------------------
namespace {

// Books up to this size replay their passes from one graph.  Small books are
// bound by launch cost (a pass takes a few microseconds, about what the host
// spends per launch); larger ones gain because consecutive kernel nodes of a
// graph skip part of the cache maintenance that separate launches do between
// passes that reread the same book (measured per pass: 4096 options 4.6 ->
// 1.9 us, 10M 43.2 -> 39.4 us, 32M 124.0 -> 122.9 us).  Past this size the
// gain no longer repays a first call's capture.
constexpr int64_t kPassGraphMaxOptions = int64_t(1) << 25;
constexpr size_t kPassGraphs = 16;
constexpr int kPassDevices = 64;

// A graph is valid for the arrays, sizes and device it was captured with, and
// for the launch settings that shaped its kernels.
struct PassKey {
    const void* ptr[7];
    int64_t n;
    int passes, type_bytes, device;
    int64_t launch[6];   // max_blocks, vector_width, nontemporal, pipeline, block_threads,
                         // lds_per_wave
    bool operator==(const PassKey& o) const {
        return memcmp(ptr, o.ptr, sizeof(ptr)) == 0 && n == o.n && passes == o.passes &&
               type_bytes == o.type_bytes && device == o.device &&
               memcmp(launch, o.launch, sizeof(launch)) == 0;
    }
};

struct PassGraph {
    PassKey key;
    hipGraphExec_t exec;
    hipEvent_t done;   // recorded after each replay: eviction waits for it
};

struct PassGraphCache {
    std::mutex mu;
    std::vector<PassGraph> lru;                     // most recently used first
    hipStream_t capture[kPassDevices] = {};
};

PassGraphCache& pass_cache() {
    static PassGraphCache* c = new PassGraphCache();   // lives until the process ends
    return *c;
}

hipError_t launch_pass(const SoaIn<float>& in, const void* type, int type_bytes, float* out,
                       int64_t n, hipStream_t st) {
    return type_bytes == 1
               ? launch_price_f32(in, static_cast<const uint8_t*>(type), out, n, st)
               : launch_price_i32type_f32(in, static_cast<const int32_t*>(type), out, n, st);
}

// The cached graph for key, captured now if absent (the cache's lock held).
PassGraph* pass_graph(PassGraphCache& c, const PassKey& key, const SoaIn<float>& in,
                      const void* type, float* out, hipError_t* e) {
    for (size_t i = 0; i < c.lru.size(); ++i)
        if (c.lru[i].key == key) {
            std::rotate(c.lru.begin(), c.lru.begin() + (std::ptrdiff_t)i,
                        c.lru.begin() + (std::ptrdiff_t)i + 1);
            return &c.lru.front();
        }
    hipStream_t& cs = c.capture[key.device];
    if (!cs && (*e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) {
        cs = nullptr;
        return nullptr;
    }
    if ((*e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return nullptr;
    hipError_t le = hipSuccess;
    for (int p = 0; p < key.passes && le == hipSuccess; ++p)
        le = launch_pass(in, type, key.type_bytes, out, key.n, cs);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(cs, &graph);
    *e = le != hipSuccess ? le : ce;
    PassGraph g{key, nullptr, nullptr};
    if (*e == hipSuccess) *e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (*e == hipSuccess) *e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
    if (*e != hipSuccess) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        return nullptr;
    }
    if (c.lru.size() == kPassGraphs) {
        PassGraph& old = c.lru.back();
        (void)hipEventSynchronize(old.done);
        (void)hipEventDestroy(old.done);
        (void)hipGraphExecDestroy(old.exec);
        c.lru.pop_back();
    }
    c.lru.insert(c.lru.begin(), g);
    return &c.lru.front();
}

}  // namespace

extern "C" {

const char* bs_status_string(int status) {
    switch (status) {
        case BS_OK: return "ok";
        case BS_ERR_INVALID: return "invalid argument";
        case BS_ERR_HIP: return "HIP runtime error";
        case BS_ERR_NOMEM: return "out of device memory";
        case BS_ERR_BUDGET: return "over the HBM budget";
        case BS_ERR_NOT_FOUND: return "unknown book or ticket";
        case BS_ERR_CANCELLED: return "cancelled";
        case BS_ERR_TIMEOUT: return "timed out";
        case BS_ERR_BUSY: return "book in use";
        case BS_ERR_NODEVICE: return "no HIP device";
        case BS_ERR_FAILED: return "request failed";
        case BS_ERR_SHUTDOWN: return "context shutting down";
        default: return "unknown status";
    }
}

const char* bs_last_error(void) { return t_last_error.c_str(); }

int bs_version(void) { return 10000 * 1 + 100 * 0 + 0; }

int bs_config_set(int key, int64_t value) {
    switch (key) {
        case BS_CFG_CHECK_POINTERS: config().check_pointers = value != 0; return BS_OK;
        case BS_CFG_MAX_BLOCKS:
            if (value < 0) return set_error(BS_ERR_INVALID, "max_blocks < 0");
            config().max_blocks = value;
            return BS_OK;
        case BS_CFG_VECTOR_WIDTH:
            if (value != 0 && value != 2 && value != 4)
                return set_error(BS_ERR_INVALID, "vector width must be 0, 2 or 4");
            config().vector_width = (int)value;
            return BS_OK;
        case BS_CFG_NONTEMPORAL: config().nontemporal = value != 0; return BS_OK;
        case BS_CFG_PIPELINE:
            config().pipeline = value < 0 ? -1 : (value != 0);
            return BS_OK;
        case BS_CFG_BLOCK_THREADS:
            if (value != 0 && value != 64 && value != 128 && value != 256)
                return set_error(BS_ERR_INVALID, "block threads must be 0, 64, 128 or 256");
            config().block_threads = (int)value;
            return BS_OK;
        case BS_CFG_LDS_PER_WAVE:
            if (value < -1 || value > 65536)
                return set_error(BS_ERR_INVALID, "LDS per wave must be -1, 0 or 1..65536 bytes");
            config().lds_per_wave = (int)value;
            return BS_OK;
        default: return set_error(BS_ERR_INVALID, "unknown config key");
    }
}

int bs_config_get(int key, int64_t* value) {
    if (!value) return set_error(BS_ERR_INVALID, "value is NULL");
    switch (key) {
        case BS_CFG_CHECK_POINTERS: *value = config().check_pointers; return BS_OK;
        case BS_CFG_MAX_BLOCKS: *value = config().max_blocks; return BS_OK;
        case BS_CFG_BLOCK_THREADS: *value = config().block_threads; return BS_OK;
        case BS_CFG_VECTOR_WIDTH: *value = config().vector_width; return BS_OK;
        case BS_CFG_NONTEMPORAL: *value = config().nontemporal; return BS_OK;
        case BS_CFG_PIPELINE: *value = config().pipeline; return BS_OK;
        case BS_CFG_LDS_PER_WAVE: *value = config().lds_per_wave; return BS_OK;
        default: return set_error(BS_ERR_INVALID, "unknown config key");
    }
}

// ---- per-option functions -----------------------------------------------------
int bs_cndf_f32(const float* x, float* out, int64_t n, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_device_range(x, n * 4, "x")) || (rc = check_device_range(out, n * 4, "out")))
        return rc;
    return done(launch_cndf_f32(x, out, n, as_stream(stream)), "bs_cndf_f32");
}

int bs_cndf_f64(const double* x, double* out, int64_t n, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_device_range(x, n * 8, "x")) || (rc = check_device_range(out, n * 8, "out")))
        return rc;
    return done(launch_cndf_f64(x, out, n, as_stream(stream)), "bs_cndf_f64");
}

int bs_cndf_scalar_f32(float x, float* out) {
    if (!out) return set_error(BS_ERR_INVALID, "out is NULL");
    float* slot = scalar_slot();
    if (!slot) return set_error(BS_ERR_NOMEM, "pinned scalar slot");
    hipError_t e = launch_cndf_scalar(x, slot, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_error(e, "bs_cndf_scalar_f32");
    *out = *slot;
    return BS_OK;
}

int bs_price_scalar_f32(float sptprice, float strike, float rate, float volatility, float time,
                        int otype, float timet, float* out) {
    (void)timet;  // unused by the reference as well (blackscholes.c:192)
    if (!out) return set_error(BS_ERR_INVALID, "out is NULL");
    float* slot = scalar_slot();
    if (!slot) return set_error(BS_ERR_NOMEM, "pinned scalar slot");
    hipError_t e =
        launch_price_scalar(sptprice, strike, rate, volatility, time, otype != 0, slot, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_error(e, "bs_price_scalar_f32");
    *out = *slot;
    return BS_OK;
}

// ---- batch pricing --------------------------------------------------------------
int bs_price_f32(const float* s, const float* k, const float* r, const float* v, const float* t,
                 const uint8_t* otype, float* price, int64_t n, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_soa(s, k, r, v, t, n)) || (rc = check_device_range(otype, n, "otype")) ||
        (rc = check_device_range(price, n * 4, "price")))
        return rc;
    return done(launch_price_f32({s, k, r, v, t}, otype, price, n, as_stream(stream)),
                "bs_price_f32");
}

int bs_price_i32type_f32(const float* s, const float* k, const float* r, const float* v,
                         const float* t, const int32_t* otype, float* price, int64_t n,
                         bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_soa(s, k, r, v, t, n)) || (rc = check_device_range(otype, n * 4, "otype")) ||
        (rc = check_device_range(price, n * 4, "price")))
        return rc;
    return done(launch_price_i32type_f32({s, k, r, v, t}, otype, price, n, as_stream(stream)),
                "bs_price_i32type_f32");
}

int bs_price_fast_f32(const float* s, const float* k, const float* r, const float* v,
                          const float* t, const uint8_t* otype, float* price, int64_t n,
                          bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_soa(s, k, r, v, t, n)) || (rc = check_device_range(otype, n, "otype")) ||
        (rc = check_device_range(price, n * 4, "price")))
        return rc;
    return done(launch_price_fast_f32({s, k, r, v, t}, otype, price, n, as_stream(stream)),
                "bs_price_fast_f32");
}

int bs_price_f64(const double* s, const double* k, const double* r, const double* v,
                 const double* t, const uint8_t* otype, double* price, int64_t n,
                 bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((rc = check_soa(s, k, r, v, t, n)) || (rc = check_device_range(otype, n, "otype")) ||
        (rc = check_device_range(price, n * 8, "price")))
        return rc;
    return done(launch_price_f64({s, k, r, v, t}, otype, price, n, as_stream(stream)),
                "bs_price_f64");
}

int bs_price_greeks_f32(const float* s, const float* k, const float* r, const float* v,
                        const float* t, const uint8_t* otype, float* price, float* delta,
                        float* vega, float* rho, float* theta, int64_t n, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    const size_t b = (size_t)n * 4;
    if ((rc = check_soa(s, k, r, v, t, n)) || (rc = check_device_range(otype, n, "otype")) ||
        (rc = check_device_range(price, b, "price")) ||
        (rc = check_device_range(delta, b, "delta")) ||
        (rc = check_device_range(vega, b, "vega")) || (rc = check_device_range(rho, b, "rho")) ||
        (rc = check_device_range(theta, b, "theta")))
        return rc;
    return done(launch_greeks_f32({s, k, r, v, t}, otype, {price, delta, vega, rho, theta}, n,
                                  as_stream(stream)),
                "bs_price_greeks_f32");
}

int bs_portfolio_f32(const float* s, const float* k, const float* r, const float* v,
                     const float* t, const uint8_t* otype, const float* weights, int64_t n,
                     int greeks, double* sums, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n))) return rc;
    if ((rc = check_device_range(sums, (greeks ? 5 : 1) * sizeof(double), "sums"))) return rc;
    if (n > 0 && ((rc = check_soa(s, k, r, v, t, n)) ||
                  (rc = check_device_range(otype, (size_t)n, "otype")) ||
                  (weights && (rc = check_device_range(weights, (size_t)n * 4, "weights")))))
        return rc;
    return done(launch_portfolio_f32({s, k, r, v, t}, otype, weights, n, greeks != 0, sums,
                                     as_stream(stream)),
                "bs_portfolio_f32");
}

int bs_portfolio_f32_noalloc(const float* s, const float* k, const float* r, const float* v,
                             const float* t, const uint8_t* otype, const float* weights,
                             int64_t n, int greeks, double* sums, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n))) return rc;
    if (n > kPfNoAllocMaxOptions)
        return set_error(BS_ERR_INVALID,
                         "bs_portfolio_f32_noalloc: n exceeds 2^25 options, the size its fixed "
                         "scratch holds (bs_portfolio_f32 takes any n)");
    if ((rc = check_device_range(sums, (greeks ? 5 : 1) * sizeof(double), "sums"))) return rc;
    if (n > 0 && ((rc = check_soa(s, k, r, v, t, n)) ||
                  (rc = check_device_range(otype, (size_t)n, "otype")) ||
                  (weights && (rc = check_device_range(weights, (size_t)n * 4, "weights")))))
        return rc;
    return done(launch_portfolio_f32_noalloc({s, k, r, v, t}, otype, weights, n, greeks != 0,
                                             sums, as_stream(stream)),
                "bs_portfolio_f32_noalloc");
}

int bs_repack_aos_f32(const void* aos, int64_t n, float* s, float* k, float* r, float* v,
                      float* t, uint8_t* otype, bs_stream_t stream) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if ((reinterpret_cast<uintptr_t>(aos) & 3) != 0)
        return set_error(BS_ERR_INVALID, "aos must be 4-byte aligned");
    if ((rc = check_device_range(aos, (size_t)n * 36, "aos")) ||
        (rc = check_soa<float>(s, k, r, v, t, n)) ||
        (rc = check_device_range(otype, n, "otype")))
        return rc;
    return done(launch_repack_aos(aos, n, s, k, r, v, t, otype, as_stream(stream)),
                "bs_repack_aos_f32");
}

// ---- host-pointer forms -------------------------------------------------------------
int bs_price_passes_f32(const float* s, const float* k, const float* r, const float* v,
                        const float* t, const void* otype, int otype_bytes, float* price,
                        int64_t n, int passes, bs_stream_t stream) {
    return guarded("bs_price_passes_f32", [&]() -> int {
        int rc;
        if ((rc = check_n(n))) return rc;
        if (passes < 1) return set_error(BS_ERR_INVALID, "passes < 1");
        if (otype_bytes != 1 && otype_bytes != 4)
            return set_error(BS_ERR_INVALID, "otype_bytes must be 1 or 4");
        if (n == 0) return BS_OK;
        if ((rc = check_soa(s, k, r, v, t, n)) ||
            (rc = check_device_range(otype, (size_t)n * otype_bytes, "otype")) ||
            (rc = check_device_range(price, (size_t)n * 4, "price")))
            return rc;
        const SoaIn<float> in{s, k, r, v, t};
        const hipStream_t st = as_stream(stream);
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return hip_error(e, "hipGetDevice");
        if (n <= kPassGraphMaxOptions && passes > 1 && dev < kPassDevices) {
            const Config& cf = config();
            const PassKey key{{s, k, r, v, t, otype, price}, n, passes, otype_bytes, dev,
                              {cf.max_blocks, cf.vector_width, cf.nontemporal, cf.pipeline,
                               cf.block_threads, cf.lds_per_wave}};
            PassGraphCache& c = pass_cache();
            std::lock_guard<std::mutex> lk(c.mu);
            PassGraph* g = pass_graph(c, key, in, otype, price, &e);
            if (!g) return hip_error(e, "bs_price_passes_f32: graph capture");
            e = hipGraphLaunch(g->exec, st);
            if (e == hipSuccess) e = hipEventRecord(g->done, st);
            return done(e, "bs_price_passes_f32");
        }
        for (int p = 0; p < passes && e == hipSuccess; ++p)
            e = launch_pass(in, otype, otype_bytes, price, n, st);
        return done(e, "bs_price_passes_f32");
    });
}

int bs_price_host_f32(const float* s, const float* k, const float* r, const float* v,
                      const float* t, const int32_t* otype, float* price, int64_t n,
                      int passes, int device) {
    int rc;
    if ((rc = check_n(n)) || n == 0) return rc;
    if (!s || !k || !r || !v || !t || !otype || !price)
        return set_error(BS_ERR_INVALID, "NULL host array");
    if (passes < 1) return set_error(BS_ERR_INVALID, "passes < 1");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_error(e, "hipSetDevice");
    // One allocation: 5 float arrays + int32 types + prices, each 256-B aligned.
    const size_t fb = ((size_t)n * 4 + 255) & ~(size_t)255;
    char* base = nullptr;
    e = hipMalloc((void**)&base, fb * 7);
    if (e != hipSuccess) return hip_error(e, "hipMalloc");
    float* d[5];
    const float* h[5] = {s, k, r, v, t};
    for (int i = 0; i < 5; ++i) d[i] = reinterpret_cast<float*>(base + fb * i);
    int32_t* dty = reinterpret_cast<int32_t*>(base + fb * 5);
    float* dout = reinterpret_cast<float*>(base + fb * 6);
    for (int i = 0; i < 5 && e == hipSuccess; ++i)
        e = hipMemcpy(d[i], h[i], (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dty, otype, (size_t)n * 4, hipMemcpyHostToDevice);
    for (int p = 0; p < passes && e == hipSuccess; ++p)
        e = launch_price_i32type_f32({d[0], d[1], d[2], d[3], d[4]}, dty, dout, n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(price, dout, (size_t)n * 4, hipMemcpyDeviceToHost);
    hipError_t ef = hipFree(base);
    if (e != hipSuccess) return hip_error(e, "bs_price_host_f32");
    if (ef != hipSuccess) return hip_error(ef, "hipFree");
    return BS_OK;
}

// ---- bs_thread drop-in -------------------------------------------------------------
static struct {
    std::mutex mu;
    int num_options = 0, nthreads = 1;
    float *sptprice = nullptr, *strike = nullptr, *rate = nullptr, *volatility = nullptr,
          *otime = nullptr, *prices = nullptr;
    int* otype = nullptr;
} g_globals;

int bs_bind_globals(int num_options, int nthreads, float* sptprice, float* strike, float* rate,
                    float* volatility, float* otime, int* otype, float* prices) {
    return guarded("bs_bind_globals", [&]() -> int {
        if (num_options < 0 || nthreads < 1)
            return set_error(BS_ERR_INVALID, "num_options < 0 or nthreads < 1");
        std::lock_guard<std::mutex> lk(g_globals.mu);
        g_globals.num_options = num_options;
        g_globals.nthreads = nthreads;
        g_globals.sptprice = sptprice;
        g_globals.strike = strike;
        g_globals.rate = rate;
        g_globals.volatility = volatility;
        g_globals.otime = otime;
        g_globals.otype = otype;
        g_globals.prices = prices;
        return BS_OK;
    });
}

int bs_thread_compat(void* tid_ptr) {
    return guarded("bs_thread_compat", [&]() -> int {
        constexpr int kNumRuns = 100;  // NUM_RUNS, blackscholes.c:87
        if (!tid_ptr) return set_error(BS_ERR_INVALID, "tid_ptr is NULL");
        const int tid = *static_cast<int*>(tid_ptr);
        int n, nt;
        float *s, *k, *r, *v, *t, *out;
        int* ty;
        {
            std::lock_guard<std::mutex> lk(g_globals.mu);
            n = g_globals.num_options;
            nt = g_globals.nthreads;
            s = g_globals.sptprice;
            k = g_globals.strike;
            r = g_globals.rate;
            v = g_globals.volatility;
            t = g_globals.otime;
            ty = g_globals.otype;
            out = g_globals.prices;
        }
        if (tid < 0 || tid >= nt) return set_error(BS_ERR_INVALID, "tid outside [0, nthreads)");
        // Block of the static partition (blackscholes.c:613-614): N/nthreads
        // options from tid*(N/nthreads); with one thread that is the whole book.
        const int64_t len = n / nt;
        const int64_t start = (int64_t)tid * len;
        if (len == 0) return BS_OK;
        int dev = 0;
        (void)hipGetDevice(&dev);
        return bs_price_host_f32(s + start, k + start, r + start, v + start, t + start, ty + start,
                                 out + start, len, kNumRuns, dev);
    });
}

// ---- device plumbing -----------------------------------------------------------------
int bs_device_count(int* count) {
    if (!count) return set_error(BS_ERR_INVALID, "count is NULL");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return hip_error(e, "hipGetDeviceCount");
    }
    return BS_OK;
}
int bs_set_device(int device) { return done(hipSetDevice(device), "hipSetDevice"); }
int bs_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(BS_ERR_INVALID, "ptr is NULL");
    return done(hipMalloc(ptr, bytes), "hipMalloc");
}
int bs_free(void* ptr) { return done(hipFree(ptr), "hipFree"); }
int bs_memcpy_h2d(void* dst, const void* src, size_t bytes, bs_stream_t stream) {
    return done(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)),
                "hipMemcpyAsync H2D");
}
int bs_memcpy_d2h(void* dst, const void* src, size_t bytes, bs_stream_t stream) {
    return done(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)),
                "hipMemcpyAsync D2H");
}
int bs_memcpy_d2d(void* dst, const void* src, size_t bytes, bs_stream_t stream) {
    return done(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)),
                "hipMemcpyAsync D2D");
}
int bs_memset(void* dst, int value, size_t bytes, bs_stream_t stream) {
    return done(hipMemsetAsync(dst, value, bytes, as_stream(stream)), "hipMemsetAsync");
}
int bs_stream_create(bs_stream_t* stream) {
    if (!stream) return set_error(BS_ERR_INVALID, "stream is NULL");
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = s;
    return done(e, "hipStreamCreate");
}
int bs_stream_destroy(bs_stream_t stream) {
    return done(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
}
int bs_stream_sync(bs_stream_t stream) {
    return done(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
}
int bs_device_sync(void) { return done(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
int bs_stream_wait_event(bs_stream_t stream, bs_event_t ev) {
    if (!ev) return set_error(BS_ERR_INVALID, "ev is NULL");
    return done(hipStreamWaitEvent(as_stream(stream), static_cast<hipEvent_t>(ev), 0),
                "hipStreamWaitEvent");
}
int bs_event_create(bs_event_t* ev) {
    if (!ev) return set_error(BS_ERR_INVALID, "ev is NULL");
    hipEvent_t e = nullptr;
    hipError_t rc = hipEventCreate(&e);
    *ev = e;
    return done(rc, "hipEventCreate");
}
int bs_event_create_timing(bs_event_t* ev) {
    if (!ev) return set_error(BS_ERR_INVALID, "ev is NULL");
    hipEvent_t e = nullptr;
    hipError_t rc = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    *ev = e;
    return done(rc, "hipEventCreateWithFlags");
}
int bs_event_destroy(bs_event_t ev) {
    return done(hipEventDestroy(static_cast<hipEvent_t>(ev)), "hipEventDestroy");
}
int bs_event_record(bs_event_t ev, bs_stream_t stream) {
    return done(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream)), "hipEventRecord");
}
int bs_event_sync(bs_event_t ev) {
    return done(hipEventSynchronize(static_cast<hipEvent_t>(ev)), "hipEventSynchronize");
}
int bs_event_elapsed_ms(bs_event_t start, bs_event_t stop, float* ms) {
    if (!ms) return set_error(BS_ERR_INVALID, "ms is NULL");
    return done(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                    static_cast<hipEvent_t>(stop)),
                "hipEventElapsedTime");
}

}  // extern "C"
