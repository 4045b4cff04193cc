// bs_internal.hpp -- shared host-side declarations of libbs_mi355.so.
#pragma once
#include <atomic>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/bs_mi355.h"

namespace bs {

constexpr int kThreads = 256;  // threads per block: 4 wave64s

// Process-wide launch settings (bs_config_set).  Atomics: any thread may set
// them while others launch; a launch reads each setting once.
struct Config {
    std::atomic<int> check_pointers{1};
    std::atomic<int64_t> max_blocks{0};  // 0: one group per work-item, no grid cap
    std::atomic<int> vector_width{0};    // 0: each kernel's measured default (2 or 4)
    std::atomic<int> nontemporal{1};
    std::atomic<int> pipeline{-1};       // -1: each kernel's measured default
    std::atomic<int> block_threads{0};   // streaming kernels' block size: 0 = each kernel's default
    std::atomic<int> lds_per_wave{-1};   // LDS reserved per wave: -1 = each kernel's default
};
Config& config();

// Runs an entry point's body so that no C++ exception leaves the C ABI: an
// allocation failure becomes BS_ERR_NOMEM, anything else BS_ERR_HIP.
int guard_failed(int status, const char* what, const char* detail) noexcept;
template <class F>
int guarded(const char* what, F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return guard_failed(BS_ERR_NOMEM, what, "out of host memory");
    } catch (const std::exception& e) {
        return guard_failed(BS_ERR_HIP, what, e.what());
    } catch (...) {
        return guard_failed(BS_ERR_HIP, what, "unknown exception");
    }
}

// Thread-local error text behind bs_last_error().
int set_error(int status, const std::string& msg) noexcept;
int hip_error(hipError_t e, const char* what);

// Validates that [p, p + bytes) is memory the GPU may access.
int check_device_range(const void* p, size_t bytes, const char* name);

template <class T>
struct SoaIn {
    const T* s;
    const T* k;
    const T* r;
    const T* v;
    const T* t;
};

struct GreeksOut {
    float* price;
    float* delta;
    float* vega;
    float* rho;
    float* theta;
};

// Kernel launchers (bs_kernels.hip).  They check nothing but alignment.
hipError_t launch_price_f32(SoaIn<float> in, const uint8_t* type, float* out, int64_t n,
                            hipStream_t st);
hipError_t launch_price_i32type_f32(SoaIn<float> in, const int32_t* type, float* out,
                                    int64_t n, hipStream_t st);
hipError_t launch_price_fast_f32(SoaIn<float> in, const uint8_t* type, float* out,
                                     int64_t n, hipStream_t st);
hipError_t launch_price_f64(SoaIn<double> in, const uint8_t* type, double* out,
                            int64_t n, hipStream_t st);
hipError_t launch_greeks_f32(SoaIn<float> in, const uint8_t* type, GreeksOut out,
                             int64_t n, hipStream_t st);
hipError_t launch_cndf_f32(const float* x, float* out, int64_t n, hipStream_t st);
// Portfolio sums (bs_portfolio.hip); w may be NULL.
hipError_t launch_portfolio_f32(SoaIn<float> in, const uint8_t* type, const float* w, int64_t n,
                                bool greeks, double* sums, hipStream_t st);
// The same sums with the device's fixed scratch (bs_portfolio.hip):
// n <= kPfNoAllocMaxOptions; no allocation.
constexpr int64_t kPfNoAllocMaxOptions = int64_t(1) << 25;
hipError_t launch_portfolio_f32_noalloc(SoaIn<float> in, const uint8_t* type, const float* w,
                                        int64_t n, bool greeks, double* sums, hipStream_t st);
hipError_t launch_cndf_f64(const double* x, double* out, int64_t n, hipStream_t st);
hipError_t launch_repack_aos(const void* aos, int64_t n, float* s, float* k, float* r,
                             float* v, float* t, uint8_t* type, hipStream_t st);
hipError_t launch_price_scalar(float s, float k, float r, float v, float t, int put,
                               float* out, hipStream_t st);
hipError_t launch_cndf_scalar(float x, float* out, hipStream_t st);

}  // namespace bs
