// Test-only: accuracy of the FP64 policy's written-out primitives
// (p3arsec_amd/csrc/bs_math.hpp: log64, recip64) on the GPU against the host
// C library's log and IEEE division, in units in the last place of the host
// result.  Inputs: random doubles over every binade (normal and subnormal),
// dense samples of [0.5, 2] and of 1 +- k ulp, the reciprocal's CNDF range
// [1, 2.4e9] and random binades in [2^-1000, 2^1000], and IEEE special values.
//
//   fp64_prims        prints the worst errors; exit 1 if log64 exceeds 2 ulp,
//                     recip64 exceeds 1 ulp or a special value differs.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "bs_math.hpp"

#define HIP_OK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
            return 2;                                                              \
        }                                                                          \
    } while (0)

__global__ void k_prims(const double* x, double* lg, double* rc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        lg[i] = bs::log64(x[i]);
        rc[i] = bs::recip64(x[i]);
    }
}

static double ulp_of(double r) {
    if (r == 0.0) return 4.9406564584124654e-324;
    const double a = fabs(r);
    return nextafter(a, INFINITY) - a;
}

static double ulp_err(double got, double ref) {
    if (isnan(ref) || isinf(ref)) return (isnan(ref) ? isnan(got) : got == ref) ? 0.0 : 1e300;
    return fabs(got - ref) / ulp_of(ref);
}

static double from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}

int main() {
    std::mt19937_64 rng(20240611);
    std::vector<double> lx, rx;   // log inputs, reciprocal inputs
    for (int i = 0; i < (1 << 22); ++i) {   // every binade, positive
        const uint64_t exp = 1 + rng() % 2046, mant = rng() & ((uint64_t(1) << 52) - 1);
        lx.push_back(from_bits((exp << 52) | mant));
    }
    for (int i = 0; i < (1 << 18); ++i) lx.push_back(from_bits(1 + (rng() & ((uint64_t(1) << 52) - 1))));   // subnormal
    std::uniform_real_distribution<double> u(0.5, 2.0);
    for (int i = 0; i < (1 << 22); ++i) lx.push_back(u(rng));
    for (int k = -4096; k <= 4096; ++k) lx.push_back(k < 0 ? 1.0 - (-k) * 1.1102230246251565e-16 : 1.0 + k * 2.220446049250313e-16);
    const double special[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, -1.0, -1e-300, 1.0, 2.0, 0.5,
                              4.9406564584124654e-324, 1.7976931348623157e308, M_SQRT1_2, M_SQRT2};
    const size_t nspecial = sizeof(special) / sizeof(special[0]);
    for (double s : special) lx.push_back(s);
    std::uniform_real_distribution<double> uc(1.0, 2.4e9);
    for (int i = 0; i < (1 << 22); ++i) rx.push_back(uc(rng));
    for (int i = 0; i < (1 << 22); ++i) {
        const uint64_t exp = 1023 - 1000 + rng() % 2001, mant = rng() & ((uint64_t(1) << 52) - 1);
        rx.push_back(from_bits((exp << 52) | mant) * (rng() & 1 ? -1.0 : 1.0));
    }
    // one buffer of inputs: log inputs then reciprocal inputs
    std::vector<double> x(lx);
    x.insert(x.end(), rx.begin(), rx.end());
    const int64_t n = (int64_t)x.size();
    double *dx, *dl, *dr;
    HIP_OK(hipMalloc(&dx, n * sizeof(double)));
    HIP_OK(hipMalloc(&dl, n * sizeof(double)));
    HIP_OK(hipMalloc(&dr, n * sizeof(double)));
    HIP_OK(hipMemcpy(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice));
    const int threads = 256;
    const int64_t blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL(k_prims, dim3((unsigned)blocks), dim3(threads), 0, 0, dx, dl, dr, n);
    HIP_OK(hipGetLastError());
    std::vector<double> lg(n), rc(n);
    HIP_OK(hipMemcpy(lg.data(), dl, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(rc.data(), dr, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipFree(dx));
    HIP_OK(hipFree(dl));
    HIP_OK(hipFree(dr));

    double log_max = 0, log_at = 0, rcp_max = 0, rcp_at = 0;
    double log_sum = 0;
    int64_t log_exact = 0;
    const size_t nl = lx.size();
    for (size_t i = 0; i < nl; ++i) {
        const double e = ulp_err(lg[i], log(x[i]));
        if (e > log_max) log_max = e, log_at = x[i];
        log_sum += e < 1e299 ? e : 0;
        log_exact += lg[i] == log(x[i]);
    }
    int special_bad = 0;
    for (size_t j = 0; j < nspecial; ++j) {
        const size_t i = nl - nspecial + j;
        const double ref = log(x[i]);
        const bool same = isnan(ref) ? isnan(lg[i]) : lg[i] == ref;
        if (!same && (isnan(ref) || isinf(ref) || ref == 0.0)) {
            printf("log64(%a) = %a, expected %a\n", x[i], lg[i], ref);
            ++special_bad;
        }
    }
    for (size_t i = nl; i < (size_t)n; ++i) {
        const double e = ulp_err(rc[i], 1.0 / x[i]);
        if (e > rcp_max) rcp_max = e, rcp_at = x[i];
    }
    printf("log64: %zu inputs, max %.3f ulp at %a, mean %.4f ulp, %.2f%% identical to the C library\n",
           nl, log_max, log_at, log_sum / nl, 100.0 * log_exact / nl);
    printf("recip64: %zu inputs, max %.3f ulp at %a\n", rx.size(), rcp_max, rcp_at);
    printf("special values differing: %d\n", special_bad);
    return log_max <= 2.0 && rcp_max <= 1.0 && special_bad == 0 ? 0 : 1;
}
