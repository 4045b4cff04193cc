// bs_portfolio.hip -- portfolio reductions over a book: the MapReduce step
// after the map (SkePU2 MapReduce, skepu2/backend/impl/mapreduce/).
//
// sums[0] = sum_i w_i * price_i and, with Greeks, sums[1..4] = the same sums of
// delta, vega, rho and theta, where each per-option value is exactly the one
// the map computes (the same per-group code, bs_ops.hpp) and w_i is the
// position (1 when no weights are given).
//
// Two kernels on the caller's stream.  k_pf_blocks: each work-item walks its
// 2-option groups grid-stride, accumulating w*value in FP64 in a fixed order;
// each wave folds its lanes with a fixed butterfly, and the block's four wave
// sums are added in wave order into one partial per block.  k_pf_final: one
// block adds the partials in a fixed order.  The grid depends only on n, so
// the sums are bit-reproducible for a given book.  The partials live in
// stream-ordered memory (hipMallocAsync/hipFreeAsync), so calls on different
// streams never share them.
//
// Traffic: 21 bytes per option read (25 with weights), 8 or 40 bytes written
// in all; the price form is bound by HBM reads and VALU together, the Greeks
// form by VALU.
#include <type_traits>

#include "bs_internal.hpp"
#include "bs_math.hpp"
#include "bs_ops.hpp"

namespace bs {
namespace {

constexpr int kPfThreads = 256;
constexpr int64_t kPfMaxBlocks = 16384;
constexpr int kPfFinalThreads = 256;

struct PfArgs {
    SoaIn<float> in;
    const uint8_t* type;
    const float* w;   // NULL: every weight is 1
    int64_t n;
    double* partial;  // [gridDim.x][NQ]
};

template <bool kGreeks>
__device__ __forceinline__ void add_option(double* acc, float s, float k, float r, float v,
                                           float t, bool put, float w) {
    const double dw = (double)w;
    if constexpr (kGreeks) {
        const Greeks32 g = bs_greeks<Ref32>(s, k, r, v, t, put);
        acc[0] += dw * (double)g.price;
        acc[1] += dw * (double)g.delta;
        acc[2] += dw * (double)g.vega;
        acc[3] += dw * (double)g.rho;
        acc[4] += dw * (double)g.theta;
    } else {
        acc[0] += dw * (double)bs_price<Ref32>(s, k, r, v, t, put);
    }
}

template <int NQ>
__device__ __forceinline__ void block_sum(double* acc, double* out) {
    __shared__ double lds[kPfThreads / 64][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = acc[q];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
        acc[q] = x;
    }
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) lds[wave][q] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x = lds[0][q];
            for (int wv = 1; wv < kPfThreads / 64; ++wv) x += lds[wv][q];
            out[q] = x;
        }
}

// VW = 2: aligned arrays, one 2-option group per step; VW = 1: any alignment.
template <int VW, bool kGreeks>
__global__ __launch_bounds__(kPfThreads) __attribute__((amdgpu_waves_per_eu(kGreeks ? 6 : 8)))
void k_pf_blocks(PfArgs a) {
    constexpr int NQ = kGreeks ? 5 : 1;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kPfThreads;
    const int64_t gid = (int64_t)blockIdx.x * kPfThreads + threadIdx.x;
    const int64_t ng = a.n / VW;
    if constexpr (VW == 2) {
        using F = vec<float, 2>;
        for (int64_t g = gid; g < ng; g += stride) {
            const F s = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.in.s) + g);
            const F k = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.in.k) + g);
            const F r = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.in.r) + g);
            const F v = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.in.v) + g);
            const F t = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.in.t) + g);
            const uint16_t ty = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(a.type) + g);
            F w = {1.0f, 1.0f};
            if (a.w) w = __builtin_nontemporal_load(reinterpret_cast<const F*>(a.w) + g);
            add_option<kGreeks>(acc, s[0], k[0], r[0], v[0], t[0], put_of(ty, 0), w[0]);
            add_option<kGreeks>(acc, s[1], k[1], r[1], v[1], t[1], put_of(ty, 1), w[1]);
        }
        if (gid == 0 && (a.n & 1)) {   // the last option of an odd book
            const int64_t i = a.n - 1;
            add_option<kGreeks>(acc, a.in.s[i], a.in.k[i], a.in.r[i], a.in.v[i], a.in.t[i],
                                a.type[i] != 0, a.w ? a.w[i] : 1.0f);
        }
    } else {
        for (int64_t i = gid; i < a.n; i += stride)
            add_option<kGreeks>(acc, a.in.s[i], a.in.k[i], a.in.r[i], a.in.v[i], a.in.t[i],
                                a.type[i] != 0, a.w ? a.w[i] : 1.0f);
    }
    block_sum<NQ>(acc, a.partial + (int64_t)blockIdx.x * NQ);
}

// One block: thread t adds partials t, t + 256, ... in order, then a fixed tree.
template <int NQ>
__global__ __launch_bounds__(kPfFinalThreads) void k_pf_final(const double* __restrict__ partial,
                                                              int64_t nblocks,
                                                              double* __restrict__ sums) {
    __shared__ double lds[kPfFinalThreads][NQ];
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += kPfFinalThreads)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += partial[b * NQ + q];
#pragma unroll
    for (int q = 0; q < NQ; ++q) lds[threadIdx.x][q] = acc[q];
    __syncthreads();
    for (int half = kPfFinalThreads / 2; half >= 1; half >>= 1) {
        if (threadIdx.x < half)
#pragma unroll
            for (int q = 0; q < NQ; ++q) lds[threadIdx.x][q] += lds[threadIdx.x + half][q];
        __syncthreads();
    }
    if (threadIdx.x < NQ) sums[threadIdx.x] = lds[0][threadIdx.x];
}

template <int VW, bool kGreeks>
hipError_t launch_pf(const PfArgs& base, double* sums, hipStream_t st) {
    constexpr int NQ = kGreeks ? 5 : 1;
    const int64_t items = VW == 2 ? base.n / 2 : base.n;
    int64_t nb = (items + kPfThreads - 1) / kPfThreads;
    if (nb > kPfMaxBlocks) nb = kPfMaxBlocks;
    if (nb < 1) nb = 1;
    PfArgs a = base;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&a.partial),
                                  (size_t)nb * NQ * sizeof(double), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_pf_blocks<VW, kGreeks>), dim3((unsigned)nb), dim3(kPfThreads), 0, st, a);
    hipLaunchKernelGGL((k_pf_final<NQ>), dim3(1), dim3(kPfFinalThreads), 0, st, a.partial, nb,
                       sums);
    e = hipGetLastError();
    const hipError_t f = hipFreeAsync(a.partial, st);
    return e != hipSuccess ? e : f;
}

}  // namespace

hipError_t launch_portfolio_f32(SoaIn<float> in, const uint8_t* type, const float* w, int64_t n,
                                bool greeks, double* sums, hipStream_t st) {
    auto al = [](const void* p, uintptr_t b) { return (reinterpret_cast<uintptr_t>(p) & (b - 1)) == 0; };
    const bool vec2 = al(in.s, 8) && al(in.k, 8) && al(in.r, 8) && al(in.v, 8) && al(in.t, 8) &&
                      al(type, 2) && (w == nullptr || al(w, 8));
    const PfArgs a{in, type, w, n, nullptr};
    if (vec2) return greeks ? launch_pf<2, true>(a, sums, st) : launch_pf<2, false>(a, sums, st);
    return greeks ? launch_pf<1, true>(a, sums, st) : launch_pf<1, false>(a, sums, st);
}

}  // namespace bs
