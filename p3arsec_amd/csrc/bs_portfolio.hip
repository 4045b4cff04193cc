// bs_portfolio.hip -- portfolio reductions over a book: the MapReduce step
// after the map (SkePU2 MapReduce, skepu2/backend/impl/mapreduce/).
//
// sums[0] = sum_i w_i * price_i and, with Greeks, sums[1..4] = the same sums of
// delta, vega, rho and theta, where each per-option value is exactly the one
// the map computes (the same per-group code, bs_ops.hpp) and w_i is the
// position (1 when no weights are given).
//
// Three kernels on the caller's stream.  k_pf_waves has the map's shape: one
// 2-option group per work-item (64-thread blocks for prices, 256 for Greeks,
// as the maps); each work-item forms w*value in FP64 and each wave folds its
// lanes with a fixed butterfly into one partial.  k_pf_chunks: each block adds
// a contiguous chunk of 2048 partials (eight per work-item, in order, then a
// fixed tree).  k_pf_final: one block adds the chunk sums the same way.  All
// grids depend only on n, so the sums are bit-reproducible for a given book.
// The partials live in stream-ordered memory (hipMallocAsync/hipFreeAsync), so
// calls on different streams never share them.
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

constexpr int kPfChunkThreads = 256;
constexpr int kPfChunkPer = 8;                                  // partials per work-item
constexpr int64_t kPfChunk = kPfChunkThreads * kPfChunkPer;     // partials per chunk block

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

// VW = 2: aligned arrays, one 2-option group per work-item; VW = 1: any
// alignment, one option per work-item.
template <int VW, bool kGreeks, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kGreeks ? 6 : 8)))
void k_pf_waves(PfArgs a) {
    constexpr int NQ = kGreeks ? 5 : 1;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    const int64_t gid = (int64_t)blockIdx.x * NT + threadIdx.x;
    if constexpr (VW == 2) {
        using F = vec<float, 2>;
        if (gid < a.n / 2) {
            const int64_t g = gid;
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
        } else if (gid == a.n / 2 && (a.n & 1)) {   // the last option of an odd book
            const int64_t i = a.n - 1;
            add_option<kGreeks>(acc, a.in.s[i], a.in.k[i], a.in.r[i], a.in.v[i], a.in.t[i],
                                a.type[i] != 0, a.w ? a.w[i] : 1.0f);
        }
    } else {
        if (gid < a.n) {
            const int64_t i = gid;
            add_option<kGreeks>(acc, a.in.s[i], a.in.k[i], a.in.r[i], a.in.v[i], a.in.t[i],
                                a.type[i] != 0, a.w ? a.w[i] : 1.0f);
        }
    }
    const int64_t wave = (int64_t)blockIdx.x * (NT / 64) + threadIdx.x / 64;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = acc[q];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
        if ((threadIdx.x & 63) == 0) a.partial[wave * NQ + q] = x;
    }
}

// Block b sums partials [b*kPfChunk, (b+1)*kPfChunk) (clipped to count): each
// work-item adds its eight (stride kPfChunkThreads, all loads issued first) in
// order, then a fixed tree over the block; out[b] is the chunk's sum.
template <int NQ>
__global__ __launch_bounds__(kPfChunkThreads) void k_pf_chunks(const double* __restrict__ in,
                                                               int64_t count,
                                                               double* __restrict__ out) {
    __shared__ double lds[NQ][kPfChunkThreads];
    const int64_t base = (int64_t)blockIdx.x * kPfChunk + threadIdx.x;
    double v[kPfChunkPer][NQ];
#pragma unroll
    for (int j = 0; j < kPfChunkPer; ++j) {
        const int64_t i = base + (int64_t)j * kPfChunkThreads;
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[j][q] = i < count ? in[i * NQ + q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double acc = v[0][q];
#pragma unroll
        for (int j = 1; j < kPfChunkPer; ++j) acc += v[j][q];
        lds[q][threadIdx.x] = acc;
    }
    __syncthreads();
    for (int half = kPfChunkThreads / 2; half >= 1; half >>= 1) {
        if (threadIdx.x < half)
#pragma unroll
            for (int q = 0; q < NQ; ++q) lds[q][threadIdx.x] += lds[q][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x < NQ) out[(int64_t)blockIdx.x * NQ + threadIdx.x] = lds[threadIdx.x][0];
}

// Reduces count partials (count >= 1) to out[0..NQ) in a fixed order: chunk
// sums until one is left.  tmp holds at least ceil(count / kPfChunk) entries.
template <int NQ>
void reduce_partials(const double* in, int64_t count, double* tmp, double* out, hipStream_t st) {
    while (count > kPfChunk) {
        const int64_t nb = (count + kPfChunk - 1) / kPfChunk;
        hipLaunchKernelGGL((k_pf_chunks<NQ>), dim3((unsigned)nb), dim3(kPfChunkThreads), 0, st,
                           in, count, tmp);
        in = tmp;
        tmp += nb * NQ;
        count = nb;
    }
    hipLaunchKernelGGL((k_pf_chunks<NQ>), dim3(1), dim3(kPfChunkThreads), 0, st, in, count, out);
}

template <int VW, bool kGreeks>
hipError_t launch_pf(const PfArgs& base, double* sums, hipStream_t st) {
    constexpr int NQ = kGreeks ? 5 : 1;
    constexpr int NT = kGreeks ? 256 : 64;
    // work-items: one per group, plus one for the last option of an odd book
    const int64_t items = VW == 2 ? base.n / 2 + (base.n & 1) : base.n;
    int64_t nb = (items + NT - 1) / NT;
    if (nb < 1) nb = 1;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    const int64_t nwaves = nb * (NT / 64);
    // partials, then room for every level of chunk sums (each level is at most
    // 1/kPfChunk of the one before)
    const int64_t room = nwaves + nwaves / kPfChunk + 16;
    PfArgs a = base;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&a.partial),
                                  (size_t)room * NQ * sizeof(double), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_pf_waves<VW, kGreeks, NT>), dim3((unsigned)nb), dim3(NT), 0, st, a);
    reduce_partials<NQ>(a.partial, nwaves, a.partial + nwaves * NQ, sums, st);
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
