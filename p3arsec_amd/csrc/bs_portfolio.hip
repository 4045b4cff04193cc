// bs_portfolio.hip -- portfolio reductions over a book: the MapReduce step
// after the map (SkePU2 MapReduce, skepu2/backend/impl/mapreduce/).
//
// sums[0] = sum_i w_i * price_i and, with Greeks, sums[1..4] = the same sums of
// delta, vega, rho and theta, where each per-option value is exactly the one
// the map computes (the same per-option code, bs_math.hpp) and w_i is the
// position (1 when no weights are given).
//
// Two kernels on the caller's stream.  k_pf_rows: wave w of the grid owns
// kPfRows consecutive rows of 64 two-option groups (the map's loads: one
// 8-byte load per array per lane, so each wave instruction reads 512
// contiguous bytes); a lane walks its column of those rows in order, the next
// row's loads issued before the current row's arithmetic, and adds w*value to
// FP64 accumulators; the wave then folds its lanes with a fixed butterfly into
// one partial.  k_pf_final: one block of 1024 work-items adds the partials
// (work-item j takes partials j, j + 1024, ... in order, then a fixed tree);
// past 65536 partials (books over 2^29 options), k_pf_chunks blocks first add
// contiguous chunks of 2048.  Every grid depends only on n, so the sums are
// bit-reproducible for a given book.
//
// Scratch (the partials and the chunk sums): bs_portfolio_f32 takes it from
// stream-ordered memory (hipMallocAsync/hipFreeAsync), so calls on different
// streams never share it; bs_portfolio_f32_noalloc uses a fixed device array
// of this file (g_pf_scratch) and allocates nothing.
//
// Traffic: 21 bytes per option read (25 with weights), 8 or 40 bytes written
// in all, plus 8 or 40 bytes of partials per kPfRows x 128 options.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "bs_internal.hpp"
#include "bs_math.hpp"
#include "bs_ops.hpp"

namespace bs {
namespace {

constexpr int kPfRows = 8;                                      // rows of 64 groups per wave
// LDS reserved per wave (caps the resident waves): price form none, Greeks
// form 8 KB (20 waves per CU; 10M options 46.0 us against 47.9 with none)
constexpr int pf_lds_per_wave(bool greeks) { return greeks ? 8192 : 0; }
constexpr int kPfChunkThreads = 256;
constexpr int kPfChunkPer = 8;                                  // partials per work-item
constexpr int64_t kPfChunk = kPfChunkThreads * kPfChunkPer;     // partials per chunk block

struct PfArgs {
    SoaIn<float> in;
    const uint8_t* type;
    const float* w;   // NULL: every weight is 1
    int64_t n;
    double* partial;  // [waves][NQ]
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

// One group's inputs: VW = 2 options of aligned arrays (8-byte loads), or
// VW = 1 option of any alignment.  Ix = uint32_t when every array of the call
// spans less than 4 GiB (32-bit offsets on the arrays' bases), else int64_t.
template <int VW>
struct PfGroup {
    using F = vec<float, VW>;
    F s, k, r, v, t, w;
    uint32_t ty;
};

template <int VW, class Ix>
__device__ __forceinline__ PfGroup<VW> pf_load(const PfArgs& a, Ix g) {
    using F = vec<float, VW>;
    PfGroup<VW> x;
    x.s = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.in.s), g));
    x.k = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.in.k), g));
    x.r = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.in.r), g));
    x.v = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.in.v), g));
    x.t = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.in.t), g));
    if constexpr (VW == 2)
        x.ty = __builtin_nontemporal_load(at(reinterpret_cast<const uint16_t*>(a.type), g));
    else
        x.ty = __builtin_nontemporal_load(at(a.type, g));
    if (a.w) x.w = __builtin_nontemporal_load(at(reinterpret_cast<const F*>(a.w), g));
    else x.w = F(1.0f);
    return x;
}

template <int VW, bool kGreeks>
__device__ __forceinline__ void pf_add(double* acc, const PfGroup<VW>& x) {
#pragma unroll
    for (int j = 0; j < VW; ++j)
        add_option<kGreeks>(acc, x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], put_of(x.ty, j), x.w[j]);
}

// Groups of the call: n / VW whole groups, plus (VW = 2, n odd) the last
// option alone, priced by the lane that would own group n / 2.
template <int VW, bool kGreeks, int NT, class Ix>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kGreeks ? 4 : 5)))
void k_pf_rows(PfArgs a) {
    constexpr int NQ = kGreeks ? 5 : 1;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    const Ix ngroups = (Ix)(VW == 2 ? a.n / 2 : a.n);
    const Ix lane = threadIdx.x & 63;
    const Ix wave = (Ix)blockIdx.x * (NT / 64) + threadIdx.x / 64;
    Ix g = wave * (64 * kPfRows) + lane;
    const Ix g_end = g + 64 * kPfRows;
    if (g < ngroups) {
        PfGroup<VW> cur = pf_load<VW>(a, g);
        for (Ix gn = g + 64; gn < g_end && gn < ngroups; gn += 64) {
            const PfGroup<VW> nxt = pf_load<VW>(a, gn);
            pf_add<VW, kGreeks>(acc, cur);
            cur = nxt;
            g = gn;
        }
        pf_add<VW, kGreeks>(acc, cur);
        g += 64;
    }
    if (VW == 2 && (a.n & 1) && g == ngroups && g < g_end) {   // the last option of an odd book
        const int64_t i = a.n - 1;
        add_option<kGreeks>(acc, a.in.s[i], a.in.k[i], a.in.r[i], a.in.v[i], a.in.t[i],
                            a.type[i] != 0, a.w ? a.w[i] : 1.0f);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = acc[q];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
        if (lane == 0) a.partial[(int64_t)wave * NQ + q] = x;
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
        if ((int)threadIdx.x < half)
#pragma unroll
            for (int q = 0; q < NQ; ++q) lds[q][threadIdx.x] += lds[q][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x < NQ) out[(int64_t)blockIdx.x * NQ + threadIdx.x] = lds[threadIdx.x][0];
}

// One block sums count <= kPfFinalMax partials: work-item j adds partials j,
// j + 1024, j + 2048, ... in order, then a fixed tree over the block.
constexpr int kPfFinalThreads = 1024;
constexpr int64_t kPfFinalMax = int64_t(kPfFinalThreads) * 64;

template <int NQ>
__global__ __launch_bounds__(kPfFinalThreads) void k_pf_final(const double* __restrict__ in,
                                                              int64_t count,
                                                              double* __restrict__ out) {
    __shared__ double lds[NQ][kPfFinalThreads];
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += kPfFinalThreads)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += in[i * NQ + q];
#pragma unroll
    for (int q = 0; q < NQ; ++q) lds[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int half = kPfFinalThreads / 2; half >= 1; half >>= 1) {
        if ((int)threadIdx.x < half)
#pragma unroll
            for (int q = 0; q < NQ; ++q) lds[q][threadIdx.x] += lds[q][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x < NQ) out[threadIdx.x] = lds[threadIdx.x][0];
}

// Reduces count partials (count >= 1) to out[0..NQ) in a fixed order: chunk
// sums while more than kPfFinalMax are left, then one final block.  tmp holds
// ceil(count / kPfChunk) entries per chunk level (pf_scratch_doubles counts
// them).
template <int NQ>
void reduce_partials(const double* in, int64_t count, double* tmp, double* out, hipStream_t st) {
    while (count > kPfFinalMax) {
        const int64_t nb = (count + kPfChunk - 1) / kPfChunk;
        hipLaunchKernelGGL((k_pf_chunks<NQ>), dim3((unsigned)nb), dim3(kPfChunkThreads), 0, st,
                           in, count, tmp);
        in = tmp;
        tmp += nb * NQ;
        count = nb;
    }
    hipLaunchKernelGGL((k_pf_final<NQ>), dim3(1), dim3(kPfFinalThreads), 0, st, in, count, out);
}

constexpr int pf_threads(bool greeks) { return greeks ? 256 : 64; }

// Waves of k_pf_rows for n options (VW-option groups plus the odd tail).
constexpr int64_t pf_waves(int64_t n, int vw) {
    const int64_t items = vw == 2 ? n / 2 + (n & 1) : n;
    int64_t waves = (items + 64 * kPfRows - 1) / (64 * kPfRows);
    return waves < 1 ? 1 : waves;
}

// Doubles of scratch one call needs: the partials (whole blocks of waves) and
// every level of chunk sums.
constexpr int64_t pf_scratch_doubles(int64_t n, int vw, bool greeks) {
    const int nq = greeks ? 5 : 1;
    const int64_t wpb = pf_threads(greeks) / 64;
    const int64_t waves = (pf_waves(n, vw) + wpb - 1) / wpb * wpb;
    int64_t room = waves, level = waves;
    while (level > kPfFinalMax) {
        level = (level + kPfChunk - 1) / kPfChunk;
        room += level;
    }
    return room * nq;
}

template <int VW, bool kGreeks>
hipError_t launch_pf(const PfArgs& a, double* sums, hipStream_t st) {
    constexpr int NQ = kGreeks ? 5 : 1;
    constexpr int NT = pf_threads(kGreeks);
    constexpr int64_t wpb = NT / 64;
    const int64_t waves = pf_waves(a.n, VW);
    const int64_t nb = (waves + wpb - 1) / wpb;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    // LDS reserved per block caps the waves resident on a CU (none is used)
    const int lw = config().lds_per_wave.load(std::memory_order_relaxed);
    const size_t lds =
        std::min<size_t>((size_t)(lw < 0 ? pf_lds_per_wave(kGreeks) : lw) * wpb, 65536);
    if (a.n < (int64_t(1) << 30))
        hipLaunchKernelGGL((k_pf_rows<VW, kGreeks, NT, uint32_t>), dim3((unsigned)nb), dim3(NT),
                           lds, st, a);
    else
        hipLaunchKernelGGL((k_pf_rows<VW, kGreeks, NT, int64_t>), dim3((unsigned)nb), dim3(NT),
                           lds, st, a);
    reduce_partials<NQ>(a.partial, nb * wpb, a.partial + nb * wpb * NQ, sums, st);
    return hipGetLastError();
}

bool pf_vec2(const PfArgs& a) {
    auto al = [](const void* p, uintptr_t b) { return (reinterpret_cast<uintptr_t>(p) & (b - 1)) == 0; };
    return al(a.in.s, 8) && al(a.in.k, 8) && al(a.in.r, 8) && al(a.in.v, 8) && al(a.in.t, 8) &&
           al(a.type, 2) && (a.w == nullptr || al(a.w, 8));
}

hipError_t launch_pf_any(const PfArgs& a, bool greeks, double* sums, hipStream_t st) {
    if (pf_vec2(a)) return greeks ? launch_pf<2, true>(a, sums, st) : launch_pf<2, false>(a, sums, st);
    return greeks ? launch_pf<1, true>(a, sums, st) : launch_pf<1, false>(a, sums, st);
}

// ---- the form that allocates nothing -----------------------------------------
// Its scratch is this fixed device array, sized for the largest call it takes
// (kPfNoAllocMaxOptions options with Greeks on unaligned arrays, the most
// partials any call of that size needs): 2.6 MB, held on every device the
// library runs on until the process ends.  In host code the array's name is
// not a device address: its address on each device comes from
// hipGetSymbolAddress, resolved once per device with that device current, and
// only that address reaches a kernel.
constexpr int64_t kPfScratchDoubles = pf_scratch_doubles(kPfNoAllocMaxOptions, 1, true);
static_assert(kPfScratchDoubles >= pf_scratch_doubles(kPfNoAllocMaxOptions, 2, true) &&
                  kPfScratchDoubles >= pf_scratch_doubles(kPfNoAllocMaxOptions, 1, false),
              "scratch sized for the worst form");
__device__ __align__(256) double g_pf_scratch[kPfScratchDoubles];

constexpr int kPfMaxDevices = 64;
std::mutex g_pf_scratch_mu;
double* g_pf_scratch_ptr[kPfMaxDevices] = {};

hipError_t pf_scratch_of(int dev, double** out) {
    if (dev < 0 || dev >= kPfMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_pf_scratch_mu);
    if (!g_pf_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_pf_scratch));
        if (e != hipSuccess) return e;
        g_pf_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_pf_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

// bs_portfolio_f32 with the device's fixed scratch: no allocation.  The caller
// has checked n <= kPfNoAllocMaxOptions.  Calls that overlap on one device
// share the scratch and race.
hipError_t launch_portfolio_f32_noalloc(SoaIn<float> in, const uint8_t* type, const float* w,
                                        int64_t n, bool greeks, double* sums, hipStream_t st) {
    if (n > kPfNoAllocMaxOptions) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    PfArgs a{in, type, w, n, nullptr};
    e = pf_scratch_of(dev, &a.partial);
    if (e != hipSuccess) return e;
    return launch_pf_any(a, greeks, sums, st);
}

hipError_t launch_portfolio_f32(SoaIn<float> in, const uint8_t* type, const float* w, int64_t n,
                                bool greeks, double* sums, hipStream_t st) {
    PfArgs a{in, type, w, n, nullptr};
    const int64_t need = pf_scratch_doubles(n, pf_vec2(a) ? 2 : 1, greeks);
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&a.partial),
                                  (size_t)need * sizeof(double), st);
    if (e != hipSuccess) return e;
    e = launch_pf_any(a, greeks, sums, st);
    const hipError_t f = hipFreeAsync(a.partial, st);
    return e != hipSuccess ? e : f;
}

}  // namespace bs
