// bs_swaptions.hip -- the swaptions map of p3arsec: HJM Monte-Carlo prices of
// a portfolio of swaptions (parsec-ff/pkgs/apps/swaptions/src).
//
// The reference maps HJM_Swaption_Blocking (HJM_Swaption_Blocking.cpp:20-222)
// over the swaptions with a FastFlow ParallelFor (HJM_Securities.cpp:311-323),
// each swaption simulating its trials serially in blocks of BLOCK_SIZE paths.
// Here every Monte-Carlo path of every swaption is one GPU work-item.
//
// Per path p of swaption i the reference draws (n-1) F uniforms from the
// counter seed_i + p (n-1) F + (s-1) F + k (time step s, factor k: its
// generator advances the counter by one per draw and draws path by path),
// maps them through Moro's inverse normal, evolves the forward curve
//   f(s, l) = f(s-1, l+1) + mu[l] dt + sqrt(dt) sum_k sigma_k[l] Z_k(s)
// for every step, discounts along the short rate f(s, 0) and prices the swap
// on the curve f(m, .) at the swaption's start m.  The payoff reads only the
// short rates before m and the curve at m, so the kernel follows each forward
// chain d (f(0,d) -> f(1,d-1) -> ...) through steps 1..min(d, m) only, with
// the reference's additions in the reference's order, and takes the swap
// curve's exponentials only up to the swap's last cash flow (past it the
// reference multiplies its discount factors by zero cash flows; a guard sends
// any path where those factors could be non-finite through the full
// computation).  The prices are the reference's up to the device exp/log
// (last-bit differences) and the order of the final sums.  With the reference's parameters (11 steps, maturity 1
// year on a 5-20 year horizon) m is 1 or 2 of the 10 steps.
//
// Generator: the reference's RanUnif (PARSEC 3.0 RanUnif.cpp, not in the
// reference tree) is the Park-Miller minimal standard generator applied to
// 1513517 c mod (2^31-1); both products are reduced here with the Mersenne
// identity 2^31 = 1 (mod 2^31-1) in 64-bit integers, which gives the same
// residues as its Schrage factorisation.
//
// Kernels (the caller's stream): k_sw_paths, one path at a time per work-item,
// kSwPer paths each, Moro's tail evaluated for a whole wave's draws at once
// (draws_compact), each block's sums of payoff and payoff^2 folded in a fixed
// order into one partial; k_sw_final, one block per swaption, adds its
// partials in a fixed order and forms the mean and standard error as the
// reference does.  Results are bit-reproducible for given inputs.  The
// arithmetic is FP64 (the reference's FTYPE); this file is compiled without
// FMA contraction so every product and sum rounds as the reference's do.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>

#include <algorithm>
#include <mutex>
#include <vector>

#include "bs_internal.hpp"

namespace bs {
namespace {

constexpr int kSwThreads = 256;
constexpr int kSwPer = 4;                 // paths per work-item
constexpr int kSwMaxSteps = 64;           // n_steps
constexpr int kSwMaxFactors = 8;          // n_factors
constexpr int kSwMaxDraws = 64;           // start index x n_factors
constexpr int kSwCompactDraws = 12;       // register forms up to this many draws compact Moro's tail
constexpr int64_t kSwCounterLimit = int64_t(1) << 42;
constexpr uint64_t kM31 = 2147483647u;

// One swaption as the kernel reads it.  pool + off holds f0[n] (the forward
// curve at t = 0), mudt[n-1] (total drift x dt), sigma[F][n-1] and pay[L]
// (the swap's cash flows per unit notional).
struct SwHead {
    int64_t seed;        // counter of path 0's first draw
    double dt, sdt, dts; // HJM step, its square root, the swap curve's step
    int32_t n, m, L, off;
    int32_t last;        // index of the swap's last cash flow (pay[l] = 0 past it)
    int32_t pad;
};
static_assert(sizeof(SwHead) == 56, "SwHead layout");

__device__ __forceinline__ uint64_t mod_m31(uint64_t x) {   // x < 2^63
    x = (x & kM31) + (x >> 31);
    x = (x & kM31) + (x >> 31);
    return x >= kM31 ? x - kM31 : x;
}

// RanUnif of counter c (0 <= c < 2^42).
__device__ __forceinline__ double ran_unif(int64_t c) {
    uint64_t x = mod_m31((uint64_t)c * 1513517u);
    x = mod_m31(x * 16807u);
    return (double)(int64_t)x * 4.656612875e-10;
}

// Moro's inverse normal (Moro 1995; Glasserman, Monte Carlo Methods in
// Financial Engineering, fig. 2.13), as PARSEC's CumNormalInv evaluates it:
// a rational function of x = u - 0.5 for |x| < 0.42, a Chebyshev series in
// log(-log(min(u, 1 - u))) outside.
__device__ __forceinline__ double moro_central(double x) {
    const double r = x * x;
    const double num = x * (((-25.44106049637 * r + 41.39119773534) * r + -18.61500062529) * r +
                            2.50662823884);
    const double den = (((3.13082909833 * r + -21.06224101826) * r + 23.08336743743) * r +
                        -8.47351093090) * r + 1.0;
    return num / den;
}

__device__ __forceinline__ double moro_tail(double u) {
    const double x = u - 0.5;
    double t = x > 0.0 ? 1.0 - u : u;
    t = log(-log(t));
    t = 0.3374754822726147 + t * (0.9761690190917186 + t * (0.1607979714918209 +
        t * (0.0276438810333863 + t * (0.0038405729373609 + t * (0.0003951896511919 +
        t * (0.0000321767881768 + t * (0.0000002888167364 + t * 0.0000003960315187)))))));
    return x < 0.0 ? -t : t;
}

__device__ __forceinline__ double moro_inv(double u) {
    const double x = u - 0.5;
    return fabs(x) < 0.42 ? moro_central(x) : moro_tail(u);
}

// The normal draws of path p: z[s][k] = Z_k(s + 1) for s < m.
template <int F, int MAXM>
__device__ __forceinline__ void draws_direct(const SwHead& h, int64_t p, double (&z)[MAXM][F]) {
    const int64_t c0 = h.seed + p * (int64_t)((h.n - 1) * F);
#pragma unroll
    for (int s = 0; s < MAXM; ++s) {
        if (s >= h.m) break;
#pragma unroll
        for (int k = 0; k < F; ++k) z[s][k] = moro_inv(ran_unif(c0 + s * F + k));
    }
}

// The same draws, with Moro's tail (16% of draws, but taken by nearly every
// wave at every draw when each lane branches on its own) compacted: the lanes
// write the uniforms that need it to their wave's LDS list (ballot + mbcnt),
// every lane then evaluates the tail for list entries lane, lane + 64, ...,
// and each owner reads its results back.  The same operations on the same
// values as draws_direct.  Every thread of the block must call it (it
// synchronises the block); wl is this wave's MAXM F 64 doubles.
template <int F, int MAXM>
__device__ __forceinline__ void draws_compact(const SwHead& h, int64_t p, double (&z)[MAXM][F],
                                              double* wl) {
    const int64_t c0 = h.seed + p * (int64_t)((h.n - 1) * F);
    const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    int pos[MAXM][F];
    bool tail[MAXM][F];
    int count = 0;   // tail entries of this wave so far (uniform)
#pragma unroll
    for (int s = 0; s < MAXM; ++s) {
        const bool live = s < h.m;   // uniform
#pragma unroll
        for (int k = 0; k < F; ++k) {
            const double u = live ? ran_unif(c0 + s * F + k) : 0.25;
            const double x = u - 0.5;
            const bool t = !(fabs(x) < 0.42);
            z[s][k] = moro_central(x);
            const uint64_t b = __ballot(t);
            pos[s][k] = count + (int)__builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
            tail[s][k] = t;
            if (t) wl[pos[s][k]] = u;
            count += __popcll(b);
        }
    }
    __syncthreads();
    for (int r = 0; r < count; r += 64) {
        const int i = r + lane;
        if (i < count) wl[i] = moro_tail(wl[i]);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXM; ++s)
#pragma unroll
        for (int k = 0; k < F; ++k)
            if (tail[s][k]) z[s][k] = wl[pos[s][k]];
    __syncthreads();   // the list is rewritten for the next path
}

// Discounted payoff of a path from its draws z.  F factors; the swaption
// starts at step m <= MAXM.
//
// kAll = true is the reference's arithmetic in full.  kAll = false skips the
// swap-curve exponentials past the last cash flow: those discount factors
// only ever multiply pay[l] = 0, so the terms they make are +0 and leave the
// fixed leg unchanged -- provided every one of them is finite.  That holds
// when the rates past the last cash flow are finite, the discount factor at
// the last cash flow is at most 1e300 and the positive exponents past it sum
// to at most 1 (the factors then stay below 1e300 e); *exact reports it, and
// the caller recomputes the path with kAll = true when it does not hold.
template <int F, int MAXM, bool kAll>
__device__ __forceinline__ double path_payoff(const SwHead& h, const double* __restrict__ q,
                                              const double (&z)[MAXM][F], bool* exact) {
    const int n = h.n, m = h.m, L = h.L, w = n - 1;
    const int last = kAll ? L : h.last;
    const double* f0 = q;
    const double* mudt = q + n;
    const double* sigma = q + n + w;
    const double* pay = q + n + w + F * w;
    double df = 1.0, sdf = 1.0, leg = 0.0, grow = 0.0;
    bool finite = true;
    const int dend = min(n, m + L);
    for (int d = 0; d < dend; ++d) {
        double x = f0[d];
        const int steps = min(d, m);
#pragma unroll
        for (int s = 1; s <= MAXM; ++s) {
            if (s > steps) break;
            const int l = d - s;
            double shock = 0.0;
#pragma unroll
            for (int k = 0; k < F; ++k) shock += sigma[k * w + l] * z[s - 1][k];
            x = x + mudt[l] + h.sdt * shock;
        }
        if (d < m) {
            df *= exp(-x * h.dt);
        } else {
            const int l = d - m;
            if (l <= last) leg += pay[l] * sdf;
            if (l < L - 1) {
                const double y = -x * h.dts;
                if (l < last) {
                    sdf *= exp(y);
                } else {
                    finite = finite && fabs(y) <= 1e300;   // false for inf and NaN
                    grow += y > 0 ? y : 0;
                }
            }
        }
    }
    // curve points past the simulated horizon are zero rates: factor exp(-0) = 1
    for (int l = n - m; l < L; ++l)
        if (l <= last) leg += pay[l] * sdf;
    *exact = kAll || (finite && grow <= 1.0 && sdf <= 1e300);
    const double payoff = leg - 1.0 > 0 ? leg - 1.0 : 0;
    return payoff * df;
}

// Block b covers swaption b / bps, paths [(b % bps) kSwThreads kSwPer, +kSwThreads kSwPer).
// The compacted forms hold the draws, their list slots and the payoff state at
// once; up to 6 draws per path, asking for 6 waves per SIMD keeps them within
// 80 VGPRs without spilling (the reference's shape: 79).
template <int F, int MAXM>
__global__ __launch_bounds__(kSwThreads)
__attribute__((amdgpu_waves_per_eu(MAXM * F <= 6 ? 6 : 1))) void k_sw_paths(const SwHead* __restrict__ heads,
                                                         const double* __restrict__ pool, int bps,
                                                         int64_t paths,
                                                         double* __restrict__ partial) {
    const int sw = (int)(blockIdx.x / (unsigned)bps);
    const int64_t chunk = blockIdx.x % (unsigned)bps;
    const SwHead h = heads[sw];
    const double* q = pool + h.off;
    constexpr bool kCompact = MAXM * F <= kSwCompactDraws;
    __shared__ double lds_tail[kCompact ? kSwThreads / 64 : 1][kCompact ? MAXM * F * 64 : 1];
    double sum = 0.0, sumsq = 0.0;
#pragma unroll 1
    for (int j = 0; j < kSwPer; ++j) {
        const int64_t p = (chunk * kSwPer + j) * kSwThreads + threadIdx.x;
        const bool valid = p < paths;
        double z[MAXM][F];
        if constexpr (kCompact) {
            // every work-item takes part (the draws synchronise the block);
            // past the last path it draws for the last one and discards them
            draws_compact<F, MAXM>(h, valid ? p : paths - 1, z, lds_tail[threadIdx.x / 64]);
        } else {
            if (valid) draws_direct<F, MAXM>(h, p, z);
        }
        if (valid) {
            bool exact;
            double v = path_payoff<F, MAXM, false>(h, q, z, &exact);
            if (!exact) v = path_payoff<F, MAXM, true>(h, q, z, &exact);
            sum += v;
            sumsq += v * v;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sum += __shfl_xor(sum, o, 64);
        sumsq += __shfl_xor(sumsq, o, 64);
    }
    __shared__ double lds[2][kSwThreads / 64];
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x / 64] = sum;
        lds[1][threadIdx.x / 64] = sumsq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = lds[0][0], b = lds[1][0];
#pragma unroll
        for (int i = 1; i < kSwThreads / 64; ++i) {
            a += lds[0][i];
            b += lds[1][i];
        }
        partial[2 * (int64_t)blockIdx.x] = a;
        partial[2 * (int64_t)blockIdx.x + 1] = b;
    }
}

// Block i adds swaption i's bps partials (work-item t takes t, t + 256, ...
// in order, then a fixed tree) and writes out[2i] = mean, out[2i+1] = standard
// error (HJM_Swaption_Blocking.cpp:212-218).
__global__ __launch_bounds__(kSwThreads) void k_sw_final(const double* __restrict__ partial,
                                                         int bps, int64_t trials,
                                                         double* __restrict__ out) {
    __shared__ double lds[2][kSwThreads];
    const double* in = partial + 2 * (int64_t)blockIdx.x * bps;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < bps; i += kSwThreads) {
        a += in[2 * i];
        b += in[2 * i + 1];
    }
    lds[0][threadIdx.x] = a;
    lds[1][threadIdx.x] = b;
    __syncthreads();
    for (int half = kSwThreads / 2; half >= 1; half >>= 1) {
        if ((int)threadIdx.x < half) {
            lds[0][threadIdx.x] += lds[0][threadIdx.x + half];
            lds[1][threadIdx.x] += lds[1][threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double sum = lds[0][0], sumsq = lds[1][0];
        out[2 * blockIdx.x] = sum / trials;
        out[2 * blockIdx.x + 1] =
            sqrt((sumsq - sum * sum / trials) / (trials - 1.0)) / sqrt((double)trials);
    }
}

template <int F, int MAXM>
void launch_paths(unsigned blocks, hipStream_t st, const SwHead* heads, const double* pool,
                  int bps, int64_t paths, double* partial) {
    hipLaunchKernelGGL((k_sw_paths<F, MAXM>), dim3(blocks), dim3(kSwThreads), 0, st, heads, pool,
                       bps, paths, partial);
}

template <int F>
void launch_f(int mmax, unsigned blocks, hipStream_t st, const SwHead* heads, const double* pool,
              int bps, int64_t paths, double* partial) {
    if (mmax <= 2)
        launch_paths<F, 2>(blocks, st, heads, pool, bps, paths, partial);
    else if (mmax <= 4)
        launch_paths<F, 4>(blocks, st, heads, pool, bps, paths, partial);
    else if constexpr (F <= 4) {
        // starts at steps 5-8: draws in registers (86-118 VGPRs), 2.1-2.5x the
        // speed of the generic form, whose draws live in scratch
        if (mmax <= 8)
            launch_paths<F, 8>(blocks, st, heads, pool, bps, paths, partial);
        else
            launch_paths<F, kSwMaxDraws / F>(blocks, st, heads, pool, bps, paths, partial);
    } else {
        launch_paths<F, kSwMaxDraws / F>(blocks, st, heads, pool, bps, paths, partial);
    }
}

using LaunchF = void (*)(int, unsigned, hipStream_t, const SwHead*, const double*, int, int64_t,
                         double*);
constexpr LaunchF kLaunchByFactors[kSwMaxFactors] = {launch_f<1>, launch_f<2>, launch_f<3>,
                                                     launch_f<4>, launch_f<5>, launch_f<6>,
                                                     launch_f<7>, launch_f<8>};

bool finite_all(const bs_swaption_t& s) {
    return std::isfinite(s.strike) && std::isfinite(s.compounding) && std::isfinite(s.maturity) &&
           std::isfinite(s.tenor) && std::isfinite(s.payment_interval) && std::isfinite(s.years);
}

// (int)(x + 0.5) as the reference rounds, or INT32_MIN when x is out of range.
int round_index(double x) {
    const double y = x + 0.5;
    return (y > -1e9 && y < 1e9) ? (int)y : INT32_MIN;
}

// Validates swaption i and appends its head and data (the per-swaption set-up
// of HJM_Swaption_Blocking.cpp:48-150 and HJM.cpp's forward curve and drifts,
// in host FP64 without contraction, as the reference computes them).
int pack_swaption(const bs_swaption_t& s, int64_t i, int64_t seed, SwHead* h,
                  std::vector<double>* pool) {
    const int n = s.n_steps, F = s.n_factors;
    char msg[160];
    auto bad = [&](const char* why) {
        snprintf(msg, sizeof msg, "swaption %lld: %s", (long long)i, why);
        return set_error(BS_ERR_INVALID, msg);
    };
    if (n < 2 || n > kSwMaxSteps) return bad("n_steps must be in [2, 64]");
    if (F < 1 || F > kSwMaxFactors) return bad("n_factors must be in [1, 8]");
    if (!s.yield || !s.factors) return bad("NULL yield or factors");
    if (!finite_all(s) || !(s.years > 0)) return bad("non-finite parameter or years <= 0");
    const double dt = (double)(s.years / n);
    const int freq = round_index(s.payment_interval / dt);
    const int L = round_index(n - s.maturity / dt);
    const int m = round_index(s.maturity / dt);
    const int points = round_index(s.tenor / dt);
    if (freq < 1) return bad("payment interval rounds to less than one step");
    if (m < 0 || m >= n) return bad("maturity outside the simulated horizon");
    if (L < 1 || points < 0 || points > L - 1) return bad("swap tenor outside the horizon");
    if (m * F > kSwMaxDraws) return bad("start step x n_factors exceeds 64");
    const double strike_cont = s.compounding == 0
                                   ? s.strike
                                   : (1 / s.compounding) * log(1 + s.strike * s.compounding);
    const int w = n - 1;
    h->seed = seed;
    h->dt = dt;
    h->sdt = sqrt(dt);
    h->dts = (double)((double)(L * dt) / L);
    h->n = n;
    h->m = m;
    h->L = L;
    h->off = (int32_t)pool->size();
    h->last = points;
    h->pad = 0;
    const size_t base = pool->size();
    pool->resize(base + n + w + (size_t)F * w + L, 0.0);
    double* f0 = pool->data() + base;
    double* mudt = f0 + n;
    double* sig = mudt + w;
    double* pay = sig + (size_t)F * w;
    // forward curve from the yield curve
    f0[0] = s.yield[0];
    for (int j = 1; j < n; ++j) f0[j] = (j + 1) * s.yield[j] - j * s.yield[j - 1];
    // drifts: sum_{l<=j} mu_k[l] = dt/2 (sum_{l<=j} sigma_k[l])^2, then the total
    std::vector<double> drift((size_t)F * w);
    for (int k = 0; k < F; ++k) {
        const double* fk = s.factors + (size_t)k * w;
        double* dk = drift.data() + (size_t)k * w;
        dk[0] = 0.5 * dt * fk[0] * fk[0];
        for (int j = 1; j < w; ++j) {
            double d = 0;
            for (int l = 0; l < j; ++l) d -= dk[l];
            double vol = 0;
            for (int l = 0; l <= j; ++l) vol += fk[l];
            dk[j] = d + 0.5 * dt * vol * vol;
        }
        for (int l = 0; l < w; ++l) sig[(size_t)k * w + l] = fk[l];
    }
    for (int l = 0; l < w; ++l) {
        double t = 0;
        for (int k = 0; k < F; ++k) t += drift[(size_t)k * w + l];
        mudt[l] = t * dt;
    }
    for (int l = 0; l < L; ++l) pay[l] = 0.0;
    for (int l = freq; l <= points; l += freq)
        pay[l] = l != points ? exp(strike_cont * s.payment_interval) - 1
                             : exp(strike_cont * s.payment_interval);
    return BS_OK;
}

// The per-call buffers (parameters, partial sums, results) come from one
// stream-ordered pool per device that keeps freed memory up to 1 GiB: the
// call ends with a stream synchronisation, at which the default pool would
// hand its memory back to the device and the next call would allocate anew.
constexpr int kSwDevices = 64;

hipError_t pool_of(int dev, hipMemPool_t* out) {
    static std::mutex mu;
    static hipMemPool_t pools[kSwDevices] = {};
    if (dev < 0 || dev >= kSwDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!pools[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t p = nullptr;
        hipError_t e = hipMemPoolCreate(&p, &props);
        if (e != hipSuccess) return e;
        uint64_t keep = uint64_t(1) << 30;
        e = hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(p);
            return e;
        }
        pools[dev] = p;
    }
    *out = pools[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace bs

using namespace bs;

extern "C" int bs_swaptions_host_f64(const bs_swaption_t* sw, int64_t n, int64_t trials,
                                     int64_t seed, int32_t block_size, double* price,
                                     double* std_error, bs_stream_t stream) {
    return guarded("bs_swaptions_host_f64", [&]() -> int {
        if (n < 0) return set_error(BS_ERR_INVALID, "n < 0");
        if (n == 0) return BS_OK;
        if (!sw || !price || !std_error) return set_error(BS_ERR_INVALID, "NULL array");
        if (n > (1 << 20)) return set_error(BS_ERR_INVALID, "more than 2^20 swaptions");
        if (trials < 1 || trials > (int64_t(1) << 40))
            return set_error(BS_ERR_INVALID, "trials must be in [1, 2^40]");
        if (block_size < 1) return set_error(BS_ERR_INVALID, "block_size < 1");
        if (seed < 0 || seed >= kSwCounterLimit)
            return set_error(BS_ERR_INVALID, "seed must be in [0, 2^42)");
        const int F = sw[0].n_factors;
        const int64_t paths = (trials + block_size - 1) / block_size * block_size;
        std::vector<SwHead> heads((size_t)n);
        std::vector<double> pool;
        int mmax = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (sw[i].n_factors != F)
                return set_error(BS_ERR_INVALID, "every swaption must have the same n_factors");
            int rc = pack_swaption(sw[i], i, seed + i, &heads[(size_t)i], &pool);
            if (rc) return rc;
            const int64_t draws = (int64_t)(sw[i].n_steps - 1) * F;
            if (paths > (kSwCounterLimit - (seed + i)) / draws)
                return set_error(BS_ERR_INVALID, "seed + paths x draws exceeds 2^42");
            mmax = std::max(mmax, heads[(size_t)i].m);
        }
        if (pool.size() > (size_t)INT32_MAX) return set_error(BS_ERR_INVALID, "too many swaptions");
        const int64_t per_block = (int64_t)kSwThreads * kSwPer;
        const int64_t bps64 = (paths + per_block - 1) / per_block;
        // one launch of at most 2^32 work-items: 2^24 blocks of 1024 paths
        if (bps64 * n > (int64_t(1) << 24))
            return set_error(BS_ERR_INVALID,
                             "more than 2^24 blocks of 1024 paths (2^34 paths) in one call");
        const int bps = (int)bps64;
        const unsigned blocks = (unsigned)(bps64 * n);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        // one device allocation: heads | pool | partials | results
        const size_t hb = ((size_t)n * sizeof(SwHead) + 255) & ~(size_t)255;
        const size_t pb = (pool.size() * sizeof(double) + 255) & ~(size_t)255;
        const size_t qb = (size_t)blocks * 2 * sizeof(double);
        const size_t ob = (size_t)n * 2 * sizeof(double);
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return hip_error(e, "hipGetDevice");
        hipMemPool_t mem = nullptr;
        e = pool_of(dev, &mem);
        if (e != hipSuccess) return hip_error(e, "hipMemPoolCreate");
        char* base = nullptr;
        e = hipMallocFromPoolAsync(reinterpret_cast<void**>(&base), hb + pb + qb + ob, mem, st);
        if (e != hipSuccess) return hip_error(e, "hipMallocFromPoolAsync");
        SwHead* dh = reinterpret_cast<SwHead*>(base);
        double* dp = reinterpret_cast<double*>(base + hb);
        double* dq = reinterpret_cast<double*>(base + hb + pb);
        double* dout = reinterpret_cast<double*>(base + hb + pb + qb);
        std::vector<double> out((size_t)n * 2);
        // heads and pool leave the host as one copy, laid out as on the device
        std::vector<char> staged(hb + pool.size() * sizeof(double));
        memcpy(staged.data(), heads.data(), (size_t)n * sizeof(SwHead));
        memcpy(staged.data() + hb, pool.data(), pool.size() * sizeof(double));
        e = hipMemcpyAsync(base, staged.data(), staged.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            kLaunchByFactors[F - 1](mmax, blocks, st, dh, dp, bps, paths, dq);
            hipLaunchKernelGGL(k_sw_final, dim3((unsigned)n), dim3(kSwThreads), 0, st, dq, bps,
                               trials, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out.data(), dout, ob, hipMemcpyDeviceToHost, st);
        const hipError_t ef = hipFreeAsync(base, st);
        const hipError_t es = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_error(e, "bs_swaptions_host_f64");
        if (ef != hipSuccess) return hip_error(ef, "hipFreeAsync");
        if (es != hipSuccess) return hip_error(es, "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i) {
            price[i] = out[(size_t)(2 * i)];
            std_error[i] = out[(size_t)(2 * i + 1)];
        }
        return BS_OK;
    });
}
