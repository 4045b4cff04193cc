// bs_kernels.hip -- gfx950 streaming kernels of the Black-Scholes map.
//
// Layout in HBM: structure of arrays, one array per field (s, k, r, v, t as
// float or double, the option type as one byte).  A work-item prices a group
// of 4 consecutive options: one 128-bit load per float array (two for
// double), one 32-bit load for the 4 type bytes and one 128-bit store, so a
// wave64 moves 1 KiB per array per instruction, fully coalesced.  Groups are
// walked grid-stride (optionally U groups per step to keep more loads in
// flight).  Loads and stores are non-temporal: every byte is touched once.
//
// The pricing arithmetic lives in bs_math.hpp; this file only moves data.
#include "bs_internal.hpp"
#include "bs_math.hpp"

namespace bs {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef int i4 __attribute__((ext_vector_type(4)));

template <bool NT, class T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

__device__ __forceinline__ bool put_of(uint32_t word, int j) {
    return ((word >> (8 * j)) & 0xffu) != 0u;
}

// ---- per-group operations --------------------------------------------------
// Each op: Group (registers of one 4-option group), load<NT>(g), finish<NT>(g,
// x) (compute + store) and scalar(i) for options outside whole groups.

template <class Math>
struct PriceF32 {
    SoaIn<float> in;
    const uint8_t* type;
    float* out;
    struct Group {
        f4 s, k, r, v, t;
        uint32_t ty;
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        Group x;
        x.s = ld<NT>(reinterpret_cast<const f4*>(in.s) + g);
        x.k = ld<NT>(reinterpret_cast<const f4*>(in.k) + g);
        x.r = ld<NT>(reinterpret_cast<const f4*>(in.r) + g);
        x.v = ld<NT>(reinterpret_cast<const f4*>(in.v) + g);
        x.t = ld<NT>(reinterpret_cast<const f4*>(in.t) + g);
        x.ty = ld<NT>(reinterpret_cast<const uint32_t*>(type) + g);
        return x;
    }
    __device__ __forceinline__ static float one(float s, float k, float r, float v, float t,
                                                bool put) {
        return bs_price<Math>(s, k, r, v, t, put);
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& x) const {
        f4 p;
#pragma unroll
        for (int j = 0; j < 4; ++j) p[j] = one(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], put_of(x.ty, j));
        st<NT>(reinterpret_cast<f4*>(out) + g, p);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = one(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

struct PriceI32F32 {
    SoaIn<float> in;
    const int32_t* type;
    float* out;
    struct Group {
        f4 s, k, r, v, t;
        i4 ty;
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        Group x;
        x.s = ld<NT>(reinterpret_cast<const f4*>(in.s) + g);
        x.k = ld<NT>(reinterpret_cast<const f4*>(in.k) + g);
        x.r = ld<NT>(reinterpret_cast<const f4*>(in.r) + g);
        x.v = ld<NT>(reinterpret_cast<const f4*>(in.v) + g);
        x.t = ld<NT>(reinterpret_cast<const f4*>(in.t) + g);
        x.ty = ld<NT>(reinterpret_cast<const i4*>(type) + g);
        return x;
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& x) const {
        f4 p;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            p[j] = bs_price<Ref32>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], x.ty[j] != 0);
        st<NT>(reinterpret_cast<f4*>(out) + g, p);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = bs_price<Ref32>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

struct PriceF64 {
    SoaIn<double> in;
    const uint8_t* type;
    double* out;
    struct Group {
        d2 s[2], k[2], r[2], v[2], t[2];
        uint32_t ty;
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        Group x;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            x.s[h] = ld<NT>(reinterpret_cast<const d2*>(in.s) + 2 * g + h);
            x.k[h] = ld<NT>(reinterpret_cast<const d2*>(in.k) + 2 * g + h);
            x.r[h] = ld<NT>(reinterpret_cast<const d2*>(in.r) + 2 * g + h);
            x.v[h] = ld<NT>(reinterpret_cast<const d2*>(in.v) + 2 * g + h);
            x.t[h] = ld<NT>(reinterpret_cast<const d2*>(in.t) + 2 * g + h);
        }
        x.ty = ld<NT>(reinterpret_cast<const uint32_t*>(type) + g);
        return x;
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& x) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            d2 p;
#pragma unroll
            for (int j = 0; j < 2; ++j)
                p[j] = bs_price<Exact64>(x.s[h][j], x.k[h][j], x.r[h][j], x.v[h][j],
                                         x.t[h][j], put_of(x.ty, 2 * h + j));
            st<NT>(reinterpret_cast<d2*>(out) + 2 * g + h, p);
        }
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = bs_price<Exact64>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

struct GreeksF32 {
    SoaIn<float> in;
    const uint8_t* type;
    GreeksOut out;
    struct Group {
        f4 s, k, r, v, t;
        uint32_t ty;
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        Group x;
        x.s = ld<NT>(reinterpret_cast<const f4*>(in.s) + g);
        x.k = ld<NT>(reinterpret_cast<const f4*>(in.k) + g);
        x.r = ld<NT>(reinterpret_cast<const f4*>(in.r) + g);
        x.v = ld<NT>(reinterpret_cast<const f4*>(in.v) + g);
        x.t = ld<NT>(reinterpret_cast<const f4*>(in.t) + g);
        x.ty = ld<NT>(reinterpret_cast<const uint32_t*>(type) + g);
        return x;
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& x) const {
        f4 p, d, ve, rh, th;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const Greeks32 o = bs_greeks<Ref32>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j],
                                                put_of(x.ty, j));
            p[j] = o.price;
            d[j] = o.delta;
            ve[j] = o.vega;
            rh[j] = o.rho;
            th[j] = o.theta;
        }
        st<NT>(reinterpret_cast<f4*>(out.price) + g, p);
        st<NT>(reinterpret_cast<f4*>(out.delta) + g, d);
        st<NT>(reinterpret_cast<f4*>(out.vega) + g, ve);
        st<NT>(reinterpret_cast<f4*>(out.rho) + g, rh);
        st<NT>(reinterpret_cast<f4*>(out.theta) + g, th);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        const Greeks32 o = bs_greeks<Ref32>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i],
                                            type[i] != 0);
        out.price[i] = o.price;
        out.delta[i] = o.delta;
        out.vega[i] = o.vega;
        out.rho[i] = o.rho;
        out.theta[i] = o.theta;
    }
};

struct CndfF32 {
    const float* x;
    float* out;
    struct Group {
        f4 x;
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        return Group{ld<NT>(reinterpret_cast<const f4*>(x) + g)};
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& a) const {
        f4 y;
#pragma unroll
        for (int j = 0; j < 4; ++j) y[j] = Ref32::cndf<false>(a.x[j], nullptr);
        st<NT>(reinterpret_cast<f4*>(out) + g, y);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = Ref32::cndf<false>(x[i], nullptr);
    }
};

struct CndfF64 {
    const double* x;
    double* out;
    struct Group {
        d2 x[2];
    };
    template <bool NT>
    __device__ __forceinline__ Group load(int64_t g) const {
        Group a;
        a.x[0] = ld<NT>(reinterpret_cast<const d2*>(x) + 2 * g);
        a.x[1] = ld<NT>(reinterpret_cast<const d2*>(x) + 2 * g + 1);
        return a;
    }
    template <bool NT>
    __device__ __forceinline__ void finish(int64_t g, const Group& a) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            d2 y;
            y[0] = Exact64::cndf<false>(a.x[h][0], nullptr);
            y[1] = Exact64::cndf<false>(a.x[h][1], nullptr);
            st<NT>(reinterpret_cast<d2*>(out) + 2 * g + h, y);
        }
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = Exact64::cndf<false>(x[i], nullptr);
    }
};

// ---- generic streaming kernels ----------------------------------------------
// Grid-stride over `ngroups` 4-option groups, U groups per step; options
// [tail_begin, n) (fewer than 4) are priced one per thread by block 0.
template <class Op, int U, bool NT>
__global__ __launch_bounds__(kThreads) void k_stream(Op op, int64_t ngroups,
                                                     int64_t tail_begin, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const int64_t gid = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    int64_t g = gid;
    for (; g + (U - 1) * stride < ngroups; g += U * stride) {
        typename Op::Group x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = op.template load<NT>(g + u * stride);
#pragma unroll
        for (int u = 0; u < U; ++u) op.template finish<NT>(g + u * stride, x[u]);
    }
    if constexpr (U > 1) {
        for (; g < ngroups; g += stride) {
            const typename Op::Group x = op.template load<NT>(g);
            op.template finish<NT>(g, x);
        }
    }
    if (gid < n - tail_begin) op.scalar(tail_begin + gid);
}

// One option per thread, any alignment.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_scalar(Op op, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
        op.scalar(i);
}

inline bool aligned(const void* p, uintptr_t a) {
    return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0;
}

inline unsigned grid_for(int64_t work_items) {
    int64_t blocks = (work_items + kThreads - 1) / kThreads;
    const int64_t cap = config().max_blocks;
    if (cap > 0 && blocks > cap) blocks = cap;
    if (blocks > 0x7fffffff) blocks = 0x7fffffff;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

template <class Op, int U, bool NT>
hipError_t launch_u(const Op& op, int64_t n, hipStream_t st) {
    const int64_t ngroups = n / 4;
    const int64_t per_thread_groups = U;
    unsigned grid = grid_for((ngroups + per_thread_groups - 1) / per_thread_groups);
    hipLaunchKernelGGL((k_stream<Op, U, NT>), dim3(grid), dim3(kThreads), 0, st, op,
                       ngroups, ngroups * 4, n);
    return hipGetLastError();
}

template <class Op>
hipError_t launch_stream(const Op& op, int64_t n, bool vector_ok, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!vector_ok) {
        hipLaunchKernelGGL((k_scalar<Op>), dim3(grid_for(n)), dim3(kThreads), 0, st, op, n);
        return hipGetLastError();
    }
    const bool nt = config().nontemporal != 0;
    switch (config().unroll) {
        case 2: return nt ? launch_u<Op, 2, true>(op, n, st) : launch_u<Op, 2, false>(op, n, st);
        case 4: return nt ? launch_u<Op, 4, true>(op, n, st) : launch_u<Op, 4, false>(op, n, st);
        default: return nt ? launch_u<Op, 1, true>(op, n, st) : launch_u<Op, 1, false>(op, n, st);
    }
}

template <class T>
bool soa_aligned(const SoaIn<T>& in, uintptr_t a) {
    return aligned(in.s, a) && aligned(in.k, a) && aligned(in.r, a) && aligned(in.v, a) &&
           aligned(in.t, a);
}

// ---- AoS -> SoA repack ------------------------------------------------------
// OptionData (blackscholes.c:89-100): 9 dwords per record; dword 6 holds the
// OptionType char in its low byte.  A block stages 256 records (2304 dwords)
// through LDS with coalesced dword loads, then writes each field array
// coalesced.  The stride of 9 dwords is odd, so the LDS reads are free of
// bank conflicts.
constexpr int kRecDwords = 9;

__global__ __launch_bounds__(kThreads) void k_repack_aos(const uint32_t* __restrict__ aos,
                                                          int64_t n, float* s, float* k,
                                                          float* r, float* v, float* t,
                                                          uint8_t* type) {
    __shared__ uint32_t tile[kThreads * kRecDwords];
    for (int64_t base = (int64_t)blockIdx.x * kThreads; base < n;
         base += (int64_t)gridDim.x * kThreads) {
        const int64_t recs = n - base < kThreads ? n - base : kThreads;
        const int64_t words = recs * kRecDwords;
        const uint32_t* src = aos + base * kRecDwords;
#pragma unroll
        for (int j = 0; j < kRecDwords; ++j) {
            const int w = threadIdx.x + j * kThreads;
            if (w < words) tile[w] = src[w];
        }
        __syncthreads();
        if (threadIdx.x < recs) {
            const uint32_t* rec = tile + threadIdx.x * kRecDwords;
            const int64_t i = base + threadIdx.x;
            s[i] = __uint_as_float(rec[0]);
            k[i] = __uint_as_float(rec[1]);
            r[i] = __uint_as_float(rec[2]);
            v[i] = __uint_as_float(rec[4]);
            t[i] = __uint_as_float(rec[5]);
            type[i] = (uint8_t)((rec[6] & 0xffu) == (uint32_t)'P' ? 1 : 0);
        }
        __syncthreads();
    }
}

__global__ void k_price_scalar(float s, float k, float r, float v, float t, int put,
                               float* out) {
    out[0] = bs_price<Ref32>(s, k, r, v, t, put != 0);
}
__global__ void k_cndf_scalar(float x, float* out) { out[0] = Ref32::cndf<false>(x, nullptr); }

}  // namespace

// ---- launchers ---------------------------------------------------------------
hipError_t launch_price_f32(SoaIn<float> in, const uint8_t* type, float* out, int64_t n,
                            hipStream_t st) {
    const bool vec = soa_aligned(in, 16) && aligned(out, 16) && aligned(type, 4);
    return launch_stream(PriceF32<Ref32>{in, type, out}, n, vec, st);
}

hipError_t launch_price_fast_f32(SoaIn<float> in, const uint8_t* type, float* out,
                                 int64_t n, hipStream_t st) {
    const bool vec = soa_aligned(in, 16) && aligned(out, 16) && aligned(type, 4);
    return launch_stream(PriceF32<Fast32>{in, type, out}, n, vec, st);
}

hipError_t launch_price_i32type_f32(SoaIn<float> in, const int32_t* type, float* out,
                                    int64_t n, hipStream_t st) {
    const bool vec = soa_aligned(in, 16) && aligned(out, 16) && aligned(type, 16);
    return launch_stream(PriceI32F32{in, type, out}, n, vec, st);
}

hipError_t launch_price_f64(SoaIn<double> in, const uint8_t* type, double* out, int64_t n,
                            hipStream_t st) {
    const bool vec = soa_aligned(in, 16) && aligned(out, 16) && aligned(type, 4);
    return launch_stream(PriceF64{in, type, out}, n, vec, st);
}

hipError_t launch_greeks_f32(SoaIn<float> in, const uint8_t* type, GreeksOut out, int64_t n,
                             hipStream_t st) {
    const bool vec = soa_aligned(in, 16) && aligned(type, 4) && aligned(out.price, 16) &&
                     aligned(out.delta, 16) && aligned(out.vega, 16) &&
                     aligned(out.rho, 16) && aligned(out.theta, 16);
    return launch_stream(GreeksF32{in, type, out}, n, vec, st);
}

hipError_t launch_cndf_f32(const float* x, float* out, int64_t n, hipStream_t st) {
    return launch_stream(CndfF32{x, out}, n, aligned(x, 16) && aligned(out, 16), st);
}

hipError_t launch_cndf_f64(const double* x, double* out, int64_t n, hipStream_t st) {
    return launch_stream(CndfF64{x, out}, n, aligned(x, 16) && aligned(out, 16), st);
}

hipError_t launch_repack_aos(const void* aos, int64_t n, float* s, float* k, float* r,
                             float* v, float* t, uint8_t* type, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_repack_aos, dim3((unsigned)blocks), dim3(kThreads), 0, st,
                       static_cast<const uint32_t*>(aos), n, s, k, r, v, t, type);
    return hipGetLastError();
}

hipError_t launch_price_scalar(float s, float k, float r, float v, float t, int put,
                               float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_price_scalar, dim3(1), dim3(1), 0, st, s, k, r, v, t, put, out);
    return hipGetLastError();
}

hipError_t launch_cndf_scalar(float x, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_cndf_scalar, dim3(1), dim3(1), 0, st, x, out);
    return hipGetLastError();
}

}  // namespace bs
