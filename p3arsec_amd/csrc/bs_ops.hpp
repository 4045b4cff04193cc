// bs_ops.hpp -- per-group operations of the streaming kernels, shared by the
// batch map (bs_kernels.hip) and the serving loop's fused kernels
// (bs_serve.hip), so both compute every option with the same code.
//
// Each op: Group (registers of one VW-option group), load<NT>(g),
// finish<NT>(g, x) (compute + store; g is the group index, int64_t or, for
// arrays under 4 GiB, uint32_t), scalar(i) for options outside whole groups,
// kElemBytes (the widest array element, for the 4 GiB test), kLdsPerWave
// (LDS bytes reserved per wave of a block, which caps the waves resident on a
// CU: 160 KB / kLdsPerWave; write-heavy streams run faster with fewer waves
// in flight), kWaves (the occupancy the register allocator is held to: 8 waves
// per SIMD = 64 VGPRs), kBlocks (default grid cap, 0 = one group per
// work-item) and kPipe (default loop form); the defaults and each launcher's
// preferred width are the fastest forms measured on MI355X for 10M options.
// Defaults measured on MI355X, 10M options, warm clocks, 300 passes on one /
// two streams, with the agent-scope stores below: price (Ref32) 64 threads,
// no LDS reserved 37.98 / 37.25 us (6 KB per wave: 38.64 / 37.90; 7 KB:
// 40.1 / 39.3); Fast32 64 threads, 6 KB 37.56 / 36.95 us (none: 39.06 /
// 37.80); Greeks 64 threads, four options per work-item, 16 KB per wave
// 65.9 us on two streams (14 KB: 65.7-65.9; 18 KB: 67.1; none: 77.1), two
// options per work-item, 11 KB per wave 67.98 / 67.22 us (none: 74.16 /
// 73.79; 9 KB: 68.18 / 68.52; 12 KB: 70.82 / 69.59); int32-type price 64
// threads 44.9 us, FP64 64 threads pipelined 85 us.  Small blocks help:
// one-wave workgroups retire and refill independently, so the tail of the
// grid is short.  Fewer resident waves help the write-heavy Greeks stream
// most: with 14 waves per CU instead of 32 its five output streams keep fewer
// DRAM pages open at once.
#pragma once
#include <type_traits>

#include "bs_internal.hpp"
#include "bs_math.hpp"

namespace bs {

template <class T, int N>
using vec = T __attribute__((ext_vector_type(N)));

// VW type bytes as one integer: uint16_t for 2 options, uint32_t for 4.
template <int VW>
using tword = typename std::conditional<VW == 4, uint32_t, uint16_t>::type;

template <bool NT, class T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
// Streaming stores.  An 8-byte value goes out as a relaxed agent-scope atomic
// store: the same global_store_dwordx2, with the memory model's agent-scope
// cache policy (sc1), which on MI355X streams the outputs faster than the
// non-temporal hint (10M options: price map on one stream 38.9 -> 37.9 us,
// Greeks 67.9 -> 67.1 us).  Other sizes keep the non-temporal store (the FP64
// map's 16-byte stores split in two ran 4 us slower).
template <bool NT, class T>
__device__ __forceinline__ void st(T* p, T v) {
    if constexpr (NT && sizeof(T) == 8) {
        __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (NT) {
        __builtin_nontemporal_store(v, p);
    } else {
        *p = v;
    }
}

// Element g of an array.  I = int64_t: 64-bit index arithmetic.  I = uint32_t
// (the launcher guarantees every array of the call spans less than 4 GiB): a
// 32-bit byte offset added to the array's base, which the compiler folds into
// the global load/store's SGPR-base + VGPR-offset form, so one offset register
// serves every array of the same element size.
template <class T, class I>
__device__ __forceinline__ T* at(T* base, I g) {
    if constexpr (__is_same(I, uint32_t)) {
        using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
        return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (uint32_t)(g * (uint32_t)sizeof(T)));
    } else {
        return base + g;
    }
}

__device__ __forceinline__ bool put_of(uint32_t word, int j) {
    return ((word >> (8 * j)) & 0xffu) != 0u;
}

template <class Math, int VW>
struct PriceF32 {
    static constexpr int kVW = VW;
    static constexpr int kElemBytes = 4;   // widest array element
    // Fast32 runs best at 26 waves per CU; Ref32's longer arithmetic needs
    // every wave slot to hide its latency
    static constexpr int kLdsPerWave = std::is_same<Math, Fast32>::value ? 6144 : 0;
    static constexpr int kWaves = 8;
    static constexpr int kThreadsPref = 64;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = true;
    using F = vec<float, VW>;
    SoaIn<float> in;
    const uint8_t* type;
    float* out;
    struct Group {
        F s, k, r, v, t;
        tword<VW> ty;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        Group x;
        x.s = ld<NT>(at(reinterpret_cast<const F*>(in.s), g));
        x.k = ld<NT>(at(reinterpret_cast<const F*>(in.k), g));
        x.r = ld<NT>(at(reinterpret_cast<const F*>(in.r), g));
        x.v = ld<NT>(at(reinterpret_cast<const F*>(in.v), g));
        x.t = ld<NT>(at(reinterpret_cast<const F*>(in.t), g));
        x.ty = ld<NT>(at(reinterpret_cast<const tword<VW>*>(type), g));
        return x;
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& x) const {
        F p;
#pragma unroll
        for (int j = 0; j < VW; ++j)
            p[j] = bs_price<Math>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], put_of(x.ty, j));
        st<NT>(at(reinterpret_cast<F*>(out), g), p);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = bs_price<Math>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

template <int VW>
struct PriceI32F32 {
    static constexpr int kVW = VW;
    static constexpr int kElemBytes = 4;   // widest array element
    static constexpr int kLdsPerWave = 0;
    static constexpr int kWaves = 7;
    static constexpr int kThreadsPref = 64;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = true;
    using F = vec<float, VW>;
    using I = vec<int, VW>;
    SoaIn<float> in;
    const int32_t* type;
    float* out;
    struct Group {
        F s, k, r, v, t;
        I ty;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        Group x;
        x.s = ld<NT>(at(reinterpret_cast<const F*>(in.s), g));
        x.k = ld<NT>(at(reinterpret_cast<const F*>(in.k), g));
        x.r = ld<NT>(at(reinterpret_cast<const F*>(in.r), g));
        x.v = ld<NT>(at(reinterpret_cast<const F*>(in.v), g));
        x.t = ld<NT>(at(reinterpret_cast<const F*>(in.t), g));
        x.ty = ld<NT>(at(reinterpret_cast<const I*>(type), g));
        return x;
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& x) const {
        F p;
#pragma unroll
        for (int j = 0; j < VW; ++j)
            p[j] = bs_price<Ref32>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], x.ty[j] != 0);
        st<NT>(at(reinterpret_cast<F*>(out), g), p);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = bs_price<Ref32>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

// fptype = double: a group is VW = 2 options (one 128-bit load per array).
template <int VW>
struct PriceF64 {
    static constexpr int kVW = 2;
    static constexpr int kElemBytes = 8;   // widest array element
    static constexpr int kLdsPerWave = 0;
    static constexpr int kWaves = 5;
    static constexpr int kThreadsPref = 64;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = true;
    using D = vec<double, 2>;
    SoaIn<double> in;
    const uint8_t* type;
    double* out;
    struct Group {
        D s, k, r, v, t;
        uint16_t ty;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        Group x;
        x.s = ld<NT>(at(reinterpret_cast<const D*>(in.s), g));
        x.k = ld<NT>(at(reinterpret_cast<const D*>(in.k), g));
        x.r = ld<NT>(at(reinterpret_cast<const D*>(in.r), g));
        x.v = ld<NT>(at(reinterpret_cast<const D*>(in.v), g));
        x.t = ld<NT>(at(reinterpret_cast<const D*>(in.t), g));
        x.ty = ld<NT>(at(reinterpret_cast<const uint16_t*>(type), g));
        return x;
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& x) const {
        D p;
#pragma unroll
        for (int j = 0; j < 2; ++j)
            p[j] = bs_price<Ref64>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j], put_of(x.ty, j));
        st<NT>(at(reinterpret_cast<D*>(out), g), p);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = bs_price<Ref64>(in.s[i], in.k[i], in.r[i], in.v[i], in.t[i], type[i] != 0);
    }
};

template <int VW>
struct GreeksF32 {
    static constexpr int kVW = VW;
    static constexpr int kElemBytes = 4;   // widest array element
    // VW = 4 (16-byte stores, the launcher's choice): 10 waves per CU; VW = 2: 14
    static constexpr int kLdsPerWave = VW == 4 ? 16384 : 11264;
    static constexpr int kWaves = VW == 4 ? 5 : 8;
    static constexpr int kThreadsPref = 64;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = true;
    using F = vec<float, VW>;
    SoaIn<float> in;
    const uint8_t* type;
    GreeksOut out;
    struct Group {
        F s, k, r, v, t;
        tword<VW> ty;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        Group x;
        x.s = ld<NT>(at(reinterpret_cast<const F*>(in.s), g));
        x.k = ld<NT>(at(reinterpret_cast<const F*>(in.k), g));
        x.r = ld<NT>(at(reinterpret_cast<const F*>(in.r), g));
        x.v = ld<NT>(at(reinterpret_cast<const F*>(in.v), g));
        x.t = ld<NT>(at(reinterpret_cast<const F*>(in.t), g));
        x.ty = ld<NT>(at(reinterpret_cast<const tword<VW>*>(type), g));
        return x;
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& x) const {
        F p, d, ve, rh, th;
#pragma unroll
        for (int j = 0; j < VW; ++j) {
            const Greeks32 o = bs_greeks<Ref32>(x.s[j], x.k[j], x.r[j], x.v[j], x.t[j],
                                                put_of(x.ty, j));
            p[j] = o.price;
            d[j] = o.delta;
            ve[j] = o.vega;
            rh[j] = o.rho;
            th[j] = o.theta;
        }
        st<NT>(at(reinterpret_cast<F*>(out.price), g), p);
        st<NT>(at(reinterpret_cast<F*>(out.delta), g), d);
        st<NT>(at(reinterpret_cast<F*>(out.vega), g), ve);
        st<NT>(at(reinterpret_cast<F*>(out.rho), g), rh);
        st<NT>(at(reinterpret_cast<F*>(out.theta), g), th);
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

template <int VW>
struct CndfF32 {
    static constexpr int kVW = VW;
    static constexpr int kElemBytes = 4;   // widest array element
    static constexpr int kLdsPerWave = 0;
    static constexpr int kWaves = 8;
    static constexpr int kThreadsPref = 256;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = false;
    using F = vec<float, VW>;
    const float* x;
    float* out;
    struct Group {
        F x;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        return Group{ld<NT>(at(reinterpret_cast<const F*>(x), g))};
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& a) const {
        F y;
#pragma unroll
        for (int j = 0; j < VW; ++j) y[j] = Ref32::cndf<false>(a.x[j], nullptr);
        st<NT>(at(reinterpret_cast<F*>(out), g), y);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = Ref32::cndf<false>(x[i], nullptr);
    }
};

template <int VW>
struct CndfF64 {
    static constexpr int kVW = 2;
    static constexpr int kElemBytes = 8;   // widest array element
    static constexpr int kLdsPerWave = 0;
    static constexpr int kWaves = 8;
    static constexpr int kThreadsPref = 256;
    static constexpr int64_t kBlocks = 0;
    static constexpr bool kPipe = false;
    using D = vec<double, 2>;
    const double* x;
    double* out;
    struct Group {
        D x;
    };
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ Group load(Ix g) const {
        return Group{ld<NT>(at(reinterpret_cast<const D*>(x), g))};
    }
    template <bool NT, class Ix = int64_t>
    __device__ __forceinline__ void finish(Ix g, const Group& a) const {
        D y;
        y[0] = Ref64::cndf<false>(a.x[0], nullptr);
        y[1] = Ref64::cndf<false>(a.x[1], nullptr);
        st<NT>(at(reinterpret_cast<D*>(out), g), y);
    }
    __device__ __forceinline__ void scalar(int64_t i) const {
        out[i] = Ref64::cndf<false>(x[i], nullptr);
    }
};

// Occupancy target of a kernel built on Op: the pipelined loop holds one
// more group of registers, so it aims two waves lower.
template <class Op, bool PIPE>
constexpr int waves_for() {
    return PIPE ? (Op::kWaves > 5 ? Op::kWaves - 2 : 3) : Op::kWaves;
}

}  // namespace bs
