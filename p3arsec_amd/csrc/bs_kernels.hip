// bs_kernels.hip -- gfx950 streaming kernels of the Black-Scholes map.
//
// Layout in HBM: structure of arrays, one array per field (s, k, r, v, t as
// float or double, the option type as one byte).  A work-item handles a group
// of VW consecutive options: one VW*4-byte load per float array, one VW-byte
// load of the type bytes and one store per output array, so every wave64
// instruction moves a contiguous 64*VW*4-byte span (fully coalesced).  VW is 4
// (128-bit accesses) or 2 (64-bit): on MI355X the 2-wide form is the faster
// one for every FP32 stream here (10M options: Greeks 71 vs 75 us, price 43
// vs 46 us).  Groups are walked grid-stride, optionally software-pipelined
// (the next group's loads issued before the current group's math).
// Loads and stores are non-temporal: every byte is touched once.
//
// The pricing arithmetic lives in bs_math.hpp; this file only moves data.
#include <algorithm>

#include "bs_internal.hpp"
#include "bs_math.hpp"
#include "bs_ops.hpp"

namespace bs {
namespace {

// ---- argument packs and their op constructors ----------------------------------
struct PriceArgs {
    SoaIn<float> in;
    const uint8_t* type;
    float* out;
};
struct PriceI32Args {
    SoaIn<float> in;
    const int32_t* type;
    float* out;
};
struct PriceF64Args {
    SoaIn<double> in;
    const uint8_t* type;
    double* out;
};
struct GreeksArgs {
    SoaIn<float> in;
    const uint8_t* type;
    GreeksOut out;
};
template <class T>
struct CndfArgs {
    const T* x;
    T* out;
};

template <class Op>
Op make_op(const PriceArgs& a) { return Op{a.in, a.type, a.out}; }
template <class Op>
Op make_op(const PriceI32Args& a) { return Op{a.in, a.type, a.out}; }
template <class Op>
Op make_op(const PriceF64Args& a) { return Op{a.in, a.type, a.out}; }
template <class Op>
Op make_op(const GreeksArgs& a) { return Op{a.in, a.type, a.out}; }
template <class Op>
Op make_op(const CndfArgs<float>& a) { return Op{a.x, a.out}; }
template <class Op>
Op make_op(const CndfArgs<double>& a) { return Op{a.x, a.out}; }


// ---- generic streaming kernels ----------------------------------------------
// Grid-stride over `ngroups` groups of VW options; the options [tail_begin, n)
// that do not fill a group are priced one per work-item by block 0.  PIPE
// issues the loads of a work-item's next group before pricing the current one
// (register double buffering), so a wave overlaps its own HBM latency with its
// arithmetic; it holds one more group of registers, so its occupancy target
// is two waves lower.  Ix is the group index type: uint32_t when every array
// spans less than 4 GiB (32-bit offsets on the arrays' SGPR bases; ngroups <
// 2^30 there, so g + stride cannot wrap), int64_t otherwise.
template <class Op, bool NT, bool PIPE, class Ix>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(waves_for<Op, PIPE>())))
void k_stream(Op op, int64_t ngroups, int64_t tail_begin, int64_t n) {
    const Ix stride = (Ix)gridDim.x * blockDim.x;
    const Ix gid = (Ix)blockIdx.x * blockDim.x + threadIdx.x;
    const Ix ng = (Ix)ngroups;
    Ix g = gid;
    if constexpr (PIPE) {
        if (g < ng) {
            typename Op::Group cur = op.template load<NT>(g);
            for (Ix gn = g + stride; gn < ng; gn += stride) {
                const typename Op::Group nxt = op.template load<NT>(gn);
                op.template finish<NT>(g, cur);
                cur = nxt;
                g = gn;
            }
            op.template finish<NT>(g, cur);
        }
    } else {
        for (; g < ng; g += stride) {
            const typename Op::Group x = op.template load<NT>(g);
            op.template finish<NT>(g, x);
        }
    }
    if ((int64_t)gid < n - tail_begin) op.scalar(tail_begin + (int64_t)gid);
}

// One option per work-item, any alignment.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_scalar(Op op, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
        op.scalar(i);
}

inline bool aligned(const void* p, uintptr_t a) {
    return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0;
}

inline unsigned grid_for(int64_t work_items, int64_t cap, int threads = kThreads) {
    int64_t blocks = (work_items + threads - 1) / threads;
    if (cap > 0 && blocks > cap) blocks = cap;
    if (blocks > 0x7fffffff) blocks = 0x7fffffff;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

template <class Op, bool NT, bool PIPE>
hipError_t launch_form(const Op& op, int64_t n, hipStream_t st) {
    constexpr int VW = Op::kVW;
    const int64_t ngroups = n / VW;
    const int64_t mb = config().max_blocks.load(std::memory_order_relaxed);
    const int bt = config().block_threads.load(std::memory_order_relaxed);
    const int64_t cap = mb > 0 ? mb : Op::kBlocks;
    const int nt = bt > 0 ? bt : Op::kThreadsPref;
    const unsigned grid = grid_for(std::max<int64_t>(ngroups, 1), cap, nt);
    // LDS reserved per block caps the waves resident on a CU (none is used)
    const int lw = config().lds_per_wave.load(std::memory_order_relaxed);
    const size_t lds = std::min<size_t>(
        (size_t)(lw < 0 ? Op::kLdsPerWave : lw) * (size_t)((nt + 63) / 64), 65536);
    if (n < (int64_t(1) << 32) / Op::kElemBytes)
        hipLaunchKernelGGL((k_stream<Op, NT, PIPE, uint32_t>), dim3(grid), dim3(nt), lds, st, op,
                           ngroups, ngroups * VW, n);
    else
        hipLaunchKernelGGL((k_stream<Op, NT, PIPE, int64_t>), dim3(grid), dim3(nt), lds, st, op,
                           ngroups, ngroups * VW, n);
    return hipGetLastError();
}

template <class Op>
hipError_t launch_op(const Op& op, int64_t n, hipStream_t st) {
    const bool nt = config().nontemporal.load(std::memory_order_relaxed) != 0;
    const int pl = config().pipeline.load(std::memory_order_relaxed);
    const bool pipe = pl < 0 ? Op::kPipe : pl != 0;
    if (pipe)
        return nt ? launch_form<Op, true, true>(op, n, st) : launch_form<Op, false, true>(op, n, st);
    return nt ? launch_form<Op, true, false>(op, n, st) : launch_form<Op, false, false>(op, n, st);
}

// Instantiates OP<VW> for the configured (or the op's preferred) width;
// `align` gives the byte alignment every pointer needs for width VW, and a
// misaligned call takes the one-option-per-work-item kernel.
template <template <int> class OP, class Args, class AlignFn>
hipError_t launch_stream(const Args& a, int64_t n, int preferred_vw, AlignFn aligned_for,
                         hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int cvw = config().vector_width.load(std::memory_order_relaxed);
    int vw = cvw > 0 ? cvw : preferred_vw;
    if (vw != 2 && vw != 4) vw = 4;
    if (vw == 4 && !aligned_for(4)) vw = 2;
    if (!aligned_for(vw)) {
        const OP<2> op = make_op<OP<2>>(a);
        hipLaunchKernelGGL((k_scalar<OP<2>>), dim3(grid_for(n, 0)), dim3(kThreads), 0, st, op, n);
        return hipGetLastError();
    }
    if (vw == 4) return launch_op(make_op<OP<4>>(a), n, st);
    return launch_op(make_op<OP<2>>(a), n, st);
}

template <class T>
bool soa_aligned(const SoaIn<T>& in, uintptr_t a) {
    return aligned(in.s, a) && aligned(in.k, a) && aligned(in.r, a) && aligned(in.v, a) &&
           aligned(in.t, a);
}

// ---- AoS -> SoA repack ------------------------------------------------------
// OptionData (blackscholes.c:89-100): 9 dwords per record; dword 6 holds the
// OptionType char in its low byte.  A block stages 256 records (2304 dwords,
// 576 16-byte words) through LDS with coalesced 16-byte non-temporal loads,
// then each thread writes one record's fields, so every field array is
// written coalesced (non-temporal: nothing is read back).  The stride of 9
// dwords between the records threads read is odd, so the LDS reads are free
// of bank conflicts.  A step that is partial, or whose records do not start
// on a 16-byte boundary (the reference's records are 4-byte aligned), loads
// dword by dword.  Measured on MI355X, 10M records (57 B each: 36 in, 21
// out): 92-94 us per pass, 6.1-6.2 TB/s; dword loads and cached stores
// 136.6 us; 4 or 2 records per thread, or other block sizes, 93-102 us.
constexpr int kRecDwords = 9;
using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
constexpr int kRepackWords4 = kThreads * kRecDwords / 4;   // 16-byte words per step

__global__ __launch_bounds__(kThreads) void k_repack_aos(const uint32_t* __restrict__ aos,
                                                          int64_t n, float* s, float* k,
                                                          float* r, float* v, float* t,
                                                          uint8_t* type) {
    __shared__ u32x4 tile4[kRepackWords4];
    const uint32_t* tile = reinterpret_cast<const uint32_t*>(tile4);
    for (int64_t base = (int64_t)blockIdx.x * kThreads; base < n;
         base += (int64_t)gridDim.x * kThreads) {
        const int64_t recs = n - base < kThreads ? n - base : kThreads;
        const uint32_t* src = aos + base * kRecDwords;
        if (recs == kThreads && (reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
            const u32x4* src4 = reinterpret_cast<const u32x4*>(src);
            for (int w = threadIdx.x; w < kRepackWords4; w += kThreads)
                tile4[w] = __builtin_nontemporal_load(src4 + w);
        } else {
            uint32_t* tw = reinterpret_cast<uint32_t*>(tile4);
            const int words = (int)recs * kRecDwords;
            for (int w = threadIdx.x; w < words; w += kThreads) tw[w] = src[w];
        }
        __syncthreads();
        if (threadIdx.x < recs) {
            const uint32_t* rec = tile + threadIdx.x * kRecDwords;
            const int64_t i = base + threadIdx.x;
            __builtin_nontemporal_store(__uint_as_float(rec[0]), s + i);
            __builtin_nontemporal_store(__uint_as_float(rec[1]), k + i);
            __builtin_nontemporal_store(__uint_as_float(rec[2]), r + i);
            __builtin_nontemporal_store(__uint_as_float(rec[4]), v + i);
            __builtin_nontemporal_store(__uint_as_float(rec[5]), t + i);
            __builtin_nontemporal_store(
                (uint8_t)((rec[6] & 0xffu) == (uint32_t)'P' ? 1 : 0), type + i);
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

namespace {
template <int VW>
using PriceRef = PriceF32<Ref32, VW>;
template <int VW>
using PriceFast = PriceF32<Fast32, VW>;

// Alignment test of a price-shaped call for width vw: floats need 4*vw bytes,
// the uint8 type array vw bytes (int32 types: 4*vw).
template <class In, class Ty, class Out>
auto price_align(const In& in, const Ty* type, const Out* out, size_t ty_elem) {
    return [=](int vw) {
        const uintptr_t fa = sizeof(typename std::remove_pointer<decltype(in.s)>::type) * vw;
        return soa_aligned(in, fa) && aligned(out, fa) && aligned(type, ty_elem * vw);
    };
}
}  // namespace

// ---- launchers ---------------------------------------------------------------
hipError_t launch_price_f32(SoaIn<float> in, const uint8_t* type, float* out, int64_t n,
                            hipStream_t st) {
    return launch_stream<PriceRef>(PriceArgs{in, type, out}, n, 2, price_align(in, type, out, 1),
                                   st);
}

hipError_t launch_price_fast_f32(SoaIn<float> in, const uint8_t* type, float* out,
                                 int64_t n, hipStream_t st) {
    return launch_stream<PriceFast>(PriceArgs{in, type, out}, n, 2, price_align(in, type, out, 1),
                                    st);
}

hipError_t launch_price_i32type_f32(SoaIn<float> in, const int32_t* type, float* out,
                                    int64_t n, hipStream_t st) {
    return launch_stream<PriceI32F32>(PriceI32Args{in, type, out}, n, 2,
                                      price_align(in, type, out, 4), st);
}

hipError_t launch_price_f64(SoaIn<double> in, const uint8_t* type, double* out, int64_t n,
                            hipStream_t st) {
    auto al = [=](int) { return soa_aligned(in, 16) && aligned(out, 16) && aligned(type, 2); };
    return launch_stream<PriceF64>(PriceF64Args{in, type, out}, n, 2, al, st);
}

hipError_t launch_greeks_f32(SoaIn<float> in, const uint8_t* type, GreeksOut out, int64_t n,
                             hipStream_t st) {
    auto al = [=](int vw) {
        const uintptr_t fa = 4 * vw;
        return soa_aligned(in, fa) && aligned(type, vw) && aligned(out.price, fa) &&
               aligned(out.delta, fa) && aligned(out.vega, fa) && aligned(out.rho, fa) &&
               aligned(out.theta, fa);
    };
    // four options per work-item: 16-byte stores of the five outputs (65.5 against
    // 66.5 us per 10M-option pass with two per work-item)
    return launch_stream<GreeksF32>(GreeksArgs{in, type, out}, n, 4, al, st);
}

hipError_t launch_cndf_f32(const float* x, float* out, int64_t n, hipStream_t st) {
    auto al = [=](int vw) { return aligned(x, 4 * vw) && aligned(out, 4 * vw); };
    return launch_stream<CndfF32>(CndfArgs<float>{x, out}, n, 4, al, st);
}

hipError_t launch_cndf_f64(const double* x, double* out, int64_t n, hipStream_t st) {
    auto al = [=](int) { return aligned(x, 16) && aligned(out, 16); };
    return launch_stream<CndfF64>(CndfArgs<double>{x, out}, n, 2, al, st);
}

hipError_t launch_repack_aos(const void* aos, int64_t n, float* s, float* k, float* r,
                             float* v, float* t, uint8_t* type, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 1048576) blocks = 1048576;   // past 2^28 records, blocks loop
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
