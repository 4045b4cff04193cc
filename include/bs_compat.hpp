// bs_compat.hpp -- source-compatible C++ front of libbs_mi355.so for code
// written against the reference's Black-Scholes entry points.
//
// The reference compiles these into its driver (no exported header):
//   fptype CNDF(fptype InputX)                          blackscholes.c:126-184
//   fptype BlkSchlsEqEuroNoDiv(fptype, ..., int otype, float timet)   :190-258
//   void BlkSchlsEqEuroNoDiv(fptype* OptionPrice, int numOptions, ...)
//                                                       blackscholes.simd.c:221-223
//   int bs_thread(void* tid_ptr)                        blackscholes.c:607-658
//   struct map : ff_Map<int> { svc: NUM_RUNS x parallel_for(0, N, ...) }
//                                                       blackscholes.c:315-348
//   ff::ParallelFor::parallel_for / parallel_for_thid   fastflow/ff/parallel_for.hpp:111-317
//   skepu2::Vector<T>(T*, size_t, bool), skepu2 Map<7>  blackscholes_skepu_omp.cpp:710-728
//   int HJM_Swaption_Blocking(FTYPE* pdSwaptionPrice, ...)
//                              swaptions/src/HJM_Swaption_Blocking.cpp:20-222
// Every call below runs on the GPU through the C ABI (include/bs_mi355.h);
// nothing here prices on the host.  Errors throw bs_compat::Error (the C ABI
// itself never throws).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "bs_mi355.h"

namespace bs_compat {

constexpr int kNumRuns = 100;  // NUM_RUNS, blackscholes.c:87

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& where)
        : std::runtime_error(where + ": " + bs_status_string(st) + ": " + bs_last_error()),
          status(st) {}
};

inline void check(int st, const char* where) {
    if (st != BS_OK) throw Error(st, where);
}

// Per-option forms (one device round trip each: for spot checks).
inline float CNDF(float x) {
    float y;
    check(bs_cndf_scalar_f32(x, &y), "CNDF");
    return y;
}

inline float BlkSchlsEqEuroNoDiv(float sptprice, float strike, float rate, float volatility,
                                 float time, int otype, float timet) {
    float p;
    check(bs_price_scalar_f32(sptprice, strike, rate, volatility, time, otype, timet, &p),
          "BlkSchlsEqEuroNoDiv");
    return p;
}

// Pointer-batch form of blackscholes.simd.c:221-223, over host arrays of
// numOptions options (the SSE version handles exactly SIMD_WIDTH lanes and
// ignores numOptions; here numOptions is honoured).
inline void BlkSchlsEqEuroNoDiv(float* OptionPrice, int numOptions, float* sptprice,
                                float* strike, float* rate, float* volatility, float* time,
                                int* otype, float /*timet*/) {
    check(bs_price_host_f32(sptprice, strike, rate, volatility, time, otype, OptionPrice,
                            numOptions, 1, 0),
          "BlkSchlsEqEuroNoDiv(batch)");
}

// HJM_Swaption_Blocking with the reference's signature (FTYPE = double,
// HJM_Swaption_Blocking.cpp:20-40): one swaption's Monte-Carlo price and
// standard error into pdSwaptionPrice[0..1], every path on the GPU.  Returns
// 1 as the reference does; a record the reference cannot price throws.  tid
// (the reference's instrumentation thread id) is unused.
inline int HJM_Swaption_Blocking(double* pdSwaptionPrice, double dStrike, double dCompounding,
                                 double dMaturity, double dTenor, double dPaymentInterval,
                                 int iN, int iFactors, double dYears, const double* pdYield,
                                 double** ppdFactors, long iRndSeed, long lTrials,
                                 int BLOCKSIZE, int /*tid*/ = 0) {
    // ppdFactors is the reference's row-pointer matrix: gather it row-major
    std::vector<double> fac((size_t)iFactors * (iN > 1 ? iN - 1 : 0));
    for (int k = 0; k < iFactors && iN > 1; ++k)
        for (int j = 0; j < iN - 1; ++j) fac[(size_t)k * (iN - 1) + j] = ppdFactors[k][j];
    const bs_swaption_t sw{dStrike, dCompounding, dMaturity, dTenor, dPaymentInterval, dYears,
                           iN, iFactors, pdYield, fac.data()};
    check(bs_swaptions_host_f64(&sw, 1, lTrials, iRndSeed, BLOCKSIZE, &pdSwaptionPrice[0],
                                &pdSwaptionPrice[1], nullptr),
          "HJM_Swaption_Blocking");
    return 1;
}

// The whole loop over the driver's parm array in one call: the FastFlow
// ParallelFor (HJM_Securities.cpp:311-323), the pthreads worker (:102-135),
// the TBB Worker (:74-97) and the SkePU Map<1> (HJM_Securities_skepu.cpp:
// 48-61, 189-200) each call HJM_Swaption_Blocking for swaption i with seed
// swaption_seed + i and store dSimSwaptionMeanPrice / dSimSwaptionStdError.
// Parm is the caller's own record: any type with the reference's field names
// (dStrike, dCompounding, dMaturity, dTenor, dPaymentInterval, iN, iFactors,
// dYears, pdYield, ppdFactors, dSimSwaptionMeanPrice, dSimSwaptionStdError).
// Every path of every swaption runs on the GPU; a record the reference cannot
// price throws.
template <class Parm>
inline void price_swaptions(Parm* swaptions, int nSwaptions, long swaption_seed, long lTrials,
                            int BLOCKSIZE) {
    std::vector<bs_swaption_t> rec((size_t)(nSwaptions > 0 ? nSwaptions : 0));
    std::vector<std::vector<double>> fac(rec.size());
    for (size_t i = 0; i < rec.size(); ++i) {
        const Parm& p = swaptions[i];
        const int w = p.iN > 1 ? p.iN - 1 : 0;
        fac[i].resize((size_t)p.iFactors * w);
        for (int k = 0; k < p.iFactors; ++k)
            for (int j = 0; j < w; ++j) fac[i][(size_t)k * w + j] = p.ppdFactors[k][j];
        rec[i] = bs_swaption_t{p.dStrike, p.dCompounding, p.dMaturity, p.dTenor,
                               p.dPaymentInterval, p.dYears, p.iN, p.iFactors, p.pdYield,
                               fac[i].data()};
    }
    std::vector<double> price(rec.size()), err(rec.size());
    check(bs_swaptions_host_f64(rec.data(), (int64_t)rec.size(), lTrials, swaption_seed,
                                BLOCKSIZE, price.data(), err.data(), nullptr),
          "price_swaptions");
    for (size_t i = 0; i < rec.size(); ++i) {
        swaptions[i].dSimSwaptionMeanPrice = price[i];
        swaptions[i].dSimSwaptionStdError = err[i];
    }
}

// bs_thread drop-in over the caller's globals (blackscholes.c:102-113).
inline void bind_globals(int numOptions, int nThreads, float* sptprice, float* strike,
                         float* rate, float* volatility, float* otime, int* otype,
                         float* prices) {
    check(bs_bind_globals(numOptions, nThreads, sptprice, strike, rate, volatility, otime,
                          otype, prices),
          "bind_globals");
}

inline int bs_thread(void* tid_ptr) {
    check(bs_thread_compat(tid_ptr), "bs_thread");
    return 0;
}

// The FastFlow driver's `map m; m.run(); m.wait();` (blackscholes.c:825-827):
// run() uploads the book once and enqueues NUM_RUNS price passes on a
// library stream; wait() waits for them and copies the prices back.
class Map {
   public:
    Map(int numOptions, const float* sptprice, const float* strike, const float* rate,
        const float* volatility, const float* otime, const int* otype, float* prices,
        int passes = kNumRuns)
        : n_(numOptions), passes_(passes), prices_(prices) {
        const float* h[5] = {sptprice, strike, rate, volatility, otime};
        check(bs_stream_create(&stream_), "Map: stream");
        const size_t fb = (size_t)n_ * sizeof(float);
        for (int i = 0; i < 5; ++i) {
            check(bs_malloc(&d_[i], fb > 0 ? fb : 1), "Map: alloc");
            if (fb) check(bs_memcpy_h2d(d_[i], h[i], fb, stream_), "Map: upload");
        }
        check(bs_malloc(&dty_, fb > 0 ? fb : 1), "Map: alloc");
        check(bs_malloc(&dout_, fb > 0 ? fb : 1), "Map: alloc");
        if (fb) check(bs_memcpy_h2d(dty_, otype, fb, stream_), "Map: upload");
    }
    ~Map() {
        if (stream_) bs_stream_sync(stream_);
        for (void* p : d_) bs_free(p);
        bs_free(dty_);
        bs_free(dout_);
        if (stream_) bs_stream_destroy(stream_);
    }
    Map(const Map&) = delete;
    Map& operator=(const Map&) = delete;

    int run() {
        if (n_ > 0)
            check(bs_price_passes_f32(static_cast<float*>(d_[0]), static_cast<float*>(d_[1]),
                                      static_cast<float*>(d_[2]), static_cast<float*>(d_[3]),
                                      static_cast<float*>(d_[4]), dty_, 4,
                                      static_cast<float*>(dout_), n_, passes_, stream_),
                  "Map::run");
        return 0;
    }
    int wait() {
        if (n_ > 0)
            check(bs_memcpy_d2h(prices_, dout_, (size_t)n_ * sizeof(float), stream_), "Map::wait");
        check(bs_stream_sync(stream_), "Map::wait");
        return 0;
    }

   private:
    int n_, passes_;
    float* prices_;
    bs_stream_t stream_ = nullptr;
    void* d_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* dty_ = nullptr;
    void* dout_ = nullptr;
};

// ---- FastFlow ParallelFor front --------------------------------------------
// The loop body the FastFlow driver hands to parallel_for_thid
// (blackscholes.c:322-330): prices[i] = BlkSchlsEqEuroNoDiv(...) for every i of
// the range, over the caller's host arrays.  The body's arrays stay resident on
// the device between calls; refresh() re-uploads them after the host changed
// an input.
class PriceBody {
   public:
    PriceBody(int numOptions, const float* sptprice, const float* strike, const float* rate,
              const float* volatility, const float* otime, const int* otype, float* prices)
        : n_(numOptions), prices_(prices) {
        h_[0] = sptprice;
        h_[1] = strike;
        h_[2] = rate;
        h_[3] = volatility;
        h_[4] = otime;
        hty_ = otype;
        check(bs_stream_create(&stream_), "PriceBody: stream");
        const size_t fb = (size_t)(n_ > 0 ? n_ : 1) * sizeof(float);
        for (void*& p : d_) check(bs_malloc(&p, fb), "PriceBody: alloc");
        check(bs_malloc(&dty_, fb), "PriceBody: alloc");
        check(bs_malloc(&dout_, fb), "PriceBody: alloc");
        refresh();
    }
    ~PriceBody() {
        if (stream_) bs_stream_sync(stream_);
        for (void* p : d_) bs_free(p);
        bs_free(dty_);
        bs_free(dout_);
        if (stream_) bs_stream_destroy(stream_);
    }
    PriceBody(const PriceBody&) = delete;
    PriceBody& operator=(const PriceBody&) = delete;

    void refresh() {
        const size_t fb = (size_t)n_ * sizeof(float);
        if (fb == 0) return;
        for (int i = 0; i < 5; ++i) check(bs_memcpy_h2d(d_[i], h_[i], fb, stream_), "PriceBody: upload");
        check(bs_memcpy_h2d(dty_, hty_, fb, stream_), "PriceBody: upload");
    }
    // Prices [first, last) on the GPU and copies those prices to the host.
    void run_range(long first, long last) {
        if (first < 0 || last > n_ || first > last) throw Error(BS_ERR_INVALID, "PriceBody: range");
        const int64_t m = last - first;
        if (m == 0) return;
        float* f[5];
        for (int i = 0; i < 5; ++i) f[i] = static_cast<float*>(d_[i]) + first;
        check(bs_price_i32type_f32(f[0], f[1], f[2], f[3], f[4], static_cast<int32_t*>(dty_) + first,
                                   static_cast<float*>(dout_) + first, m, stream_),
              "PriceBody: price");
        check(bs_memcpy_d2h(prices_ + first, static_cast<float*>(dout_) + first,
                            (size_t)m * sizeof(float), stream_),
              "PriceBody: download");
        check(bs_stream_sync(stream_), "PriceBody: sync");
    }

   private:
    int n_;
    float* prices_;
    const float* h_[5];
    const int* hty_;
    bs_stream_t stream_ = nullptr;
    void* d_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* dty_ = nullptr;
    void* dout_ = nullptr;
};

// ff::ParallelFor's interface for the Black-Scholes loop: the same
// constructor and parallel_for / parallel_for_thid / parallel_for_static
// signatures; the body must be a PriceBody (any other body is a compile-time
// error: nothing runs on the host).  The iteration space is priced in one GPU
// launch, so grain, nw, spinwait and the worker count have nothing to set;
// step must be 1 (the reference's loops use 1).
class ParallelFor {
   public:
    explicit ParallelFor(long maxnw = -1, bool spinwait = false, bool spinbarrier = false) {
        (void)maxnw;
        (void)spinwait;
        (void)spinbarrier;
    }
    template <class F>
    void parallel_for(long first, long last, const F& f, long nw = -1) {
        parallel_for(first, last, 1, 0, f, nw);
    }
    template <class F>
    void parallel_for(long first, long last, long step, const F& f, long nw = -1) {
        parallel_for(first, last, step, 0, f, nw);
    }
    template <class F>
    void parallel_for(long first, long last, long step, long grain, const F& f, long nw = -1) {
        static_assert(std::is_same<F, PriceBody>::value,
                      "bs_compat::ParallelFor runs the Black-Scholes body (PriceBody) only");
        (void)grain;
        (void)nw;
        if (step != 1) throw Error(BS_ERR_INVALID, "ParallelFor: step must be 1");
        const_cast<F&>(f).run_range(first, last);
    }
    template <class F>
    void parallel_for_thid(long first, long last, long step, long grain, const F& f, long nw = -1) {
        parallel_for(first, last, step, grain, f, nw);
    }
    template <class F>
    void parallel_for_static(long first, long last, long step, long grain, const F& f,
                             long nw = -1) {
        parallel_for(first, last, step, grain, f, nw);
    }
};

// ---- SkePU2 Vector / Map front ------------------------------------------------
namespace skepu {

// skepu2::BackendSpec{Backend::Type::...}: the GPU types are accepted; the
// CPU and OpenMP types are refused, since this library never prices on the
// host.
struct Backend {
    enum class Type { CPU, OpenMP, CUDA, OpenCL };
};
struct BackendSpec {
    Backend::Type type;
    explicit BackendSpec(Backend::Type t) : type(t) {}
    void setCPUThreads(size_t) {}
    void setGPUThreads(size_t) {}
    void setGPUBlocks(size_t) {}
    void setDevices(size_t) {}
};

// skepu2::Vector<T>(T* p, size_t n, bool dealloc): wraps caller memory with a
// lazily synchronised device copy, as SkePU's CUDA back-end does
// (map_cu.inl:15-58): a map uploads an input only when the host copy is newer
// and marks its output's device copy newer; host reads download it again only
// then.  Non-const host access invalidates the device copy.
template <class T>
class Vector {
   public:
    Vector(T* p, size_t n, bool dealloc = false) : h_(p), n_(n), own_(dealloc) {}
    explicit Vector(size_t n) : h_(new T[n]()), n_(n), own_(true) {}
    ~Vector() {
        if (d_) bs_free(d_);
        if (own_) delete[] h_;
    }
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;

    size_t size() const { return n_; }
    T& operator[](size_t i) {
        updateHost();
        invalidateDeviceData();
        return h_[i];
    }
    const T& operator[](size_t i) const {
        updateHost();
        return h_[i];
    }
    T* data() {
        updateHost();
        invalidateDeviceData();
        return h_;
    }
    void flush() const { updateHost(); }
    void updateHost() const {
        if (host_valid_) return;
        check(bs_memcpy_d2h(h_, d_, n_ * sizeof(T), nullptr), "Vector::updateHost");
        check(bs_stream_sync(nullptr), "Vector::updateHost");
        host_valid_ = true;
        ++downloads_;
    }
    void invalidateDeviceData() { device_valid_ = false; }

    // Device side (used by Map): read access uploads a stale device copy;
    // write access marks the device copy as the newer one.
    T* device_read() {
        alloc();
        if (!device_valid_ && n_ > 0) {
            check(bs_memcpy_h2d(d_, h_, n_ * sizeof(T), nullptr), "Vector::updateDevice");
            ++uploads_;
        }
        device_valid_ = true;
        return static_cast<T*>(d_);
    }
    T* device_write() {
        alloc();
        device_valid_ = true;
        host_valid_ = false;
        return static_cast<T*>(d_);
    }
    int uploads() const { return uploads_; }
    int downloads() const { return downloads_; }

   private:
    void alloc() {
        if (!d_) check(bs_malloc(&d_, (n_ > 0 ? n_ : 1) * sizeof(T)), "Vector: alloc");
    }
    T* h_;
    size_t n_;
    bool own_;
    void* d_ = nullptr;
    mutable bool host_valid_ = true;
    bool device_valid_ = false;
    mutable int downloads_ = 0;
    int uploads_ = 0;
};

// The reference's skepu2 Map<7> over the Black-Scholes user function
// (blackscholes_skepu_omp.cpp:718-728): map(prices, sptprice, strike, rate,
// volatility, otime, otype, data).  The OptionData vector is accepted and not
// read, as the user function does not read it for pricing.
class Map {
   public:
    void setBackend(const BackendSpec& spec) {
        if (spec.type == Backend::Type::CPU || spec.type == Backend::Type::OpenMP)
            throw Error(BS_ERR_INVALID, "skepu::Map: only GPU back-ends are available");
    }
    template <class Any>
    void operator()(Vector<float>& prices, Vector<float>& sptprice, Vector<float>& strike,
                    Vector<float>& rate, Vector<float>& volatility, Vector<float>& otime,
                    Vector<int>& otype, Vector<Any>& /*data*/) {
        const size_t n = prices.size();
        if (sptprice.size() != n || strike.size() != n || rate.size() != n ||
            volatility.size() != n || otime.size() != n || otype.size() != n)
            throw Error(BS_ERR_INVALID, "skepu::Map: size mismatch");
        if (n == 0) return;
        check(bs_price_i32type_f32(sptprice.device_read(), strike.device_read(), rate.device_read(),
                                   volatility.device_read(), otime.device_read(),
                                   otype.device_read(), prices.device_write(), (int64_t)n,
                                   nullptr),
              "skepu::Map");
    }
};

}  // namespace skepu
}  // namespace bs_compat
