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
// Every call below runs on the GPU through the C ABI (include/bs_mi355.h);
// nothing here prices on the host.  Errors throw bs_compat::Error (the C ABI
// itself never throws).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>

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
        for (int j = 0; j < passes_ && n_ > 0; ++j)
            check(bs_price_i32type_f32(static_cast<float*>(d_[0]), static_cast<float*>(d_[1]),
                                       static_cast<float*>(d_[2]), static_cast<float*>(d_[3]),
                                       static_cast<float*>(d_[4]), static_cast<int32_t*>(dty_),
                                       static_cast<float*>(dout_), n_, stream_),
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

}  // namespace bs_compat
