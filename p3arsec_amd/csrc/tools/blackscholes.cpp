// blackscholes -- the PARSEC blackscholes command line on the MI355X.
//
//   blackscholes <nthreads> <inputFile> <outputFile> [--check] [--passes N]
//
// Same input and output formats as the reference driver
// (parsec-ff/pkgs/apps/blackscholes/src/blackscholes.c:664-960): the input is
// "N" then N lines "s strike r divq v t type divs DGrefval"; the output is
// "N" then N prices printed "%.18f".  The AoS records are uploaded as they
// were read and repacked to SoA on the device (bs_repack_aos_f32, the
// reference's repack at blackscholes.c:747-767), then NUM_RUNS = 100 passes of
// the price map run on the GPU (the reference's ROI, blackscholes.c:782-913).
// <nthreads> is accepted for command-line compatibility; the GPU map does not
// use host threads (the text formats are read and written on up to 16).
// --check runs the reference's ERR_CHK test (blackscholes.c:333-340): prices
// off the DGrefval column by 1e-4 or more, counted and reported once per pass
// as the reference's build with ERR_CHK does (parsec_io::err_chk).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "bs_mi355.h"
#include "parsec_io.hpp"

namespace {

using parsec_io::OptionData;

int fail(const char* what, int rc) {
    fprintf(stderr, "ERROR: %s: %s (%s)\n", what, bs_status_string(rc), bs_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    printf("PARSEC Benchmark Suite\n");
    fflush(nullptr);
    int passes = 100;
    bool check = false;
    std::vector<const char*> pos;
    for (int i = 1; i < argc; ++i) {
        if (strcmp(argv[i], "--check") == 0) check = true;
        else if (strcmp(argv[i], "--passes") == 0 && i + 1 < argc) passes = atoi(argv[++i]);
        else pos.push_back(argv[i]);
    }
    if (pos.size() != 3 || passes < 1) {
        printf("Usage:\n\t%s <nthreads> <inputFile> <outputFile> [--check] [--passes N]\n",
               argv[0]);
        return 1;
    }
    int nThreads = atoi(pos[0]);
    const char* inputFile = pos[1];
    const char* outputFile = pos[2];

    // the text formats are read and written on the host's threads (the
    // conversions are the reference's; see parsec_io.hpp)
    const int io_threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<OptionData> data;
    std::string io_err;
    if (!parsec_io::read_options(inputFile, &data, io_threads, &io_err)) {
        printf("ERROR: %s\n", io_err.c_str());
        return 1;
    }
    const int numOptions = (int)data.size();
    if (nThreads > numOptions) {
        printf("WARNING: Not enough work, reducing number of threads to match number of options.\n");
        nThreads = numOptions;
    }
    printf("Num of Options: %d\n", numOptions);
    printf("Num of Runs: %d\n", passes);
    printf("Size of data: %zu\n", (size_t)numOptions * (sizeof(OptionData) + sizeof(int)));

    std::vector<float> prices((size_t)numOptions);
    const size_t n = (size_t)numOptions;
    if (n > 0) {
        int rc;
        bs_stream_t st = nullptr;
        void *aos = nullptr, *soa[5] = {}, *type = nullptr, *out = nullptr;
        if ((rc = bs_set_device(0)) || (rc = bs_stream_create(&st))) return fail("device", rc);
        if ((rc = bs_malloc(&aos, n * sizeof(OptionData)))) return fail("alloc", rc);
        for (auto& p : soa)
            if ((rc = bs_malloc(&p, n * 4))) return fail("alloc", rc);
        if ((rc = bs_malloc(&type, n)) || (rc = bs_malloc(&out, n * 4))) return fail("alloc", rc);
        if ((rc = bs_memcpy_h2d(aos, data.data(), n * sizeof(OptionData), st)))
            return fail("upload", rc);
        float* f[5];
        for (int j = 0; j < 5; ++j) f[j] = static_cast<float*>(soa[j]);
        if ((rc = bs_repack_aos_f32(aos, (int64_t)n, f[0], f[1], f[2], f[3], f[4],
                                    static_cast<uint8_t*>(type), st)))
            return fail("repack", rc);
        if ((rc = bs_stream_sync(st))) return fail("repack", rc);
        // ROI: the passes only
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = bs_price_passes_f32(f[0], f[1], f[2], f[3], f[4], type, 1,
                                      static_cast<float*>(out), (int64_t)n, passes, st)))
            return fail("price", rc);
        if ((rc = bs_stream_sync(st))) return fail("price", rc);
        const double roi = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        printf("roi.time|%.6f\n", roi);
        if ((rc = bs_memcpy_d2h(prices.data(), out, n * 4, st)) || (rc = bs_stream_sync(st)))
            return fail("download", rc);
        bs_free(aos);
        for (void* p : soa) bs_free(p);
        bs_free(type);
        bs_free(out);
        bs_stream_destroy(st);
    }

    if (!parsec_io::write_prices(outputFile, prices.data(), numOptions, io_threads, &io_err)) {
        printf("ERROR: %s\n", io_err.c_str());
        return 1;
    }
    if (check) {
        const long long numError = parsec_io::err_chk(data.data(), prices.data(), numOptions,
                                                      passes, stdout);
        printf("Num Errors: %lld\n", numError);
    }
    return 0;
}
