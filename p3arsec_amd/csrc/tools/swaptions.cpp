// swaptions -- the PARSEC `swaptions` command line on the GPU.
//
//   swaptions -ns <swaptions> -sm <trials> [-nt <threads>] [-sd <seed>]
//
// The reference's driver (parsec-ff/pkgs/apps/swaptions/src/HJM_Securities.cpp:
// 153-378): the same options and checks, the same portfolio (seed 1979 by
// default; horizons and strikes drawn from its generator, the three-factor
// volatility table, the yield curve 0.10 + 0.005 j), and the same output
// lines, "Swaption i: [SwaptionPrice: %.10lf StdError: %.10lf]" on stderr.
// The pricing (its ROI, HJM_Securities.cpp:301-345) is one call of
// bs_swaptions_host_f64: every path of every swaption on the GPU; -nt is
// checked as the reference checks it and otherwise unused.  The ROI's wall
// time is printed as "roi.time|<seconds>".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "bs_mi355.h"

namespace {

constexpr int kSteps = 11;      // iN
constexpr int kFactors = 3;     // iFactors
constexpr int kBlockSize = 16;  // BLOCK_SIZE
constexpr int kMaxThreads = 1024;

// HJM_Securities.cpp:231-264
constexpr double kFactorVols[kFactors][kSteps - 1] = {
    {.01, .01, .01, .01, .01, .01, .01, .01, .01, .01},
    {.009048, .008187, .007408, .006703, .006065, .005488, .004966, .004493, .004066, .003679},
    {.001000, .000750, .000500, .000250, .000000, -.000250, -.000500, -.000750, -.001000,
     -.001250}};

// The driver's generator on counter *s (advanced by one): Park-Miller's
// minimal standard generator of 1513517 c mod (2^31 - 1), scaled.  For a
// negative counter both remainders are <= 0; the sign correction leaves the
// same residue in [0, 2^31 - 1) as the reference's Schrage steps.
double ran_unif(long* s) {
    const long long m = 2147483647LL;
    long long x = *s;
    *s = (long)(x + 1);
    x = (x * 1513517LL) % m;
    x = (x * 16807LL) % m;
    if (x < 0) x += m;
    return (double)x * 4.656612875e-10;
}

void usage(const char* name) {
    fprintf(stderr, "Usage: %s OPTION [OPTIONS]...\n", name);
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "\t-ns [number of swaptions (should be > number of threads]\n");
    fprintf(stderr, "\t-sm [number of simulations]\n");
    fprintf(stderr, "\t-nt [number of threads]\n");
    fprintf(stderr, "\t-sd [random number seed]\n");
}

}  // namespace

int main(int argc, char** argv) {
    printf("PARSEC Benchmark Suite\n");
    fflush(NULL);
    int trials = 102400, nthreads = 1, nswaptions = 1;
    long seed = 1979;
    if (argc == 1) {
        usage(argv[0]);
        return 1;
    }
    for (int j = 1; j < argc; j++) {
        const bool has = j + 1 < argc;
        if (!strcmp("-sm", argv[j]) && has) trials = atoi(argv[++j]);
        else if (!strcmp("-nt", argv[j]) && has) nthreads = atoi(argv[++j]);
        else if (!strcmp("-ns", argv[j]) && has) nswaptions = atoi(argv[++j]);
        else if (!strcmp("-sd", argv[j]) && has) seed = atoi(argv[++j]);
        else {
            fprintf(stderr, "Error: Unknown option: %s\n", argv[j]);
            usage(argv[0]);
            return 1;
        }
    }
    if (nswaptions < nthreads) {
        fprintf(stderr, "Error: Fewer swaptions than threads.\n");
        usage(argv[0]);
        return 1;
    }
    printf("Number of Simulations: %d,  Number of threads: %d Number of swaptions: %d\n", trials,
           nthreads, nswaptions);
    if (nthreads < 1 || nthreads > kMaxThreads) {
        fprintf(stderr, "Number of threads must be between 1 and %d.\n", kMaxThreads);
        return 1;
    }
    const long base = (long)(2147483647L * ran_unif(&seed));
    std::vector<double> yield(kSteps), factors(kFactors * (kSteps - 1));
    yield[0] = .1;
    for (int j = 1; j < kSteps; ++j) yield[j] = yield[j - 1] + .005;
    for (int k = 0; k < kFactors; ++k)
        for (int j = 0; j < kSteps - 1; ++j) factors[k * (kSteps - 1) + j] = kFactorVols[k][j];
    std::vector<bs_swaption_t> sw((size_t)nswaptions);
    for (auto& s : sw) {
        s.years = 5.0 + ((int)(60 * ran_unif(&seed))) * 0.25;
        s.strike = 0.1 + ((int)(49 * ran_unif(&seed))) * 0.1;
        s.compounding = 0;
        s.maturity = 1.0;
        s.tenor = 2.0;
        s.payment_interval = 1.0;
        s.n_steps = kSteps;
        s.n_factors = kFactors;
        s.yield = yield.data();
        s.factors = factors.data();
    }
    std::vector<double> price((size_t)nswaptions), err((size_t)nswaptions);
    const auto t0 = std::chrono::steady_clock::now();
    const int st = bs_swaptions_host_f64(sw.data(), nswaptions, trials, base, kBlockSize,
                                         price.data(), err.data(), nullptr);
    const double roi = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st != BS_OK) {
        fprintf(stderr, "bs_swaptions_host_f64: %s: %s\n", bs_status_string(st), bs_last_error());
        return 1;
    }
    printf("roi.time|%.6f\n", roi);
    for (int i = 0; i < nswaptions; i++)
        fprintf(stderr, "Swaption %d: [SwaptionPrice: %.10lf StdError: %.10lf] \n", i, price[i],
                err[i]);
    return 0;
}
