// Calls bs_compat::HJM_Swaption_Blocking the way the reference's workers call
// HJM_Swaption_Blocking (HJM_Securities.cpp:122-131): a parm-like record with
// a row-pointer factor matrix.  Usage: swaption_compat_demo <years> <strike>
// <seed> <trials>; prints "price stderr" with 17 significant digits.
// swaption_compat_demo map <n> <base seed> <trials>: the driver's portfolio
// loop through bs_compat::price_swaptions over a record with the reference's
// field names (seed 1979 draws, as HJM_Securities.cpp:198,275-296); prints
// "i price stderr" per swaption.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bs_compat.hpp"

// The reference's parm fields (HJM_Securities.cpp:275-296), in an order of
// our own: price_swaptions reads them by name.
struct parm {
    int Id;
    double dSimSwaptionMeanPrice, dSimSwaptionStdError;
    double dStrike, dCompounding, dMaturity, dTenor, dPaymentInterval;
    int iN;
    double dYears;
    int iFactors;
    double* pdYield;
    double** ppdFactors;
};

static double ran_unif(long* s) {
    const long long m = 2147483647LL;
    long long x = *s;
    *s = (long)(x + 1);
    x = (x * 1513517LL) % m;
    x = (x * 16807LL) % m;
    if (x < 0) x += m;
    return (double)x * 4.656612875e-10;
}

static int map_mode(int n, long trials) {
    const int iN = 11, iFactors = 3;
    static double vols[3][10] = {
        {.01, .01, .01, .01, .01, .01, .01, .01, .01, .01},
        {.009048, .008187, .007408, .006703, .006065, .005488, .004966, .004493, .004066, .003679},
        {.001, .00075, .0005, .00025, 0, -.00025, -.0005, -.00075, -.001, -.00125}};
    static double* rows[3] = {vols[0], vols[1], vols[2]};
    static double yield[11];
    yield[0] = .1;
    for (int j = 1; j < iN; ++j) yield[j] = yield[j - 1] + .005;
    long seed = 1979;
    const long base = (long)(2147483647L * ran_unif(&seed));
    parm* sw = new parm[n];
    for (int i = 0; i < n; ++i) {
        sw[i] = parm{};
        sw[i].Id = i;
        sw[i].iN = iN;
        sw[i].iFactors = iFactors;
        sw[i].dYears = 5.0 + ((int)(60 * ran_unif(&seed))) * 0.25;
        sw[i].dStrike = 0.1 + ((int)(49 * ran_unif(&seed))) * 0.1;
        sw[i].dMaturity = 1.0;
        sw[i].dTenor = 2.0;
        sw[i].dPaymentInterval = 1.0;
        sw[i].pdYield = yield;
        sw[i].ppdFactors = rows;
    }
    try {
        bs_compat::price_swaptions(sw, n, base, trials, 16);
    } catch (const bs_compat::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        delete[] sw;
        return 1;
    }
    for (int i = 0; i < n; ++i)
        printf("%d %.17g %.17g\n", i, sw[i].dSimSwaptionMeanPrice, sw[i].dSimSwaptionStdError);
    delete[] sw;
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && !strcmp(argv[1], "map")) return map_mode(atoi(argv[2]), atol(argv[3]));
    if (argc != 5) return 2;
    const int iN = 11, iFactors = 3;
    double vols[3][10] = {
        {.01, .01, .01, .01, .01, .01, .01, .01, .01, .01},
        {.009048, .008187, .007408, .006703, .006065, .005488, .004966, .004493, .004066, .003679},
        {.001, .00075, .0005, .00025, 0, -.00025, -.0005, -.00075, -.001, -.00125}};
    double* rows[3] = {vols[0], vols[1], vols[2]};
    double yield[11];
    yield[0] = .1;
    for (int j = 1; j < iN; ++j) yield[j] = yield[j - 1] + .005;
    double out[2];
    try {
        const int ok = bs_compat::HJM_Swaption_Blocking(out, atof(argv[2]), 0, 1.0, 2.0, 1.0, iN,
                                                        iFactors, atof(argv[1]), yield, rows,
                                                        atol(argv[3]), atol(argv[4]), 16, 0);
        printf("%d %.17g %.17g\n", ok, out[0], out[1]);
    } catch (const bs_compat::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
