// Calls bs_compat::HJM_Swaption_Blocking the way the reference's workers call
// HJM_Swaption_Blocking (HJM_Securities.cpp:122-131): a parm-like record with
// a row-pointer factor matrix.  Usage: swaption_compat_demo <years> <strike>
// <seed> <trials>; prints "price stderr" with 17 significant digits.
#include <stdio.h>
#include <stdlib.h>

#include "bs_compat.hpp"

int main(int argc, char** argv) {
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
