/*
 * swaptions_oracle.c -- CPU oracle for the swaptions (HJM Monte-Carlo) map.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The product library has no CPU pricing path.
 *
 * What the reference tree holds (parsec-ff/pkgs/apps/swaptions/src):
 *   HJM_Swaption_Blocking.cpp:20-222  the per-swaption Monte-Carlo loop
 *   HJM_Securities.cpp:153-296        the driver: seeds and parameters
 *   HJM_Securities.cpp:311-323        the FastFlow ParallelFor over swaptions
 * What it does not hold: the routines that loop calls, which come from the
 * PARSEC 3.0 tarball (pkgs/apps/swaptions/src: RanUnif.cpp, CumNormalInv.cpp,
 * HJM.cpp, HJM_SimPath_Forward_Blocking.cpp, MaxFunction.cpp, HJM_type.h).
 * They are restated here from the published algorithms they implement:
 *   swo_ranunif          Park & Miller's minimal standard generator
 *                        (CACM 31(10), 1988) with Schrage's factorisation,
 *                        applied to a hashed counter: u(c) =
 *                        MINSTD(1513517 c mod (2^31-1)) * 4.656612875e-10,
 *                        the seed advancing by one per call;
 *   swo_cum_normal_inv   Moro's inverse normal ("The Full Monte", Risk 1995;
 *                        Glasserman, Monte Carlo Methods in Financial
 *                        Engineering, fig. 2.13);
 *   swo_yield_to_forward f_0 = y_0, f_i = (i+1) y_i - i y_{i-1};
 *   swo_drifts           the discrete arbitrage-free HJM drift:
 *                        sum_{l<=j} mu_k[l] = dt/2 (sum_{l<=j} sigma_k[l])^2;
 *   swo_simpath_forward_blocking, swo_discount_factors_blocking
 *                        the Musiela-shifted forward recursion
 *                        f(j,l) = f(j-1,l+1) + mu[l] dt + sqrt(dt) sum_k
 *                        sigma_k[l] Z_k(j) and the running products of
 *                        exp(-r dt), over BLOCKSIZE paths at a time.
 * Those restatements are not pinned against PARSEC's own sources (absent);
 * the tests pin them against what the mathematics fixes instead.  Parity
 * unpinned: the generator's counter hash (1513517), its scale
 * (4.656612875e-10) and the draw order (path, time step, factor), which no
 * check can tie to PARSEC's sources or outputs here.  Pinned: Moro
 * against the exact inverse normal, Schrage's step against direct modular
 * arithmetic and Park-Miller's published check value, the forward curve and
 * drift identities, and deep in-the-money prices against their closed form
 * (the discounted bond prices are martingales under the drift above).
 *
 * Contents
 *   swo_swaption_blocking  HJM_Swaption_Blocking, FTYPE = double, blocked as
 *                          the reference blocks it, every time step simulated
 *                          and the sums accumulated path by path in order.
 *   swo_swaption_chain     the same arithmetic reorganised per forward-rate
 *                          chain, simulating only the steps the payoff reads
 *                          (the form the GPU kernel uses); bit-identical to
 *                          swo_swaption_blocking, which a test checks.
 *   swo_chain_payoffs      the per-path payoffs of the chain form from
 *                          caller-supplied normal draws (tests of the model
 *                          apart from the generator).
 *   swo_portfolio          the driver's parameter generation.
 *   swo_price_batch        the FastFlow ParallelFor over swaptions: nthreads
 *                          contiguous static blocks, pinned threads.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SWO_M31 2147483647LL

/* RanUnif: one draw from counter *s, which then advances by one. */
double swo_ranunif(int64_t *s)
{
    int64_t x = *s;
    *s = x + 1;
    x = (x * 1513517LL) % SWO_M31;          /* hash of the counter */
    const int64_t q = x / 127773LL;          /* Schrage: 16807 x mod M */
    x = 16807LL * (x - q * 127773LL) - q * 2836LL;
    if (x < 0) x += SWO_M31;
    return (double)x * 4.656612875e-10;
}

/* The same draw by direct 64-bit modular arithmetic (the check of Schrage's
   factorisation above). */
double swo_ranunif_direct(int64_t c)
{
    int64_t x = (c * 1513517LL) % SWO_M31;
    x = (x * 16807LL) % SWO_M31;
    return (double)x * 4.656612875e-10;
}

/* Park-Miller minimal standard iteration x <- 16807 x mod (2^31-1) by
   Schrage's factorisation, `steps` times from x (the published check:
   10000 steps from 1 give 1043618065). */
int64_t swo_minstd_iterate(int64_t x, int64_t steps)
{
    for (int64_t i = 0; i < steps; i++) {
        const int64_t q = x / 127773LL;
        x = 16807LL * (x - q * 127773LL) - q * 2836LL;
        if (x < 0) x += SWO_M31;
    }
    return x;
}

/* Moro's inverse of the standard normal distribution function. */
static const double kMoroA[4] = {2.50662823884, -18.61500062529, 41.39119773534,
                                 -25.44106049637};
static const double kMoroB[4] = {-8.47351093090, 23.08336743743, -21.06224101826,
                                 3.13082909833};
static const double kMoroC[9] = {0.3374754822726147, 0.9761690190917186,
                                 0.1607979714918209, 0.0276438810333863,
                                 0.0038405729373609, 0.0003951896511919,
                                 0.0000321767881768, 0.0000002888167364,
                                 0.0000003960315187};

double swo_cum_normal_inv(double u)
{
    const double x = u - 0.5;
    if (fabs(x) < 0.42) {                    /* central region: rational */
        const double r = x * x;
        const double num = x * (((kMoroA[3] * r + kMoroA[2]) * r + kMoroA[1]) * r + kMoroA[0]);
        const double den = (((kMoroB[3] * r + kMoroB[2]) * r + kMoroB[1]) * r + kMoroB[0]) * r + 1.0;
        return num / den;
    }
    double t = x > 0.0 ? 1.0 - u : u;        /* tails: Chebyshev in log(-log) */
    t = log(-log(t));
    t = kMoroC[0] + t * (kMoroC[1] + t * (kMoroC[2] + t * (kMoroC[3] + t * (kMoroC[4] +
        t * (kMoroC[5] + t * (kMoroC[6] + t * (kMoroC[7] + t * kMoroC[8])))))));
    return x < 0.0 ? -t : t;
}

void swo_yield_to_forward(double *fwd, int n, const double *yield)
{
    fwd[0] = yield[0];
    for (int i = 1; i < n; i++) fwd[i] = (i + 1) * yield[i] - i * yield[i - 1];
}

/* drifts: F rows of n-1; factors: F rows of n-1 (row-major). */
void swo_drifts(double *total, double *drifts, int n, int F, double years,
                const double *factors)
{
    const double dt = (double)(years / n);
    const int w = n - 1;
    for (int k = 0; k < F; k++)
        drifts[k * w] = 0.5 * dt * factors[k * w] * factors[k * w];
    for (int k = 0; k < F; k++)
        for (int j = 1; j < w; j++) {
            double d = 0;
            for (int l = 0; l < j; l++) d -= drifts[k * w + l];
            double vol = 0;
            for (int l = 0; l <= j; l++) vol += factors[k * w + l];
            drifts[k * w + j] = d + 0.5 * dt * vol * vol;
        }
    for (int l = 0; l < w; l++) {
        double t = 0;
        for (int k = 0; k < F; k++) t += drifts[k * w + l];
        total[l] = t;
    }
}

/* One block of B paths: path[j][B*l + b] = f(j, l) of path b (n x n*B). */
static void swo_simpath_forward_blocking(double *path, int n, int F, double years,
                                         const double *fwd, const double *total,
                                         const double *factors, int64_t *seed, int B,
                                         double *z /* F x n*B */)
{
    const double dt = (double)(years / n);
    const double sdt = sqrt(dt);
    const int w = n - 1, nb = n * B;
    for (int b = 0; b < B; b++)
        for (int l = 0; l < n; l++) {
            path[B * l + b] = fwd[l];
            for (int j = 1; j < n; j++) path[(size_t)j * nb + B * l + b] = 0;
        }
    /* draws in the generator's order: path, then time step, then factor */
    for (int b = 0; b < B; b++)
        for (int j = 1; j < n; j++)
            for (int k = 0; k < F; k++) z[(size_t)k * nb + B * j + b] = swo_ranunif(seed);
    for (int k = 0; k < F; k++)
        for (int b = 0; b < B; b++)
            for (int j = 1; j < n; j++)
                z[(size_t)k * nb + B * j + b] = swo_cum_normal_inv(z[(size_t)k * nb + B * j + b]);
    for (int b = 0; b < B; b++)
        for (int j = 1; j < n; j++)
            for (int l = 0; l <= n - (j + 1); l++) {
                double shock = 0;
                for (int k = 0; k < F; k++)
                    shock += factors[k * w + l] * z[(size_t)k * nb + B * j + b];
                path[(size_t)j * nb + B * l + b] =
                    path[(size_t)(j - 1) * nb + B * (l + 1) + b] + total[l] * dt + sdt * shock;
            }
}

/* df[i*B + b] = prod_{j<i} exp(-rate[j*B + b] dt), i < n. */
static void swo_discount_factors_blocking(double *df, int n, double years,
                                          const double *rate, int B, double *ex)
{
    const double dt = (double)(years / n);
    const int m = (n - 1) * B;
    for (int j = 0; j < m; j++) ex[j] = -rate[j] * dt;
    for (int j = 0; j < m; j++) ex[j] = exp(ex[j]);
    for (int i = 0; i < n * B; i++) df[i] = 1.0;
    for (int i = 1; i < n; i++)
        for (int b = 0; b < B; b++)
            for (int j = 0; j < i; j++) df[i * B + b] *= ex[j * B + b];
}

/* Derived sizes of one swaption (HJM_Swaption_Blocking.cpp:48-127), or -1
   when the reference would index out of its arrays or loop forever. */
typedef struct {
    double dt, strike_cont, swap_years;
    int freq, swap_len, start, swap_points;
} swo_shape;

static int swo_make_shape(swo_shape *s, double strike, double compounding, double maturity,
                          double tenor, double interval, int n, double years)
{
    if (n < 2) return -1;
    s->dt = (double)(years / n);
    s->freq = (int)(interval / s->dt + 0.5);
    s->strike_cont = compounding == 0 ? strike : (1 / compounding) * log(1 + strike * compounding);
    s->swap_len = (int)(n - maturity / s->dt + 0.5);
    s->start = (int)(maturity / s->dt + 0.5);
    s->swap_points = (int)(tenor / s->dt + 0.5);
    s->swap_years = (double)(s->swap_len * s->dt);
    if (s->freq < 1 || s->swap_len < 1 || s->start < 0 || s->start >= n ||
        s->swap_points > s->swap_len - 1 || s->swap_points < 0)
        return -1;
    return 0;
}

static void swo_swap_payoffs(double *pay, const swo_shape *s, double interval)
{
    for (int i = 0; i < s->swap_len; i++) pay[i] = 0.0;
    for (int i = s->freq; i <= s->swap_points; i += s->freq)
        pay[i] = i != s->swap_points ? exp(s->strike_cont * interval) - 1
                                     : exp(s->strike_cont * interval);
}

static void swo_finish(double *out, double sum, double sumsq, int64_t trials)
{
    out[0] = sum / trials;
    out[1] = sqrt((sumsq - sum * sum / trials) / (trials - 1.0)) / sqrt((double)trials);
}

/* HJM_Swaption_Blocking.cpp:20-222.  factors: F rows of n-1.  out[0] = mean
   price, out[1] = standard error.  Returns 1, or 0 for unusable sizes. */
int swo_swaption_blocking(double *out, double strike, double compounding, double maturity,
                          double tenor, double interval, int n, int F, double years,
                          const double *yield, const double *factors, int64_t seed,
                          int64_t trials, int B)
{
    swo_shape s;
    if (B < 1 || trials < 1 || F < 1 || swo_make_shape(&s, strike, compounding, maturity, tenor,
                                                       interval, n, years) != 0)
        return 0;
    const int w = n - 1, L = s.swap_len;
    double *path = malloc(sizeof(double) * (size_t)n * n * B);
    double *z = malloc(sizeof(double) * (size_t)F * n * B);
    double *fwd = malloc(sizeof(double) * n);
    double *drifts = malloc(sizeof(double) * (size_t)F * w);
    double *total = malloc(sizeof(double) * w);
    double *pdf = malloc(sizeof(double) * (size_t)n * B);
    double *prate = malloc(sizeof(double) * (size_t)n * B);
    double *srate = malloc(sizeof(double) * (size_t)L * B);
    double *sdf = malloc(sizeof(double) * (size_t)L * B);
    double *ex = malloc(sizeof(double) * (size_t)n * B);
    double *pay = malloc(sizeof(double) * L);
    int ok = path && z && fwd && drifts && total && pdf && prate && srate && sdf && ex && pay;
    if (ok) {
        swo_swap_payoffs(pay, &s, interval);
        swo_yield_to_forward(fwd, n, yield);
        swo_drifts(total, drifts, n, F, years, factors);
        double sum = 0.0, sumsq = 0.0;
        const size_t nb = (size_t)n * B;
        for (int64_t p = 0; p < trials; p += B) {
            swo_simpath_forward_blocking(path, n, F, years, fwd, total, factors, &seed, B, z);
            for (int i = 0; i < n; i++)
                for (int b = 0; b < B; b++) prate[B * i + b] = path[(size_t)i * nb + b];
            swo_discount_factors_blocking(pdf, n, years, prate, B, ex);
            for (int i = 0; i < L; i++)
                for (int b = 0; b < B; b++)
                    srate[i * B + b] = path[(size_t)s.start * nb + (size_t)i * B + b];
            swo_discount_factors_blocking(sdf, L, s.swap_years, srate, B, ex);
            for (int b = 0; b < B; b++) {
                double leg = 0.0;
                for (int i = 0; i < L; i++) leg += pay[i] * sdf[i * B + b];
                const double payoff = leg - 1.0 > 0 ? leg - 1.0 : 0;
                const double disc = payoff * pdf[s.start * B + b];
                sum += disc;
                sumsq += disc * disc;
            }
        }
        swo_finish(out, sum, sumsq, trials);
    }
    free(path); free(z); free(fwd); free(drifts); free(total); free(pdf);
    free(prate); free(srate); free(sdf); free(ex); free(pay);
    return ok ? 1 : 0;
}

/* The same arithmetic, path by path, one forward-rate chain at a time: chain
   d starts at f(0, d) and takes the steps s = 1..min(d, start) in order,
   exactly the additions the blocked recursion makes to reach f(s, d - s).
   Steps after the swaption's start and the exponentials of short rates past
   it never reach the payoff and are skipped. */
typedef struct {
    swo_shape s;
    int n, F;
    double sdt, dts;
    const double *factors;
    double *fwd, *drifts, *total, *mudt, *pay;
} swo_setup;

static void swo_setup_free(swo_setup *u)
{
    free(u->fwd); free(u->drifts); free(u->total); free(u->mudt); free(u->pay);
}

static int swo_setup_make(swo_setup *u, double strike, double compounding, double maturity,
                          double tenor, double interval, int n, int F, double years,
                          const double *yield, const double *factors)
{
    memset(u, 0, sizeof(*u));
    if (F < 1 || swo_make_shape(&u->s, strike, compounding, maturity, tenor, interval, n,
                                years) != 0)
        return 0;
    const int w = n - 1, L = u->s.swap_len;
    u->n = n;
    u->F = F;
    u->factors = factors;
    u->fwd = malloc(sizeof(double) * n);
    u->drifts = malloc(sizeof(double) * (size_t)F * w);
    u->total = malloc(sizeof(double) * w);
    u->mudt = malloc(sizeof(double) * w);
    u->pay = malloc(sizeof(double) * L);
    if (!u->fwd || !u->drifts || !u->total || !u->mudt || !u->pay) {
        swo_setup_free(u);
        return 0;
    }
    swo_swap_payoffs(u->pay, &u->s, interval);
    swo_yield_to_forward(u->fwd, n, yield);
    swo_drifts(u->total, u->drifts, n, F, years, factors);
    u->sdt = sqrt(u->s.dt);
    u->dts = (double)(u->s.swap_years / L);
    for (int l = 0; l < w; l++) u->mudt[l] = u->total[l] * u->s.dt;
    return 1;
}

/* Discounted payoff of one path; z[(s-1) F + k] = Z_k(s), s = 1..start. */
static double swo_chain_payoff(const swo_setup *u, const double *z)
{
    const int n = u->n, F = u->F, w = n - 1, m = u->s.start, L = u->s.swap_len;
    double df = 1.0, sdf = 1.0, leg = 0.0;
    for (int d = 0; d < n && d < m + L; d++) {
        double x = u->fwd[d];
        const int steps = d < m ? d : m;
        for (int st = 1; st <= steps; st++) {
            const int l = d - st;
            double shock = 0;
            for (int k = 0; k < F; k++) shock += u->factors[k * w + l] * z[(st - 1) * F + k];
            x = x + u->mudt[l] + u->sdt * shock;
        }
        if (d < m) {
            df *= exp(-x * u->s.dt);
        } else {
            const int l = d - m;
            leg += u->pay[l] * sdf;
            if (l < L - 1) sdf *= exp(-x * u->dts);
        }
    }
    for (int l = n - m; l < L; l++) {       /* rates past the horizon are 0 */
        leg += u->pay[l] * sdf;
        if (l < L - 1) sdf *= exp(-0.0 * u->dts);
    }
    const double payoff = leg - 1.0 > 0 ? leg - 1.0 : 0;
    return payoff * df;
}

/* Counters: path p draws seed + p (n-1) F + (s-1) F + k. */
int swo_swaption_chain(double *out, double strike, double compounding, double maturity,
                       double tenor, double interval, int n, int F, double years,
                       const double *yield, const double *factors, int64_t seed,
                       int64_t trials, int B)
{
    swo_setup u;
    if (B < 1 || trials < 1 ||
        !swo_setup_make(&u, strike, compounding, maturity, tenor, interval, n, F, years, yield,
                        factors))
        return 0;
    const int w = n - 1, m = u.s.start;
    double *z = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * F);
    if (!z) {
        swo_setup_free(&u);
        return 0;
    }
    const int64_t paths = (trials + B - 1) / B * B;
    double sum = 0.0, sumsq = 0.0;
    for (int64_t p = 0; p < paths; p++) {
        const int64_t c0 = seed + p * (int64_t)w * F;
        for (int st = 1; st <= m; st++)
            for (int k = 0; k < F; k++) {
                int64_t c = c0 + (int64_t)(st - 1) * F + k;
                z[(st - 1) * F + k] = swo_cum_normal_inv(swo_ranunif(&c));
            }
        const double disc = swo_chain_payoff(&u, z);
        sum += disc;
        sumsq += disc * disc;
    }
    swo_finish(out, sum, sumsq, trials);
    free(z);
    swo_setup_free(&u);
    return 1;
}

/* The model with caller-supplied normal draws: payoff[p] for path p reads
   z[p * start * F ...].  Lets the tests check the HJM arithmetic with
   independent normals, apart from the generator.  Returns the start step
   (draws per path = start x F), or -1 for unusable sizes. */
int swo_chain_payoffs(double strike, double compounding, double maturity, double tenor,
                      double interval, int n, int F, double years, const double *yield,
                      const double *factors, const double *z, int64_t paths, double *payoff)
{
    swo_setup u;
    if (!swo_setup_make(&u, strike, compounding, maturity, tenor, interval, n, F, years, yield,
                        factors))
        return -1;
    const int m = u.s.start;
    if (z && payoff)
        for (int64_t p = 0; p < paths; p++)
            payoff[p] = swo_chain_payoff(&u, z + (size_t)p * m * F);
    swo_setup_free(&u);
    return m;
}

/* HJM_Securities.cpp:198,275-296: the base seed of the swaptions (one draw
   from *seed) and each swaption's horizon and strike (two draws each). */
void swo_portfolio(int n, int64_t *seed, int64_t *swaption_seed, double *years, double *strike)
{
    *swaption_seed = (int64_t)(2147483647LL * swo_ranunif(seed));
    for (int i = 0; i < n; i++) {
        years[i] = 5.0 + ((int)(60 * swo_ranunif(seed))) * 0.25;
        strike[i] = 0.1 + ((int)(49 * swo_ranunif(seed))) * 0.1;
    }
}

/* ------------------------------------------------------------------------ */
/* FastFlow ParallelFor over swaptions (HJM_Securities.cpp:311-323): static
   contiguous blocks, swaption i priced with seed + i. */
typedef struct {
    int begin, end, cpu;
    const double *strike, *compounding, *maturity, *tenor, *interval, *years;
    int n, F;
    const double *yield, *factors;
    int64_t seed, trials;
    int B, chain;
    double *price, *stderr_;
    int ok;
} swo_block;

static void *swo_worker(void *arg)
{
    swo_block *b = (swo_block *)arg;
    if (b->cpu >= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(b->cpu, &set);
        pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
    }
    b->ok = 1;
    for (int i = b->begin; i < b->end; i++) {
        double o[2];
        int (*price_one)(double *, double, double, double, double, double, int, int, double,
                         const double *, const double *, int64_t, int64_t, int) =
            b->chain ? swo_swaption_chain : swo_swaption_blocking;
        if (!price_one(o, b->strike[i], b->compounding[i], b->maturity[i], b->tenor[i],
                       b->interval[i], b->n, b->F, b->years[i], b->yield + (size_t)i * b->n,
                       b->factors + (size_t)i * b->F * (b->n - 1), b->seed + i, b->trials,
                       b->B))
            b->ok = 0;
        b->price[i] = o[0];
        b->stderr_[i] = o[1];
    }
    return NULL;
}

/* Prices ns swaptions (yield: ns rows of n; factors: ns blocks of F x (n-1))
   on nthreads threads (<= 0: every CPU this process may use), each by
   swo_swaption_blocking (chain = 0: every time step, as the reference) or by
   swo_swaption_chain (chain = 1: only the steps the payoff reads, the same
   bits).  Returns the wall time in seconds, or a negative value on error. */
double swo_price_batch(int ns, const double *strike, const double *compounding,
                       const double *maturity, const double *tenor, const double *interval,
                       const double *years, int n, int F, const double *yield,
                       const double *factors, int64_t seed, int64_t trials, int B, int chain,
                       double *price, double *stderr_, int nthreads, int *used_threads)
{
    cpu_set_t allowed;
    int cpus[1024], ncpu = 0;
    CPU_ZERO(&allowed);
    if (sched_getaffinity(0, sizeof(allowed), &allowed) == 0)
        for (int c = 0; c < CPU_SETSIZE && ncpu < 1024; c++)
            if (CPU_ISSET(c, &allowed)) cpus[ncpu++] = c;
    if (ncpu == 0) { cpus[0] = -1; ncpu = 1; }
    if (nthreads <= 0) nthreads = ncpu;
    if (nthreads > ns) nthreads = ns;
    if (nthreads < 1) nthreads = 1;
    if (used_threads) *used_threads = nthreads;
    swo_block *blk = calloc((size_t)nthreads, sizeof(swo_block));
    pthread_t *tid = calloc((size_t)nthreads, sizeof(pthread_t));
    if (!blk || !tid) { free(blk); free(tid); return -1.0; }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const int chunk = ns / nthreads, extra = ns % nthreads;
    int start = 0;
    for (int i = 0; i < nthreads; i++) {
        const int len = chunk + (i < extra ? 1 : 0);
        blk[i] = (swo_block){start, start + len, cpus[i % ncpu], strike, compounding, maturity,
                             tenor, interval, years, n, F, yield, factors, seed, trials, B,
                             chain, price, stderr_, 0};
        start += len;
        pthread_create(&tid[i], NULL, swo_worker, &blk[i]);
    }
    int ok = 1;
    for (int i = 0; i < nthreads; i++) {
        pthread_join(tid[i], NULL);
        ok = ok && blk[i].ok;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(blk);
    free(tid);
    if (!ok) return -1.0;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
