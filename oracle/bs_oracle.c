/*
 * bs_oracle.c -- CPU oracle for the Black-Scholes map.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The product library (libbs_mi355.so) never links it
 * and has no CPU pricing path.
 *
 * Contents
 *   bso_cndf_f32 / bso_price_f32
 *       Restatement, in fptype=float, of the reference's CNDF
 *       (parsec-ff/pkgs/apps/blackscholes/src/blackscholes.c:126-184) and
 *       BlkSchlsEqEuroNoDiv (blackscholes.c:190-258).  The reference mixes
 *       float variables with double literals; every step below keeps the same
 *       promotion (float op double -> double, stored back to float), the same
 *       evaluation order and the float libm overloads expf/logf/sqrtf that a
 *       C++ build of blackscholes.c binds to.  Built with the reference's
 *       flags (no fast-math, no FMA contraction) this is bit-exact with the
 *       compiled reference; the committed golden vectors pin that.
 *   bso_cndf_f64 / bso_price_f64
 *       The same algorithm with fptype=double (every variable double,
 *       exp/log/sqrt in double).  The reference only builds this by editing
 *       blackscholes.c:85; the restatement is pinned by the FP64 closed form
 *       (|A&S - Phi| <= 7.5e-8), put-call parity and the Hull known answers.
 *   bso_price_exact / bso_greeks_exact
 *       FP64 closed form (Hull, "Options, Futures and Other Derivatives",
 *       ch. on Black-Scholes-Merton) with Phi(x) = erfc(-x/sqrt2)/2.
 *   bso_greeks_fd
 *       Central finite differences, in FP64, of bso_price_f64 (the A&S-based
 *       price the adjoint kernel differentiates), relative bump h.
 *   bso_batch_f32
 *       The FastFlow ParallelFor driver semantics (blackscholes.c:315-348,
 *       parallel_for_internals.hpp:494-518): `passes` passes over the book,
 *       nthreads contiguous static blocks (the first N % nthreads blocks one
 *       option longer), a barrier after every pass, one pinned thread per
 *       core.  Used as the CPU baseline and to price large books for tests.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* Abramowitz & Stegun 26.2.17 constants (Hull's CNDF), as double literals.  */
/* The reference spells them at blackscholes.c:124,156,164-172.             */
#define AS_INV_SQRT_2PI 0.39894228040143270286
#define AS_P 0.2316419
#define AS_B1 0.319381530
#define AS_B2 (-0.356563782)
#define AS_B3 1.781477937
#define AS_B4 (-1.821255978)
#define AS_B5 1.330274429

/* blackscholes.c:126-184, fptype = float.                                   */
float bso_cndf_f32(float x)
{
    int negative = 0;
    if (x < 0.0) {           /* blackscholes.c:143-147 */
        x = -x;
        negative = 1;
    }
    float ax = x;
    /* n(x): expf of a float argument built as (-0.5f * x) * x (:152),
       then a double multiply stored to float (:153-154). */
    float pdf = expf(-0.5f * x * x);
    pdf = pdf * AS_INV_SQRT_2PI;
    /* k = 1 / (1 + p x), every step promoted to double and stored to float
       (:156-158). */
    float k = AS_P * ax;
    k = 1.0 + k;
    k = 1.0 / k;
    float k2 = k * k;         /* powers in float (:159-162) */
    float k3 = k2 * k;
    float k4 = k3 * k;
    float k5 = k4 * k;
    /* terms: float * double -> float; summed in float in the order
       ((b2 k^2 + b3 k^3) + b4 k^4) + b5 k^5, then + b1 k (:164-173). */
    float t1 = k * AS_B1;
    float acc = k2 * AS_B2;
    float t = k3 * AS_B3;
    acc = acc + t;
    t = k4 * AS_B4;
    acc = acc + t;
    t = k5 * AS_B5;
    acc = acc + t;
    t1 = acc + t1;
    float tail = t1 * pdf;    /* float product (:174) */
    float y = 1.0 - tail;     /* double subtract, stored to float (:175) */
    if (negative)
        y = 1.0 - y;          /* (:179-181) */
    return y;
}

/* blackscholes.c:190-258, fptype = float; timet is unused there too.       */
float bso_price_f32(float s, float k, float r, float v, float t, int otype)
{
    float sqrt_t = sqrtf(t);                 /* :224 */
    float log_sk = logf(s / k);              /* :226, float divide first */
    float half_var = v * v;                  /* :231 */
    half_var = half_var * 0.5;               /* :232, exact in any width */
    float d1 = r + half_var;                 /* :234-236 */
    d1 = d1 * t;
    d1 = d1 + log_sk;
    float den = v * sqrt_t;                  /* :238 */
    d1 = d1 / den;                           /* :239 */
    float d2 = d1 - den;                     /* :240 */
    float n1 = bso_cndf_f32(d1);             /* :245-246 */
    float n2 = bso_cndf_f32(d2);
    float fv = k * expf(-(r) * (t));         /* :248 */
    float price;
    if (otype == 0) {
        price = (s * n1) - (fv * n2);        /* :250 */
    } else {
        float m1 = (1.0 - n1);               /* :252-254 */
        float m2 = (1.0 - n2);
        price = (fv * m2) - (s * m1);
    }
    return price;
}

/* fptype = double: the same statements with every variable double.         */
double bso_cndf_f64(double x)
{
    int negative = 0;
    if (x < 0.0) {
        x = -x;
        negative = 1;
    }
    double ax = x;
    double pdf = exp(-0.5f * x * x);
    pdf = pdf * AS_INV_SQRT_2PI;
    double k = AS_P * ax;
    k = 1.0 + k;
    k = 1.0 / k;
    double k2 = k * k;
    double k3 = k2 * k;
    double k4 = k3 * k;
    double k5 = k4 * k;
    double t1 = k * AS_B1;
    double acc = k2 * AS_B2;
    double t = k3 * AS_B3;
    acc = acc + t;
    t = k4 * AS_B4;
    acc = acc + t;
    t = k5 * AS_B5;
    acc = acc + t;
    t1 = acc + t1;
    double tail = t1 * pdf;
    double y = 1.0 - tail;
    if (negative)
        y = 1.0 - y;
    return y;
}

double bso_price_f64(double s, double k, double r, double v, double t, int otype)
{
    double sqrt_t = sqrt(t);
    double log_sk = log(s / k);
    double half_var = v * v;
    half_var = half_var * 0.5;
    double d1 = r + half_var;
    d1 = d1 * t;
    d1 = d1 + log_sk;
    double den = v * sqrt_t;
    d1 = d1 / den;
    double d2 = d1 - den;
    double n1 = bso_cndf_f64(d1);
    double n2 = bso_cndf_f64(d2);
    double fv = k * exp(-(r) * (t));
    if (otype == 0)
        return (s * n1) - (fv * n2);
    return (fv * (1.0 - n2)) - (s * (1.0 - n1));
}

/* ------------------------------------------------------------------------ */
/* FP64 closed forms (exact normal CDF).                                     */
double bso_phi_exact(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

static double pdf_exact(double x) { return exp(-0.5 * x * x) * 0.39894228040143267794; }

double bso_price_exact(double s, double k, double r, double v, double t, int otype)
{
    double srt = v * sqrt(t);
    double d1 = (log(s / k) + (r + 0.5 * v * v) * t) / srt;
    double d2 = d1 - srt;
    double disc = k * exp(-r * t);
    if (otype == 0)
        return s * bso_phi_exact(d1) - disc * bso_phi_exact(d2);
    return disc * bso_phi_exact(-d2) - s * bso_phi_exact(-d1);
}

/* out = {price, delta, vega, rho, theta}; vega and rho per unit (not per %),
   theta = -dV/dT per year. */
void bso_greeks_exact(double s, double k, double r, double v, double t, int otype,
                      double *out)
{
    double sq = sqrt(t);
    double srt = v * sq;
    double d1 = (log(s / k) + (r + 0.5 * v * v) * t) / srt;
    double d2 = d1 - srt;
    double disc = exp(-r * t);
    double kd = k * disc;
    double nd1 = pdf_exact(d1);
    double decay = -s * nd1 * v / (2.0 * sq);
    out[2] = s * nd1 * sq;
    if (otype == 0) {
        out[0] = s * bso_phi_exact(d1) - kd * bso_phi_exact(d2);
        out[1] = bso_phi_exact(d1);
        out[3] = k * t * disc * bso_phi_exact(d2);
        out[4] = decay - r * kd * bso_phi_exact(d2);
    } else {
        out[0] = kd * bso_phi_exact(-d2) - s * bso_phi_exact(-d1);
        out[1] = bso_phi_exact(d1) - 1.0;
        out[3] = -k * t * disc * bso_phi_exact(-d2);
        out[4] = decay + r * kd * bso_phi_exact(-d2);
    }
}

/* Central differences of the A&S-based FP64 price, bump h relative to each
   input: out = {price, d/dS, d/dv, d/dr, -d/dT}. */
void bso_greeks_fd(double s, double k, double r, double v, double t, int otype,
                   double h, double *out)
{
    double hs = h * s, hv = h * v, hr = h * r, ht = h * t;
    out[0] = bso_price_f64(s, k, r, v, t, otype);
    out[1] = (bso_price_f64(s + hs, k, r, v, t, otype) -
              bso_price_f64(s - hs, k, r, v, t, otype)) / (2.0 * hs);
    out[2] = (bso_price_f64(s, k, r, v + hv, t, otype) -
              bso_price_f64(s, k, r, v - hv, t, otype)) / (2.0 * hv);
    out[3] = (bso_price_f64(s, k, r + hr, v, t, otype) -
              bso_price_f64(s, k, r - hr, v, t, otype)) / (2.0 * hr);
    out[4] = -(bso_price_f64(s, k, r, v, t + ht, otype) -
               bso_price_f64(s, k, r, v, t - ht, otype)) / (2.0 * ht);
}

/* ------------------------------------------------------------------------ */
/* Vector helpers for the Python tests.  Each element is a pure function of   */
/* its inputs, so the helpers split [0, n) into contiguous blocks over up to  */
/* 64 threads (one per online CPU) with results independent of the split.    */
enum { VEC_CNDF32, VEC_CNDF64, VEC_PRICE64, VEC_EXACT, VEC_GREEKS_EXACT, VEC_GREEKS_FD };

typedef struct {
    int op;
    int64_t n, begin, end;
    const float *x32, *s, *k, *r, *v, *t;
    const double *x64, *ds, *dk, *dr, *dv, *dt;
    const uint8_t *type;
    double h;
    float *out32;
    double *out;
} vec_job;

static void vec_range(const vec_job *j)
{
    double g[5];
    for (int64_t i = j->begin; i < j->end; i++) {
        switch (j->op) {
        case VEC_CNDF32: j->out32[i] = bso_cndf_f32(j->x32[i]); break;
        case VEC_CNDF64: j->out[i] = bso_cndf_f64(j->x64[i]); break;
        case VEC_PRICE64:
            j->out[i] = bso_price_f64(j->ds[i], j->dk[i], j->dr[i], j->dv[i], j->dt[i],
                                      j->type[i] != 0);
            break;
        case VEC_EXACT:
            j->out[i] = bso_price_exact(j->s[i], j->k[i], j->r[i], j->v[i], j->t[i],
                                        j->type[i] != 0);
            break;
        case VEC_GREEKS_EXACT:
        case VEC_GREEKS_FD:
            if (j->op == VEC_GREEKS_EXACT)
                bso_greeks_exact(j->s[i], j->k[i], j->r[i], j->v[i], j->t[i],
                                 j->type[i] != 0, g);
            else
                bso_greeks_fd(j->s[i], j->k[i], j->r[i], j->v[i], j->t[i], j->type[i] != 0,
                              j->h, g);
            for (int q = 0; q < 5; q++) j->out[q * j->n + i] = g[q];   /* 5 x n, row per quantity */
            break;
        }
    }
}

static void *vec_worker(void *arg)
{
    vec_range((const vec_job *)arg);
    return NULL;
}

static void run_vec(vec_job job)
{
    long cpus = sysconf(_SC_NPROCESSORS_ONLN);
    int64_t nt = cpus < 1 ? 1 : (cpus > 64 ? 64 : cpus);
    if (nt > job.n / 4096 + 1) nt = job.n / 4096 + 1;   /* small inputs: one thread */
    vec_job jobs[64];
    pthread_t tids[64];
    int started[64] = {0};
    int64_t chunk = job.n / nt, extra = job.n % nt, start = 0;
    for (int64_t i = 0; i < nt; i++) {
        jobs[i] = job;
        jobs[i].begin = start;
        jobs[i].end = start + chunk + (i < extra ? 1 : 0);
        start = jobs[i].end;
        started[i] = i > 0 && pthread_create(&tids[i], NULL, vec_worker, &jobs[i]) == 0;
        if (i > 0 && !started[i]) vec_range(&jobs[i]);   /* no thread: run it here */
    }
    vec_range(&jobs[0]);
    for (int64_t i = 1; i < nt; i++)
        if (started[i]) pthread_join(tids[i], NULL);
}

void bso_cndf_f32_vec(int64_t n, const float *x, float *out)
{
    run_vec((vec_job){.op = VEC_CNDF32, .n = n, .x32 = x, .out32 = out});
}

void bso_cndf_f64_vec(int64_t n, const double *x, double *out)
{
    run_vec((vec_job){.op = VEC_CNDF64, .n = n, .x64 = x, .out = out});
}

void bso_price_f64_vec(int64_t n, const double *s, const double *k, const double *r,
                       const double *v, const double *t, const uint8_t *type,
                       double *out)
{
    run_vec((vec_job){.op = VEC_PRICE64, .n = n, .ds = s, .dk = k, .dr = r, .dv = v, .dt = t,
                      .type = type, .out = out});
}

void bso_price_exact_vec(int64_t n, const float *s, const float *k, const float *r,
                         const float *v, const float *t, const uint8_t *type,
                         double *out)
{
    run_vec((vec_job){.op = VEC_EXACT, .n = n, .s = s, .k = k, .r = r, .v = v, .t = t,
                      .type = type, .out = out});
}

/* out is 5 x n (row per quantity). */
void bso_greeks_exact_vec(int64_t n, const float *s, const float *k, const float *r,
                          const float *v, const float *t, const uint8_t *type,
                          double *out)
{
    run_vec((vec_job){.op = VEC_GREEKS_EXACT, .n = n, .s = s, .k = k, .r = r, .v = v, .t = t,
                      .type = type, .out = out});
}

void bso_greeks_fd_vec(int64_t n, const float *s, const float *k, const float *r,
                       const float *v, const float *t, const uint8_t *type, double h,
                       double *out)
{
    run_vec((vec_job){.op = VEC_GREEKS_FD, .n = n, .s = s, .k = k, .r = r, .v = v, .t = t,
                      .type = type, .h = h, .out = out});
}

/* ------------------------------------------------------------------------ */
/* Static-block multi-pass batch (FastFlow ParallelFor semantics).           */
typedef struct {
    int64_t begin, end;
    int passes;
    int cpu;                       /* -1: no pinning */
    const float *s, *k, *r, *v, *t;
    const uint8_t *type;
    float *out;
    pthread_barrier_t *bar;
} bso_block;

static void *block_worker(void *arg)
{
    bso_block *b = (bso_block *)arg;
    if (b->cpu >= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(b->cpu, &set);
        pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
    }
    pthread_barrier_wait(b->bar);           /* start of the timed region */
    for (int p = 0; p < b->passes; p++) {
        for (int64_t i = b->begin; i < b->end; i++)
            b->out[i] = bso_price_f32(b->s[i], b->k[i], b->r[i], b->v[i], b->t[i],
                                      b->type[i] != 0);
        pthread_barrier_wait(b->bar);       /* per-pass barrier */
    }
    return NULL;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Prices the book `passes` times on `nthreads` threads (<=0: every CPU this
   process may run on).  Returns the wall time of the passes in seconds (the
   region between the start barrier and the last pass barrier), or a negative
   value on error.  *used_threads receives the thread count. */
double bso_batch_f32(int64_t n, const float *s, const float *k, const float *r,
                     const float *v, const float *t, const uint8_t *type, float *out,
                     int nthreads, int passes, int pin, int *used_threads)
{
    cpu_set_t allowed;
    int cpus[1024];
    int ncpu = 0;
    CPU_ZERO(&allowed);
    if (sched_getaffinity(0, sizeof(allowed), &allowed) == 0) {
        for (int c = 0; c < CPU_SETSIZE && ncpu < 1024; c++)
            if (CPU_ISSET(c, &allowed)) cpus[ncpu++] = c;
    }
    if (ncpu == 0) {
        cpus[0] = -1;
        ncpu = 1;
    }
    if (nthreads <= 0) nthreads = ncpu;
    if (n > 0 && nthreads > n) nthreads = (int)n;
    if (nthreads < 1) nthreads = 1;
    if (used_threads) *used_threads = nthreads;

    bso_block *blocks = (bso_block *)calloc((size_t)nthreads, sizeof(bso_block));
    pthread_t *tids = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    pthread_barrier_t bar;
    if (!blocks || !tids) return -1.0;
    pthread_barrier_init(&bar, NULL, (unsigned)nthreads + 1);

    int64_t chunk = n / nthreads, extra = n % nthreads, start = 0;
    for (int i = 0; i < nthreads; i++) {
        int64_t len = chunk + (i < extra ? 1 : 0);
        blocks[i] = (bso_block){start, start + len, passes,
                                pin ? cpus[i % ncpu] : -1,
                                s, k, r, v, t, type, out, &bar};
        start += len;
        pthread_create(&tids[i], NULL, block_worker, &blocks[i]);
    }
    pthread_barrier_wait(&bar);
    double t0 = now_s();
    for (int p = 0; p < passes; p++) pthread_barrier_wait(&bar);
    double t1 = now_s();
    for (int i = 0; i < nthreads; i++) pthread_join(tids[i], NULL);
    pthread_barrier_destroy(&bar);
    free(blocks);
    free(tids);
    return t1 - t0;
}

/* ------------------------------------------------------------------------ */
/* The PARSEC input format (read by blackscholes.c:701-734): N, then one line
   "s strike r divq v t type divs DGrefval" per option, floats with 9
   significant digits so that fscanf("%f") reads back the same float32;
   divq, divs and DGrefval are 0, type is 'P' for puts and 'C' for calls.
   Returns 0, or -1 if the file cannot be written. */
int bso_write_parsec_input(const char *path, int64_t n, const float *s, const float *k,
                           const float *r, const float *v, const float *t,
                           const uint8_t *type)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    int ok = fprintf(f, "%lld\n", (long long)n) > 0;
    for (int64_t i = 0; i < n && ok; i++)
        ok = fprintf(f, "%.9g %.9g %.9g 0 %.9g %.9g %c 0 0\n", (double)s[i], (double)k[i],
                     (double)r[i], (double)v[i], (double)t[i], type[i] ? 'P' : 'C') > 0;
    if (fclose(f) != 0) ok = 0;
    return ok ? 0 : -1;
}
