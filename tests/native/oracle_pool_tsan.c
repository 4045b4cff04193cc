/* Test-only: the oracle's static-block thread pool (oracle/bs_oracle.c,
   bso_batch_f32) under ThreadSanitizer.  Prices a ragged book on 1 and on 7
   threads over 3 passes and checks the outputs are bit-identical; the build
   (tests/native/Makefile) adds -fsanitize=thread, so a race between the
   blocks, the barrier or the join is reported and fails the run. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

double bso_batch_f32(int64_t n, const float *s, const float *k, const float *r,
                     const float *v, const float *t, const uint8_t *type, float *out,
                     int nthreads, int passes, int pin, int *used_threads);

int main(void)
{
    const int64_t n = 100003;
    float *in = (float *)malloc(sizeof(float) * 5 * n);
    uint8_t *type = (uint8_t *)malloc(n);
    float *a = (float *)malloc(sizeof(float) * n), *b = (float *)malloc(sizeof(float) * n);
    uint32_t x = 12345u;
    for (int64_t i = 0; i < 5 * n; i++) {
        x = x * 1664525u + 1013904223u;
        in[i] = (float)(x >> 8) * (1.0f / 16777216.0f);
    }
    for (int64_t i = 0; i < n; i++) {
        in[i] = 10.0f + 100.0f * in[i];             /* spot */
        in[n + i] = 10.0f + 100.0f * in[n + i];     /* strike */
        in[2 * n + i] = 0.1f * in[2 * n + i];       /* rate */
        in[3 * n + i] = 0.05f + 0.6f * in[3 * n + i];   /* volatility */
        in[4 * n + i] = 0.05f + 2.0f * in[4 * n + i];   /* time */
        type[i] = (uint8_t)(i % 3 == 0);
    }
    int u1 = 0, u7 = 0;
    if (bso_batch_f32(n, in, in + n, in + 2 * n, in + 3 * n, in + 4 * n, type, a, 1, 1, 0, &u1) < 0 ||
        bso_batch_f32(n, in, in + n, in + 2 * n, in + 3 * n, in + 4 * n, type, b, 7, 3, 0, &u7) < 0) {
        fprintf(stderr, "batch failed\n");
        return 1;
    }
    if (u1 != 1 || u7 != 7 || memcmp(a, b, sizeof(float) * n) != 0) {
        fprintf(stderr, "thread counts %d %d or outputs differ\n", u1, u7);
        return 1;
    }
    printf("oracle pool: %lld options, 7 threads x 3 passes, bit-identical to 1 thread\n", (long long)n);
    free(in); free(type); free(a); free(b);
    return 0;
}
