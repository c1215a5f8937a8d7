/*
 * synth.c -- synthetic block-Jacobian and vector generator used as bench and
 * test input (not a reference algorithm).
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled with -ffp-contract=off: every value is
 * produced by integer hashing plus single IEEE roundings, so the device
 * generator in the product library reproduces it bit for bit.
 */
#include "oracle.h"

#include <stddef.h>

extern int orc_g_threads;

/* ------------------------------------------------------------------------- */
/* Synthetic generator (input data, not a reference algorithm)               */
/* ------------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Irwin-Hall sum of four 32-bit uniforms, recentred and rescaled to unit
 * variance: every step is exact or a single IEEE rounding. */
static inline double g_at(uint64_t seed, uint64_t idx)
{
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    const uint64_t h1 = mix64(seed + G * (2 * idx + 1));
    const uint64_t h2 = mix64(seed + G * (2 * idx + 2));
    const uint64_t S = (h1 & 0xffffffffULL) + (h1 >> 32) + (h2 & 0xffffffffULL) + (h2 >> 32);
    const double u = (double)S * 2.3283064365386963e-10; /* 2^-32, exact */
    return (u - 2.0) * 1.7320508075688772;
}

double orc_synth_value(uint64_t seed, int64_t blk, int64_t elem)
{
    return g_at(seed, (uint64_t)blk * 2097152ULL + (uint64_t)elem);
}

void orc_fill_synthetic(const orc_pattern *p, uint64_t seed, double s_diag, double s_off, double *blocks)
{
    const int n = p->n;
    const size_t nn = (size_t)n * n;
#pragma omp parallel for num_threads(orc_g_threads) schedule(static)
    for (int i = 0; i < p->nrows; ++i) {
        for (int b = p->row_ptr[i]; b < p->row_ptr[i + 1]; ++b) {
            const int is_diag = (p->kind[b] == 0 && p->col[b] == i);
            double *B = blocks + (size_t)b * nn;
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) {
                    const int64_t e = (int64_t)c * n + r;
                    const double g = orc_synth_value(seed, b, e);
                    double val = g * (is_diag ? s_diag : s_off);
                    if (is_diag && r == c) val = val + 2.0;
                    B[e] = val;
                }
        }
    }
}

void orc_fill_vector(int64_t len, uint64_t seed, double scale, double *v)
{
    const uint64_t sv = mix64(seed ^ 0xD1B54A32D192ED03ULL);
#pragma omp parallel for num_threads(orc_g_threads) schedule(static)
    for (int64_t k = 0; k < len; ++k) v[k] = g_at(sv, (uint64_t)k) * scale;
}
