/*
 * vmsynth.c -- seeded synthetic inputs for the voltage-beamforming path.
 *
 * A counter-based generator (SplitMix64 finaliser over seed and element
 * index) so that any element can be produced independently: the same seed
 * yields byte-identical arrays on any host, in any thread order.  Gaussian
 * deviates use the Box-Muller transform on consecutive uniform pairs.
 *
 * Seeds used by the workloads: data 1, filter 2, Jones 3, geometry 4,
 * beams 5, flags 6, surrogate network 7.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t draw(uint64_t seed, uint64_t idx)
{
    return mix64(seed * 0xD1B54A32D192ED03ULL + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}

static inline double unif(uint64_t seed, uint64_t idx)
{
    return (double)(draw(seed, idx) >> 11) * 0x1.0p-53; /* [0, 1) */
}

static inline void gauss2(uint64_t seed, uint64_t pair, double *z0, double *z1)
{
    double u1 = unif(seed, 2 * pair), u2 = unif(seed, 2 * pair + 1);
    double r = sqrt(-2.0 * log(1.0 - u1));
    double t = 2.0 * M_PI * u2;
    *z0 = r * cos(t);
    *z1 = r * sin(t);
}

static inline int8_t q8(double x)
{
    double r = round(x);
    if (r > 127.0) r = 127.0;
    if (r < -127.0) r = -127.0;
    return (int8_t)r;
}

/* n uniform deviates in [0,1) for indices offset .. offset+n-1 */
void vmsynth_uniform(uint64_t seed, uint64_t offset, size_t n, double *out)
{
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) out[j] = unif(seed, offset + j);
}

/* n standard normal deviates (n even or odd), pair index offset */
void vmsynth_normal(uint64_t seed, uint64_t offset, size_t n, double *out)
{
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < (n + 1) / 2; j++) {
        double a, b;
        gauss2(seed, offset + j, &a, &b);
        out[2 * j] = a;
        if (2 * j + 1 < n) out[2 * j + 1] = b;
    }
}

/*
 * MWAX 8+8-bit voltages: buffer of nblocks voltage blocks, each
 * [ninput][spb] (re,im) int8 pairs (the layout mwalib's read_second
 * delivers).  Sample (block b, input i, t) draws Gaussian pair number
 * ((first_block + b) * ninput + i) * spb + t, scaled by sigma, rounded and
 * clipped to [-127, 127].  Blocks with index < zero_blocks are zero-filled
 * (a fresh margin).
 */
void vmsynth_mwax_voltages(uint64_t seed, uint64_t first_block, int nblocks,
                           int ninput, int spb, double sigma, int zero_blocks,
                           int8_t *out)
{
    size_t per_block = (size_t)ninput * spb;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < nblocks; b++)
        for (int i = 0; i < ninput; i++) {
            int8_t *dst = out + 2 * ((size_t)b * per_block + (size_t)i * spb);
            uint64_t base = ((first_block + (uint64_t)b) * ninput + i) * (uint64_t)spb;
            for (int t = 0; t < spb; t++) {
                if (b < zero_blocks) {
                    dst[2 * t] = dst[2 * t + 1] = 0;
                    continue;
                }
                double z0, z1;
                gauss2(seed, base + t, &z0, &z1);
                dst[2 * t] = q8(sigma * z0);
                dst[2 * t + 1] = q8(sigma * z1);
            }
        }
}

/*
 * Legacy 4+4-bit recombined voltages [s][c][i]: each nibble is a two's
 * complement integer in [-8, 7] drawn as round(sigma * N(0,1)) clipped.
 * High nibble real, low nibble imaginary.
 */
void vmsynth_legacy_int4(uint64_t seed, size_t n, double sigma, uint8_t *out)
{
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
        double z0, z1;
        gauss2(seed, j, &z0, &z1);
        int r = (int)round(sigma * z0), i = (int)round(sigma * z1);
        r = r < -8 ? -8 : (r > 7 ? 7 : r);
        i = i < -8 ? -8 : (i > 7 ? 7 : i);
        out[j] = (uint8_t)(((r & 0xf) << 4) | (i & 0xf));
    }
}

static double bessel_i0(double x)
{
    /* power series, converges for the |x| <= 10 used here */
    double sum = 1.0, term = 1.0, q = x * x / 4.0;
    for (int k = 1; k < 200; k++) {
        term *= q / ((double)k * k);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

/*
 * Kaiser-windowed sinc low-pass prototype of ntaps_total coefficients with
 * cutoff 1/K (one fine channel), window parameter beta, scaled so that the
 * largest coefficient is `peak` and rounded to integers.  A seeded dither of
 * +-0.5 LSB before rounding breaks the exact symmetry, so that a reversed
 * filter is distinguishable from the original.
 */
void vmsynth_kaiser_filter(uint64_t seed, int ntaps_total, int K, double beta,
                           double peak, double *out)
{
    double mx = 0.0;
    double centre = 0.5 * (ntaps_total - 1);
    for (int n = 0; n < ntaps_total; n++) {
        double t = (n - centre) / (double)K;
        double s = (t == 0.0) ? 1.0 : sin(M_PI * t) / (M_PI * t);
        double r = 2.0 * n / (ntaps_total - 1) - 1.0;
        double w = bessel_i0(beta * sqrt(fmax(0.0, 1.0 - r * r))) / bessel_i0(beta);
        out[n] = s * w;
        if (fabs(out[n]) > mx) mx = fabs(out[n]);
    }
    for (int n = 0; n < ntaps_total; n++) {
        double d = unif(seed, (uint64_t)n) - 0.5;
        double v = round(out[n] / mx * peak + d);
        if (v > peak) v = peak;
        if (v < -peak) v = -peak;
        out[n] = v;
    }
}
