/* cpu_pipeline.c -- the CPU baseline of the beamforming path.
 *
 * TEST INFRASTRUCTURE, like oracle.c: the benchmark's CPU leg and the tests
 * call it, nothing in the product does.  It runs the oracle's restatement of
 * the reference arithmetic stage by stage on the host cores and times each
 * stage under the stopwatch name the reference gives it
 * (src/metadata.c:127-139): pfb-wola, pfb-round, pfb-fft, pfb-pack, delay,
 * calc, ipfb; plus the VDIF requantisation (src/beam_vdif.c), which the
 * reference times inside its "write" stopwatch.
 *
 * Stages reuse oracle.c's functions where those are already batched
 * (orc_wola, orc_beamform, orc_calc_phi) and batch its per-row functions
 * (orc_fpga_round / orc_int2float, orc_fft, orc_vdif_requant) with OpenMP.
 * The inverse PFB is the reference's direct form in the reference's own
 * precision (src/pfb.cpp:842-907: FP32 products and sums over a precomputed
 * filter-times-twiddle table, cu_load_ipfb_filter, src/pfb.cpp:1025-1046),
 * not the FP64 checker orc_ipfb, which evaluates a sine and a cosine per
 * multiply-add.
 *
 * Parity: orc_cpu_pfb equals orc_pfb_forward (FP64 out) to FP32 rounding and
 * orc_cpu_ipfb equals orc_ipfb to FP32 accuracy for n >= P*K, both checked
 * in tests/test_oracle_pins.py.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

static double wall(void) { return omp_get_wtime(); }

/* Forward PFB of nspectra spectra x ninput inputs (input as orc_wola takes
 * it), one stage at a time: t[0..3] = seconds in pfb-wola, pfb-round,
 * pfb-fft, pfb-pack.  vout: complex float [m][k'][out_idx[i]] with k' = k +
 * K/2 mod K, Y/K, re/im swapped unless imag_first (orc_pfb_forward's layout
 * and values, src/pfb.cpp:297-343).  Returns 0, or -1 if memory ran out. */
int orc_cpu_pfb(const int8_t *htr, int ninput, int nspectra, int K, int P, const int32_t *h,
                const int32_t *out_idx, int fpga, int imag_first, float *vout, double *t)
{
    const size_t nrow = (size_t)nspectra * ninput, row = 2 * (size_t)K;
    int32_t *b = (int32_t *)malloc(nrow * row * sizeof(int32_t));
    float *bf = (float *)malloc(nrow * row * sizeof(float));
    double *Y = (double *)malloc(nrow * row * sizeof(double));
    if (!b || !bf || !Y) {
        free(b);
        free(bf);
        free(Y);
        return -1;
    }
    double t0 = wall();
    orc_wola(htr, ninput, nspectra, K, P, h, b); /* pfb-wola */
    double t1 = wall();
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < nrow; r++) { /* pfb-round */
        if (fpga) orc_fpga_round(b + r * row, row, bf + r * row);
        else orc_int2float(b + r * row, row, bf + r * row);
    }
    double t2 = wall();
#pragma omp parallel
    {
        double *y = (double *)malloc(row * sizeof(double));
#pragma omp for schedule(static)
        for (size_t r = 0; r < nrow; r++) { /* pfb-fft */
            for (size_t n = 0; n < row; n++) y[n] = bf[r * row + n];
            orc_fft(y, Y + r * row, K);
        }
        free(y);
    }
    double t3 = wall();
#pragma omp parallel for collapse(2) schedule(static)
    for (int m = 0; m < nspectra; m++) /* pfb-pack */
        for (int i = 0; i < ninput; i++) {
            const double *Yr = Y + ((size_t)m * ninput + i) * row;
            for (int k = 0; k < K; k++) {
                const int kp = (k + K / 2) % K;
                const double re = Yr[2 * k] / K, im = Yr[2 * k + 1] / K;
                const size_t x = ((size_t)m * K + kp) * ninput + out_idx[i];
                vout[2 * x] = (float)(imag_first ? re : im);
                vout[2 * x + 1] = (float)(imag_first ? im : re);
            }
        }
    double t4 = wall();
    t[0] = t1 - t0;
    t[1] = t2 - t1;
    t[2] = t3 - t2;
    t[3] = t4 - t3;
    free(b);
    free(bf);
    free(Y);
    return 0;
}

/* The inverse PFB's filter-times-twiddle table as cu_load_ipfb_filter builds
 * it (src/pfb.cpp:1025-1046): ft[ch * F + f] = h[f] e^{2 pi i ch / K}, the
 * roots of unity of src/filter.c:21-48, products in FP64 stored as FP32. */
void orc_cpu_ipfb_table(const double *h, int K, int P, float *ft_re, float *ft_im)
{
    const int F = P * K;
    for (int ch = 0; ch < K; ch++) {
        const double th = 2.0 * M_PI * (double)ch / (double)K;
        const double c = cos(th), s = sin(th);
        for (int f = 0; f < F; f++) {
            ft_re[(size_t)ch * F + f] = (float)(h[f] * c);
            ft_im[(size_t)ch * F + f] = (float)(h[f] * s);
        }
    }
}

/* Outputs n in [n0, n0 + nout) of the direct-form inverse PFB
 * (src/pfb.cpp:842-907), n0 >= P*K so that every filter index is in range.
 * in: complex float [row][k][pol], row = m + P as the reference packs it
 * (cu_invert_pfb); rows up to (n0 + nout - 1)/K + P must exist.  out:
 * complex float [n - n0][pol].  FP32 arithmetic as the reference's kernel. */
void orc_cpu_ipfb(const float *in, long n0, long nout, int K, int P, int npol,
                  const float *ft_re, const float *ft_im, float *out)
{
    const long F = (long)P * K, M = K;
#pragma omp parallel for schedule(static)
    for (long n = n0; n < n0 + nout; n++)
        for (int pol = 0; pol < npol; pol++) {
            float orr = 0.0f, oi = 0.0f;
            const long m0 = (n - F) / M + 1;
            for (long m = m0; m < m0 + P; m++) {
                const long f = n - m * M;
                for (int k = 0; k < K; k++) {
                    const long tw = ((long)(k + K / 2) * n) % K;
                    const size_t ft = (size_t)(F * tw + f);
                    const size_t i = ((size_t)(m + P) * K + k) * npol + pol;
                    orr += in[2 * i] * ft_re[ft] - in[2 * i + 1] * ft_im[ft];
                    oi += in[2 * i] * ft_im[ft] + in[2 * i + 1] * ft_re[ft];
                }
            }
            out[((size_t)(n - n0) * npol + pol) * 2] = orr / K;
            out[((size_t)(n - n0) * npol + pol) * 2 + 1] = oi / K;
        }
}

/* VDIF requantisation of n floats (orc_vdif_requant) over the threads. */
void orc_cpu_vdif(const float *f, size_t n, uint8_t *out)
{
    const size_t chunk = 1 << 16;
    const long nchunk = (long)((n + chunk - 1) / chunk);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nchunk; c++) {
        const size_t a = (size_t)c * chunk;
        orc_vdif_requant(f + a, (a + chunk < n ? chunk : n - a), out + a);
    }
}

/* Geometric-delay phases of npointing pointings (orc_calc_phi, one pointing
 * per thread): phi complex [p][a][c]. */
void orc_cpu_delay(const double *E, const double *N, const double *H, const double *L,
                   int nant, const double *unit, int npointing, const double *freq_hz,
                   int nchan, double Href, double *phi)
{
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < npointing; p++)
        orc_calc_phi(E, N, H, L, nant, unit + 3 * (size_t)p, 1, freq_hz, nchan, Href,
                     phi + (size_t)p * nant * nchan * 2);
}
