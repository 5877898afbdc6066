/*
 * oracle.c -- CPU restatement of VCSBeam's voltage-beamforming arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py; never by the product.
 *
 * Integer stages are exact.  Floating-point stages run in FP64.  Each
 * function names the reference file:line it follows (paths relative to the
 * MWA-vcsbeam source tree).  OpenMP parallelises the outer loops so that the
 * same code doubles as the CPU baseline.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define VSPVB 64000 /* samples per MWAX voltage block, include/vcsbeam.h.in:238 */

static inline size_t mwax_idx(size_t s, size_t i, size_t ni)
{
    /* vMWAX_IDX, include/vcsbeam.h.in:239-240 */
    return ((s / VSPVB) * ni + i) * VSPVB + (s % VSPVB);
}

/* ---------------------------------------------------------------- a2 ---- */
void orc_wola(const int8_t *htr, int ninput, int nspectra, int K, int P,
              const int32_t *h, int32_t *b)
{
    /* src/pfb.cpp:121-179: mprime = m + 500 - P (hard-coded 500, line 153),
     * h_idx = K*P - K*p - n - 1, x_idx = vMWAX_IDX(mprime*M + p*K + n). */
#pragma omp parallel for collapse(2) schedule(static)
    for (int m = 0; m < nspectra; m++)
        for (int i = 0; i < ninput; i++) {
            for (int n = 0; n < K; n++) {
                int64_t X = 0, Y = 0;
                for (int p = 0; p < P; p++) {
                    int hv = h[K * P - K * p - n - 1];
                    size_t s = (size_t)(m + 500 - P) * K + (size_t)p * K + n;
                    size_t xi = mwax_idx(s, i, ninput);
                    X += (int64_t)hv * htr[2 * xi];
                    Y += (int64_t)hv * htr[2 * xi + 1];
                }
                size_t bi = ((size_t)m * ninput + i) * K + n;
                b[2 * bi] = (int32_t)(uint32_t)X;     /* int atomicAdd wraps */
                b[2 * bi + 1] = (int32_t)(uint32_t)Y;
            }
        }
}

/* ---------------------------------------------------------------- a3 ---- */
void orc_fpga_round(const int32_t *b, size_t n, float *out)
{
    /* src/pfb.cpp:225-240 */
    for (size_t j = 0; j < n; j++) {
        int32_t X = b[j];
        if (X > 0) X += 0x2000;
        X >>= 14; /* arithmetic shift == floor division by 2^14 */
        out[j] = (float)X;
    }
}

void orc_int2float(const int32_t *b, size_t n, float *out)
{
    /* src/pfb.cpp:266: (float)(*pint) * scale, scale = 1.0/16384.0 (double) */
    for (size_t j = 0; j < n; j++)
        out[j] = (float)((double)(float)b[j] * (1.0 / 16384.0));
}

/* ---------------------------------------------------------------- a4 ---- */
void orc_dft(const double *in, double *out, int K)
{
    for (int k = 0; k < K; k++) {
        double sr = 0.0, si = 0.0;
        for (int n = 0; n < K; n++) {
            /* exact argument reduction: (k*n) mod K */
            double th = -2.0 * M_PI * (double)((long)k * n % K) / (double)K;
            double c = cos(th), s = sin(th);
            sr += in[2 * n] * c - in[2 * n + 1] * s;
            si += in[2 * n] * s + in[2 * n + 1] * c;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
}

void orc_fft(const double *in, double *out, int K)
{
    /* textbook iterative radix-2 decimation-in-time (Cooley-Tukey) */
    int lg = 0;
    while ((1 << lg) < K) lg++;
    for (int n = 0; n < K; n++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (n & (1 << b)) r |= 1 << (lg - 1 - b);
        out[2 * r] = in[2 * n];
        out[2 * r + 1] = in[2 * n + 1];
    }
    for (int len = 2; len <= K; len <<= 1) {
        int half = len / 2;
        for (int j = 0; j < half; j++) {
            double th = -2.0 * M_PI * (double)j / (double)len;
            double wr = cos(th), wi = sin(th);
            for (int st = 0; st < K; st += len) {
                double *a = out + 2 * (st + j), *c = out + 2 * (st + j + half);
                double tr = c[0] * wr - c[1] * wi, ti = c[0] * wi + c[1] * wr;
                c[0] = a[0] - tr;
                c[1] = a[1] - ti;
                a[0] += tr;
                a[1] += ti;
            }
        }
    }
}

/* --------------------------------------------------------------- a1-a5 -- */
void orc_pfb_forward(const int8_t *htr, int ninput, int nspectra, int K, int P,
                     const int32_t *h, const int32_t *out_idx, int fpga,
                     int imag_first, int use_fft, double *vout)
{
#pragma omp parallel
    {
        int32_t *bi = (int32_t *)malloc(sizeof(int32_t) * 2 * K);
        float *bf = (float *)malloc(sizeof(float) * 2 * K);
        double *y = (double *)malloc(sizeof(double) * 2 * K);
        double *Y = (double *)malloc(sizeof(double) * 2 * K);
#pragma omp for collapse(2) schedule(static)
        for (int m = 0; m < nspectra; m++)
            for (int i = 0; i < ninput; i++) {
                /* a2 (src/pfb.cpp:121-179) */
                for (int n = 0; n < K; n++) {
                    int64_t X = 0, Yi = 0;
                    for (int p = 0; p < P; p++) {
                        int hv = h[K * P - K * p - n - 1];
                        size_t s = (size_t)(m + 500 - P) * K + (size_t)p * K + n;
                        size_t xi = mwax_idx(s, i, ninput);
                        X += (int64_t)hv * htr[2 * xi];
                        Yi += (int64_t)hv * htr[2 * xi + 1];
                    }
                    bi[2 * n] = (int32_t)(uint32_t)X;
                    bi[2 * n + 1] = (int32_t)(uint32_t)Yi;
                }
                /* a3 */
                if (fpga) orc_fpga_round(bi, 2 * K, bf);
                else orc_int2float(bi, 2 * K, bf);
                for (int n = 0; n < 2 * K; n++) y[n] = bf[n];
                /* a4 */
                if (use_fft) orc_fft(y, Y, K);
                else orc_dft(y, Y, K);
                /* a5 (src/pfb.cpp:297-343) */
                int iout = out_idx[i];
                for (int k = 0; k < K; k++) {
                    int kp = (k + K / 2) % K;
                    double re = Y[2 * k] / K, im = Y[2 * k + 1] / K;
                    size_t X_idx = ((size_t)m * K + kp) * ninput + iout;
                    if (imag_first) {
                        vout[2 * X_idx] = re;
                        vout[2 * X_idx + 1] = im;
                    } else {
                        vout[2 * X_idx] = im;
                        vout[2 * X_idx + 1] = re;
                    }
                }
            }
        free(bi);
        free(bf);
        free(y);
        free(Y);
    }
}

uint8_t orc_demote(double x)
{
    /* DEMOTE(x) = INT_TO_UINT8((int)round(x)) with CLIP to [-8, 7]
     * (include/vcsbeam.h.in:453-456) */
    int v = (int)round(x);
    if (v < -8) v = -8;
    if (v > 7) v = 7;
    return (uint8_t)(v < 0 ? v + 0x10 : v);
}

void orc_pack_int4(const double *vout, size_t nsamples, uint8_t *packed)
{
    /* PACK_NIBBLES(r,i) = (DEMOTE(r) << 4) + DEMOTE(i); the reference passes
     * the stored (first, second) pair, src/pfb.cpp:327-331 */
    for (size_t j = 0; j < nsamples; j++)
        packed[j] = (uint8_t)((orc_demote(vout[2 * j]) << 4) +
                              orc_demote(vout[2 * j + 1]));
}

/* ---------------------------------------------------------------- a8 ---- */
void orc_unpack_int4(uint8_t byte, double *re, double *im)
{
    /* REAL_NIBBLE_TO_UINT8 = (X>>4)&0xf, IMAG = X&0xf, UINT8_TO_INT
     * (include/vcsbeam.h.in:997-1003) */
    int r = (byte >> 4) & 0xf, i = byte & 0xf;
    *re = (double)(r >= 8 ? r - 16 : r);
    *im = (double)(i >= 8 ? i - 16 : i);
}

/* ---------------------------------------------------------------- a9 ---- */
void orc_incoh(const uint8_t *v, int nsample, int nchan, int ninput, float *I)
{
    /* src/form_beam.cpp:60-87; every term is an integer <= 128 and every
     * partial sum < 2^24, so the float atomics of the reference are exact */
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < nsample; s++)
        for (int c = 0; c < nchan; c++) {
            long acc = 0;
            for (int i = 0; i < ninput; i++) {
                double re, im;
                orc_unpack_int4(v[((size_t)s * nchan + c) * ninput + i], &re, &im);
                acc += (long)(re * re + im * im);
            }
            I[(size_t)s * nchan + c] = (float)acc;
        }
}

/* --------------------------------------------------------------- a13 ---- */
typedef struct { double x, y; } cplx;
static inline cplx cmul(cplx a, cplx b) { cplx r = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.x + b.x, a.y + b.y}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.x - b.x, a.y - b.y}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.x, -a.y}; return r; }

void orc_calc_J(const double *D, const double *B, int npointing, int nant,
                int nchan, double *J)
{
    const cplx *Dc = (const cplx *)D, *Bc = (const cplx *)B;
    cplx *Jc = (cplx *)J;
    for (int p = 0; p < npointing; p++)
        for (int a = 0; a < nant; a++)
            for (int c = 0; c < nchan; c++) {
                const cplx *M1 = Dc + ((size_t)a * nchan + c) * 4;  /* D_IDX */
                const cplx *M2 = Bc + ((size_t)p * nant + a) * 4;   /* PB_IDX */
                cplx Ji[4];
                /* mult2x2d, src/jones.c:316-330 */
                Ji[0] = cadd(cmul(M1[0], M2[0]), cmul(M1[1], M2[2]));
                Ji[1] = cadd(cmul(M1[0], M2[1]), cmul(M1[1], M2[3]));
                Ji[2] = cadd(cmul(M1[2], M2[0]), cmul(M1[3], M2[2]));
                Ji[3] = cadd(cmul(M1[2], M2[1]), cmul(M1[3], M2[3]));
                /* norm2x2, src/jones.c:404-424 */
                double F = 0.0;
                for (int q = 0; q < 4; q++) F += Ji[q].x * Ji[q].x + Ji[q].y * Ji[q].y;
                F = sqrt(F);
                cplx *out = Jc + (((size_t)p * nant + a) * nchan + c) * 4; /* J_IDX */
                if (F == 0.0) {
                    for (int q = 0; q < 4; q++) out[q].x = out[q].y = 0.0;
                    continue;
                }
                for (int q = 0; q < 4; q++) { Ji[q].x /= F; Ji[q].y /= F; }
                /* inv2x2S, src/jones.c:294-306 (reciprocal_complex 215-219) */
                cplx det = csub(cmul(Ji[0], Ji[3]), cmul(Ji[1], Ji[2]));
                double sc = 1.0 / (det.x * det.x + det.y * det.y);
                cplx inv = {sc * det.x, -sc * det.y}, ninv = {-inv.x, -inv.y};
                out[0] = cmul(inv, Ji[3]);
                out[1] = cmul(ninv, Ji[1]);
                out[2] = cmul(ninv, Ji[2]);
                out[3] = cmul(inv, Ji[0]);
            }
}

/* --------------------------------------------------------------- a14 ---- */
#define ORC_C 299792458.0 /* SPEED_OF_LIGHT_IN_VACUUM_M_PER_S (mwalib) */

static cplx orc_phase(double E, double N, double H, double L,
                      const double *unit, double f, double Href)
{
    /* src/geometry.c:82-99 */
    double w = E * unit[0] + N * unit[1] + (H - Href) * unit[2];
    double dt = (w - L) / ORC_C;
    /* 2 pi dt f, reduced exactly in cycles: the product dt*f is formed with
     * an error-free FMA split so that the fractional cycle is accurate */
    double cyc = dt * f;
    double err = fma(dt, f, -cyc);
    double fr = cyc - nearbyint(cyc);
    fr += err;
    cplx r = {cos(2.0 * M_PI * fr), sin(2.0 * M_PI * fr)};
    return r;
}

void orc_calc_phi(const double *E, const double *N, const double *H,
                  const double *L, int nant, const double *unit, int npointing,
                  const double *freq_hz, int nchan, double Href, double *phi)
{
    cplx *ph = (cplx *)phi;
    for (int p = 0; p < npointing; p++)
        for (int a = 0; a < nant; a++)
            for (int c = 0; c < nchan; c++)
                ph[((size_t)p * nant + a) * nchan + c] = /* PHI_IDX */
                    orc_phase(E[a], N[a], H[a], L[a], unit + 3 * p, freq_hz[c], Href);
}

double orc_array_factor(const double *E, const double *N, const double *H,
                        const double *L, int nant, const double *unit1,
                        const double *unit2, double freq_hz, double Href)
{
    /* src/geometry.c:440-467 */
    cplx acc = {0.0, 0.0};
    for (int a = 0; a < nant; a++) {
        cplx w = orc_phase(E[a], N[a], H[a], L[a], unit1, freq_hz, Href);
        cplx psi = orc_phase(E[a], N[a], H[a], L[a], unit2, freq_hz, Href);
        acc = cadd(acc, cmul(cconj(w), psi));
    }
    double af = hypot(acc.x, acc.y) / nant;
    return af * af;
}

/* ----------------------------------------------------------- a10 + a11 -- */
void orc_beamform(const void *v, int datatype, const double *J,
                  const double *phi, const uint32_t *polQ,
                  const uint32_t *polP, int npointing, int nant, int nchan,
                  int nsample, int ninput, double invw, int nstokes,
                  double *S, double *e)
{
    const cplx *Jc = (const cplx *)J, *ph = (const cplx *)phi;
#pragma omp parallel for collapse(3) schedule(static)
    for (int p = 0; p < npointing; p++)
        for (int s = 0; s < nsample; s++)
            for (int c = 0; c < nchan; c++) {
                cplx Ex = {0, 0}, Ey = {0, 0}, Nxy = {0, 0};
                double Nxx = 0.0, Nyy = 0.0;
                for (int a = 0; a < nant; a++) {
                    cplx vq, vp;
                    size_t iq = ((size_t)s * nchan + c) * ninput + polQ[a];
                    size_t ip = ((size_t)s * nchan + c) * ninput + polP[a];
                    if (datatype == 0) {
                        orc_unpack_int4(((const uint8_t *)v)[iq], &vq.x, &vq.y);
                        orc_unpack_int4(((const uint8_t *)v)[ip], &vp.x, &vp.y);
                    } else {
                        vq = ((const cplx *)v)[iq];
                        vp = ((const cplx *)v)[ip];
                    }
                    /* vmApplyJ_kernel, src/form_beam.cpp:175-178 */
                    const cplx *Jm = Jc + (((size_t)p * nant + a) * nchan + c) * 4;
                    cplx JvQ = cadd(cmul(Jm[0], vq), cmul(Jm[1], vp));
                    cplx JvP = cadd(cmul(Jm[2], vq), cmul(Jm[3], vp));
                    /* vmBeamform_kernel, src/form_beam.cpp:297-320 */
                    cplx f = ph[((size_t)p * nant + a) * nchan + c];
                    cplx ex = cmul(f, JvQ), ey = cmul(f, JvP);
                    Ex = cadd(Ex, ex);
                    Ey = cadd(Ey, ey);
                    Nxx += ex.x * ex.x + ex.y * ex.y;
                    Nyy += ey.x * ey.x + ey.y * ey.y;
                    Nxy = cadd(Nxy, cmul(ex, cconj(ey)));
                }
                /* src/form_beam.cpp:322-339 */
                double bnXX = Ex.x * Ex.x + Ex.y * Ex.y - Nxx;
                double bnYY = Ey.x * Ey.x + Ey.y * Ey.y - Nyy;
                cplx bnXY = csub(cmul(Ex, cconj(Ey)), Nxy);
                if (S) {
                    size_t base = (((size_t)p * nsample + s) * nstokes) * nchan + c;
                    S[base] = invw * (bnXX + bnYY);
                    if (nstokes == 4) {
                        S[base + (size_t)nchan] = invw * (bnXX - bnYY);
                        S[base + 2 * (size_t)nchan] = 2.0 * invw * bnXY.x;
                        S[base + 3 * (size_t)nchan] = -2.0 * invw * bnXY.y;
                    }
                }
                if (e) {
                    size_t eb = (((size_t)p * nsample + s) * nchan + c) * 2;
                    e[2 * eb] = Ex.x;
                    e[2 * eb + 1] = Ex.y;
                    e[2 * eb + 2] = Ey.x;
                    e[2 * eb + 3] = Ey.y;
                }
            }
}

/* --------------------------------------------------------------- a12 ---- */
void orc_renormalise(const float *S, int npointing, int nstep, int nstokes,
                     int nchan, float *offsets, float *scales, uint8_t *Sscaled,
                     uint8_t *valid)
{
    for (int p = 0; p < npointing; p++)
        for (int st = 0; st < nstokes; st++)
            for (int c = 0; c < nchan; c++) {
#define CI(s) ((((size_t)p * nstep + (s)) * nstokes + st) * nchan + c)
                float mn = S[CI(0)], mx = S[CI(0)];
                for (int s = 0; s < nstep; s++) {
                    float val = S[CI(s)];
                    mn = val < mn ? val : mn;
                    mx = val > mx ? val : mx;
                }
                /* src/form_beam.cpp:436-437: scale=(max-min)/256.0, offset=min */
                float scale = (float)((mx - mn) / 256.0);
                float offset = mn;
                for (int s = 0; s < nstep; s++) {
                    /* src/form_beam.cpp:443-444: (uint8_t)(val + 0.5), the
                     * sum in double */
                    float val = (S[CI(s)] - offset) / scale;
                    double r = (double)val + 0.5;
                    int ok = (scale != 0.0f) && (r < 256.0) && (r >= 0.0);
                    Sscaled[CI(s)] = ok ? (uint8_t)r : 0;
                    if (valid) valid[CI(s)] = (uint8_t)ok;
                }
                scales[((size_t)p * nstokes + st) * nchan + c] = scale;
                offsets[((size_t)p * nstokes + st) * nchan + c] = offset;
#undef CI
            }
}

/* --------------------------------------------------------------- a16 ---- */
void orc_ipfb_gather(const double *ring, int file_no, int nsamples, int K,
                     int P, int npol, double *in)
{
    /* src/pfb.cpp:950-984 */
    int start_s = (file_no % 2 == 0 ? 2 * nsamples - P : nsamples - P);
    for (int s_in = 0; s_in < nsamples + P; s_in++) {
        int s = (start_s + s_in) % (2 * nsamples);
        for (int ch = 0; ch < K; ch++)
            for (int pol = 0; pol < npol; pol++) {
                size_t i = ((size_t)s_in * K + ch) * npol + pol;
                size_t j = ((size_t)s * K + ch) * npol + pol;
                if (file_no == 0 && s_in < P) {
                    in[2 * i] = 0.0;
                    in[2 * i + 1] = 0.0;
                } else {
                    in[2 * i] = ring[2 * j];
                    in[2 * i + 1] = ring[2 * j + 1];
                }
            }
    }
}

void orc_ipfb(const double *in, int nsamples, int K, int P, int npol,
              const double *h, double *out, uint8_t *mask)
{
    /* ipfb_kernel, src/pfb.cpp:842-907, with the filter-twiddle products
     * ft = h[f] * e^{2 pi i tw / K} (cu_load_ipfb_filter, 1025-1046) kept in
     * FP64.  m0 uses C truncating division (line 870). */
    const int F = P * K, M = K;
    const long nout = (long)nsamples * K;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < nout; n++) {
        int ok = !(n < F && (n % K) != 0);
        if (mask) mask[n] = (uint8_t)ok;
        for (int pol = 0; pol < npol; pol++) {
            double orr = 0.0, oi = 0.0;
            if (ok) {
                long m0 = (n - F) / M + 1; /* truncation toward zero */
                for (long m = m0; m < m0 + P; m++) {
                    long f = n - m * M;
                    for (int k = 0; k < K; k++) {
                        long tw = ((long)(k + K / 2) * n) % K;
                        double th = 2.0 * M_PI * (double)tw / (double)K;
                        double fr = h[f] * cos(th), fi = h[f] * sin(th);
                        size_t i = ((size_t)(m + P) * K + k) * npol + pol;
                        orr += in[2 * i] * fr - in[2 * i + 1] * fi;
                        oi += in[2 * i] * fi + in[2 * i + 1] * fr;
                    }
                }
            }
            out[((size_t)n * npol + pol) * 2] = orr / K;
            out[((size_t)n * npol + pol) * 2 + 1] = oi / K;
        }
    }
}

/* --------------------------------------------------------------- a20 ---- */
void orc_vdif_requant(const float *f, size_t n, uint8_t *out)
{
    for (size_t j = 0; j < n; j++) {
        float x = f[j];
        x = (x > 127.0f) ? 127.0f : x;   /* float2int8_trunc order: max */
        x = (x < -126.0f) ? -126.0f : x; /* then min */
        int8_t q = (int8_t)rintf(x);     /* round half to even (default mode) */
        out[j] = (uint8_t)(q ^ 0x80);    /* to_offset_binary */
    }
}

/* VDIF 1.0 frame headers and frames, as src/beam_vdif.c:52-88 and 196-205
 * produce them through vdifio (createVDIFHeader, setVDIFEpochMJD,
 * setVDIFFrameMJDSec, nextVDIFHeader).  vdifio is not part of the reference
 * tree; the header follows the VDIF 1.0 specification (vlbi.org, 2009):
 *   word 0: invalid (bit 31), legacy (30), seconds since the epoch (29-0)
 *   word 1: epoch as half-years since 2000-01-01 (29-24), frame (23-0)
 *   word 2: version (31-29), log2 channels (28-24), frame bytes / 8 (23-0)
 *   word 3: complex (31), bits - 1 (30-26), thread (25-16), station (15-0,
 *           two ASCII characters, the first in the high byte)
 * Dates: days since 1970-01-01 by the era decomposition of the proleptic
 * Gregorian calendar (H. Hinnant, "chrono-compatible low-level date
 * algorithms"), MJD = days + 40587. */
static long orc_days_from_civil(long y, long m, long d)
{
    y -= m <= 2;
    const long era = (y >= 0 ? y : y - 399) / 400;
    const long yoe = y - era * 400;
    const long doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
    const long doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
    return era * 146097 + doe - 719468;
}

long orc_vdif_epoch_mjd(int e)
{
    return orc_days_from_civil(2000 + e / 2, (e & 1) ? 7 : 1, 1) + 40587;
}

void orc_vdif_header(int payload_bytes, int threadid, int bits, int nchan, int iscomplex,
                     const char *station, long mjd, long sec_of_day, uint32_t *w)
{
    int e = 0;
    while (e < 63 && orc_vdif_epoch_mjd(e + 1) <= mjd) e++;
    const long secs = (mjd - orc_vdif_epoch_mjd(e)) * 86400 + sec_of_day;
    int lg = 0;
    while ((1 << lg) < nchan) lg++;
    memset(w, 0, 8 * sizeof(uint32_t));
    w[0] = (uint32_t)secs & 0x3FFFFFFFu;
    w[1] = (uint32_t)e << 24;
    w[2] = ((uint32_t)lg << 24) | (uint32_t)((payload_bytes + 32) / 8);
    w[3] = ((uint32_t)(iscomplex != 0) << 31) | ((uint32_t)(bits - 1) << 26) |
           ((uint32_t)threadid << 16) | ((uint32_t)(uint8_t)station[0] << 8) |
           (uint32_t)(uint8_t)station[1];
}

void orc_vdif_frames(const float *in, size_t nframes, int payload_bytes, uint32_t *hdr,
                     int frame_rate, uint8_t *out)
{
    /* one frame at a time: header copy, requantised payload, next header */
    for (size_t f = 0; f < nframes; f++) {
        uint8_t *o = out + f * (size_t)(32 + payload_bytes);
        memcpy(o, hdr, 32);
        orc_vdif_requant(in + f * (size_t)payload_bytes, (size_t)payload_bytes, o + 32);
        uint32_t frame = (hdr[1] & 0x00FFFFFFu) + 1, sec = hdr[0] & 0x3FFFFFFFu;
        if ((int)frame >= frame_rate) {
            frame = 0;
            sec++;
        }
        hdr[0] = (hdr[0] & 0xC0000000u) | (sec & 0x3FFFFFFFu);
        hdr[1] = (hdr[1] & 0xFF000000u) | frame;
    }
}

/* ------------------------------------------------ primary beam (a19) --- */
#define ORC_C 299792458.0
#define ORC_DIP_SPACING 1.1
#define ORC_DIP_HEIGHT 0.278

void orc_dipole_beam(const double *az, const double *za, const double *freq, size_t n,
                     const double *g, const double *tau, double *J)
{
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n; j++) {
        const double sa = sin(az[j]), ca = cos(az[j]), sz = sin(za[j]), cz = cos(za[j]);
        const double l = sz * sa, m = sz * ca, f = freq[j];
        const double k = 2.0 * M_PI * f / ORC_C;
        double afr[2] = {0, 0}, afi[2] = {0, 0};
        for (int P = 0; P < 2; P++)
            for (int d = 0; d < 16; d++) {
                const double E = ((d % 4) - 1.5) * ORC_DIP_SPACING;
                const double N = ((d / 4) - 1.5) * ORC_DIP_SPACING;
                const double ph = k * (E * l + N * m) - 2.0 * M_PI * f * tau[P * 16 + d];
                const double gr = g[2 * (P * 16 + d)], gi = g[2 * (P * 16 + d) + 1];
                const double c = cos(ph), s = sin(ph);
                afr[P] += gr * c - gi * s;
                afi[P] += gr * s + gi * c;
            }
        const double gp = 2.0 * sin(k * ORC_DIP_HEIGHT * cz);
        const double n0 = 16.0 * 2.0 * sin(k * ORC_DIP_HEIGHT);
        const double sc = gp / n0;
        const double el[2][2] = {{cz * sa, ca}, {cz * ca, -sa}};
        for (int P = 0; P < 2; P++)
            for (int q = 0; q < 2; q++) {
                J[8 * j + 4 * P + 2 * q] = sc * afr[P] * el[P][q];
                J[8 * j + 4 * P + 2 * q + 1] = sc * afi[P] * el[P][q];
            }
    }
}

void orc_mlp_features(const double *az, const double *za, const double *freq, size_t n,
                      double *X)
{
    for (size_t j = 0; j < n; j++) {
        const double sa = sin(az[j]), ca = cos(az[j]), sz = sin(za[j]), cz = cos(za[j]);
        const double l = sz * sa, m = sz * ca;
        const double k = 2.0 * M_PI * freq[j] / ORC_C;
        const double px = k * 0.5 * ORC_DIP_SPACING * l, py = k * 0.5 * ORC_DIP_SPACING * m;
        const double gh = k * ORC_DIP_HEIGHT * cz;
        double *x = X + 16 * j;
        x[0] = l;
        x[1] = m;
        x[2] = cz;
        x[3] = (freq[j] - 200e6) / 30e6;
        x[4] = sa;
        x[5] = ca;
        x[6] = sin(px);
        x[7] = cos(px);
        x[8] = sin(3.0 * px);
        x[9] = cos(3.0 * px);
        x[10] = sin(py);
        x[11] = cos(py);
        x[12] = sin(3.0 * py);
        x[13] = cos(3.0 * py);
        x[14] = sin(gh);
        x[15] = cos(gh);
    }
}

static double orc_gelu(double x, double *dgelu)
{
    const double a = 0.7978845608028654, c = 0.044715;
    const double u = a * (x + c * x * x * x);
    const double t = tanh(u);
    if (dgelu) *dgelu = 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * a * (1.0 + 3.0 * c * x * x);
    return 0.5 * x * (1.0 + t);
}

/* forward keeping every layer's pre-activation z_l and output a_l */
static void orc_mlp_run(const double *X, int B, int nl, const int *dims, const double *W,
                        const double *b, double **z, double **a)
{
    size_t wo = 0, bo = 0;
    for (int L = 0; L < nl; L++) {
        const int ni = dims[L], no = dims[L + 1];
        const double *in = L == 0 ? X : a[L - 1];
        for (int r = 0; r < B; r++)
            for (int o = 0; o < no; o++) {
                double acc = b[bo + o];
                for (int i = 0; i < ni; i++) acc += W[wo + (size_t)o * ni + i] * in[(size_t)r * ni + i];
                z[L][(size_t)r * no + o] = acc;
                a[L][(size_t)r * no + o] = L < nl - 1 ? orc_gelu(acc, NULL) : acc;
            }
        wo += (size_t)ni * no;
        bo += no;
    }
}

void orc_mlp_forward(const double *X, int B, int nl, const int *dims, const double *W,
                     const double *b, double *Y)
{
    double *z[16] = {0}, *a[16] = {0};
    if (nl < 1 || nl > 16) return;
    for (int L = 0; L < nl; L++) {
        z[L] = (double *)malloc(sizeof(double) * B * dims[L + 1]);
        a[L] = (double *)malloc(sizeof(double) * B * dims[L + 1]);
    }
    orc_mlp_run(X, B, nl, dims, W, b, z, a);
    memcpy(Y, a[nl - 1], sizeof(double) * B * dims[nl]);
    for (int L = 0; L < nl; L++) {
        free(z[L]);
        free(a[L]);
    }
}

double orc_mlp_grad(const double *X, const double *T, int B, int nl, const int *dims,
                    const double *W, const double *b, double *dW, double *db)
{
    double *z[16], *a[16];
    size_t woff[17], boff[17];
    woff[0] = boff[0] = 0;
    for (int L = 0; L < nl; L++) {
        z[L] = (double *)malloc(sizeof(double) * B * dims[L + 1]);
        a[L] = (double *)malloc(sizeof(double) * B * dims[L + 1]);
        woff[L + 1] = woff[L] + (size_t)dims[L] * dims[L + 1];
        boff[L + 1] = boff[L] + dims[L + 1];
    }
    orc_mlp_run(X, B, nl, dims, W, b, z, a);
    const int nout = dims[nl];
    const double inv = 1.0 / ((double)B * nout);
    double loss = 0.0;
    double *g = (double *)malloc(sizeof(double) * B * nout); /* dLoss/dz of the current layer */
    for (size_t j = 0; j < (size_t)B * nout; j++) {
        const double d = a[nl - 1][j] - T[j];
        loss += d * d;
        g[j] = 2.0 * d * inv;
    }
    for (int L = nl - 1; L >= 0; L--) {
        const int ni = dims[L], no = dims[L + 1];
        const double *in = L == 0 ? X : a[L - 1];
        for (int o = 0; o < no; o++) {
            double sb = 0.0;
            for (int r = 0; r < B; r++) sb += g[(size_t)r * no + o];
            db[boff[L] + o] = sb;
            for (int i = 0; i < ni; i++) {
                double sw = 0.0;
                for (int r = 0; r < B; r++) sw += g[(size_t)r * no + o] * in[(size_t)r * ni + i];
                dW[woff[L] + (size_t)o * ni + i] = sw;
            }
        }
        if (L > 0) {
            double *gp = (double *)malloc(sizeof(double) * B * ni);
            for (int r = 0; r < B; r++)
                for (int i = 0; i < ni; i++) {
                    double acc = 0.0;
                    for (int o = 0; o < no; o++)
                        acc += g[(size_t)r * no + o] * W[woff[L] + (size_t)o * ni + i];
                    double dg;
                    orc_gelu(z[L - 1][(size_t)r * ni + i], &dg);
                    gp[(size_t)r * ni + i] = acc * dg;
                }
            free(g);
            g = gp;
        }
    }
    free(g);
    for (int L = 0; L < nl; L++) {
        free(z[L]);
        free(a[L]);
    }
    return loss * inv;
}

double orc_rel_err(const double *x, const double *ref, size_t n)
{
    long double num = 0.0L, den = 0.0L;
    for (size_t j = 0; j < n; j++) {
        long double d = (long double)x[j] - ref[j];
        num += d * d;
        den += (long double)ref[j] * ref[j];
    }
    return den > 0 ? (double)sqrtl(num / den) : (double)sqrtl(num);
}

#ifdef _OPENMP
#include <omp.h>
#endif
/* number of OpenMP threads the oracle's parallel loops use */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
