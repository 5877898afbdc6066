/*
 * oracle.h -- CPU restatement of VCSBeam's voltage-beamforming arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product library never links or calls it.
 *
 * Every function restates one step of the reference (MWA-vcsbeam) and cites
 * the file:line it follows.  Integer steps are exact; floating-point steps are
 * carried out in FP64 ("ideal" arithmetic), which is the yardstick the GPU
 * kernels are measured against.
 *
 * Parity pins (what checks each function against something other than its
 * own formula) are listed in the tests; see tests/test_oracle_*.py.
 */
#ifndef VCSBEAM_ORACLE_H
#define VCSBEAM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* a1+a2: integer weighted overlap-add on MWAX-layout char2 voltages.
 *   htr : (re,im) int8 pairs, MWAX layout [block][ninput][64000] (vMWAX_IDX,
 *         include/vcsbeam.h.in:238-240); block 0 is the one-block margin.
 *   h   : ncoeffs = K*P integer taps (already cast to int, src/pfb.cpp:459-460)
 *   b   : out int32 (re,im) pairs, layout [m][i][n] (src/pfb.cpp:161)
 * b[m][i][n] = sum_p h[K*P - K*p - n - 1] * x_i[(m + 500 - P)*K + p*K + n]
 * (src/pfb.cpp:121-179).  Accumulates in int64 and wraps to int32 like the
 * reference's int atomicAdd. */
void orc_wola(const int8_t *htr, int ninput, int nspectra, int K, int P,
              const int32_t *h, int32_t *b);

/* a3: FPGA rounding y = (X>0 ? X+0x2000 : X) >> 14, as float
 * (src/pfb.cpp:202-242). n = number of int32 values. */
void orc_fpga_round(const int32_t *b, size_t n, float *out);

/* a3: non-FPGA demotion y = (float)X * 2^-14 (src/pfb.cpp:256-268). */
void orc_int2float(const int32_t *b, size_t n, float *out);

/* a4: unnormalised forward DFT Y[k] = sum_n y[n] e^{-2 pi i k n / K} in FP64,
 * direct O(K^2) evaluation (src/pfb.cpp:478-493, 629-649 use cuFFT/hipFFT
 * GPUFFT_FORWARD).  in/out: K complex (interleaved re,im). */
void orc_dft(const double *in, double *out, int K);

/* a4 (fast path for the CPU baseline): radix-2 iterative FFT, FP64, K power
 * of two.  Same definition as orc_dft. */
void orc_fft(const double *in, double *out, int K);

/* a1..a5 combined: the whole forward fine PFB in ideal arithmetic.
 *   fpga    : 1 -> FPGA rounding (PFB_EMULATE_FPGA), 0 -> int2float
 *   imag_first: 1 -> PFB_IMAG_PART_FIRST set (store (re,im)); 0 -> default
 *               (store (im,re)), src/pfb.cpp:326-340
 *   out_idx : i_output_idx[mwax input] = vcs order (src/pfb.cpp:404-421)
 *   vout    : complex double [m][k'][vcs(i)] (v_IDX, include/vcsbeam.h.in:234)
 *             k' = (k + K/2) % K, value = Y/K (swapped unless imag_first)
 *   use_fft : 0 -> direct DFT, 1 -> radix-2 FFT (both FP64) */
void orc_pfb_forward(const int8_t *htr, int ninput, int nspectra, int K, int P,
                     const int32_t *h, const int32_t *out_idx, int fpga,
                     int imag_first, int use_fft, double *vout);

/* a5: PACK_NIBBLES / DEMOTE (include/vcsbeam.h.in:453-457) applied to the
 * complex double output of orc_pfb_forward: byte = (DEMOTE(first)<<4) +
 * DEMOTE(second) where (first, second) are the stored (x, y) components. */
void orc_pack_int4(const double *vout, size_t nsamples, uint8_t *packed);
uint8_t orc_demote(double x);

/* a8: 4+4-bit legacy unpack (include/vcsbeam.h.in:997-1003): high nibble is
 * the real part, low nibble the imaginary part, two's complement. */
void orc_unpack_int4(uint8_t byte, double *re, double *im);

/* a9: incoherent beam I[s][c] = sum_i |v(s,c,i)|^2 over all inputs
 * (src/form_beam.cpp:60-87), v legacy 4+4 [s][c][i]. */
void orc_incoh(const uint8_t *v, int nsample, int nchan, int ninput, float *I);

/* a13: J^{-1} = inv(norm(D*B)) per pointing/antenna/channel (src/jones.c:97-157)
 *   D [a][c][2][2] complex, B [p][a][2][2] complex, J [p][a][c][2][2] complex */
void orc_calc_J(const double *D, const double *B, int npointing, int nant,
                int nchan, double *J);

/* a14: phi[p][a][c] = exp(i 2 pi f_c (w - L)/c) with w = E uE + N uN +
 * (H - Href) uH (src/geometry.c:40-135). unit: [p][3] = (uE, uN, uH). */
void orc_calc_phi(const double *E, const double *N, const double *H,
                  const double *L, int nant, const double *unit, int npointing,
                  const double *freq_hz, int nchan, double Href, double *phi);

/* Array factor |sum conj(w) psi|^2 / N^2 (src/geometry.c:440-467). */
double orc_array_factor(const double *E, const double *N, const double *H,
                        const double *L, int nant, const double *unit1,
                        const double *unit2, double freq_hz, double Href);

/* a10+a11: tied-array beamformer in FP64 (src/form_beam.cpp:122-371).
 *   datatype 0: v is legacy 4+4 uint8 [s][c][i]
 *   datatype 1: v is complex double [s][c][i]
 *   J [p][a][c][2][2], phi [p][a][c], polQ/polP [a] (src/jones.c:40-64)
 *   S out [p][s][st][c] (C_IDX), e out complex [p][s][c][pol] (B_IDX); either
 *   may be NULL.  Auto-correlations subtracted; V sign as in
 *   src/form_beam.cpp:330-333; invw = 1/num_not_flagged. */
void orc_beamform(const void *v, int datatype, const double *J,
                  const double *phi, const uint32_t *polQ,
                  const uint32_t *polP, int npointing, int nant, int nchan,
                  int nsample, int ninput, double invw, int nstokes,
                  double *S, double *e);

/* a12: renormalise_channels_kernel (src/form_beam.cpp:408-449).
 * S [p][s][st][c] -> Sscaled uint8 same layout, offsets/scales [p][st][c].
 * valid[...] (may be NULL) receives 0 where the reference result is
 * undefined (value/scale + 0.5 >= 256 or scale == 0), 1 elsewhere. */
void orc_renormalise(const float *S, int npointing, int nstep, int nstokes,
                     int nchan, float *offsets, float *scales, uint8_t *Sscaled,
                     uint8_t *valid);

/* a16: inverse PFB (src/pfb.cpp:842-907, 940-1003), direct form in FP64.
 *   in  : complex [nsamples + P][K][npol] (the IPFB input buffer: P samples of
 *         history then the current second)
 *   h   : synthesis filter, F = K*P doubles (already scaled)
 *   out : [nsamples*K][npol][2] (re, im)
 *   mask: [nsamples*K] set to 0 where the reference reads outside its filter
 *         (n < F and n % K != 0, truncating m0), 1 elsewhere. */
void orc_ipfb(const double *in, int nsamples, int K, int P, int npol,
              const double *h, double *out, uint8_t *mask);

/* Ring-buffer side of the IPFB (src/pfb.cpp:950-984): builds the IPFB input
 * buffer from the 2-second ring data_buffer_fine [2*nsamples][K][npol]
 * (complex double) for file_no, zeroing the history when file_no == 0. */
void orc_ipfb_gather(const double *ring, int file_no, int nsamples, int K,
                     int P, int npol, double *in);

/* a20: VDIF requantisation float2int8_trunc + to_offset_binary
 * (src/beam_vdif.c:29-38, 254-260): clamp to [-126,127], rint, ^0x80. */
void orc_vdif_requant(const float *f, size_t n, uint8_t *out);

/* a20: VDIF 1.0 frame header of the VDIF writer (src/beam_vdif.c:196-205
 * through vdifio, restated from the VDIF 1.0 specification) and whole frames
 * (src/beam_vdif.c:52-88: header, requantised payload, next header; hdr is
 * advanced in place, the frame counter wrapping at frame_rate). */
long orc_vdif_epoch_mjd(int e);
void orc_vdif_header(int payload_bytes, int threadid, int bits, int nchan, int iscomplex,
                     const char *station, long mjd, long sec_of_day, uint32_t *w);
void orc_vdif_frames(const float *in, size_t nframes, int payload_bytes, uint32_t *hdr,
                     int frame_rate, uint8_t *out);

/* a19: project-defined analytic dipole-array primary beam (the reference
 * has no analytic beam; its FEE model, hyperbeam + HDF5, is not available,
 * src/primary_beam.c:479-532 is the call site it stands in for).  Tile: 16
 * dual-pol short dipoles on a 4x4 grid (spacing 1.1 m, E = ((d%4)-1.5)*1.1,
 * N = ((d/4)-1.5)*1.1), 0.278 m above a ground plane.  With l = sin za sin az,
 * m = sin za cos az, k = 2 pi f / c:
 *   AF_P = sum_d g[P][d] exp(i k (E_d l + N_d m) - i 2 pi f tau[P][d])
 *   GP   = 2 sin(k h cos za),   N0 = 16 * 2 sin(k h)
 *   X (E-W) dipole: (E_theta, E_phi) = (cos za sin az,  cos az)
 *   Y (N-S) dipole: (E_theta, E_phi) = (cos za cos az, -sin az)
 *   J = GP / N0 [[AF_X Xt, AF_X Xp], [AF_Y Yt, AF_Y Yp]]
 * g: complex dipole gains [2][16] as (re, im); tau: dipole delays [2][16]
 * in seconds.  J: [n][4] complex (J00 J01 J10 J11) as (re, im). */
void orc_dipole_beam(const double *az, const double *za, const double *freq, size_t n,
                     const double *g, const double *tau, double *J);

/* The surrogate's 16 input features of (az, za, f): l, m, cos za, f_n,
 * sin az, cos az, sin/cos(psi_x), sin/cos(3 psi_x), sin/cos(psi_y),
 * sin/cos(3 psi_y), sin/cos(k h cos za), with f_n = (f - 200 MHz)/30 MHz,
 * psi_x = k 0.55 l, psi_y = k 0.55 m (half the dipole spacing). X: [n][16]. */
void orc_mlp_features(const double *az, const double *za, const double *freq, size_t n,
                      double *X);

/* MLP forward / gradient in FP64 (the gradient check of the GPU training
 * kernels).  nl linear layers, dims[0..nl]; layer l has W_l [dims[l+1]][dims[l]]
 * (row-major) and b_l [dims[l+1]], packed one after another in W and b;
 * GELU (tanh form, 0.5 x (1 + tanh(0.7978845608 (x + 0.044715 x^3)))) after
 * every layer but the last.  orc_mlp_forward: Y [B][dims[nl]].
 * orc_mlp_grad: loss = mean over B x dims[nl] of (Y - T)^2; dW, db packed
 * like W, b.  Returns the loss. */
void orc_mlp_forward(const double *X, int B, int nl, const int *dims, const double *W,
                     const double *b, double *Y);
double orc_mlp_grad(const double *X, const double *T, int B, int nl, const int *dims,
                    const double *W, const double *b, double *dW, double *db);

/* Number of OpenMP threads the parallel loops use. */
int orc_num_threads(void);

/* cpu_pipeline.c: the CPU baseline, stage by stage (test infrastructure) */
void orc_set_threads(int n);
int orc_cpu_pfb(const int8_t *htr, int ninput, int nspectra, int K, int P, const int32_t *h,
                const int32_t *out_idx, int fpga, int imag_first, float *vout, double *t);
void orc_cpu_ipfb_table(const double *h, int K, int P, float *ft_re, float *ft_im);
void orc_cpu_ipfb(const float *in, long n0, long nout, int K, int P, int npol,
                  const float *ft_re, const float *ft_im, float *out);
void orc_cpu_vdif(const float *f, size_t n, uint8_t *out);
void orc_cpu_delay(const double *E, const double *N, const double *H, const double *L,
                   int nant, const double *unit, int npointing, const double *freq_hz,
                   int nchan, double Href, double *phi);

/* FP64 norm-wise relative error ||x - ref|| / ||ref|| over n values. */
double orc_rel_err(const double *x, const double *ref, size_t n);

#ifdef __cplusplus
}
#endif
#endif
