// ipfb.hip -- inverse (synthesis) PFB and VDIF requantisation on gfx950.
//
// Reference behaviour: MWA-vcsbeam src/pfb.cpp:842-907 (ipfb_kernel, direct
// form: 12 x 128 complex MACs per output sample), 940-1102 (host), and
// src/beam_vdif.c:29-38, 254-260 (float2int8_trunc, to_offset_binary).
//
// MI355X design: the synthesis sum
//   x[n] = (1/K) sum_m f[n - mK] sum_k X_k[m+P] e^{2 pi i ((k+K/2) n mod K)/K}
// factorises for n = jK + r into a 128-point inverse DFT per fine sample,
//   Z_m[r] = (-1)^r sum_k X_k[m+P] e^{2 pi i k r / K},
// followed by a P-tap polyphase FIR along m:
//   x[jK + r] = (1/K) sum_{q<P} f[qK + r] Z_{j-q}[r].
// The inverse DFT reuses the forward FFT of fft128.h via conj(FFT(conj(x))).
// Work per output sample drops from 4*P*K real MACs to ~4*P + ~21 flops; the
// kernel is HBM-bound (8 B in + 8 B out per output sample per pol).
//
// The reference's truncating division for m0 (src/pfb.cpp:870) selects the
// wrong tap set for n < P*K with n % K != 0 and then reads outside its filter
// table; those outputs are undefined there.  This implementation computes the
// mathematically intended sum for every n.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fft128.h"
#include "vm_internal.h"

using vmfft::cf;

#define HIPCHECK(x)                                                               \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) VM_FATAL("HIP: %s (%s)", hipGetErrorString(e_), #x); \
    } while (0)

#define TJ 32     // output fine samples (of 128 coarse samples each) per workgroup
#define MAXP 16   // largest supported number of synthesis taps
#define ROW 129

static __device__ cf g_tw128i[KCH];

__global__ void k_init_tw128i()
{
    int j = threadIdx.x;
    double s, c;
    sincospi(-2.0 * (double)j / KCH, &s, &c);
    g_tw128i[j] = cf{(float)c, (float)s};
}

static void ensure_twiddles_i()
{
    static int done = 0;
    if (done) return;
    hipLaunchKernelGGL(k_init_tw128i, dim3(1), dim3(KCH), 0, 0);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    done = 1;
}

// in: complex [s_in][k][pol] for s_in in [0, ns + P), element type T
// (double2 for the host-buffer path, float2 for the device path).
// hK: polyphase filter f[qK + r] / K as float, length P*K.
// out: float [n][pol][2], n in [0, ns*K).
// grid (ceil(ns/TJ), npol), block 256.
template <typename T>
__global__ __launch_bounds__(256) void k_ipfb(const T *__restrict__ in, int ns, int P, int npol,
                                              const float *__restrict__ hK,
                                              float *__restrict__ out)
{
    __shared__ cf Z[(TJ + MAXP - 1) * ROW];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * TJ;
    const int pol = blockIdx.y;
    const int nrows = TJ + P - 1; // s_in = j0 + 1 + row
    const int jmax = min(TJ, ns - j0);

    // load conj(X) for rows whose s_in exists
    for (int idx = t; idx < nrows * KCH; idx += 256) {
        int row = idx >> 7, k = idx & 127;
        int s_in = j0 + 1 + row;
        cf v = {0.f, 0.f};
        if (s_in < ns + P) {
            T x = in[((size_t)s_in * KCH + k) * npol + pol];
            v = cf{(float)x.x, -(float)x.y};
        }
        Z[row * ROW + k] = v;
    }
    __syncthreads();

    // inverse DFTs, 32 rows per pass
    const int f = t & 31, l = t >> 5;
    for (int base = 0; base < nrows; base += 32) {
        cf *row = &Z[min(base + f, nrows - 1) * ROW];
        // rows past nrows duplicate the last row's work but never store
        cf keep[2][8];
        vmfft::fft_stage1(row, l, g_tw128i);
        vmfft::fft_stage2(row, l, keep);
        __syncthreads();
        if (base + f < nrows) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int k2 = 0; k2 < 8; k2++) {
                    int r = l + 8 * h + 16 * k2;
                    cf y = keep[h][k2];
                    float sgn = (r & 1) ? -1.f : 1.f; // (-1)^r, conj of FFT
                    row[r] = cf{sgn * y.x, -sgn * y.y};
                }
        }
        __syncthreads();
    }

    // polyphase FIR: out[(j0+jj)K + r] = sum_q hK[qK + r] Z[jj - q + P - 1][r]
    const int r = t & 127;
    float hr[MAXP];
#pragma unroll
    for (int q = 0; q < MAXP; q++) hr[q] = (q < P) ? hK[q * KCH + r] : 0.f;
    for (int jj = t >> 7; jj < jmax; jj += 2) {
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int q = 0; q < MAXP; q++) {
            if (q < P) {
                cf z = Z[(jj - q + P - 1) * ROW + r];
                ar = fmaf(hr[q], z.x, ar);
                ai = fmaf(hr[q], z.y, ai);
            }
        }
        size_t n = (size_t)(j0 + jj) * KCH + r;
        *(float2 *)&out[(n * npol + pol) * 2] = make_float2(ar, ai);
    }
}

// Gather the device-resident IPFB input from the beamformer output d_e
// ([c][s][pol] complex float for pointing 0) and the P-sample history
// ([P][c][pol]); zero history when file_no == 0.
__global__ void k_ipfb_gather(const cf *__restrict__ e, const cf *__restrict__ hist, int ns, int P,
                              int npol, int file_no, cf *__restrict__ in)
{
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)(ns + P) * KCH * npol;
    if (idx >= total) return;
    int pol = (int)(idx % npol);
    int k = (int)((idx / npol) % KCH);
    int s_in = (int)(idx / ((size_t)npol * KCH));
    cf v;
    if (s_in < P) v = (file_no == 0) ? cf{0.f, 0.f} : hist[((size_t)s_in * KCH + k) * npol + pol];
    else v = e[((size_t)k * ns + (s_in - P)) * npol + pol];
    in[idx] = v;
}

// Keep the last P fine samples of this second as the next second's history.
__global__ void k_ipfb_save_hist(const cf *__restrict__ e, int ns, int P, int npol,
                                 cf *__restrict__ hist)
{
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)P * KCH * npol;
    if (idx >= total) return;
    int pol = (int)(idx % npol);
    int k = (int)((idx / npol) % KCH);
    int q = (int)(idx / ((size_t)npol * KCH));
    hist[idx] = e[((size_t)k * ns + (ns - P + q)) * npol + pol];
}

// VDIF requantisation: clamp to [-126, 127], round half to even, int8,
// offset binary (XOR 0x80).  Four values per lane.
__global__ void k_vdif_requant(const float *__restrict__ in, uint8_t *__restrict__ out, size_t n)
{
    size_t j = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (j >= n) return;
    if (j + 4 <= n) {
        float4 v = *(const float4 *)(in + j);
        float a[4] = {v.x, v.y, v.z, v.w};
        uint32_t w = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float x = fminf(a[q], 127.0f);
            x = fmaxf(x, -126.0f);
            int iq = (int)rintf(x);
            w |= (uint32_t)((uint8_t)((int8_t)iq ^ 0x80)) << (8 * q);
        }
        *(uint32_t *)(out + j) = w;
    } else {
        for (size_t q = j; q < n; q++) {
            float x = fmaxf(fminf(in[q], 127.0f), -126.0f);
            out[q] = (uint8_t)((int8_t)(int)rintf(x) ^ 0x80);
        }
    }
}

// =================================================================== host ==
extern "C" {

void malloc_ipfb(struct gpu_ipfb_arrays *g, pfb_filter *filter, int nsamples, int npol,
                 int npointing)
{
    // src/pfb.cpp:1054-1085.  d_in_real is the device staging buffer of the
    // IPFB input (complex double, nsamples + ntaps fine samples); d_ft_real
    // holds the polyphase filter f/K as float.
    ensure_twiddles_i();
    memset(g, 0, sizeof(*g));
    g->ntaps = filter->ntaps;
    if (filter->nchans != KCH) VM_FATAL("malloc_ipfb: only %d channels are supported", KCH);
    if (g->ntaps > MAXP) VM_FATAL("malloc_ipfb: at most %d synthesis taps supported", MAXP);
    size_t in_bytes = (size_t)npointing * (nsamples + g->ntaps) * filter->nchans * npol *
                      sizeof(gpuDoubleComplex);
    g->in_size = (int)in_bytes;
    g->ft_size = filter->ncoeffs * (int)sizeof(float);
    g->out_size = npointing * nsamples * filter->nchans * npol * 2 * (int)sizeof(float);
    HIPCHECK(hipMalloc((void **)&g->d_in_real, in_bytes));
    HIPCHECK(hipMalloc((void **)&g->d_ft_real, g->ft_size));
    HIPCHECK(hipMalloc((void **)&g->d_out, g->out_size));
    g->ft_real = (float *)malloc(g->ft_size);
    if (!g->ft_real) VM_FATAL("out of host memory");
}

void free_ipfb(struct gpu_ipfb_arrays *g)
{
    free(g->ft_real);
    if (g->d_in_real) HIPCHECK(hipFree(g->d_in_real));
    if (g->d_in_imag) HIPCHECK(hipFree(g->d_in_imag));
    if (g->d_ft_real) HIPCHECK(hipFree(g->d_ft_real));
    if (g->d_out) HIPCHECK(hipFree(g->d_out));
    memset(g, 0, sizeof(*g));
}

void cu_load_ipfb_filter(pfb_filter *filter, struct gpu_ipfb_arrays *g)
{
    // src/pfb.cpp:1025-1046 precomputes filter x twiddle products; in the
    // factorised form only f[n]/K is needed (1/K = 2^-7 exactly)
    for (int f = 0; f < filter->ncoeffs; f++)
        g->ft_real[f] = (float)(filter->coeffs[f] / (double)filter->nchans);
    HIPCHECK(hipMemcpy(g->d_ft_real, g->ft_real, g->ft_size, hipMemcpyHostToDevice));
}

void cu_invert_pfb(gpuDoubleComplex *data_buffer_fine, int file_no, int npointing, int nsamples,
                   int nchan, int npol, int sizeof_buffer, struct gpu_ipfb_arrays *g,
                   float *data_buffer_vdif)
{
    // src/pfb.cpp:940-1003: the input is the last ntaps samples of the
    // previous second followed by this second, taken from the 2-second ring
    (void)sizeof_buffer;
    if (npointing > 1) VM_FATAL("PFB inversion currently only supports a single pointing");
    if (nchan != KCH) VM_FATAL("cu_invert_pfb: only %d channels are supported", KCH);
    int P = g->ntaps;
    size_t per_s = (size_t)nchan * npol;
    int start_s = (file_no % 2 == 0 ? 2 * nsamples - P : nsamples - P);
    double2 *din = (double2 *)g->d_in_real;
    // contiguous pieces of the ring: s_in in [0, nsamples + P)
    int s_in = 0;
    while (s_in < nsamples + P) {
        int s = (start_s + s_in) % (2 * nsamples);
        int run = min(nsamples + P - s_in, 2 * nsamples - s);
        HIPCHECK(hipMemcpy(din + (size_t)s_in * per_s, data_buffer_fine + (size_t)s * per_s,
                           (size_t)run * per_s * sizeof(double2), hipMemcpyHostToDevice));
        s_in += run;
    }
    if (file_no == 0) HIPCHECK(hipMemset(din, 0, (size_t)P * per_s * sizeof(double2)));
    dim3 grid((nsamples + TJ - 1) / TJ, npol);
    hipLaunchKernelGGL(k_ipfb<double2>, grid, dim3(256), 0, 0, din, nsamples, P, npol,
                       g->d_ft_real, g->d_out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpy(data_buffer_vdif, g->d_out, g->out_size, hipMemcpyDeviceToHost));
}

// Device-resident path: e straight from the beamformer output.
void vm_ipfb_device_launch(const void *d_in_f32, int ns, int P, int npol, const float *d_hK,
                           float *d_out, hipStream_t stream)
{
    dim3 grid((ns + TJ - 1) / TJ, npol);
    hipLaunchKernelGGL(k_ipfb<float2>, grid, dim3(256), 0, stream, (const float2 *)d_in_f32, ns, P,
                       npol, d_hK, d_out);
    HIPCHECK(hipGetLastError());
}

void vmInvertPFBDevice(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                       float *data_buffer_vdif)
{
    ensure_twiddles_i();
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    int ns = vm->fine_sample_rate, P = g->ntaps, npol = (int)vm->obs_metadata->num_ant_pols;
    size_t in_elems = (size_t)(ns + P) * KCH * npol;
    size_t hist_elems = (size_t)P * KCH * npol;
    size_t need = (in_elems + hist_elems) * sizeof(cf);
    if (d->ipfb_out_bytes < need) {
        if (d->d_ipfb_out) HIPCHECK(hipFree(d->d_ipfb_out));
        HIPCHECK(hipMalloc(&d->d_ipfb_out, need));
        HIPCHECK(hipMemsetAsync(d->d_ipfb_out, 0, need, s)); // same stream as its users
        d->ipfb_out_bytes = need;
    }
    cf *din = (cf *)d->d_ipfb_out;
    cf *hist = din + in_elems;
    vm_timer_start(d, VM_T_IPFB);
    hipLaunchKernelGGL(k_ipfb_gather, dim3((unsigned)((in_elems + 255) / 256)), dim3(256), 0, s,
                       (const cf *)vm->d_e, hist, ns, P, npol, file_no, din);
    HIPCHECK(hipGetLastError());
    vm_ipfb_device_launch(din, ns, P, npol, g->d_ft_real, g->d_out, s);
    hipLaunchKernelGGL(k_ipfb_save_hist, dim3((unsigned)((hist_elems + 255) / 256)), dim3(256), 0,
                       s, (const cf *)vm->d_e, ns, P, npol, hist);
    HIPCHECK(hipGetLastError());
    vm_timer_stop(d, VM_T_IPFB);
    if (data_buffer_vdif)
        HIPCHECK(hipMemcpyAsync(data_buffer_vdif, g->d_out, g->out_size, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
}

void vmRequantVDIFDevice(const float *d_in, uint8_t *d_out, size_t n, gpuStream_t stream)
{
    size_t lanes = (n + 3) / 4;
    vm_timer_start_on(vm_dev_global(), VM_T_VDIF, stream);
    hipLaunchKernelGGL(k_vdif_requant, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, d_in, d_out, n);
    HIPCHECK(hipGetLastError());
    vm_timer_stop_on(vm_dev_global(), VM_T_VDIF, stream);
}

} // extern "C"
