// beamform_mfma.hip -- precision many-beam tied-array beamformer on MFMA.
//
// Reference arithmetic: MWA-vcsbeam src/form_beam.cpp:122-204 and 239-371
// (e = sum_a phi_a J^-1_a v_a, auto-correlation-subtracted Stokes).
//
// With many beams the sum over antennas is a dense complex GEMM per fine
// channel c:
//     E[b, pol][s] = sum_i W_c[b, pol][i] v_c[i][s]
// with i the RF input (VCS order), W = phi * J^-1 routed to the input's
// polarisation, M = 2 x beams, K = 2 x tiles, N = samples.  It runs as a
// real GEMM on v_mfma_f32_32x32x16_{f16,bf16}: the K axis interleaves
// (re, im) of each input (k' = 2i + {0,1}) and each beam owns four output
// rows (Xre, Xim, Yre, Yim), so one MFMA accumulator lane holds a complete
// (Ex, Ey) pair and the Stokes parameters are formed in registers.
//
// Precision form: voltages are converted to FP16 (or BF16) per channel with a
// power-of-two scale; weights are split W = W_hi + W_lo into two FP16 (BF16)
// parts and both products accumulate in FP32.  Phases and Jones inverses are
// produced in FP64 (geometry.hip) and demoted only at the split.  The
// beam-independent auto-correlation sums N_xx, N_yy, N_xy are computed once
// per (channel, sample) in FP32 from the FP32 voltages.
//
// Tiling: workgroup 256 lanes (4 wavefronts) computes 128 rows (32 beams) x
// 128 samples; each wavefront a 64 x 64 block as 2 x 2 MFMA tiles; K steps of
// 16 staged through double-buffered LDS with padded rows (conflict-free
// 16-byte fragment reads).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define HIPCHECK(x)                                                               \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) VM_FATAL("HIP: %s (%s)", hipGetErrorString(e_), #x); \
    } while (0)

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

#define MT 128              // rows per workgroup
#define NT 128              // samples per workgroup
#define KS 16               // k' per step
#define A_ROWB 80           // LDS bytes per A row: hi 32 + lo 32 + pad 16
#define B_ROWB 48           // LDS bytes per B column: 32 + pad 16

struct cf { float x, y; };

template <bool BF16> struct mtype;
template <> struct mtype<false> {
    typedef _Float16 t;
    typedef half8 v8;
    static __device__ __forceinline__ uint16_t bits(float x)
    {
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    }
    static __device__ __forceinline__ float back(uint16_t b)
    {
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ f16v mfma(v8 a, v8 b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct mtype<true> {
    typedef __bf16 t;
    typedef bf8 v8;
    static __device__ __forceinline__ uint16_t bits(float x)
    {
        return __builtin_bit_cast(uint16_t, (__bf16)x);
    }
    static __device__ __forceinline__ float back(uint16_t b)
    {
        return (float)__builtin_bit_cast(__bf16, b);
    }
    static __device__ __forceinline__ f16v mfma(v8 a, v8 b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// ------------------------------------------------------------ weights -----
// Wm layout: [c][kstep][row][32 x 16-bit]: for each k' step of 16, the row's
// 16 hi parts then 16 lo parts (64 bytes).  Rows: 4*beam + {Xre, Xim, Yre, Yim}.
// One lane per (p, a, c) writes the 4 rows x 4 k' (its two inputs) entries.
template <bool BF16>
__global__ void k_mfma_weights(const double2 *__restrict__ J, const double2 *__restrict__ phi,
                               const uint32_t *__restrict__ polQ, const uint32_t *__restrict__ polP,
                               int npointing, int nant, int nchan, int nrows_pad, int nkstep,
                               uint16_t *__restrict__ Wm)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)npointing * nant * nchan;
    if (j >= total) return;
    int c = (int)(j % nchan);
    int a = (int)((j / nchan) % nant);
    int p = (int)(j / ((size_t)nchan * nant));
    double2 f = phi[j];
    const double2 *Jm = J + j * 4;
    double2 w[4]; // phi*J00, phi*J01, phi*J10, phi*J11
#pragma unroll
    for (int q = 0; q < 4; q++)
        w[q] = make_double2(f.x * Jm[q].x - f.y * Jm[q].y, f.x * Jm[q].y + f.y * Jm[q].x);
    const uint32_t in_idx[2] = {polQ[a], polP[a]};
#pragma unroll
    for (int src = 0; src < 2; src++) { // input Q (J_x0) or P (J_x1)
        uint32_t i = in_idx[src];
        double2 wx = w[src], wy = w[2 + src];
        // rows Xre, Xim, Yre, Yim; entries at k' = 2i (times v.re), 2i+1 (times v.im)
        double vals[4][2] = {{wx.x, -wx.y}, {wx.y, wx.x}, {wy.x, -wy.y}, {wy.y, wy.x}};
        int kp0 = 2 * (int)i;
        int kstep = kp0 / KS, kin = kp0 % KS; // kin even, kin+1 in the same step
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int row = 4 * p + r;
            uint16_t *dst = Wm + (((size_t)c * nkstep + kstep) * nrows_pad + row) * 32;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                float hi = (float)vals[r][e];
                uint16_t hb = mtype<BF16>::bits(hi);
                float lo = (float)(vals[r][e] - (double)mtype<BF16>::back(hb));
                dst[kin + e] = hb;
                dst[16 + kin + e] = mtype<BF16>::bits(lo);
            }
        }
    }
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

// ---------------------------------------------- auto-correlation terms ----
// N[c][s] = (Nxx, Nyy, Re Nxy, Im Nxy) with Jf = J^-1 of pointing 0 as float
// [c][a][4] complex; also the per-channel max |re|,|im| (as float bits) for
// the FP16 scale.  grid (ceil(ns/256), nchan), block 256.
__global__ __launch_bounds__(256) void k_nterms(const cf *__restrict__ v, int ns, int ni,
                                                const cf *__restrict__ Jf,
                                                const uint32_t *__restrict__ polQ,
                                                const uint32_t *__restrict__ polP, int nant,
                                                float4 *__restrict__ N, uint32_t *__restrict__ vmax)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    const bool ok = s < ns;
    const cf *vc = v + (size_t)c * ni * ns;
    float nxx = 0.f, nyy = 0.f, xyr = 0.f, xyi = 0.f, mx = 0.f;
    for (int a = 0; a < nant; a++) {
        const cf *w = Jf + ((size_t)c * nant + a) * 4;
        cf w00 = w[0], w01 = w[1], w10 = w[2], w11 = w[3];
        cf q = {0.f, 0.f}, p = {0.f, 0.f};
        if (ok) {
            q = vc[(size_t)polQ[a] * ns + s];
            p = vc[(size_t)polP[a] * ns + s];
        }
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)), fmaxf(fabsf(p.x), fabsf(p.y))));
        float axr = w00.x * q.x - w00.y * q.y + w01.x * p.x - w01.y * p.y;
        float axi = w00.x * q.y + w00.y * q.x + w01.x * p.y + w01.y * p.x;
        float ayr = w10.x * q.x - w10.y * q.y + w11.x * p.x - w11.y * p.y;
        float ayi = w10.x * q.y + w10.y * q.x + w11.x * p.y + w11.y * p.x;
        nxx += axr * axr + axi * axi;
        nyy += ayr * ayr + ayi * ayi;
        xyr += axr * ayr + axi * ayi;
        xyi += axi * ayr - axr * ayi;
    }
    if (ok) N[(size_t)c * ns + s] = make_float4(nxx, nyy, xyr, xyi);
    // wave-level max then one atomic per wavefront
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&vmax[c], __float_as_uint(mx));
}

// J^-1 of pointing 0 as float [c][a][4]
__global__ void k_jf(const double2 *__restrict__ J, int nant, int nchan, cf *__restrict__ Jf)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)nant * nchan) return;
    int c = (int)(j % nchan), a = (int)(j / nchan);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double2 m = J[((size_t)a * nchan + c) * 4 + q];
        Jf[((size_t)c * nant + a) * 4 + q] = cf{(float)m.x, (float)m.y};
    }
}

// per-beam J^-1 equal to pointing 0's? (flags[0] set to 1 on any difference)
__global__ void k_jcheck(const double2 *__restrict__ J, size_t per_p, int npointing,
                         uint32_t *__restrict__ flag)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= per_p * (npointing - 1)) return;
    double2 a = J[per_p + j], b = J[j % per_p];
    if (a.x != b.x || a.y != b.y) flag[0] = 1u;
}

// ------------------------------------------------------------- GEMM -------
// grid (ceil(ns/NT), nrows_pad/MT, nchan), block 256.
// v: VM_FLT [c][i][s]; Wm as above; N [c][s] float4; vmax [c] float bits.
// S: [p][st][c][s_total] (device layout), s offset soffset.
template <bool BF16>
__global__ __launch_bounds__(256) void k_beamform_mfma(
    const cf *__restrict__ v, int ns, int ni, const uint16_t *__restrict__ Wm, int nrows_pad,
    int nkstep, const float4 *__restrict__ N, const uint32_t *__restrict__ vmax, int npointing,
    int nchan, float invw, int nstokes, int soffset, int ns_total, float *__restrict__ S)
{
    typedef mtype<BF16> M;
    __shared__ __attribute__((aligned(16))) uint8_t ldsA[2][MT * A_ROWB];
    __shared__ __attribute__((aligned(16))) uint8_t ldsB[2][NT * B_ROWB];

    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int s0 = blockIdx.x * NT;
    const int r0 = blockIdx.y * MT;
    const int c = blockIdx.z;

    // per-channel power-of-two scale: largest |component| -> [2^13, 2^14)
    float vm = __uint_as_float(vmax[c]);
    int ex = 0;
    if (vm > 0.f) frexpf(vm, &ex);
    const float scale = ldexpf(1.0f, 14 - ex);
    const float unscale2 = ldexpf(1.0f, 2 * (ex - 14));

    const cf *vc = v + (size_t)c * ni * ns;
    const uint16_t *Wc = Wm + (size_t)c * nkstep * nrows_pad * 32;

    // global -> register staging for one k-step
    // A: 128 rows x 64 B = 8 KB: lane t copies 32 B (row t/2, half t%2)
    // B: 8 inputs x 128 samples: lane t takes sample s0 + t%128, inputs 4*(t/128)..+3
    const int a_row = t >> 1, a_half = t & 1;
    const int b_col = t & 127, b_grp = t >> 7;
    const bool b_ok = (s0 + b_col) < ns;
    uint4 ra0, ra1;
    cf rb[4];

    auto gload = [&](int ks) {
        const uint4 *src = (const uint4 *)(Wc + ((size_t)ks * nrows_pad + r0 + a_row) * 32) + 2 * a_half;
        ra0 = src[0];
        ra1 = src[1];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int i = ks * (KS / 2) + 4 * b_grp + q;
            rb[q] = b_ok ? vc[(size_t)i * ns + s0 + b_col] : cf{0.f, 0.f};
        }
    };
    auto lstore = [&](int buf) {
        uint4 *da = (uint4 *)(ldsA[buf] + a_row * A_ROWB + a_half * 32);
        da[0] = ra0;
        da[1] = ra1;
        uint32_t pk[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            pk[q] = (uint32_t)M::bits(rb[q].x * scale) | ((uint32_t)M::bits(rb[q].y * scale) << 16);
        *(uint4 *)(ldsB[buf] + b_col * B_ROWB + b_grp * 16) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    };

    f16v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0.f;

    const int wr = w >> 1, wc = w & 1;
    const int lr = lane & 31, lk = lane >> 5;

    gload(0);
    lstore(0);
    __syncthreads();
    for (int ks = 0; ks < nkstep; ks++) {
        const int buf = ks & 1;
        if (ks + 1 < nkstep) gload(ks + 1);
        typename M::v8 Ah[2], Al[2], Bf[2];
#pragma unroll
        for (int rb2 = 0; rb2 < 2; rb2++) {
            const uint8_t *pa = ldsA[buf] + (64 * wr + 32 * rb2 + lr) * A_ROWB + 16 * lk;
            Ah[rb2] = __builtin_bit_cast(typename M::v8, *(const uint4 *)pa);
            Al[rb2] = __builtin_bit_cast(typename M::v8, *(const uint4 *)(pa + 32));
        }
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const uint8_t *pb = ldsB[buf] + (64 * wc + 32 * cb + lr) * B_ROWB + 16 * lk;
            Bf[cb] = __builtin_bit_cast(typename M::v8, *(const uint4 *)pb);
        }
#pragma unroll
        for (int rb2 = 0; rb2 < 2; rb2++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                acc[rb2][cb] = M::mfma(Ah[rb2], Bf[cb], acc[rb2][cb]);
                acc[rb2][cb] = M::mfma(Al[rb2], Bf[cb], acc[rb2][cb]);
            }
        if (ks + 1 < nkstep) lstore(buf ^ 1);
        __syncthreads();
    }

    // epilogue: lane holds rows 8*jg + 4*lk + {0..3} = beam's (Xre, Xim, Yre, Yim)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const int s = s0 + 64 * wc + 32 * cb + lr;
        if (s >= ns) continue;
        const float4 n = N[(size_t)c * ns + s];
        const int so = s + soffset;
#pragma unroll
        for (int rb2 = 0; rb2 < 2; rb2++)
#pragma unroll
            for (int jg = 0; jg < 4; jg++) {
                const int row = r0 + 64 * wr + 32 * rb2 + 8 * jg + 4 * lk;
                const int p = row >> 2;
                if (p >= npointing) continue;
                float xr = acc[rb2][cb][4 * jg + 0], xi = acc[rb2][cb][4 * jg + 1];
                float yr = acc[rb2][cb][4 * jg + 2], yi = acc[rb2][cb][4 * jg + 3];
                // |E|^2 terms carry scale^2; N terms are unscaled
                float bnXX = (xr * xr + xi * xi) * unscale2 - n.x;
                float bnYY = (yr * yr + yi * yi) * unscale2 - n.y;
                float xyr = (xr * yr + xi * yi) * unscale2 - n.z;
                float xyi = (xi * yr - xr * yi) * unscale2 - n.w;
                float *Sp = S + (size_t)p * nstokes * nchan * ns_total;
                Sp[(size_t)c * ns_total + so] = invw * (bnXX + bnYY);
                if (nstokes == 4) {
                    Sp[((size_t)1 * nchan + c) * ns_total + so] = invw * (bnXX - bnYY);
                    Sp[((size_t)2 * nchan + c) * ns_total + so] = 2.0f * invw * xyr;
                    Sp[((size_t)3 * nchan + c) * ns_total + so] = -2.0f * invw * xyi;
                }
            }
    }
}

// =================================================================== host ==
extern "C" {

static void ensure_buf(void **p, size_t *cap, size_t need)
{
    if (*cap >= need && *p) return;
    if (*p) HIPCHECK(hipFree(*p));
    HIPCHECK(hipMalloc(p, need));
    *cap = need;
}

void vmBeamformChunkMFMA(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const bool bf16 = vm->beamform_form == VM_BF_MFMA_BF16;
    if (vm->datatype != VM_FLT)
        VM_FATAL("MFMA beamformer needs VM_FLT voltages (forward PFB output); got %d",
                 (int)vm->datatype);
    const int nant = (int)vm->obs_metadata->num_ants, ni = (int)vm->obs_metadata->num_rf_inputs;
    const int nchan = vm->nfine_chan, np = (int)vm->npointing;
    if (ni % (KS / 2) != 0) VM_FATAL("MFMA beamformer: RF inputs (%d) must be a multiple of 8", ni);
    const int nkstep = 2 * ni / KS;
    const int nrows_pad = ((4 * np + MT - 1) / MT) * MT;
    const int chunk = vm->chunk_to_load % vm->chunks_per_second;
    const int ns = vm->fine_sample_rate / vm->chunks_per_second;

    // scratch: Wm | N | vmax | Jf | jflag (allocated once, reused)
    size_t wm_bytes = (size_t)nchan * nkstep * nrows_pad * 32 * sizeof(uint16_t);
    size_t n_bytes = (size_t)nchan * ns * sizeof(float4);
    size_t vmax_bytes = 256 * ((nchan * 4 + 255) / 256);
    size_t jf_bytes = (size_t)nchan * nant * 4 * sizeof(cf);
    size_t need = wm_bytes + n_bytes + vmax_bytes + jf_bytes + 256;
    if (d->Wm_bytes < need) {
        ensure_buf(&d->d_Wm, &d->Wm_bytes, need);
        d->weights_dirty = 1;
    }
    uint8_t *base = (uint8_t *)d->d_Wm;
    uint16_t *Wm = (uint16_t *)base;
    float4 *N = (float4 *)(base + wm_bytes);
    uint32_t *vmax = (uint32_t *)(base + wm_bytes + n_bytes);
    cf *Jf = (cf *)(base + wm_bytes + n_bytes + vmax_bytes);
    uint32_t *jflag = (uint32_t *)(base + wm_bytes + n_bytes + vmax_bytes + jf_bytes);

    if (d->weights_dirty) {
        // beam-independent J^-1 is required for the shared N terms
        size_t per_p = (size_t)nant * nchan * 4;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, jflag, (size_t)1, 0u);
        if (np > 1)
            hipLaunchKernelGGL(k_jcheck, dim3((unsigned)((per_p * (np - 1) + 255) / 256)), dim3(256),
                               0, st, (const double2 *)vm->d_J, per_p, np, jflag);
        uint32_t hflag = 0;
        HIPCHECK(hipMemcpyAsync(&hflag, jflag, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (hflag)
            VM_FATAL("MFMA beamformer: J^-1 differs between pointings (per-beam primary beam); "
                     "the shared auto-correlation terms need a beam-independent J^-1");
        size_t wm_u32 = wm_bytes / 4;
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((wm_u32 + 255) / 256)), dim3(256), 0, st,
                           (uint32_t *)Wm, wm_u32, 0u);
        size_t total = (size_t)np * nant * nchan;
        if (bf16)
            hipLaunchKernelGGL(k_mfma_weights<true>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                               0, st, (const double2 *)vm->d_J, (const double2 *)vm->gdelays.d_phi,
                               vm->d_polQ_idxs, vm->d_polP_idxs, np, nant, nchan, nrows_pad, nkstep, Wm);
        else
            hipLaunchKernelGGL(k_mfma_weights<false>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                               0, st, (const double2 *)vm->d_J, (const double2 *)vm->gdelays.d_phi,
                               vm->d_polQ_idxs, vm->d_polP_idxs, np, nant, nchan, nrows_pad, nkstep, Wm);
        hipLaunchKernelGGL(k_jf, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256), 0, st,
                           (const double2 *)vm->d_J, nant, nchan, Jf);
        HIPCHECK(hipGetLastError());
        d->weights_dirty = 0;
    }
    // auto-correlation terms and the per-channel scale of this chunk
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, vmax, (size_t)nchan, 0u);
    vm_timer_start(d, VM_T_NTERMS);
    hipLaunchKernelGGL(k_nterms, dim3((ns + 255) / 256, nchan), dim3(256), 0, st,
                       (const cf *)vm->d_v, ns, ni, (const cf *)Jf, vm->d_polQ_idxs,
                       vm->d_polP_idxs, nant, N, vmax);
    HIPCHECK(hipGetLastError());
    vm_timer_stop(d, VM_T_NTERMS);
    float invw = (float)(1.0 / (double)vm->num_not_flagged);
    dim3 grid((ns + NT - 1) / NT, nrows_pad / MT, nchan);
    vm_timer_start(d, VM_T_BF_MFMA);
    if (bf16)
        hipLaunchKernelGGL(k_beamform_mfma<true>, grid, dim3(256), 0, st, (const cf *)vm->d_v, ns,
                           ni, Wm, nrows_pad, nkstep, N, vmax, np, nchan, invw, vm->out_nstokes,
                           chunk * ns, vm->fine_sample_rate, (float *)vm->d_S);
    else
        hipLaunchKernelGGL(k_beamform_mfma<false>, grid, dim3(256), 0, st, (const cf *)vm->d_v, ns,
                           ni, Wm, nrows_pad, nkstep, N, vmax, np, nchan, invw, vm->out_nstokes,
                           chunk * ns, vm->fine_sample_rate, (float *)vm->d_S);
    HIPCHECK(hipGetLastError());
    vm_timer_stop(d, VM_T_BF_MFMA);
    HIPCHECK(hipStreamSynchronize(st));
}

} // extern "C"
