// beamform.hip -- tied-array beamformer (plain FP32 form), incoherent beam,
// Stokes renormalisation and the host-layout pulls.
//
// Reference behaviour: MWA-vcsbeam src/form_beam.cpp:60-87 (incoh_beam),
// 122-204 (vmApplyJ_kernel), 239-371 (vmBeamform_kernel), 408-449
// (renormalise_channels_kernel), 463-804 (host).
//
// MI355X design:
//  * J^-1 v and phi * (J^-1 v) are fused: vmApplyJChunk folds the delay phase
//    into the Jones inverse once per second, W[p][c][a] = phi[p][a][c] J[p][a][c]
//    (products in FP64, stored as complex float); vmBeamformChunk then does
//    e = sum_a W v, the auto-correlation sums and Stokes in one pass, with no
//    Jv_Q / Jv_P intermediates (5.2 GB per beam-second in the reference).
//  * One lane per (channel, sample), PB beams per lane; the antenna loop reads
//    weights with wave-uniform (scalar) loads and voltages with coalesced loads
//    along the sample axis of the VM_FLT layout.
//  * Device outputs use sample-innermost layouts (d_S in 64-sample blocks
//    [c][s/64][p][st][s%64], d_e [p][c][s][pol]) so stores are coalesced;
//    vmPullS / vmPullE transpose to the reference host layouts.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define HIPCHECK(x)                                                               \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) VM_FATAL("HIP: %s (%s)", hipGetErrorString(e_), #x); \
    } while (0)

struct cf { float x, y; };

__device__ __forceinline__ void unpack4(uint8_t b, float &re, float &im)
{
    // include/vcsbeam.h.in:997-1003: high nibble real, low nibble imaginary
    int r = (b >> 4) & 0xf, i = b & 0xf;
    re = (float)(r >= 8 ? r - 16 : r);
    im = (float)(i >= 8 ? i - 16 : i);
}

// ------------------------------------------------------------ weights -----
// W[p][c][a][q] = phi[p][a][c] * J[p][a][c][q] (FP64 product, float result)
__global__ void k_build_weights(const double2 *__restrict__ J, const double2 *__restrict__ phi,
                                int npointing, int nant, int nchan, cf *__restrict__ W)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)npointing * nant * nchan;
    if (j >= total) return;
    int c = (int)(j % nchan);
    int a = (int)((j / nchan) % nant);
    int p = (int)(j / ((size_t)nchan * nant));
    double2 f = phi[j]; // PHI_IDX(p,a,c) == j
    const double2 *Jm = J + j * 4; // J_IDX(p,a,c,0,0)
    cf *out = W + (((size_t)p * nchan + c) * nant + a) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double2 m = Jm[q];
        out[q] = cf{(float)(f.x * m.x - f.y * m.y), (float)(f.x * m.y + f.y * m.x)};
    }
}

// ------------------------------------------------------ FP32 beamformer ----
// grid (ceil(ns/256), nchan, ceil(npointing/PB)), block 256.
// VM_FLT voltages v[c][i][s] (complex float) or VM_INT4 legacy [s][c][i].
template <int PB, bool INT4>
__global__ __launch_bounds__(256) void k_beamform_fp32(
    const void *__restrict__ v, int ns, int nchan, int ni, const cf *__restrict__ W,
    const uint32_t *__restrict__ polQ, const uint32_t *__restrict__ polP, int nant,
    int npointing, float invw, int nstokes, int soffset, int ns_total, float *__restrict__ S,
    cf *__restrict__ e)
{
    extern __shared__ uint8_t lds[];
    const int t = threadIdx.x;
    const int s = blockIdx.x * 256 + t;
    const int c = blockIdx.y;
    const int p0 = blockIdx.z * PB;
    const bool valid = s < ns;

    if (INT4) {
        // stage this block's 256 samples x ni bytes, coalesced per (s, c) row
        const uint8_t *vb = (const uint8_t *)v;
        const int s0 = blockIdx.x * 256;
        const int nrow = min(256, ns - s0);
        const int dw = ni / 4;
        for (int j = t; j < nrow * dw; j += 256) {
            int r = j / dw, q = j % dw;
            ((uint32_t *)lds)[r * dw + q] =
                ((const uint32_t *)(vb + ((size_t)(s0 + r) * nchan + c) * ni))[q];
        }
        __syncthreads();
    }

    float ex[PB][2], ey[PB][2], nxx[PB], nyy[PB], nxy[PB][2];
#pragma unroll
    for (int b = 0; b < PB; b++) {
        ex[b][0] = ex[b][1] = ey[b][0] = ey[b][1] = 0.f;
        nxx[b] = nyy[b] = nxy[b][0] = nxy[b][1] = 0.f;
    }
    const cf *vf = (const cf *)v + (size_t)c * ni * ns;
    for (int a = 0; a < nant; a++) {
        const uint32_t iq = polQ[a], ip = polP[a];
        float qr = 0.f, qi = 0.f, pr = 0.f, pi = 0.f;
        if (valid) {
            if (INT4) {
                unpack4(lds[t * ni + iq], qr, qi);
                unpack4(lds[t * ni + ip], pr, pi);
            } else {
                cf q = vf[(size_t)iq * ns + s], pp = vf[(size_t)ip * ns + s];
                qr = q.x; qi = q.y; pr = pp.x; pi = pp.y;
            }
        }
#pragma unroll
        for (int b = 0; b < PB; b++) {
            int p = min(p0 + b, npointing - 1);
            const cf *w = W + (((size_t)p * nchan + c) * nant + a) * 4;
            cf w00 = w[0], w01 = w[1], w10 = w[2], w11 = w[3];
            // ex_a = W00 vq + W01 vp ; ey_a = W10 vq + W11 vp
            float axr = w00.x * qr - w00.y * qi + w01.x * pr - w01.y * pi;
            float axi = w00.x * qi + w00.y * qr + w01.x * pi + w01.y * pr;
            float ayr = w10.x * qr - w10.y * qi + w11.x * pr - w11.y * pi;
            float ayi = w10.x * qi + w10.y * qr + w11.x * pi + w11.y * pr;
            ex[b][0] += axr; ex[b][1] += axi;
            ey[b][0] += ayr; ey[b][1] += ayi;
            nxx[b] += axr * axr + axi * axi;
            nyy[b] += ayr * ayr + ayi * ayi;
            // ex * conj(ey)
            nxy[b][0] += axr * ayr + axi * ayi;
            nxy[b][1] += axi * ayr - axr * ayi;
        }
    }
    if (!valid) return;
    const int so = s + soffset;
#pragma unroll
    for (int b = 0; b < PB; b++) {
        int p = p0 + b;
        if (p >= npointing) break;
        // src/form_beam.cpp:322-339 (auto-correlations removed; V sign as code)
        float bnXX = ex[b][0] * ex[b][0] + ex[b][1] * ex[b][1] - nxx[b];
        float bnYY = ey[b][0] * ey[b][0] + ey[b][1] * ey[b][1] - nyy[b];
        float xyr = ex[b][0] * ey[b][0] + ex[b][1] * ey[b][1] - nxy[b][0];
        float xyi = ex[b][1] * ey[b][0] - ex[b][0] * ey[b][1] - nxy[b][1];
        float *Sp = S + VM_S_IDX(p, 0, c, so, npointing, nstokes, VM_S_NBLK(ns_total));
        Sp[0] = invw * (bnXX + bnYY);
        if (nstokes == 4) {
            Sp[VM_SBLK] = invw * (bnXX - bnYY);
            Sp[2 * VM_SBLK] = 2.0f * invw * xyr;
            Sp[3 * VM_SBLK] = -2.0f * invw * xyi;
        }
        if (e) {
            cf *ep = e + (((size_t)p * nchan + c) * ns_total + so) * 2;
            ep[0] = cf{ex[b][0], ex[b][1]};
            ep[1] = cf{ey[b][0], ey[b][1]};
        }
    }
}

// ----------------------------------------------------- incoherent beam ----
// I[s][c] = sum_i |v|^2 over all inputs, exact integer arithmetic
// (src/form_beam.cpp:60-87).  16 lanes per (s, c) row, 16 bytes per lane
// per load, so one wavefront load covers 4 rows = 1 KB.  With the other
// nibble masked off, a byte is (16 re) or (16 im) as a signed int8, so
// v_dot4_i32_i8 of the masked word with itself gives 256 sum(re^2) (resp.
// im^2) over its 4 bytes, exactly; the row sum is shifted down at the end.
__device__ __forceinline__ int incoh_sq4(uint32_t w)
{
    const uint32_t hi = w & 0xF0F0F0F0u, lo = (w << 4) & 0xF0F0F0F0u;
    return __builtin_amdgcn_sdot4((int)hi, (int)hi, __builtin_amdgcn_sdot4((int)lo, (int)lo, 0, false),
                                  false);
}

// Rows are taken 64 at a time per workgroup: 4 per 16-lane group, each
// lane with its 4 row loads in flight together.
__global__ __launch_bounds__(256) void k_incoh(const uint8_t *__restrict__ data, int nsample,
                                               int nchan, int ninput, float *__restrict__ incoh)
{
    const int g = threadIdx.x & 15, rl = threadIdx.x >> 4;
    const size_t nrows = (size_t)nsample * nchan;
    const size_t stride = (size_t)gridDim.x * 64;
    for (size_t r0 = (size_t)blockIdx.x * 64 + rl; r0 < nrows; r0 += stride) {
        int acc[4] = {0, 0, 0, 0}; // 256 x each row's sum
        if ((ninput & 15) == 0) {
            for (int j = 16 * g; j < ninput; j += 256) {
                uint4 w[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const size_t row = r0 + 16 * k;
                    w[k] = row < nrows ? *(const uint4 *)(data + row * ninput + j)
                                       : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int k = 0; k < 4; k++)
                    acc[k] += incoh_sq4(w[k].x) + incoh_sq4(w[k].y) + incoh_sq4(w[k].z) +
                              incoh_sq4(w[k].w);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const size_t row = r0 + 16 * k;
                if (row < nrows)
                    for (int j = g; j < ninput; j += 16)
                        acc[k] += incoh_sq4((uint32_t)data[row * ninput + j]);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 16);
            const size_t row = r0 + 16 * k;
            if (g == 0 && row < nrows) incoh[row] = (float)(acc[k] >> 8);
        }
    }
}

// --------------------------------------------------------- renormalise ----
// 8-bit renormalisation of each channel (src/form_beam.cpp:408-449): over
// the samples of a channel, scale = (max - min) / 256, offset = min and
// q = (uint8)((S - offset) / scale + 0.5), the + 0.5 in double as the
// reference has it.  The reference's cast is undefined at S = max (256.5)
// and for scale = 0; those saturate here (NaN -> 0).
__device__ __forceinline__ uint8_t renorm_q(float S, float offset, float scale)
{
    const float val = (S - offset) / scale;
    double r = (double)val + 0.5;
    r = (r != r) ? 0.0 : fmin(fmax(r, 0.0), 255.0);
    return (uint8_t)r;
}

// Incoherent beam I[s][c] (the reference layout, one pointing, Stokes I):
// three short passes that each spread over the whole device.
// 1. partial min/max per (sample split, channel): block 64 channels x 4
//    sample lanes, grid (ceil(nchan / 64), nsplit); split q holds samples
//    [q sps, min((q + 1) sps, nsample)) and writes part[q][c] = (min, max)
//    for every c < nchan, so part needs no clearing.
__global__ __launch_bounds__(256) void k_incoh_minmax(const float *__restrict__ I, int nsample,
                                                      int nchan, int sps, float2 *__restrict__ part)
{
    __shared__ float smn[4][64], smx[4][64];
    const int cl = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cl, q = blockIdx.y;
    const int s0 = q * sps, s1 = min(s0 + sps, nsample);
    float mn = INFINITY, mx = -INFINITY;
    if (c < nchan)
        for (int s = s0 + sl; s < s1; s += 4) {
            const float v = I[(size_t)s * nchan + c];
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
    smn[sl][cl] = mn;
    smx[sl][cl] = mx;
    __syncthreads();
    if (sl == 0 && c < nchan) {
#pragma unroll
        for (int k = 1; k < 4; k++) {
            mn = fminf(mn, smn[k][cl]);
            mx = fmaxf(mx, smx[k][cl]);
        }
        part[(size_t)q * nchan + c] = make_float2(mn, mx);
    }
}

// 2. per channel: the splits' extremes -> offset and scale; block 16
//    channels x 16 lanes, each lane folding every 16th split
__global__ __launch_bounds__(256) void k_incoh_scale(const float2 *__restrict__ part, int nsplit,
                                                     int nchan, float *__restrict__ offsets,
                                                     float *__restrict__ scales)
{
    __shared__ float smn[16][17], smx[16][17];
    const int cl = threadIdx.x & 15, ql = threadIdx.x >> 4;
    const int c = blockIdx.x * 16 + cl;
    float mn = INFINITY, mx = -INFINITY;
    if (c < nchan) {
        int q = ql;
        for (; q + 48 < nsplit; q += 64) { // four loads in flight
            const float2 m0 = part[(size_t)q * nchan + c], m1 = part[(size_t)(q + 16) * nchan + c];
            const float2 m2 = part[(size_t)(q + 32) * nchan + c];
            const float2 m3 = part[(size_t)(q + 48) * nchan + c];
            mn = fminf(mn, fminf(fminf(m0.x, m1.x), fminf(m2.x, m3.x)));
            mx = fmaxf(mx, fmaxf(fmaxf(m0.y, m1.y), fmaxf(m2.y, m3.y)));
        }
        for (; q < nsplit; q += 16) {
            const float2 m = part[(size_t)q * nchan + c];
            mn = fminf(mn, m.x);
            mx = fmaxf(mx, m.y);
        }
    }
    smn[ql][cl] = mn;
    smx[ql][cl] = mx;
    __syncthreads();
    if (ql == 0 && c < nchan) {
        for (int k = 1; k < 16; k++) {
            mn = fminf(mn, smn[k][cl]);
            mx = fmaxf(mx, smx[k][cl]);
        }
        scales[c] = (mx - mn) / 256.0f;
        offsets[c] = mn;
    }
}

// 3. quantise, 4 consecutive elements per lane (one 4-byte store)
__global__ __launch_bounds__(256) void k_incoh_quant(const float *__restrict__ I, size_t n,
                                                     int nchan, const float *__restrict__ offsets,
                                                     const float *__restrict__ scales,
                                                     uint8_t *__restrict__ Iq)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
    for (size_t j = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; j < n; j += stride) {
        uint8_t q[4] = {0, 0, 0, 0};
        const int m = (int)min((size_t)4, n - j);
        int c = (int)(j % (size_t)nchan); // one division per 4 elements
        for (int k = 0; k < m; k++) {
            q[k] = renorm_q(I[j + k], offsets[c], scales[c]);
            if (++c == nchan) c = 0;
        }
        if (m == 4 && ((uintptr_t)(Iq + j) & 3) == 0)
            *(uchar4 *)(Iq + j) = make_uchar4(q[0], q[1], q[2], q[3]);
        else
            for (int k = 0; k < m; k++) Iq[j + k] = q[k];
    }
}

// Device layout d_S (VM_S_IDX): min/max per (p, st, c), one workgroup each.
// Results go to row (p*nstokes + st)*ctot + coff + c of offsets/scales: ctot =
// nchan, coff = 0 for one coarse channel; a spliced band of ctot channels
// takes this channel's nchan at coff.
__global__ __launch_bounds__(256) void k_minmax_dev(const float *__restrict__ S, int nstep,
                                                    int npointing, int nstokes, int nchan,
                                                    int ctot, int coff,
                                                    float *__restrict__ offsets,
                                                    float *__restrict__ scales)
{
    __shared__ float smin[256], smax[256];
    const size_t row = blockIdx.x; // (p*nstokes + st)*nchan + c
    const int c = (int)(row % nchan), st = (int)(row / nchan % nstokes);
    const int p = (int)(row / nchan / nstokes);
    const int nblk = VM_S_NBLK(nstep);
    const float *r0 = S + VM_S_IDX(p, st, c, 0, npointing, nstokes, nblk);
    const size_t bstride = (size_t)npointing * nstokes * VM_SBLK; // next sample block
    float mn = r0[0], mx = r0[0];
    // lane: 4 consecutive samples (q = lane % 16) of block b = lane / 16 +
    // 16 k; four blocks' loads in flight per lane (the row is 40 KB at most
    // 10 000 samples, so one workgroup has it all in flight in a few rounds)
    const int q = threadIdx.x & 15, b0 = threadIdx.x >> 4;
    const int nfull = nstep / VM_SBLK; // blocks with 64 samples
    int b = b0;
    for (; b + 48 < nfull; b += 64) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            v[u] = *(const float4 *)(r0 + (size_t)(b + 16 * u) * bstride + 4 * q);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            mn = fminf(mn, fminf(fminf(v[u].x, v[u].y), fminf(v[u].z, v[u].w)));
            mx = fmaxf(mx, fmaxf(fmaxf(v[u].x, v[u].y), fmaxf(v[u].z, v[u].w)));
        }
    }
    for (; b < nfull; b += 16) {
        const float4 v = *(const float4 *)(r0 + (size_t)b * bstride + 4 * q);
        mn = fminf(mn, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
        mx = fmaxf(mx, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
    }
    // the last, partial block sample by sample
    for (int s = nfull * VM_SBLK + threadIdx.x; s < nstep; s += 256) {
        const float val = r0[(size_t)nfull * bstride + s % VM_SBLK];
        mn = fminf(mn, val);
        mx = fmaxf(mx, val);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t orow = (size_t)(p * nstokes + st) * ctot + coff + c;
        scales[orow] = (smax[0] - smin[0]) / 256.0f;
        offsets[orow] = smin[0];
    }
}

// Quantise + transpose: d_S (VM_S_IDX) -> Ss[p][s][st][ctot] (uint8), this
// channel's nchan at coff (offsets/scales rows as in k_minmax_dev).
// block 256, tile 64 channels x 64 samples (one sample block of S, so each
// channel's 64 samples are one contiguous 256-byte row); grid
// (ceil(ns/64), ceil(nc/64), p*nst).  Lanes read 4 samples of a channel row
// (16 lanes a row) and write 4 channels of an output row (16 lanes a row).
__global__ __launch_bounds__(256) void k_quant_dev(const float *__restrict__ S, int nstep,
                                                   int npointing, int nstokes, int nchan,
                                                   int ctot, int coff,
                                                   const float *__restrict__ offsets,
                                                   const float *__restrict__ scales,
                                                   uint8_t *__restrict__ Ss)
{
    __shared__ __attribute__((aligned(16))) uint8_t T[64][68]; // [channel][sample], rows 4-byte aligned
    const int pst = blockIdx.z, p = pst / nstokes, st = pst % nstokes;
    const int c0 = blockIdx.y * 64, s0 = blockIdx.x * 64;
    const int l4 = threadIdx.x & 15, ry = threadIdx.x >> 4;
    // row c of the tile: S + VM_S_IDX(p, st, c, s0) + (c - c0) * cstr
    const size_t cstr = (size_t)VM_S_NBLK(nstep) * npointing * nstokes * VM_SBLK;
    const float *base = S + VM_S_IDX(p, st, c0, s0, npointing, nstokes, VM_S_NBLK(nstep));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = ry + 16 * k, c = c0 + r;
        uint32_t w = 0;
        if (c < nchan) {
            // samples past nstep are the padding of the last block: read, not stored
            const float4 v = *(const float4 *)(base + (size_t)r * cstr + 4 * l4);
            const size_t row = (size_t)pst * ctot + coff + c;
            const float off = offsets[row], sc = scales[row];
            w = (uint32_t)renorm_q(v.x, off, sc) | ((uint32_t)renorm_q(v.y, off, sc) << 8) |
                ((uint32_t)renorm_q(v.z, off, sc) << 16) | ((uint32_t)renorm_q(v.w, off, sc) << 24);
        }
        *(uint32_t *)&T[r][4 * l4] = w;
    }
    __syncthreads();
    const bool vec = ((ctot | coff) & 3) == 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = ry + 16 * k, s = s0 + r; // output row: sample s, channels c0 + 4 l4 ..
        if (s >= nstep) continue;
        const int c = c0 + 4 * l4;
        uint8_t *dst = Ss + (((size_t)p * nstep + s) * nstokes + st) * ctot + coff + c;
        if (vec && c + 3 < nchan) {
            *(uchar4 *)dst = make_uchar4(T[c - c0][r], T[c - c0 + 1][r], T[c - c0 + 2][r],
                                         T[c - c0 + 3][r]);
        } else {
            for (int e = 0; e < 4; e++)
                if (c + e < nchan) dst[e] = T[c - c0 + e][r];
        }
    }
}

// ------------------------------------------------------------ pulls -------
// d_S (VM_S_IDX) -> [p][s][st][c] float
__global__ __launch_bounds__(256) void k_S_to_host(const float *__restrict__ S, int nstep,
                                                   int npointing, int nstokes, int nchan,
                                                   float *__restrict__ out)
{
    __shared__ float T[64][65];
    const int pst = blockIdx.z, p = pst / nstokes, st = pst % nstokes;
    const int c0 = blockIdx.y * 64, s0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        int c = c0 + r, s = s0 + tx;
        T[r][tx] = (c < nchan && s < nstep)
                       ? S[VM_S_IDX(p, st, c, s, npointing, nstokes, VM_S_NBLK(nstep))]
                       : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        int s = s0 + r, c = c0 + tx;
        if (s < nstep && c < nchan) out[(((size_t)p * nstep + s) * nstokes + st) * nchan + c] = T[tx][r];
    }
}

// e[p][c][s][pol] complex float -> [p][s][c][pol] complex double
__global__ __launch_bounds__(256) void k_e_to_host(const cf *__restrict__ e, int nstep, int nchan,
                                                   double2 *__restrict__ out)
{
    __shared__ cf T[64][2 * 64 + 1];
    const int p = blockIdx.z;
    const int c0 = blockIdx.y * 64, s0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        int c = c0 + r, s = s0 + tx;
        bool ok = c < nchan && s < nstep;
        size_t b = (((size_t)p * nchan + c) * nstep + s) * 2;
        T[r][2 * tx] = ok ? e[b] : cf{0.f, 0.f};
        T[r][2 * tx + 1] = ok ? e[b + 1] : cf{0.f, 0.f};
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        int s = s0 + r, c = c0 + tx;
        if (s < nstep && c < nchan) {
            size_t o = (((size_t)p * nstep + s) * nchan + c) * 2;
            cf a = T[tx][2 * r], b = T[tx][2 * r + 1];
            out[o] = make_double2(a.x, a.y);
            out[o + 1] = make_double2(b.x, b.y);
        }
    }
}

// =================================================================== host ==
extern "C" {

static hipStream_t vm_stream(vcsbeam_context *vm) { return (hipStream_t)vm_dev(vm)->stream; }

static void *stage_buffer(vcsbeam_context *vm, size_t bytes)
{
    vm_device_state *d = vm_dev(vm);
    if (d->stage_bytes < bytes) {
        if (d->d_stage) HIPCHECK(hipFree(d->d_stage));
        HIPCHECK(hipMalloc(&d->d_stage, bytes));
        d->stage_bytes = bytes;
    }
    return d->d_stage;
}

// The incoherent beam on device buffers, all on stream st: the detection
// sum, then the three renormalisation passes with their (min, max)
// partials in part (room for cap pairs; at most 256 sample splits, and at
// least nchan pairs are needed).
static void incoh_launch(const uint8_t *d_data, float *d_incoh, unsigned nsample, int nchan,
                         int ninput, float *d_offsets, float *d_scales, uint8_t *d_Iscaled,
                         float2 *part, size_t cap, hipStream_t st)
{
    const size_t rows = (size_t)nsample * nchan;
    size_t nblk = (rows + 63) / 64; // 64 rows per workgroup and pass
    if (nblk > 4096) nblk = 4096;
    hipLaunchKernelGGL(k_incoh, dim3((unsigned)(nblk > 0 ? nblk : 1)), dim3(256), 0, st, d_data,
                       (int)nsample, nchan, ninput, d_incoh);
    if (nsample == 0) return;
    size_t nsplit = cap / (size_t)nchan;
    if (nsplit > 256) nsplit = 256;
    if (nsplit > nsample) nsplit = nsample;
    const int sps = (int)((nsample + nsplit - 1) / nsplit);
    nsplit = (nsample + sps - 1) / sps; // no empty split
    hipLaunchKernelGGL(k_incoh_minmax, dim3((unsigned)((nchan + 63) / 64), (unsigned)nsplit),
                       dim3(256), 0, st, (const float *)d_incoh, (int)nsample, nchan, sps, part);
    hipLaunchKernelGGL(k_incoh_scale, dim3((unsigned)((nchan + 15) / 16)), dim3(256), 0, st,
                       (const float2 *)part, (int)nsplit, nchan, d_offsets, d_scales);
    const size_t nq = (rows + 1023) / 1024;
    hipLaunchKernelGGL(k_incoh_quant, dim3((unsigned)(nq < 8192 ? nq : 8192)), dim3(256), 0, st,
                       (const float *)d_incoh, rows, nchan, (const float *)d_offsets,
                       (const float *)d_scales, d_Iscaled);
}

void cu_form_incoh_beam(uint8_t *data, uint8_t *d_data, size_t data_size, float *d_incoh,
                        unsigned int nsample, int nchan, int ninput, float *offsets,
                        float *d_offsets, float *scales, float *d_scales, uint8_t *Iscaled,
                        uint8_t *d_Iscaled, size_t Iscaled_size)
{
    // src/form_beam.cpp:463-505 (all copies complete before return)
    if (data_size < (size_t)nsample * nchan * ninput || Iscaled_size > (size_t)nsample * nchan)
        VM_FATAL("cu_form_incoh_beam: %zu data bytes for %u x %d x %d samples, %zu scaled bytes",
                 data_size, nsample, nchan, ninput, Iscaled_size);
    HIPCHECK(hipMemcpy(d_data, data, data_size, hipMemcpyHostToDevice));
    if (nchan <= 0 || ninput <= 0) VM_FATAL("cu_form_incoh_beam: %d channels, %d inputs", nchan, ninput);
    // the renormalisation's partial extremes: a library buffer per device,
    // kept between calls and grown when a call needs more
    static void *part[VM_MAX_DEVICES];
    static size_t part_bytes[VM_MAX_DEVICES];
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= VM_MAX_DEVICES) VM_FATAL("cu_form_incoh_beam: device %d", dev);
    const size_t need = (size_t)256 * (size_t)nchan * sizeof(float2);
    if (part_bytes[dev] < need) {
        if (part[dev]) HIPCHECK(hipFree(part[dev]));
        HIPCHECK(hipMalloc(&part[dev], need));
        part_bytes[dev] = need;
    }
    vm_device_state *g = vm_dev_global();
    // the timer covers the whole device part: the sum and the
    // renormalisation
    vm_timer_start_on(g, VM_T_INCOH, NULL);
    incoh_launch(d_data, d_incoh, nsample, nchan, ninput, d_offsets, d_scales, d_Iscaled,
                 (float2 *)part[dev], part_bytes[dev] / sizeof(float2), 0);
    HIPCHECK(hipGetLastError());
    vm_timer_stop_on(g, VM_T_INCOH, NULL);
    HIPCHECK(hipMemcpy(offsets, d_offsets, nchan * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(scales, d_scales, nchan * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(Iscaled, d_Iscaled, Iscaled_size, hipMemcpyDeviceToHost));
}

// ------------------------------------------------ no-allocation form ----
// cu_form_incoh_beam for callers that cannot allocate device memory: the
// renormalisation's partial extremes live in this fixed array, one copy per
// device.  Host code reaches it only through the address that
// hipGetSymbolAddress gives on each device (kept in g_scratch), never
// through the array's name.
static __device__ __align__(256) float2 scratch[VM_INCOH_SCRATCH_PAIRS];
static void *g_scratch[VM_MAX_DEVICES];

// this device's address of scratch, looked up on the first call on the
// device (with that device current); NULL on failure
static float2 *incoh_scratch(int *dev_out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= VM_MAX_DEVICES) return NULL;
    *dev_out = dev;
    void *p = __atomic_load_n(&g_scratch[dev], __ATOMIC_ACQUIRE);
    if (!p) {
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NULL;
        __atomic_store_n(&g_scratch[dev], p, __ATOMIC_RELEASE);
    }
    return (float2 *)p;
}

int vmFormIncohBeamNoAlloc(const uint8_t *d_data, size_t data_size, float *d_incoh,
                           unsigned int nsample, int nchan, int ninput, float *d_offsets,
                           float *d_scales, uint8_t *d_Iscaled, size_t Iscaled_size,
                           gpuStream_t stream)
{
    if (nchan <= 0 || nchan > VM_INCOH_SCRATCH_PAIRS || ninput <= 0) {
        fprintf(stderr, "vmFormIncohBeamNoAlloc: %d channels of %d inputs (at most %d channels)\n",
                nchan, ninput, VM_INCOH_SCRATCH_PAIRS);
        return -1;
    }
    const size_t rows = (size_t)nsample * (size_t)nchan;
    if (data_size < rows * (size_t)ninput || Iscaled_size < rows) {
        fprintf(stderr, "vmFormIncohBeamNoAlloc: %zu data bytes and %zu scaled bytes for %u x %d "
                        "x %d samples\n", data_size, Iscaled_size, nsample, nchan, ninput);
        return -1;
    }
    int dev = -1;
    float2 *part = incoh_scratch(&dev);
    if (!part) {
        fprintf(stderr, "vmFormIncohBeamNoAlloc: no scratch address on device %d\n", dev);
        return -1;
    }
    incoh_launch(d_data, d_incoh, nsample, nchan, ninput, d_offsets, d_scales, d_Iscaled, part,
                 VM_INCOH_SCRATCH_PAIRS, (hipStream_t)stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "vmFormIncohBeamNoAlloc: %s\n", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

// Rebuild W = phi * J^-1 when J or phi changed (once per second).
void vm_build_weights(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    int nant = (int)vm->obs_metadata->num_ants, nchan = vm->nfine_chan;
    size_t need = (size_t)vm->npointing * nchan * nant * 4 * sizeof(cf);
    if (d->d_W && d->W_gen == d->weights_gen && d->W_bytes >= need) return;
    if (d->W_bytes < need) {
        if (d->d_W) HIPCHECK(hipFree(d->d_W));
        HIPCHECK(hipMalloc(&d->d_W, need));
        d->W_bytes = need;
    }
    vm_expand_J(vm); // the weights read J^-1 per pointing
    size_t total = (size_t)vm->npointing * nant * nchan;
    hipLaunchKernelGGL(k_build_weights, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       vm_stream(vm), (const double2 *)vm->d_J, (const double2 *)vm->gdelays.d_phi,
                       (int)vm->npointing, nant, nchan, (cf *)d->d_W);
    HIPCHECK(hipGetLastError());
    d->W_gen = d->weights_gen;
}

void vmApplyJChunk(vcsbeam_context *vm)
{
    // The J^-1 v product is folded into the beamformer; here the combined
    // per-second weights are refreshed (the MFMA form builds its own split
    // weights inside vmBeamformChunkMFMA).
    if (vm->beamform_form != VM_BF_FP32) return;
    vm_build_weights(vm);
    HIPCHECK(hipStreamSynchronize(vm_stream(vm)));
}

extern "C++" template <int PB>
static void beamform_fp32_launch_pb(const void *d_v, vcsbeam_datatype dt, int ns, int nchan,
                                    int ni, const void *d_W, const uint32_t *d_polQ,
                                    const uint32_t *d_polP, int nant, int npointing, float invw,
                                    int nstokes, int soffset, int ns_total, float *d_S, void *d_e,
                                    hipStream_t stream)
{
    dim3 grid((ns + 255) / 256, nchan, (npointing + PB - 1) / PB), block(256);
    if (dt == VM_INT4) {
        size_t lds = (size_t)256 * ni;
        if (lds > 65536) VM_FATAL("too many RF inputs (%d) for the INT4 beamformer", ni);
        hipLaunchKernelGGL((k_beamform_fp32<PB, true>), grid, block, lds, stream, d_v, ns, nchan,
                           ni, (const cf *)d_W, d_polQ, d_polP, nant, npointing, invw, nstokes,
                           soffset, ns_total, d_S, (cf *)d_e);
    } else if (dt == VM_FLT) {
        hipLaunchKernelGGL((k_beamform_fp32<PB, false>), grid, block, 0, stream, d_v, ns, nchan,
                           ni, (const cf *)d_W, d_polQ, d_polP, nant, npointing, invw, nstokes,
                           soffset, ns_total, d_S, (cf *)d_e);
    } else {
        VM_FATAL("FP32 beamformer: voltage datatype %d; it reads VM_FLT or VM_INT4 (VM_HALF, "
                 "from PFB_COMPLEX_FLOAT16, is the MFMA forms' input)", (int)dt);
    }
    HIPCHECK(hipGetLastError());
}

// Four pointings per thread share each voltage load; a single pointing takes
// the one-pointing form, which would otherwise compute the same pointing four
// times (the kernel clamps pointings past the last).  Each pointing's
// arithmetic is the same in both forms.
void vm_beamform_fp32_launch(const void *d_v, vcsbeam_datatype dt, int ns, int nchan, int ni,
                             const void *d_W, const uint32_t *d_polQ, const uint32_t *d_polP,
                             int nant, int npointing, float invw, int nstokes, int soffset,
                             int ns_total, float *d_S, void *d_e, hipStream_t stream)
{
    if (npointing == 1)
        beamform_fp32_launch_pb<1>(d_v, dt, ns, nchan, ni, d_W, d_polQ, d_polP, nant, npointing,
                                   invw, nstokes, soffset, ns_total, d_S, d_e, stream);
    else
        beamform_fp32_launch_pb<4>(d_v, dt, ns, nchan, ni, d_W, d_polQ, d_polP, nant, npointing,
                                   invw, nstokes, soffset, ns_total, d_S, d_e, stream);
}

void vmBeamformChunk(vcsbeam_context *vm)
{
    if (vm->beamform_form != VM_BF_FP32) {
        vmBeamformChunkMFMA(vm);
        return;
    }
    vm_build_weights(vm);
    int chunk = vm->chunk_to_load % vm->chunks_per_second;
    int ns = vm->fine_sample_rate / vm->chunks_per_second;
    float invw = (float)(1.0 / (double)vm->num_not_flagged);
    vm_timer_start(vm->dev, VM_T_BF_FP32);
    vm_beamform_fp32_launch(vm->d_v, vm->datatype, ns, vm->nfine_chan,
                            (int)vm->obs_metadata->num_rf_inputs, vm->dev->d_W, vm->d_polQ_idxs,
                            vm->d_polP_idxs, (int)vm->obs_metadata->num_ants, (int)vm->npointing,
                            invw, vm->out_nstokes, chunk * ns, vm->fine_sample_rate,
                            (float *)vm->d_S, vm->d_e, vm_stream(vm));
    vm_timer_stop(vm->dev, VM_T_BF_FP32);
    HIPCHECK(hipStreamSynchronize(vm_stream(vm)));
}

void vmBeamformSecond(vcsbeam_context *vm)
{
    // src/form_beam.cpp:587-627
    for (unsigned int chunk = 0; chunk < vm->chunks_per_second; chunk++) {
        if (vm->obs_metadata->mwa_version == VCSLegacyRecombined) {
            vmPushChunk(vm);
        } else {
            vmUploadForwardPFBChunk(vm);
            logger_start_stopwatch(vm->log, "pfb", chunk == 0);
            vmWOLAChunk(vm);
            vmFPGARoundingChunk(vm);
            vmFFTChunk(vm);
            vmPackChunk(vm);
            logger_stop_stopwatch(vm->log, "pfb");
        }
        logger_start_stopwatch(vm->log, "calc", chunk == 0);
        vmApplyJChunk(vm);
        vmBeamformChunk(vm);
        logger_stop_stopwatch(vm->log, "calc");
        vm->chunk_to_load++;
    }
    HIPCHECK(hipStreamSynchronize(vm_stream(vm)));
    vm->v->locked = false;
}

void vmPullE(vcsbeam_context *vm)
{
    // Synchronous (the reference's asynchronous copy races with the host
    // reader, app/make_mwa_tied_array_beam.c:280-283)
    int ns = vm->fine_sample_rate, nchan = vm->nfine_chan;
    void *st = stage_buffer(vm, vm->e_size_bytes);
    dim3 grid((ns + 63) / 64, (nchan + 63) / 64, vm->npointing);
    hipLaunchKernelGGL(k_e_to_host, grid, dim3(256), 0, vm_stream(vm), (const cf *)vm->d_e, ns,
                       nchan, (double2 *)st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(vm->e, st, vm->e_size_bytes, hipMemcpyDeviceToHost, vm_stream(vm)));
    HIPCHECK(hipStreamSynchronize(vm_stream(vm)));
}

void vmPullS(vcsbeam_context *vm)
{
    int ns = vm->fine_sample_rate, nchan = vm->nfine_chan;
    void *st = stage_buffer(vm, vm->S_size_bytes);
    dim3 grid((ns + 63) / 64, (nchan + 63) / 64, vm->npointing * vm->out_nstokes);
    hipLaunchKernelGGL(k_S_to_host, grid, dim3(256), 0, vm_stream(vm), (const float *)vm->d_S, ns,
                       (int)vm->npointing, vm->out_nstokes, nchan, (float *)st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(vm->S, st, vm->S_size_bytes, hipMemcpyDeviceToHost, vm_stream(vm)));
    HIPCHECK(hipStreamSynchronize(vm_stream(vm)));
}

void vmSendSToFits(vcsbeam_context *vm, mpi_psrfits *mpfs)
{
    // src/form_beam.cpp:653-673: renormalise per (p, st, c) to uint8 and copy
    // the sub-integration arrays into the PSRFITS structs
    int ns = vm->fine_sample_rate, nchan = vm->nfine_chan, nst = vm->out_nstokes;
    hipStream_t s = vm_stream(vm);
    vm_timer_start(vm_dev(vm), VM_T_RENORM);
    hipLaunchKernelGGL(k_minmax_dev, dim3(vm->npointing * nst * nchan), dim3(256), 0, s,
                       (const float *)vm->d_S, ns, (int)vm->npointing, nst, nchan, nchan, 0,
                       vm->d_offsets, vm->d_scales);
    HIPCHECK(hipGetLastError());
    dim3 grid((ns + 63) / 64, (nchan + 63) / 64, vm->npointing * nst);
    hipLaunchKernelGGL(k_quant_dev, grid, dim3(256), 0, s, (const float *)vm->d_S, ns,
                       (int)vm->npointing, nst, nchan, nchan, 0,
                       vm->d_offsets, vm->d_scales, vm->d_Cscaled);
    HIPCHECK(hipGetLastError());
    vm_timer_stop(vm_dev(vm), VM_T_RENORM);
    HIPCHECK(hipMemcpyAsync(vm->offsets, vm->d_offsets, vm->offsets_size, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(vm->scales, vm->d_scales, vm->scales_size, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(vm->Cscaled, vm->d_Cscaled,
                            (size_t)vm->npointing * ns * nchan * nst, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (!mpfs) return;
    for (unsigned int p = 0; p < vm->npointing; p++) {
        vm_psrfits_sub *sub = &mpfs[p].coarse_chan_pf.sub;
        if (sub->dat_offsets)
            memcpy(sub->dat_offsets, vm->offsets + (size_t)p * nchan * nst, nchan * nst * sizeof(float));
        if (sub->dat_scales)
            memcpy(sub->dat_scales, vm->scales + (size_t)p * nchan * nst, nchan * nst * sizeof(float));
        if (sub->data)
            memcpy(sub->data, vm->Cscaled + (size_t)p * ns * nchan * nst, (size_t)ns * nchan * nst);
    }
}

void vmSpliceSToFits(vcsbeam_context **vms, int ncoarse, vm_psrfits *pf)
{
    // One process holds every coarse channel: each context's Stokes are
    // renormalised into its slot of one spliced device buffer (the layout
    // the reference's MPI_Igather builds on the writer rank), then one
    // download per pointing.
    if (ncoarse <= 0 || !vms || !pf) VM_FATAL("vmSpliceSToFits: no channels");
    vcsbeam_context *v0 = vms[0];
    const int ns = v0->fine_sample_rate, nchan = v0->nfine_chan, nst = v0->out_nstokes;
    const unsigned np = v0->npointing;
    for (int g = 0; g < ncoarse; g++) {
        const vcsbeam_context *v = vms[g];
        if (v->fine_sample_rate != ns || v->nfine_chan != nchan || v->out_nstokes != nst ||
            v->npointing != np || !v->d_S)
            VM_FATAL("vmSpliceSToFits: channel %d differs in shape or has no Stokes", g);
    }
    const int ctot = nchan * ncoarse;
    for (unsigned p = 0; p < np; p++)
        if (!pf[p].sub.data || !pf[p].sub.dat_offsets || !pf[p].sub.dat_scales ||
            pf[p].sub.bytes_per_subint != ns * nst * ctot)
            VM_FATAL("vmSpliceSToFits: pointing %u's buffers do not hold %d x %d x %d bytes", p, ns,
                     nst, ctot);
    static uint8_t *d_data = NULL;
    static float *d_stat = NULL;
    static size_t data_cap = 0, stat_cap = 0;
    const size_t data_bytes = (size_t)np * ns * nst * ctot, stat_n = (size_t)np * nst * ctot;
    if (data_cap < data_bytes) {
        if (d_data) HIPCHECK(hipFree(d_data));
        HIPCHECK(hipMalloc((void **)&d_data, data_bytes));
        data_cap = data_bytes;
    }
    if (stat_cap < 2 * stat_n * sizeof(float)) {
        if (d_stat) HIPCHECK(hipFree(d_stat));
        HIPCHECK(hipMalloc((void **)&d_stat, 2 * stat_n * sizeof(float)));
        stat_cap = 2 * stat_n * sizeof(float);
    }
    float *d_offs = d_stat, *d_scl = d_stat + stat_n;
    vm_timer_start(vm_dev(v0), VM_T_RENORM);
    for (int g = 0; g < ncoarse; g++) {
        hipStream_t s = vm_stream(vms[g]);
        hipLaunchKernelGGL(k_minmax_dev, dim3(np * nst * nchan), dim3(256), 0, s,
                           (const float *)vms[g]->d_S, ns, (int)np, nst, nchan, ctot, g * nchan,
                           d_offs, d_scl);
        HIPCHECK(hipGetLastError());
        dim3 grid((ns + 63) / 64, (nchan + 63) / 64, np * nst);
        hipLaunchKernelGGL(k_quant_dev, grid, dim3(256), 0, s, (const float *)vms[g]->d_S, ns,
                           (int)np, nst, nchan, ctot, g * nchan, d_offs, d_scl, d_data);
        HIPCHECK(hipGetLastError());
    }
    for (int g = 0; g < ncoarse; g++) HIPCHECK(hipStreamSynchronize(vm_stream(vms[g])));
    vm_timer_stop(vm_dev(v0), VM_T_RENORM);
    for (unsigned p = 0; p < np; p++) {
        HIPCHECK(hipMemcpy(pf[p].sub.data, d_data + (size_t)p * ns * nst * ctot,
                           (size_t)ns * nst * ctot, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pf[p].sub.dat_offsets, d_offs + (size_t)p * nst * ctot,
                           (size_t)nst * ctot * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pf[p].sub.dat_scales, d_scl + (size_t)p * nst * ctot,
                           (size_t)nst * ctot * sizeof(float), hipMemcpyDeviceToHost));
    }
}

void prepare_data_buffer_fine(gpuDoubleComplex *data_buffer_fine, vcsbeam_context *vm,
                              uintptr_t timestep_idx)
{
    // src/form_beam.cpp:753-780: copy e into the even/odd half of the
    // 2-second ring; rows [s][c][pol] are contiguous, so copy per pointing
    size_t nchan = vm->nfine_chan, npol = vm->obs_metadata->num_ant_pols;
    int file_no = (int)(timestep_idx % 2);
    size_t offset = (size_t)(file_no % 2) * vm->fine_sample_rate;
    size_t per_s = nchan * npol;
    for (unsigned int p = 0; p < vm->npointing; p++) {
        const gpuDoubleComplex *src = vm->e + (size_t)p * vm->fine_sample_rate * per_s;
        gpuDoubleComplex *dst = data_buffer_fine + ((size_t)p * vm->fine_sample_rate + offset) * per_s;
        memcpy(dst, src, (size_t)vm->fine_sample_rate * per_s * sizeof(gpuDoubleComplex));
    }
}

} // extern "C"
