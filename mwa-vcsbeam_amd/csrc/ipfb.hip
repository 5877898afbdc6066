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
#include <time.h>
#include <stdio.h>

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

// VDIF requantisation of one value: clamp to [-126, 127], round half to
// even, int8, offset binary (src/beam_vdif.c:29-38, 254-260)
__device__ __forceinline__ uint8_t vdif_q1(float v)
{
    const float x = fmaxf(fminf(v, 127.0f), -126.0f);
    return (uint8_t)((int8_t)(int)rintf(x) ^ 0x80);
}

// The inverse DFTs of the nrows rows in Z (conj(X) in, Z out) and the
// polyphase FIR into out (and, when out8 is not NULL, the same values
// requantised for VDIF at the same [n][pol][re, im] positions).
// The inverse DFTs of the nrows rows in Z, by the 256 threads of one
// polarisation (every thread of the workgroup passes the same barriers).
// tw: the twiddle table (g_tw128i, or a copy of it in LDS).
__device__ __forceinline__ void ipfb_idft(cf *Z, int nrows, const cf *tw = g_tw128i)
{
    const int t = threadIdx.x & 255;
    // inverse DFTs, 32 rows per pass
    const int f = t & 31, l = t >> 5;
    for (int base = 0; base < nrows; base += 32) {
        cf *row = &Z[min(base + f, nrows - 1) * ROW];
        // rows past nrows duplicate the last row's work but never store
        cf keep[2][8];
        vmfft::fft_stage1(row, l, tw);
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
}

// PC: the number of taps when known at compile time (0: P at run time, up
// to MAXP)
template <int PC = 0>
__device__ __forceinline__ void ipfb_rows(cf *Z, int nrows, int Prt, int jmax, int j0, int pol,
                                          int npol, const float *__restrict__ hK,
                                          float *__restrict__ out, uint8_t *__restrict__ out8)
{
    constexpr int NQ = PC ? PC : MAXP;
    const int P = PC ? PC : Prt;
    const int t = threadIdx.x & 255; // the 256 threads of this polarisation
    ipfb_idft(Z, nrows);
    // polyphase FIR: out[(j0+jj)K + r] = sum_q hK[qK + r] Z[jj - q + P - 1][r]
    const int r = t & 127;
    float hr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) hr[q] = (q < P) ? hK[q * KCH + r] : 0.f;
    for (int jj = t >> 7; jj < jmax; jj += 2) {
        float ar = 0.f, ai = 0.f;
        const cf *zc = Z + (jj + P - 1) * ROW + r;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            if (PC || q < P) {
                const cf z = zc[-q * ROW];
                ar = fmaf(hr[q], z.x, ar);
                ai = fmaf(hr[q], z.y, ai);
            }
        }
        size_t n = (size_t)(j0 + jj) * KCH + r;
        *(float2 *)&out[(n * npol + pol) * 2] = make_float2(ar, ai);
        if (out8)
            *(uchar2 *)&out8[(n * npol + pol) * 2] = make_uchar2(vdif_q1(ar), vdif_q1(ai));
    }
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

    // load conj(X) for rows whose s_in exists, eight loads in flight per lane
    for (int i0 = t; i0 < nrows * KCH; i0 += 8 * 256) {
        cf v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = i0 + 256 * u;
            const int row = idx >> 7, k = idx & 127;
            const int s_in = j0 + 1 + row;
            v[u] = cf{0.f, 0.f};
            if (idx < nrows * KCH && s_in < ns + P) {
                const T x = in[((size_t)s_in * KCH + k) * npol + pol];
                v[u] = cf{(float)x.x, -(float)x.y};
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = i0 + 256 * u;
            if (idx < nrows * KCH) Z[(idx >> 7) * ROW + (idx & 127)] = v[u];
        }
    }
    __syncthreads();

    ipfb_rows<0>(Z, nrows, P, jmax, j0, pol, npol, hK, out, (uint8_t *)nullptr);
}

// Device-resident form: the fine channels straight from the beamformer
// output e ([k][s][pol] complex float, pointing 0), the rows before the
// second's first sample from the previous second's last P samples in
// hist_in ([P][k][pol]; zeros when first), and this second's last P samples
// copied to hist_out for the next second (the workgroups that load them
// write them; duplicates carry equal values).  A workgroup of 256 x npol
// threads (npol <= 2) holds both polarisations, so each lane loads a
// sample's pols together (16 bytes) and lanes walk the samples of a fine
// channel; each 256-thread half then runs one polarisation's inverse DFTs
// and FIR.  It takes 33 - P output fine samples, so that with their P - 1
// earlier ones its rows are exactly one 32-row pass of inverse DFTs.
// out8: optional VDIF bytes.  grid ceil(ns / (33 - P)), block 256 x npol.
#define IPFB_DEV_ROWS 32
// One coarse channel's inverse PFB on the device: its beamformed voltages,
// history in and out, synthesis filter and outputs (out8 may be NULL).
struct ipfb_job {
    const cf *e;
    const cf *hist_in;
    cf *hist_out;
    const float *hK;
    float *out;
    uint8_t *out8;
};

// Workgroups are dispatched to the 8 XCDs in turn (linear id L to XCD L % 8),
// each XCD with its own L2.  Neighbouring tiles share P - 1 input rows, so
// the tiles are renumbered to give each XCD one contiguous run of them: the
// shared rows are then fetched from HBM once, into that XCD's L2.  A
// bijection of [0, n): XCD x (of q + 1 tiles for x < n % 8, else q) takes
// tiles [x q + min(x, n % 8), ...).
__device__ __forceinline__ int xcd_contiguous(int L, int n)
{
    const int q = n >> 3, r = n & 7, x = L & 7, k = L >> 3;
    return x * q + min(x, r) + k;
}

// BATCH: the job is jobs[y] (one coarse channel per grid row), otherwise
// job1.  (x, y): the tile after the XCD renumbering.  The general form (any
// tap count up to MAXP, one or two polarisations); k_ipfb_run takes the
// 12-tap, two-polarisation case.
template <bool BATCH>
__global__ __launch_bounds__(512) void k_ipfb_dev(ipfb_job job1, const ipfb_job *__restrict__ jobs,
                                                  int ns, int P, int npol, int first)
{
    __shared__ cf Z[2][IPFB_DEV_ROWS * ROW];
    const int tile = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bx = tile % gridDim.x, by = tile / gridDim.x;
    const ipfb_job job = BATCH ? jobs[by] : job1;
    const cf *__restrict__ e = job.e;
    const cf *__restrict__ hist_in = job.hist_in;
    cf *__restrict__ hist_out = job.hist_out;
    const float *__restrict__ hK = job.hK;
    float *__restrict__ out = job.out;
    uint8_t *__restrict__ out8 = job.out8;
    const int t = threadIdx.x;
    const int tj = IPFB_DEV_ROWS - (P - 1);
    const int j0 = bx * tj;
    const int nrows = IPFB_DEV_ROWS; // s_in = j0 + 1 + row
    const int jmax = min(tj, ns - j0);
    const int nel = nrows * KCH;      // (row, k) pairs, each with npol values
    const int nthr = 256 * npol;
    for (int i0 = t; i0 < nel; i0 += 8 * nthr) {
        cf v[8][2];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = i0 + nthr * u;
            const int row = idx % nrows, k = idx / nrows; // lanes along samples
            const int s_in = j0 + 1 + row;
            v[u][0] = v[u][1] = cf{0.f, 0.f};
            if (idx < nel && s_in < ns + P) {
                const cf *src = s_in < P ? (first ? nullptr : hist_in + ((size_t)s_in * KCH + k) * npol)
                                         : e + ((size_t)k * ns + (s_in - P)) * npol;
                if (src) {
                    if (npol == 2) {
                        const float4 w = *(const float4 *)src;
                        v[u][0] = cf{w.x, w.y};
                        v[u][1] = cf{w.z, w.w};
                    } else {
                        v[u][0] = src[0];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = i0 + nthr * u;
            if (idx >= nel) continue;
            const int row = idx % nrows, k = idx / nrows;
            const int s_in = j0 + 1 + row;
#pragma unroll
            for (int pol = 0; pol < 2; pol++) {
                if (pol >= npol) break;
                if (s_in >= ns && s_in < ns + P) // e sample ns - P + (s_in - ns)
                    hist_out[((size_t)(s_in - ns) * KCH + k) * npol + pol] = v[u][pol];
                Z[pol][row * ROW + k] = cf{v[u][pol].x, -v[u][pol].y};
            }
        }
    }
    __syncthreads();
    const int pol = t >> 8;
    // each 256-thread half: one polarisation's inverse DFTs and FIR (both
    // halves pass the same barriers inside ipfb_rows)
    ipfb_rows<0>(Z[pol], nrows, P, jmax, j0, pol, npol, hK, out, out8);
}

// Streaming form for 12 taps and two polarisations (the MWA synthesis
// filter).  A workgroup walks a run of npass consecutive passes of one coarse
// channel.  Its first pass is the tile of k_ipfb_dev: 32 rows, the 11 rows
// before its first output and 21 outputs.  Every later pass loads the next 32
// rows, inverse-transforms them and gives 32 outputs, the 11 rows before them
// carried in registers from the pass before.  So every row after a run's
// first 11 is transformed once, where k_ipfb_dev transforms 11 of every 32
// twice.  The inverse DFTs and the FIR sums are those of k_ipfb_dev, term for
// term, so both give the same bits.
#define IPFB_RUN_P 12
#define IPFB_RUN_T0 (IPFB_DEV_ROWS - (IPFB_RUN_P - 1)) // outputs of a run's first pass
#define IPFB_MAX_PASSES 64

// The job's pointers are generic; accesses through GLOBAL(T, p) are global
// memory instructions, which (unlike flat ones) do not count against the LDS
// wait counter, so waiting for LDS reads does not wait for them.
#define GLOBAL(T, p) ((__attribute__((address_space(1))) T *)(p))
typedef float vf4 __attribute__((ext_vector_type(4))); // native vectors: no
typedef float vf2 __attribute__((ext_vector_type(2))); // generic-reference copy operators

// Raw buffer accesses: 32-bit offsets from a base held in scalar registers
// (fewer vector registers and no 64-bit address arithmetic per access), and
// an offset past the buffer's size reads zeros.
typedef __amdgpu_buffer_rsrc_t ipfb_rsrc_t;
typedef uint32_t ipfb_v4u __attribute__((ext_vector_type(4)));
typedef uint32_t ipfb_v2u __attribute__((ext_vector_type(2)));
#define IPFB_NO_ROW 0x7ffffff0u // an offset past every buffer here
__device__ __forceinline__ ipfb_rsrc_t ipfb_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000 /* raw, 32-bit format */);
}

// Rows s_in = s0 + row (row < 32) of both polarisations into v: one
// (row, k) pair per lane and u, lanes along the samples of a fine channel
// (16-byte loads of e[k][s][pol] through re, the channel's e); history rows
// s_in < P from hist_in ([P][k][pol], through rh; zeros when first), rows
// past ns + P - 1 zeros.  Only a pass with s0 < P (a run at the second's
// start) reads history rows: it loads from both buffers, each lane's other
// offset out of range (zeros), and ORs the bits.
__device__ __forceinline__ void ipfb_run_load(float4 v[8], ipfb_rsrc_t re, ipfb_rsrc_t rh,
                                              int first, int s0, int ns, int t)
{
    constexpr int P = IPFB_RUN_P;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int idx = t + 512 * u;
        const int row = idx & (IPFB_DEV_ROWS - 1), k = idx / IPFB_DEV_ROWS;
        const int s_in = s0 + row;
        const uint32_t oe = s_in >= P && s_in < ns + P
                                ? ((uint32_t)k * (uint32_t)ns + (uint32_t)(s_in - P)) * 16u
                                : IPFB_NO_ROW;
        ipfb_v4u x = __builtin_amdgcn_raw_buffer_load_b128(re, (int)oe, 0, 0);
        if (s0 < P) {
            const uint32_t oh =
                s_in < P && !first ? ((uint32_t)s_in * KCH + (uint32_t)k) * 16u : IPFB_NO_ROW;
            x |= __builtin_amdgcn_raw_buffer_load_b128(rh, (int)oh, 0, 0);
        }
        const vf4 f = __builtin_bit_cast(vf4, x);
        v[u] = make_float4(f.x, f.y, f.z, f.w);
    }
}

// v (loaded by ipfb_run_load for the same s0 and t) as conj(X) into Z, and
// the second's last P rows into hist_out ([P][k][pol]) for the next second.
__device__ __forceinline__ void ipfb_run_put(const float4 v[8], cf (*Z)[IPFB_DEV_ROWS * ROW],
                                             cf *__restrict__ hist_out, int s0, int ns, int t)
{
    constexpr int P = IPFB_RUN_P;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int idx = t + 512 * u;
        const int row = idx & (IPFB_DEV_ROWS - 1), k = idx / IPFB_DEV_ROWS;
        const int s_in = s0 + row;
        if (s_in >= ns && s_in < ns + P)
            *GLOBAL(vf4, &hist_out[((size_t)(s_in - ns) * KCH + k) * 2]) =
                vf4{v[u].x, v[u].y, v[u].z, v[u].w};
        Z[0][row * ROW + k] = cf{v[u].x, -v[u].y};
        Z[1][row * ROW + k] = cf{v[u].z, -v[u].w};
    }
}

// grid (runs, jobs), block 512; a run is IPFB_RUN_T0 + 32 (npass - 1)
// outputs.  The FIR lanes are (r, pol) pairs, pol fastest, so a wavefront's
// stores (8-byte complex floats, 2 VDIF bytes) are contiguous runs; the
// first four wavefronts take the first half of a pass's outputs, the other
// four the second half.
template <bool BATCH>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_ipfb_run(ipfb_job job1, const ipfb_job *__restrict__ jobs,
                                                  int ns, int npass, int first)
{
    constexpr int P = IPFB_RUN_P, H = P - 1, NR = IPFB_DEV_ROWS, T0 = IPFB_RUN_T0;
    __shared__ cf Z[2][NR * ROW];
    __shared__ cf TW[KCH];
    __shared__ float HK[P * KCH];
    const int tile = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bx = tile % gridDim.x, by = tile / gridDim.x;
    const ipfb_job job = BATCH ? jobs[by] : job1;
    const int t = threadIdx.x;
    // twiddles and taps in LDS (read after the first pass's first barrier):
    // the taps are held in registers during the FIR only
    if (t < KCH) TW[t] = g_tw128i[t];
    for (int i = t; i < P * KCH; i += 512) HK[i] = *GLOBAL(const float, job.hK + i);
    const int fpol = t & 1, r = (t >> 1) & (KCH - 1), g = t >> 8;
    const int J0 = bx * (T0 + NR * (npass - 1)); // the run's first output
    // e, out and out8 of one channel are under 2 GB (checked at launch)
    const ipfb_rsrc_t re = ipfb_rsrc(job.e, (uint32_t)ns * KCH * 16u);
    const ipfb_rsrc_t rh = ipfb_rsrc(job.hist_in, (uint32_t)P * KCH * 16u);
    const ipfb_rsrc_t ro = ipfb_rsrc(job.out, (uint32_t)ns * KCH * 16u);
    const ipfb_rsrc_t ro8 = ipfb_rsrc(job.out8, (uint32_t)ns * KCH * 4u);
    cf carry[H];
    float4 v[8];
    ipfb_run_load(v, re, rh, first, J0 + 1, ns, t);
    for (int k = 0; k < npass; k++) {
        const int jb = k == 0 ? J0 : J0 + T0 + NR * (k - 1); // the pass's first output
        if (jb >= ns) break;
        // Z row 0 is s_in = s0; a pass's outputs jb + o take s_in jb + o + 1 .. jb + o + P
        const int s0 = k == 0 ? jb + 1 : jb + P;
        ipfb_run_put(v, Z, job.hist_out, s0, ns, t);
        __syncthreads();
        ipfb_idft(Z[t >> 8], NR, TW); // ends on a barrier
        // the next pass's rows are requested after the FIR (while the FIR's
        // registers are free) and are in flight across the pass's last
        // barrier; rows from ns + P on read as zeros without a memory
        // access, so after a run's last pass nothing is requested
        const int jn = J0 + T0 + NR * k; // the next pass's first output
        const int sn = k + 1 < npass && jn < ns ? jn + P : ns + P;
        // output o takes rows o .. o + H of the pass's combined rows: the H
        // rows before its first output, then its new rows; combined row c is
        // Z row c + zb (the first pass: all 32 rows are in Z)
        const int nnew = k == 0 ? T0 : NR;
        const int zb = NR - nnew - H;
        const int jlim = min(nnew, ns - jb);
        const cf *Zp = Z[fpol] + r;
        // the first four wavefronts take outputs [0, oh), the others
        // [oh, nnew): 16 and 16, or 11 and 10 in a run's first pass
        const int oh = (nnew + 1) >> 1;
        const int o0 = g ? oh : 0, o1 = min(g ? nnew : oh, jlim);
        if (o0 < o1) {
            float hr[P];
#pragma unroll
            for (int q = 0; q < P; q++) hr[q] = HK[q * KCH + r];
            cf w[P]; // w[i]: combined row o + i
            if (g == 0 && k > 0) {
#pragma unroll
                for (int i = 0; i < H; i++) w[i] = carry[i];
            } else {
#pragma unroll
                for (int i = 0; i < H; i++) w[i] = Zp[(o0 + i + zb) * ROW];
            }
#pragma unroll
            for (int oo = 0; oo < 16; oo++) {
                const int o = o0 + oo;
                if (o >= o1) break;
                w[H] = Zp[(o + H + zb) * ROW];
                float ar = 0.f, ai = 0.f;
#pragma unroll
                for (int q = 0; q < P; q++) {
                    ar = fmaf(hr[q], w[H - q].x, ar);
                    ai = fmaf(hr[q], w[H - q].y, ai);
                }
                const uint32_t n = ((uint32_t)(jb + o) * KCH + (uint32_t)r) * 2u + (uint32_t)fpol;
                __builtin_amdgcn_raw_buffer_store_b64(
                    __builtin_bit_cast(ipfb_v2u, vf2{ar, ai}), ro, (int)(n * 8u), 0, 0);
                if (job.out8)
                    __builtin_amdgcn_raw_buffer_store_b16(
                        (uint16_t)((uint16_t)vdif_q1(ar) | ((uint16_t)vdif_q1(ai) << 8)), ro8,
                        (int)(n * 2u), 0, 0);
#pragma unroll
                for (int i = 0; i < H; i++) w[i] = w[i + 1];
            }
        }
        // the next pass's first H combined rows: this pass's last H rows
        if (g == 0) {
#pragma unroll
            for (int i = 0; i < H; i++) carry[i] = Zp[(NR - H + i) * ROW];
        }
        ipfb_run_load(v, re, rh, first, sn, ns, t);
        __syncthreads();
    }
}

// VDIF requantisation: clamp to [-126, 127], round half to even, int8,
// offset binary (XOR 0x80).  Four values per lane.
__device__ __forceinline__ uint32_t vdif_q4(float4 v)
{
    const float a[4] = {v.x, v.y, v.z, v.w};
    uint32_t w = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float x = fminf(a[q], 127.0f);
        x = fmaxf(x, -126.0f);
        int iq = (int)rintf(x);
        w |= (uint32_t)((uint8_t)((int8_t)iq ^ 0x80)) << (8 * q);
    }
    return w;
}

__global__ void k_vdif_requant(const float *__restrict__ in, uint8_t *__restrict__ out, size_t n)
{
    size_t j = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (j >= n) return;
    if (j + 4 <= n) {
        *(uint32_t *)(out + j) = vdif_q4(*(const float4 *)(in + j));
    } else {
        for (size_t q = j; q < n; q++) {
            float x = fmaxf(fminf(in[q], 127.0f), -126.0f);
            out[q] = (uint8_t)((int8_t)(int)rintf(x) ^ 0x80);
        }
    }
}

// VDIF frames: nframes x (32-byte header + pw payload words), one lane per
// output 32-bit word.  Frame f's header is the template t with the seconds
// (word 0, bits 0-29) and frame number (word 1, bits 0-23) of frame fr0 + f
// of a second of frame_rate frames counted from sec0 (the VDIF 1.0 header;
// nextVDIFHeader's increment, src/beam_vdif.c:79); its payload word q is
// samples 4 q .. 4 q + 3 of the frame requantised like k_vdif_requant.
__global__ void k_vdif_frames(const float *__restrict__ in, size_t nframes, int pw, uint4 tlo,
                              uint4 thi, uint32_t sec0, uint32_t fr0, uint32_t frame_rate,
                              uint32_t *__restrict__ out)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t fw = 8 + (size_t)pw;
    if (j >= nframes * fw) return;
    const size_t f = j / fw;
    const int q = (int)(j - f * fw);
    uint32_t w;
    if (q >= 8) {
        w = vdif_q4(*(const float4 *)(in + (f * pw + (q - 8)) * 4));
    } else {
        const uint64_t g = (uint64_t)fr0 + f;
        const uint32_t sec = sec0 + (uint32_t)(g / frame_rate), fn = (uint32_t)(g % frame_rate);
        const uint32_t t[8] = {tlo.x, tlo.y, tlo.z, tlo.w, thi.x, thi.y, thi.z, thi.w};
        w = t[q];
        if (q == 0) w = (w & 0xC0000000u) | (sec & 0x3FFFFFFFu);
        if (q == 1) w = (w & 0xFF000000u) | (fn & 0x00FFFFFFu);
    }
    out[j] = w;
}

// ------------------------------------------------ VDIF headers (host) ----
// The VDIF 1.0 frame header as vdifio builds it for src/beam_vdif.c:216-225
// (createVDIFHeader, setVDIFEpochMJD, setVDIFFrameMJDSec, setVDIFFrameNumber):
//   word 0: invalid (31), legacy (30), seconds from the reference epoch (29-0)
//   word 1: reference epoch in half-years from 2000 (29-24), frame number (23-0)
//   word 2: version (31-29), log2 channels (28-24), frame bytes / 8 (23-0)
//   word 3: complex (31), bits per sample - 1 (30-26), thread (25-16),
//           station: two ASCII characters, first in bits 15-8 (15-0)
//   words 4-7: extended user data (version 0: zero)
// vdifio itself is not available here; these follow the specification.

// MJD of a Gregorian calendar date (integer day number arithmetic)
static int64_t vm_mjd_of_date(int64_t y, int64_t m, int64_t d)
{
    const int64_t a = (14 - m) / 12, yy = y + 4800 - a, mm = m + 12 * a - 3;
    const int64_t jdn = d + (153 * mm + 2) / 5 + 365 * yy + yy / 4 - yy / 100 + yy / 400 - 32045;
    return jdn - 2400001;
}

// MJD at which VDIF reference epoch e (half-years since 2000-01-01) starts
static int64_t vm_vdif_epoch_mjd(int e)
{
    return vm_mjd_of_date(2000 + e / 2, (e % 2) ? 7 : 1, 1);
}

static void vm_vdif_create(uint32_t *w, int dataarraylength, int threadid, int bits, int nchan,
                           int iscomplex, const char *stationid)
{
    int lg = 0;
    while ((1 << lg) < nchan) lg++;
    memset(w, 0, 32);
    w[2] = ((uint32_t)(lg & 0x1F) << 24) | (uint32_t)(((dataarraylength + 32) / 8) & 0xFFFFFF);
    const uint32_t st = ((uint32_t)(uint8_t)stationid[0] << 8) |
                        (stationid[0] ? (uint32_t)(uint8_t)stationid[1] : 0u);
    w[3] = ((uint32_t)(iscomplex ? 1 : 0) << 31) | ((uint32_t)((bits - 1) & 0x1F) << 26) |
           ((uint32_t)(threadid & 0x3FF) << 16) | (st & 0xFFFF);
}

static void vm_vdif_set_epoch_mjd(uint32_t *w, int64_t mjd)
{
    int e = 0;
    while (e < 63 && vm_vdif_epoch_mjd(e + 1) <= mjd) e++;
    w[1] = (w[1] & 0x00FFFFFFu) | ((uint32_t)e << 24);
}

static void vm_vdif_set_mjdsec(uint32_t *w, int64_t mjdsec)
{
    const int e = (int)((w[1] >> 24) & 0x3F);
    const int64_t sec = mjdsec - vm_vdif_epoch_mjd(e) * 86400;
    w[0] = (w[0] & 0xC0000000u) | (uint32_t)(sec & 0x3FFFFFFF);
}

// the header after n more frames of a second of frame_rate frames
static void vm_vdif_advance(uint32_t *w, uint64_t n, uint32_t frame_rate)
{
    const uint64_t g = (uint64_t)(w[1] & 0x00FFFFFFu) + n;
    const uint32_t sec = (w[0] & 0x3FFFFFFFu) + (uint32_t)(g / frame_rate);
    w[0] = (w[0] & 0xC0000000u) | (sec & 0x3FFFFFFFu);
    w[1] = (w[1] & 0xFF000000u) | (uint32_t)(g % frame_rate);
}

static void vdif_frames_launch(const float *d_in, size_t nframes, int payload_bytes,
                               const uint32_t *hdr, uint32_t frame_rate, uint8_t *d_out,
                               hipStream_t stream)
{
    if (payload_bytes <= 0 || payload_bytes % 4 || frame_rate == 0)
        VM_FATAL("VDIF frames: payload of %d bytes (a positive multiple of 4) at %u frames/s",
                 payload_bytes, frame_rate);
    const size_t words = nframes * (8 + (size_t)payload_bytes / 4);
    if (!words) return;
    const uint4 tlo = make_uint4(hdr[0], hdr[1], hdr[2], hdr[3]);
    const uint4 thi = make_uint4(hdr[4], hdr[5], hdr[6], hdr[7]);
    hipLaunchKernelGGL(k_vdif_frames, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, stream,
                       d_in, nframes, payload_bytes / 4, tlo, thi, hdr[0] & 0x3FFFFFFFu,
                       hdr[1] & 0x00FFFFFFu, frame_rate, (uint32_t *)d_out);
    HIPCHECK(hipGetLastError());
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
    if (nsamples <= P || (size_t)(nsamples + P) * per_s * sizeof(double2) > (size_t)g->in_size ||
        (size_t)nsamples * per_s * 2 * sizeof(float) > (size_t)g->out_size)
        VM_FATAL("cu_invert_pfb: %d samples x %d pols exceed what malloc_ipfb allocated", nsamples,
                 npol);
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


// The device-resident inverse PFB of pointing 0 (one fused kernel), with the
// VDIF bytes too when d_vdif is not NULL.  The P-sample history lives in two
// buffers used in turn: second file_no reads buffer (file_no + 1) % 2 (the
// previous second's tail) and writes buffer file_no % 2.
// Checks one context's arrays, makes its history buffers and fills its job
// for second file_no.
static void ipfb_job_for(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                         uint8_t *d_vdif, const char *who, ipfb_job &job)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    int ns = vm->fine_sample_rate, P = g->ntaps, npol = (int)vm->obs_metadata->num_ant_pols;
    if (vm->nfine_chan != KCH || !vm->d_e || ns <= P || P > MAXP ||
        (size_t)ns * KCH * npol * 2 * sizeof(float) > (size_t)g->out_size)
        VM_FATAL("%s: %d samples x %d pols of %d channels do not fit the arrays malloc_ipfb "
                 "allocated", who, ns, npol, vm->nfine_chan);
    if (npol < 1 || npol > 2) VM_FATAL("%s: %d polarisations", who, npol);
    const size_t hist_elems = (size_t)P * KCH * npol;
    const size_t need = 2 * hist_elems * sizeof(cf);
    if (d->ipfb_out_bytes < need) {
        if (d->d_ipfb_out) HIPCHECK(hipFree(d->d_ipfb_out));
        HIPCHECK(hipMalloc(&d->d_ipfb_out, need));
        HIPCHECK(hipMemsetAsync(d->d_ipfb_out, 0, need, s)); // same stream as its users
        d->ipfb_out_bytes = need;
    }
    cf *hist = (cf *)d->d_ipfb_out;
    job.e = (const cf *)vm->d_e;
    job.hist_in = hist + (size_t)((file_no + 1) & 1) * hist_elems;
    job.hist_out = hist + (size_t)(file_no & 1) * hist_elems;
    job.hK = g->d_ft_real;
    job.out = g->d_out;
    job.out8 = d_vdif;
}

// Passes per k_ipfb_run workgroup: 0 chooses them (ipfb_run_passes), else
// vmSetInversePFBPasses's value.
static int g_ipfb_npass = 0;

void vmSetInversePFBPasses(int npass)
{
    if (npass < 0 || npass > IPFB_MAX_PASSES)
        VM_FATAL("vmSetInversePFBPasses: %d passes (0 to %d)", npass, IPFB_MAX_PASSES);
    g_ipfb_npass = npass;
}

// Passes per run for njob channels of ns outputs when slots workgroups are
// resident at once: the fewest that minimise (rounds of resident
// workgroups) x (passes per workgroup), the longest chain of passes one slot
// runs.  A run's first pass gives 21 outputs, every later one 32, so long
// runs do less work per output, while short runs balance better over the
// slots (24 channels of 10000 samples on 512 slots: 8 passes, 41 runs per
// channel, 984 workgroups in two rounds).
static int ipfb_run_passes(int ns, int njob, int slots)
{
    int best = 1;
    long bestc = -1;
    for (int np = 1; np <= IPFB_MAX_PASSES; np++) {
        const long span = IPFB_RUN_T0 + (long)IPFB_DEV_ROWS * (np - 1);
        const long nwg = (long)njob * ((ns + span - 1) / span);
        const long c = ((nwg + slots - 1) / slots) * np;
        if (bestc < 0 || c < bestc) {
            bestc = c;
            best = np;
        }
    }
    return best;
}

// Launch over njob jobs (njob 1: the job by value; more: jobs_dev[y]).
static void ipfb_launch(const ipfb_job &job1, const ipfb_job *jobs_dev, int njob, int ns, int P,
                        int npol, int file_no, hipStream_t s)
{
    const int first = file_no == 0 ? 1 : 0;
    if (P == IPFB_RUN_P && npol == 2) {
        // k_ipfb_run addresses one channel's e and outputs with 32-bit offsets
        if ((size_t)ns * KCH * 16 >= (size_t)IPFB_NO_ROW)
            VM_FATAL("inverse PFB: %d samples per second exceed the 2 GB a channel may take", ns);
        static int slots = 0;
        if (!slots) {
            int dev = 0, ncu = 0, bpc = 0;
            HIPCHECK(hipGetDevice(&dev));
            HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_ipfb_run<true>, 512, 0));
            slots = ncu * (bpc < 1 ? 1 : bpc);
        }
        const int npass = g_ipfb_npass ? g_ipfb_npass : ipfb_run_passes(ns, njob, slots);
        const int span = IPFB_RUN_T0 + IPFB_DEV_ROWS * (npass - 1);
        const dim3 grid((unsigned)((ns + span - 1) / span), (unsigned)njob);
        if (njob == 1)
            hipLaunchKernelGGL((k_ipfb_run<false>), grid, dim3(512), 0, s, job1, jobs_dev, ns,
                               npass, first);
        else
            hipLaunchKernelGGL((k_ipfb_run<true>), grid, dim3(512), 0, s, job1, jobs_dev, ns,
                               npass, first);
    } else {
        const int tj = IPFB_DEV_ROWS - (P - 1);
        const dim3 grid((unsigned)((ns + tj - 1) / tj), (unsigned)njob);
        if (njob == 1)
            hipLaunchKernelGGL((k_ipfb_dev<false>), grid, dim3(256 * npol), 0, s, job1, jobs_dev,
                               ns, P, npol, first);
        else
            hipLaunchKernelGGL((k_ipfb_dev<true>), grid, dim3(256 * npol), 0, s, job1, jobs_dev,
                               ns, P, npol, first);
    }
    HIPCHECK(hipGetLastError());
}

static void ipfb_device(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                        float *data_buffer_vdif, uint8_t *d_vdif, const char *who)
{
    ensure_twiddles_i();
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    ipfb_job job;
    ipfb_job_for(vm, g, file_no, d_vdif, who, job);
    vm_timer_start(d, VM_T_IPFB);
    ipfb_launch(job, nullptr, 1, vm->fine_sample_rate, g->ntaps,
                (int)vm->obs_metadata->num_ant_pols, file_no, s);
    vm_timer_stop(d, VM_T_IPFB);
    if (data_buffer_vdif)
        HIPCHECK(hipMemcpyAsync(data_buffer_vdif, g->d_out, g->out_size, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
}

void vmInvertPFBDevice(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                       float *data_buffer_vdif)
{
    ipfb_device(vm, g, file_no, data_buffer_vdif, NULL, "vmInvertPFBDevice");
}

void vmInvertPFBDeviceVDIF(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                           uint8_t *d_vdif)
{
    if (!d_vdif) VM_FATAL("vmInvertPFBDeviceVDIF: no VDIF output buffer");
    ipfb_device(vm, g, file_no, NULL, d_vdif, "vmInvertPFBDeviceVDIF");
}

void vmInvertPFBDeviceVDIFBatch(vcsbeam_context **vm, struct gpu_ipfb_arrays **g, int nctx,
                                int file_no, uint8_t **d_vdif)
{
    const char *who = "vmInvertPFBDeviceVDIFBatch";
    if (nctx < 1 || !vm || !g || !d_vdif) VM_FATAL("%s: %d contexts", who, nctx);
    ensure_twiddles_i();
    const int ns = vm[0]->fine_sample_rate, P = g[0]->ntaps;
    const int npol = (int)vm[0]->obs_metadata->num_ant_pols;
    ipfb_job *jobs = (ipfb_job *)malloc(sizeof(ipfb_job) * nctx);
    if (!jobs) VM_FATAL("%s: out of host memory", who);
    for (int k = 0; k < nctx; k++) {
        if (!d_vdif[k]) VM_FATAL("%s: no VDIF output buffer for context %d", who, k);
        if (vm[k]->fine_sample_rate != ns || g[k]->ntaps != P ||
            (int)vm[k]->obs_metadata->num_ant_pols != npol)
            VM_FATAL("%s: context %d has %d samples, %d taps, %d pols; context 0 %d, %d, %d", who,
                     k, vm[k]->fine_sample_rate, g[k]->ntaps,
                     (int)vm[k]->obs_metadata->num_ant_pols, ns, P, npol);
        ipfb_job_for(vm[k], g[k], file_no, d_vdif[k], who, jobs[k]);
        // each context's e and history were written on its own stream
        HIPCHECK(hipStreamSynchronize((hipStream_t)vm_dev(vm[k])->stream));
    }
    vm_device_state *d = vm_dev(vm[0]);
    hipStream_t s = (hipStream_t)d->stream;
    const size_t jbytes = sizeof(ipfb_job) * (size_t)nctx;
    if (d->ipfb_jobs_bytes < jbytes) {
        if (d->d_ipfb_jobs) HIPCHECK(hipFree(d->d_ipfb_jobs));
        HIPCHECK(hipMalloc(&d->d_ipfb_jobs, jbytes));
        d->ipfb_jobs_bytes = jbytes;
    }
    HIPCHECK(hipMemcpyAsync(d->d_ipfb_jobs, jobs, jbytes, hipMemcpyHostToDevice, s));
    vm_timer_start(d, VM_T_IPFB);
    ipfb_launch(jobs[0], (const ipfb_job *)d->d_ipfb_jobs, nctx, ns, P, npol, file_no, s);
    vm_timer_stop(d, VM_T_IPFB);
    HIPCHECK(hipStreamSynchronize(s));
    free(jobs);
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


// ---------------------------------------------------------- VDIF output --

void vmVDIFFramesDevice(const float *d_in, size_t nframes, int payload_bytes, vdif_header *hdr,
                        int frame_rate, uint8_t *d_out, gpuStream_t stream)
{
    if (!hdr || frame_rate <= 0) VM_FATAL("vmVDIFFramesDevice: no header or frame rate %d", frame_rate);
    vm_timer_start_on(vm_dev_global(), VM_T_VDIF, stream);
    vdif_frames_launch(d_in, nframes, payload_bytes, hdr->w, (uint32_t)frame_rate, d_out,
                       (hipStream_t)stream);
    vm_timer_stop_on(vm_dev_global(), VM_T_VDIF, stream);
    vm_vdif_advance(hdr->w, nframes, (uint32_t)frame_rate);
}

void vmPopulateVDIFHeader(vcsbeam_context *vm, beam_geom *beam_geom_vals, double mjd_start,
                          double sec_offset)
{
    // src/beam_vdif.c:150-246
    if (!vm->obs_metadata || !vm->coarse_chan_idxs_to_process)
        VM_FATAL("vmPopulateVDIFHeader: no observation bound to the context");
    const CoarseChannel *cc = &vm->obs_metadata->metafits_coarse_chans[vm->coarse_chan_idxs_to_process[0]];
    if (vm->vf) {
        for (unsigned p = 0; p < vm->npointing; p++) {
            free(vm->vf[p].b_scales);
            free(vm->vf[p].b_offsets);
        }
        free(vm->vf);
    }
    vm->vf = (struct vdifinfo *)calloc(vm->npointing, sizeof(struct vdifinfo));
    if (!vm->vf) VM_FATAL("out of host memory");
    // UTC of the scheduled start (MJD 40587 = 1970-01-01)
    char time_utc[24];
    {
        const double days = vm->obs_metadata->sched_start_mjd - 40587.0;
        const time_t t = (time_t)llround(days * 86400.0);
        struct tm ts;
        gmtime_r(&t, &ts);
        strftime(time_utc, sizeof(time_utc), "%Y-%m-%dT%H:%M:%S", &ts);
    }
    for (unsigned p = 0; p < vm->npointing; p++) {
        struct vdifinfo *v = &vm->vf[p];
        v->bits = 8;
        v->iscomplex = 1;
        v->nchan = 2; // one channel per polarisation
        v->samples_per_frame = 128;
        v->sample_rate = vm->fine_sample_rate * vm->nfine_chan;
        v->BW = vm->obs_metadata->coarse_chan_width_hz / 1e6;
        v->dataarraylength = (int)(v->nchan * (v->iscomplex + 1) * v->samples_per_frame);
        v->frame_length = v->dataarraylength + VDIF_HEADER_BYTES;
        v->threadid = 0;
        strcpy(v->stationid, "mw");
        v->frame_rate = vm->fine_sample_rate;
        v->block_size = (size_t)v->frame_length * v->frame_rate;
        v->sizeof_beam = v->samples_per_frame * v->nchan * (v->iscomplex + 1);
        v->sizeof_buffer = (size_t)v->frame_rate * v->sizeof_beam;

        vm_vdif_create(vm->vhdr.w, v->dataarraylength, v->threadid, v->bits, v->nchan,
                       v->iscomplex, v->stationid);
        const uint64_t start_day = (uint64_t)beam_geom_vals->intmjd;
        const uint64_t start_sec = (uint64_t)roundf((float)(beam_geom_vals->fracmjd * 86400.0));
        vm_vdif_set_epoch_mjd(vm->vhdr.w, (int64_t)start_day);
        vm_vdif_set_mjdsec(vm->vhdr.w, (int64_t)(start_day * 86400 + start_sec));
        vm->vhdr.w[1] &= 0xFF000000u; // frame 0

        snprintf(v->exp_name, sizeof(v->exp_name), "%s",
                 vm->obs_metadata->project_id ? vm->obs_metadata->project_id : "");
        snprintf(v->scan_name, 17, "%d", (int)vm->obs_metadata->obs_id);
        v->b_scales = (float *)calloc(v->nchan, sizeof(float));
        v->b_offsets = (float *)calloc(v->nchan, sizeof(float));
        v->got_scales = 1;
        strncpy(v->telescope, "MWA", 24);
        strncpy(v->obs_mode, "PSR", 8);
        dec2hms(v->ra_str, beam_geom_vals[p].mean_ra * R2D / 15.0, 0);
        dec2hms(v->dec_str, beam_geom_vals[p].mean_dec * R2D, 1);
        strncpy(v->date_obs, time_utc, sizeof(v->date_obs));
        v->MJD_start = mjd_start;
        v->sec_offset = sec_offset;
        v->MJD_epoch = beam_geom_vals->intmjd + beam_geom_vals->fracmjd;
        v->fctr = cc->chan_centre_hz / 1e6;
        v->npol = 2;
        strncpy(v->source, "unset", 24);
        snprintf(v->basefilename, sizeof(v->basefilename), "%s_%s_%s_%s_ch%03ld", v->exp_name,
                 v->scan_name, v->ra_str, v->dec_str, (long)cc->rec_chan_number);
    }
}

void vdif_write_data(struct vdifinfo *vf, int8_t *output)
{
    // src/beam_vdif.c:98-143: append the frames; write the DSPSR ASCII
    // header (4096 bytes of "KEY value" lines, NUL-padded)
    char fname[1100];
    snprintf(fname, sizeof(fname), "%s.vdif", vf->basefilename);
    FILE *fs = fopen(fname, "a");
    if (!fs) VM_FATAL("vdif_write_data: cannot open %s", fname);
    if (fwrite(output, vf->block_size, 1, fs) != 1) VM_FATAL("vdif_write_data: short write to %s", fname);
    fclose(fs);
    char hdr[4096];
    memset(hdr, 0, sizeof(hdr));
    int n = 0;
#define VM_KV(k, fmt, ...)                                                                   \
    n += snprintf(hdr + n, sizeof(hdr) - (size_t)n, "%-12s " fmt "\n", k, __VA_ARGS__)
    VM_KV("HDR_VERSION", "%s", "0.2");
    VM_KV("MWA_CAPTURE_VERSION", "%s", "0.1");
    VM_KV("MWA_SAMPLE_VERSION", "%s", "0.1");
    VM_KV("VCSBEAM_VERSION", "%s", vmLibraryVersion());
    VM_KV("TELESCOPE", "%s", vf->telescope);
    VM_KV("MODE", "%s", vf->obs_mode);
    VM_KV("INSTRUMENT", "%s", "VDIF");
    VM_KV("DATAFILE", "%s", fname);
    VM_KV("MJD_START", "%f", vf->MJD_start);
    VM_KV("MJD_EPOCH", "%f", vf->MJD_epoch);
    VM_KV("SEC_OFFSET", "%f", vf->sec_offset);
    VM_KV("SOURCE", "%s", vf->source);
    VM_KV("RA", "%s", vf->ra_str);
    VM_KV("DEC", "%s", vf->dec_str);
    VM_KV("FREQ", "%f", vf->fctr);
    VM_KV("BW", "%f", vf->BW);
    VM_KV("TSAMP", "%f", 1e6 / (float)vf->sample_rate);
    VM_KV("NBIT", "%d", vf->bits);
    VM_KV("NDIM", "%d", vf->iscomplex + 1);
    VM_KV("NPOL", "%d", vf->npol);
#undef VM_KV
    if (n >= (int)sizeof(hdr)) VM_FATAL("vdif_write_data: header exceeds 4096 bytes");
    snprintf(fname, sizeof(fname), "%s.hdr", vf->basefilename);
    fs = fopen(fname, "w");
    if (!fs) VM_FATAL("vdif_write_data: cannot open %s", fname);
    if (fwrite(hdr, sizeof(hdr), 1, fs) != 1) VM_FATAL("vdif_write_data: short write to %s", fname);
    fclose(fs);
}

void vdif_write_second(struct vdifinfo *vf, vdif_header *vhdr, float *data_buffer_vdif)
{
    // src/beam_vdif.c:52-88, the framing on the GPU.  The reference clamps
    // data_buffer_vdif in place as a side effect; here it is left unchanged.
    if (vf->frame_length != vf->dataarraylength + VDIF_HEADER_BYTES ||
        (size_t)vf->frame_length * vf->frame_rate != vf->block_size ||
        (size_t)vf->dataarraylength != vf->sizeof_beam)
        VM_FATAL("vdif_write_second: inconsistent frame description");
    static float *d_in = NULL;
    static uint8_t *d_out = NULL;
    static size_t in_cap = 0, out_cap = 0;
    const size_t in_bytes = vf->sizeof_buffer * sizeof(float), out_bytes = vf->block_size;
    if (in_cap < in_bytes) {
        if (d_in) HIPCHECK(hipFree(d_in));
        HIPCHECK(hipMalloc((void **)&d_in, in_bytes));
        in_cap = in_bytes;
    }
    if (out_cap < out_bytes) {
        if (d_out) HIPCHECK(hipFree(d_out));
        HIPCHECK(hipMalloc((void **)&d_out, out_bytes));
        out_cap = out_bytes;
    }
    int8_t *frames = (int8_t *)malloc(out_bytes);
    if (!frames) VM_FATAL("out of host memory");
    HIPCHECK(hipMemcpy(d_in, data_buffer_vdif, in_bytes, hipMemcpyHostToDevice));
    vmVDIFFramesDevice(d_in, (size_t)vf->frame_rate, vf->dataarraylength, vhdr, vf->frame_rate,
                       d_out, NULL);
    HIPCHECK(hipMemcpy(frames, d_out, out_bytes, hipMemcpyDeviceToHost));
    vdif_write_data(vf, frames);
    free(frames);
}

} // extern "C"
