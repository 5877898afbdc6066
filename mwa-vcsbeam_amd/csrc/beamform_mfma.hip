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
// Precision form: voltages are converted to FP16 (or BF16) with a power-of-two
// scale per (channel, 32-sample block); weights carry a power-of-two scale per
// channel and are split W = W_hi + W_lo into two FP16 (BF16) parts, both
// products accumulating in FP32.  Phases and Jones inverses are produced in
// FP64 (geometry.hip) and demoted only at the split.  The beam-independent
// auto-correlation sums N_xx, N_yy, N_xy are formed once per (channel,
// sample) in FP32 from the FP32 voltages.
//
// Data flow per chunk:
//   k_bf_prep        FP32 voltages -> N terms, block scales and the FP16 B
//                    operand, written in MFMA fragment order (one pass).
//   k_beamform_mfma  the GEMM + Stokes epilogue.  Both operands are stored
//                    fragment-native (each 64-lane x 16-byte fragment is one
//                    contiguous kilobyte), so every wavefront loads its own
//                    fragments straight into registers with fully coalesced
//                    16-byte loads: no LDS staging, no bank conflicts, no
//                    conversion work in the GEMM.  A wavefront holds an
//                    (MB x 32) x (NB x 32) register tile; fragments of the
//                    next two k-steps are in flight while one is multiplied.
// Weights (k_mfma_weights) are rebuilt only when phases or Jones change.
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

#define KS 16   // k' (real) per MFMA step = 8 RF inputs
#define SB 32   // samples per B fragment block
#define RBK 32  // rows per A fragment block

struct cf { float x, y; };

template <bool BF16> struct mtype;
template <> struct mtype<false> {
    typedef half8 v8;
    static __device__ __forceinline__ uint16_t bits(float x)
    {
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    }
    static __device__ __forceinline__ float back(uint16_t b)
    {
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ f16v mfma(uint4 a, uint4 b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8, a),
                                                      __builtin_bit_cast(v8, b), c, 0, 0, 0);
    }
};
template <> struct mtype<true> {
    typedef bf8 v8;
    static __device__ __forceinline__ uint16_t bits(float x)
    {
        return __builtin_bit_cast(uint16_t, (__bf16)x);
    }
    static __device__ __forceinline__ float back(uint16_t b)
    {
        return (float)__builtin_bit_cast(__bf16, b);
    }
    static __device__ __forceinline__ f16v mfma(uint4 a, uint4 b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8, a),
                                                       __builtin_bit_cast(v8, b), c, 0, 0, 0);
    }
};

// exponent e with x * 2^e in [2^13, 2^14) for the largest component x > 0
__device__ __forceinline__ int scale_exp(float mx)
{
    int ex = 0;
    if (mx > 0.f) frexpf(mx, &ex);
    return 14 - ex;
}

// ------------------------------------------------------------ weights -----
// Operand A in fragment order: Af[c][rb][ks][h][lane] (uint4), rb a block of
// 32 GEMM rows, ks a step of 16 k', h = 0 hi part, 1 lo part.  Lane l holds
// row 32 rb + l%32 and k' = 16 ks + 8 (l/32) + j, j = 0..7, i.e. RF inputs
// i = 8 ks + 4 (l/32) + j/2 with j even the coefficient of v.re and j odd of
// v.im.  Rows: 4 beam + {Xre, Xim, Yre, Yim}; rows of beams >= npointing
// are zero.  One lane per (c, rb, ks, l); inv[i] = 2 antenna + pol of VCS
// input i; weights are scaled by 2^ew[c] (from the channel's largest
// |J^-1| component, wmax).  Lanes of beams p >= 1 also compare J^-1 with
// beam 0's and raise jflag on a difference (the shared auto-correlation
// terms need a beam-independent J^-1).
template <bool BF16>
__global__ __launch_bounds__(256) void k_mfma_weights(
    const double2 *__restrict__ J, const double2 *__restrict__ phi, const uint16_t *__restrict__ inv,
    const uint32_t *__restrict__ wmax, int npointing, int nant, int nchan, int nrb, int nkstep,
    uint4 *__restrict__ Af, uint32_t *__restrict__ jflag)
{
    typedef mtype<BF16> M;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nchan * nrb * nkstep * 64;
    if (j >= total) return;
    const int lane = (int)(j & 63);
    const size_t frag = j >> 6;
    const int ks = (int)(frag % nkstep);
    const int rb = (int)((frag / nkstep) % nrb);
    const int c = (int)(frag / ((size_t)nkstep * nrb));
    const int row = RBK * rb + (lane & 31);
    const int p = row >> 2, comp = row & 3; // comp: 0 Xre, 1 Xim, 2 Yre, 3 Yim
    const double ws = ldexp(1.0, scale_exp(__uint_as_float(wmax[c])));
    uint32_t hi[4], lo[4];
    bool differs = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint16_t hb[2] = {0, 0}, lb[2] = {0, 0};
        if (p < npointing) {
            const int code = inv[8 * ks + 4 * (lane >> 5) + q];
            const int a = code >> 1, pol = code & 1; // pol 0: Q input, 1: P input
            const size_t pac = ((size_t)p * nant + a) * nchan + c;
            const double2 f = phi[pac];
            const double2 m = J[pac * 4 + 2 * (comp >> 1) + pol]; // J_x0/J_x1 or J_y0/J_y1
            if (p > 0) {
                const double2 m0 = J[((size_t)a * nchan + c) * 4 + 2 * (comp >> 1) + pol];
                differs |= (m.x != m0.x) || (m.y != m0.y);
            }
            const double wr = (f.x * m.x - f.y * m.y) * ws, wi = (f.x * m.y + f.y * m.x) * ws;
            // Xre/Yre rows: (Re w, -Im w) against (v.re, v.im); Xim/Yim: (Im w, Re w)
            const double vv[2] = {(comp & 1) ? wi : wr, (comp & 1) ? wr : -wi};
#pragma unroll
            for (int e2 = 0; e2 < 2; e2++) {
                hb[e2] = M::bits((float)vv[e2]);
                lb[e2] = M::bits((float)(vv[e2] - (double)M::back(hb[e2])));
            }
        }
        hi[q] = (uint32_t)hb[0] | ((uint32_t)hb[1] << 16);
        lo[q] = (uint32_t)lb[0] | ((uint32_t)lb[1] << 16);
    }
    uint4 *dst = Af + frag * 128 + lane;
    dst[0] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    dst[64] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    if (differs) atomicOr(jflag, 1u);
}

// inv[vcs input] = 2*antenna + pol from the P/Q index lists
__global__ void k_inv_pol(const uint32_t *__restrict__ polQ, const uint32_t *__restrict__ polP,
                          int nant, uint16_t *__restrict__ inv)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nant) return;
    inv[polQ[a]] = (uint16_t)(2 * a);
    inv[polP[a]] = (uint16_t)(2 * a + 1);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

// J^-1 of pointing 0 as float [c][a][4] and the channel's largest |component|
// (float bits, atomicMax: non-negative floats order like their bit patterns)
__global__ void k_jf(const double2 *__restrict__ J, int nant, int nchan, cf *__restrict__ Jf,
                     uint32_t *__restrict__ wmax)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)nant * nchan) return;
    int c = (int)(j % nchan), a = (int)(j / nchan);
    double mx = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double2 m = J[((size_t)a * nchan + c) * 4 + q];
        Jf[((size_t)c * nant + a) * 4 + q] = cf{(float)m.x, (float)m.y};
        mx = fmax(mx, fmax(fabs(m.x), fabs(m.y)));
    }
    atomicMax(&wmax[c], __float_as_uint((float)mx));
}

// ------------------------------------------------------- voltage prep -----
// One workgroup per (32-sample block sb, channel c), one pass over the block:
// lane (sl, g) loads sample 32 sb + sl of both inputs of antennas g, g+8, ...
// (APT antennas at most) into registers, forms the auto-correlation sums
// with Jf = J^-1 (float) and the block's largest |component|.  The 8 partial
// sums per sample are reduced in LDS and written as N[c][s] = (Nxx, Nyy,
// Re Nxy, Im Nxy); the voltages, scaled by 2^es, are placed in LDS in
// fragment order and written out as B fragments: Bf[c][sb][ks][lane]
// (uint4), lane l holding sample 32 sb + l%32 and inputs 8 ks + 4 (l/32) +
// {0..3} as (re, im) pairs.  Samples >= ns are zero.  sexp[c][sb] = es.
// v: VM_FLT [c][i][s].  Needs nant <= 8 APT.
template <bool BF16, int APT>
__global__ __launch_bounds__(256) void k_bf_prep(const cf *__restrict__ v, int ns, int ni, int nsb,
                                                 int nkstep, const cf *__restrict__ Jf,
                                                 const uint32_t *__restrict__ polQ,
                                                 const uint32_t *__restrict__ polP, int nant,
                                                 uint4 *__restrict__ Bf, int *__restrict__ sexp,
                                                 float4 *__restrict__ N)
{
    typedef mtype<BF16> M;
    __shared__ float4 red[8][SB];
    __shared__ float wmx[4];
    __shared__ uint4 frag[2 * APT * 64]; // nkstep <= 2 APT fragments of 64 lanes
    const int sb = blockIdx.x, c = blockIdx.y;
    const int t = threadIdx.x, sl = t & 31, g = t >> 5;
    const int s = SB * sb + sl;
    const bool ok = s < ns;
    const cf *vc = v + (size_t)c * ni * ns;

    for (int k = t; k < nkstep * 64; k += 256) frag[k] = make_uint4(0u, 0u, 0u, 0u);

    cf q[APT], p[APT];
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = g + 8 * j;
        q[j] = p[j] = cf{0.f, 0.f};
        if (ok && a < nant) {
            q[j] = vc[(size_t)polQ[a] * ns + s];
            p[j] = vc[(size_t)polP[a] * ns + s];
        }
    }
    float nxx = 0.f, nyy = 0.f, xyr = 0.f, xyi = 0.f, mx = 0.f;
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = g + 8 * j;
        if (a >= nant) break;
        const cf *w = Jf + ((size_t)c * nant + a) * 4;
        const cf w00 = w[0], w01 = w[1], w10 = w[2], w11 = w[3];
        const cf qq = q[j], pp = p[j];
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(qq.x), fabsf(qq.y)), fmaxf(fabsf(pp.x), fabsf(pp.y))));
        const float axr = w00.x * qq.x - w00.y * qq.y + w01.x * pp.x - w01.y * pp.y;
        const float axi = w00.x * qq.y + w00.y * qq.x + w01.x * pp.y + w01.y * pp.x;
        const float ayr = w10.x * qq.x - w10.y * qq.y + w11.x * pp.x - w11.y * pp.y;
        const float ayi = w10.x * qq.y + w10.y * qq.x + w11.x * pp.y + w11.y * pp.x;
        nxx += axr * axr + axi * axi;
        nyy += ayr * ayr + ayi * ayi;
        xyr += axr * ayr + axi * ayi;
        xyi += axi * ayr - axr * ayi;
    }
    red[g][sl] = make_float4(nxx, nyy, xyr, xyi);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if ((t & 63) == 0) wmx[t >> 6] = mx;
    __syncthreads();
    if (g == 0 && ok) {
        float4 acc = red[0][sl];
#pragma unroll
        for (int k = 1; k < 8; k++) {
            const float4 r = red[k][sl];
            acc.x += r.x; acc.y += r.y; acc.z += r.z; acc.w += r.w;
        }
        N[(size_t)c * ns + s] = acc;
    }
    const int es = scale_exp(fmaxf(fmaxf(wmx[0], wmx[1]), fmaxf(wmx[2], wmx[3])));
    if (t == 0) sexp[(size_t)c * nsb + sb] = es;
    const float scale = ldexpf(1.0f, es);

    // place (re, im) of input i, sample sl at dword (i%4) of lane
    // sl + 32 ((i%8)/4) of fragment i/8
    uint32_t *fd = (uint32_t *)frag;
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = g + 8 * j;
        if (a >= nant) break;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = (int)(h ? polP[a] : polQ[a]);
            const cf x = h ? p[j] : q[j];
            fd[(i >> 3) * 256 + (sl + 32 * ((i >> 2) & 1)) * 4 + (i & 3)] =
                (uint32_t)M::bits(x.x * scale) | ((uint32_t)M::bits(x.y * scale) << 16);
        }
    }
    __syncthreads();
    uint4 *dst = Bf + ((size_t)c * nsb + sb) * nkstep * 64;
    for (int k = t; k < nkstep * 64; k += 256) dst[k] = frag[k];
}

// ------------------------------------------------------------- GEMM -------
// Fragments are fetched with buffer loads: the per-channel operand base and
// every wave-uniform offset live in scalar registers, the only vector
// address is the lane's 16-byte slot.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
#define RSRC_DWORD3 0x00020000 /* raw buffer, 32-bit data format (gfx9 family) */

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             RSRC_DWORD3);
}

__device__ __forceinline__ uint4 bload(rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

template <int MB, int NB> struct stage {
    uint4 a[MB][2];
    uint4 b[NB];
};

// Operand offsets (bytes): A block m of this wavefront at aoff + m astr,
// k-step ks at + 2048 ks (hi) and + 1024 more (lo); B block n at boff[n],
// k-step ks at + 1024 ks.
template <int MB, int NB>
__device__ __forceinline__ void bf_load(stage<MB, NB> &f, rsrc_t rA, rsrc_t rB, int vl, int aoff,
                                        int astr, const int (&boff)[NB], int ks)
{
#pragma unroll
    for (int m = 0; m < MB; m++) {
        f.a[m][0] = bload(rA, vl, aoff + m * astr + 2048 * ks);
        f.a[m][1] = bload(rA, vl, aoff + m * astr + 2048 * ks + 1024);
    }
#pragma unroll
    for (int n = 0; n < NB; n++) f.b[n] = bload(rB, vl, boff[n] + 1024 * ks);
}

// Scheduling pattern for one k-step: each of the NL fragment loads is
// issued between MFMAs (one load, then NM/NL MFMAs), so the in-order issue
// of the loads never holds the matrix pipe idle.
template <int NL, int NM>
__device__ __forceinline__ void interleave_loads()
{
#pragma unroll
    for (int i = 0; i < NL; i++) {
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);       // one VMEM read
        __builtin_amdgcn_sched_group_barrier(0x008, NM / NL, 0); // then MFMAs
    }
    if constexpr (NM % NL != 0) __builtin_amdgcn_sched_group_barrier(0x008, NM % NL, 0);
}

template <bool BF16, int MB, int NB>
__device__ __forceinline__ void bf_mma(f16v (&acc)[MB][NB], const stage<MB, NB> &f)
{
    // independent hi products first, then the lo products
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) acc[m][n] = mtype<BF16>::mfma(f.a[m][0], f.b[n], acc[m][n]);
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) acc[m][n] = mtype<BF16>::mfma(f.a[m][1], f.b[n], acc[m][n]);
}

// Stokes (I, Q, U, V) of the beam in accumulator rows 4 jg .. 4 jg + 3 =
// (Xre, Xim, Yre, Yim): |E|^2 terms carry the operand scales squared
// (unscale2 removes them), the auto-correlation terms nn are unscaled.
__device__ __forceinline__ void stokes_of(const f16v &a, int jg, float unscale2, float4 nn,
                                          float invw, float (&st)[4])
{
    const float xr = a[4 * jg + 0], xi = a[4 * jg + 1];
    const float yr = a[4 * jg + 2], yi = a[4 * jg + 3];
    const float bnXX = (xr * xr + xi * xi) * unscale2 - nn.x;
    const float bnYY = (yr * yr + yi * yi) * unscale2 - nn.y;
    const float bxyr = (xr * yr + xi * yi) * unscale2 - nn.z;
    const float bxyi = (xi * yr - xr * yi) * unscale2 - nn.w;
    st[0] = invw * (bnXX + bnYY);
    st[1] = invw * (bnXX - bnYY);
    st[2] = 2.0f * invw * bxyr;
    st[3] = -2.0f * invw * bxyi;
}

// 1-D grid over (channel, tile); a workgroup of WR x WC wavefronts covers
// (WR MB 32) rows x (WC NB 32) samples of one channel; NST register stages
// keep NST-1 k-steps of fragments in flight while one is multiplied.
// Tile order: order 0 walks the channels in turn, row tiles fastest.  Orders
// 1-4 are XCD-aware: the 8 workgroups dispatched together (one per XCD) take
// the same tile of 8 different channels, so each XCD walks its own channels;
// within a channel tiles go in bands of gr row tiles, column by column (gr =
// all row tiles for order 1, 1 for order 2, 2 for order 3, 4 for order 4), so
// the workgroups resident on an XCD at one time share a small band of weight
// rows and a small run of voltage columns in its L2.
// Epilogue EPI 0 stores each Stokes value from the accumulator lane that
// holds it (4-byte stores, 32 samples per row); EPI 1 passes each 32 x 32
// block through a 4 KB per-wavefront LDS tile so that every lane stores four
// consecutive samples of one (beam, Stokes) row with one 16-byte store.  OCC
// is the minimum number of wavefronts per SIMD the registers must allow.
// S: [p][st][c][s_total] (device layout), sample offset soffset.
template <bool BF16, int MB, int NB, int WR, int WC, int NST, int EPI, int OCC>
__global__ __launch_bounds__(64 * WR * WC, OCC) void k_beamform_mfma(
    const uint4 *__restrict__ Bf, const uint4 *__restrict__ Af, const int *__restrict__ sexp,
    const uint32_t *__restrict__ wmax, const float4 *__restrict__ N, int ns, int nsb, int nkstep,
    int nrb, int npointing, int nchan, int ntx, int nty, int order, float invw, int nstokes,
    int soffset, int ns_total, float *__restrict__ S)
{
    const int ntile = ntx * nty;
    const int L = blockIdx.x;
    int c, tile;
    if (order != 0 && (nchan & 7) == 0) {
        const int g = L >> 3, x = L & 7;
        c = (g / ntile) * 8 + x;
        tile = g % ntile;
    } else {
        c = L / ntile;
        tile = L % ntile;
    }
    int gr = order == 2 ? 1 : order == 3 ? 2 : order == 4 ? 4 : nty;
    if (nty % gr) gr = 1;
    const int band = tile / (gr * ntx), rem = tile % (gr * ntx);
    const int tx = rem / gr, ty = band * gr + rem % gr;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = w / WC, wc = w % WC;
    const int rb0 = (ty * WR + wr) * MB;
    const int sb0 = (tx * WC + wc) * NB;

    const int astr = nkstep * 2048;              // bytes per 32-row block
    const int bstr = nkstep * 1024;              // bytes per 32-sample block
    const rsrc_t rA = make_rsrc(Af + (size_t)c * nrb * nkstep * 128, (uint32_t)nrb * astr);
    const rsrc_t rB = make_rsrc(Bf + (size_t)c * nsb * nkstep * 64, (uint32_t)nsb * bstr);
    const int vl = lane * 16;
    const int aoff = rb0 * astr;
    int boff[NB];
#pragma unroll
    for (int n = 0; n < NB; n++) {
        // blocks past the last sample block re-read it (in bounds); their
        // columns are never stored
        boff[n] = min(sb0 + n, nsb - 1) * bstr;
    }

    f16v acc[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[m][n][q] = 0.f;

    // The scheduling barriers pin each stage's loads ahead of the previous
    // stage's MFMAs, so the loads stay NST-1 k-steps ahead of their use.
    // Steady state has no branches; loads past the last k-step re-read it.
#define VM_SB() __builtin_amdgcn_sched_barrier(0)
#define VM_LD(f, k) bf_load<MB, NB>(f, rA, rB, vl, aoff, astr, boff, min((k), kl))
#define VM_MM(f) bf_mma<BF16, MB, NB>(acc, f)
#define VM_STEP(fl, k, fm)                                                                    \
    VM_SB();                                                                                  \
    VM_LD(fl, k);                                                                             \
    VM_MM(fm);                                                                                \
    interleave_loads<2 * MB + NB, 2 * MB * NB>();                                            \
    VM_SB()
    stage<MB, NB> f0, f1, f2, f3;
    const int kl = nkstep - 1;
    int ks = 0;
    if constexpr (NST == 4) {
        VM_LD(f0, 0); VM_LD(f1, 1); VM_LD(f2, 2);
        for (; ks + 4 <= nkstep; ks += 4) {
            VM_STEP(f3, ks + 3, f0);
            VM_STEP(f0, ks + 4, f1);
            VM_STEP(f1, ks + 5, f2);
            VM_STEP(f2, ks + 6, f3);
        }
        VM_SB();
        if (ks < nkstep) VM_MM(f0);
        if (ks + 1 < nkstep) VM_MM(f1);
        if (ks + 2 < nkstep) VM_MM(f2);
    } else if constexpr (NST == 3) {
        VM_LD(f0, 0); VM_LD(f1, 1);
        for (; ks + 3 <= nkstep; ks += 3) {
            VM_STEP(f2, ks + 2, f0);
            VM_STEP(f0, ks + 3, f1);
            VM_STEP(f1, ks + 4, f2);
        }
        VM_SB();
        if (ks < nkstep) VM_MM(f0);
        if (ks + 1 < nkstep) VM_MM(f1);
    } else {
        VM_LD(f0, 0);
        for (; ks + 2 <= nkstep; ks += 2) {
            VM_STEP(f1, ks + 1, f0);
            VM_STEP(f0, ks + 2, f1);
        }
        VM_SB();
        if (ks < nkstep) VM_MM(f0);
    }
#undef VM_STEP
#undef VM_MM
#undef VM_LD
#undef VM_SB

    // epilogue: lane holds rows 8*jg + 4*(lane/32) + {0..3} of each 32-row
    // block = one beam's (Xre, Xim, Yre, Yim), column lane%32
    const int ew = scale_exp(__uint_as_float(wmax[c]));
    const int lr = lane & 31, lk = lane >> 5;
    const size_t pstride = (size_t)nstokes * nchan * ns_total;
    const size_t ststride = (size_t)nchan * ns_total;
    if constexpr (EPI == 0) {
#pragma unroll
        for (int n = 0; n < NB; n++) {
            const int sb = sb0 + n;
            const int s = SB * sb + lr;
            if (s >= ns) continue;
            const float4 nn = N[(size_t)c * ns + s];
            const float unscale2 = ldexpf(1.0f, -2 * (sexp[(size_t)c * nsb + sb] + ew));
            float *Sc = S + (size_t)c * ns_total + s + soffset;
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int jg = 0; jg < 4; jg++) {
                    const int p = 8 * (rb0 + m) + 2 * jg + lk;
                    if (p >= npointing) continue;
                    float st4[4];
                    stokes_of(acc[m][n], jg, unscale2, nn, invw, st4);
                    float *Sp = Sc + (size_t)p * pstride;
                    Sp[0] = st4[0];
                    if (nstokes == 4) {
                        Sp[ststride] = st4[1];
                        Sp[2 * ststride] = st4[2];
                        Sp[3 * ststride] = st4[3];
                    }
                }
        }
    } else {
        // LDS rows r = (st*4 + jg)*2 + lk of 32 samples: the two lane halves
        // write rows 32 floats apart and each 16-lane read group covers two
        // whole rows, so neither side has bank conflicts
        __shared__ __attribute__((aligned(16))) float stg[WR * WC][32 * 32];
        float *wl = stg[w];
        const bool vec = ((ns_total | soffset) & 3) == 0;
        const int q = lane & 7;
#pragma unroll
        for (int n = 0; n < NB; n++) {
            const int sb = sb0 + n;
            const int s = SB * sb + lr;
            float4 nn = make_float4(0.f, 0.f, 0.f, 0.f);
            if (s < ns) nn = N[(size_t)c * ns + s];
            const float unscale2 = ldexpf(1.0f, -2 * (sexp[(size_t)c * nsb + min(sb, nsb - 1)] + ew));
            const int s4 = SB * sb + 4 * q;
            float *Sc = S + (size_t)c * ns_total + soffset + s4;
#pragma unroll
            for (int m = 0; m < MB; m++) {
#pragma unroll
                for (int jg = 0; jg < 4; jg++) {
                    float st4[4];
                    stokes_of(acc[m][n], jg, unscale2, nn, invw, st4);
#pragma unroll
                    for (int st = 0; st < 4; st++) wl[((st * 4 + jg) * 2 + lk) * 32 + lr] = st4[st];
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = 8 * i + (lane >> 3);
                    const int st = r >> 3, bb = 2 * ((r >> 1) & 3) + (r & 1);
                    const int p = 8 * (rb0 + m) + bb;
                    const float4 val = *(const float4 *)(wl + r * 32 + 4 * q);
                    if (st >= nstokes || p >= npointing || s4 >= ns) continue;
                    float *dst = Sc + (size_t)p * pstride + (size_t)st * ststride;
                    if (vec && s4 + 3 < ns) {
                        *(float4 *)dst = val;
                    } else {
                        dst[0] = val.x;
                        if (s4 + 1 < ns) dst[1] = val.y;
                        if (s4 + 2 < ns) dst[2] = val.z;
                        if (s4 + 3 < ns) dst[3] = val.w;
                    }
                }
            }
        }
    }
}

// ------------------------------------------------- persistent GEMM ------
// k_beamform_mfma_pt: one workgroup per CU walks its share of the tiles
// (L = blockIdx.x, + gridDim.x, ...), decoded in the same orders as above.
// Each wavefront keeps two accumulator sets: the k-loop of tile i runs in
// `acc` while the Stokes epilogue of tile i-1 drains `accp`, one 32 x 32 block
// per four k-steps, its VALU, LDS and store instructions placed between the
// MFMAs of that k-step so they issue in the matrix pipe's shadow.  The last
// k-step of a tile writes its lo products straight into `accp`.  Fragments
// go through a 4-stage register ring whose last loads of each tile already
// fetch the next tile's first three k-steps.  Needs nkstep % 4 == 0 and
// nkstep / 4 >= MB NB (each block's epilogue gets a k-step group of its own).
__device__ __forceinline__ void bf_decode_tile(int L, int nchan, int ntx, int nty, int order,
                                               int &c, int &tx, int &ty)
{
    const int ntile = ntx * nty;
    int tile;
    if (order != 0 && (nchan & 7) == 0) {
        const int g = L >> 3, x = L & 7;
        c = (g / ntile) * 8 + x;
        tile = g % ntile;
    } else {
        c = L / ntile;
        tile = L % ntile;
    }
    int gr = order == 2 ? 1 : order == 3 ? 2 : order == 4 ? 4 : nty;
    if (nty % gr) gr = 1;
    const int band = tile / (gr * ntx), rem = tile % (gr * ntx);
    tx = rem / gr;
    ty = band * gr + rem % gr;
}

// where a wavefront's operands of one tile are
template <int NB> struct bf_tile {
    int c, rb0, sb0, aoff;
    int boff[NB];
    rsrc_t rA, rB;
};

template <int MB, int NB, int WR, int WC>
__device__ __forceinline__ void bf_tile_at(bf_tile<NB> &T, int L, int w, const uint4 *Af,
                                           const uint4 *Bf, int nchan, int ntx, int nty,
                                           int order, int nrb, int nsb, int nkstep)
{
    int c, tx, ty;
    bf_decode_tile(L, nchan, ntx, nty, order, c, tx, ty);
    const int wr = w / WC, wc = w % WC;
    const int astr = nkstep * 2048, bstr = nkstep * 1024;
    T.c = c;
    T.rb0 = (ty * WR + wr) * MB;
    T.sb0 = (tx * WC + wc) * NB;
    T.aoff = T.rb0 * astr;
#pragma unroll
    for (int n = 0; n < NB; n++) T.boff[n] = min(T.sb0 + n, nsb - 1) * bstr;
    T.rA = make_rsrc(Af + (size_t)c * nrb * nkstep * 128, (uint32_t)nrb * astr);
    T.rB = make_rsrc(Bf + (size_t)c * nsb * nkstep * 64, (uint32_t)nsb * bstr);
}

// what the epilogue of a finished tile needs besides its accumulators
template <int NB> struct bf_epi {
    int c, rb0, sb0;
    float4 nn[NB]; // auto-correlation terms at this lane's sample of each block
    float us[NB];  // operand scales removed from |E|^2
};

template <int NB>
__device__ __forceinline__ void bf_epi_load(bf_epi<NB> &E, const bf_tile<NB> &T, int lane,
                                            const float4 *__restrict__ N,
                                            const int *__restrict__ sexp,
                                            const uint32_t *__restrict__ wmax, int ns, int nsb)
{
    E.c = T.c;
    E.rb0 = T.rb0;
    E.sb0 = T.sb0;
    const int ew = scale_exp(__uint_as_float(wmax[T.c]));
#pragma unroll
    for (int n = 0; n < NB; n++) {
        const int sb = T.sb0 + n, s = SB * sb + (lane & 31);
        E.nn[n] = s < ns ? N[(size_t)T.c * ns + s] : make_float4(0.f, 0.f, 0.f, 0.f);
        E.us[n] = ldexpf(1.0f, -2 * (sexp[(size_t)T.c * nsb + min(sb, nsb - 1)] + ew));
    }
}

struct bf_out {
    float *S;
    size_t pstride, ststride;
    int ns, ns_total, soffset, npointing, nstokes;
    float invw;
    bool vec;
};

// Part q (0..7) of the epilogue of accumulator block (m, n): parts 0-3 form
// the Stokes of beam row group jg = q and write them to the wavefront's LDS
// tile; part 4 reads the 4 x 16 bytes this lane stores; parts 5-7 store them.
template <int NB>
__device__ __forceinline__ void bf_epi_part(int q, int m, int n, const f16v &a,
                                            const bf_epi<NB> &E, float *wl, int lane,
                                            const bf_out &O, float4 (&v)[4])
{
    const int lr = lane & 31, lk = lane >> 5, qq = lane & 7;
    if (q < 4) {
        float st4[4];
        stokes_of(a, q, E.us[n], E.nn[n], O.invw, st4);
#pragma unroll
        for (int st = 0; st < 4; st++) wl[((st * 4 + q) * 2 + lk) * 32 + lr] = st4[st];
    } else if (q == 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = *(const float4 *)(wl + (8 * i + (lane >> 3)) * 32 + 4 * qq);
    } else {
        const int i0 = q == 5 ? 0 : q == 6 ? 2 : 3, i1 = q == 5 ? 2 : q == 6 ? 3 : 4;
#pragma unroll
        for (int i = i0; i < i1; i++) {
            const int r = 8 * i + (lane >> 3);
            const int st = r >> 3, bb = 2 * ((r >> 1) & 3) + (r & 1);
            const int p = 8 * (E.rb0 + m) + bb;
            const int s4 = SB * (E.sb0 + n) + 4 * qq;
            if (st < O.nstokes && p < O.npointing && s4 < O.ns) {
                float *dst = O.S + (size_t)p * O.pstride + (size_t)st * O.ststride +
                             (size_t)E.c * O.ns_total + O.soffset + s4;
                if (O.vec && s4 + 3 < O.ns) {
                    *(float4 *)dst = v[i];
                } else {
                    dst[0] = v[i].x;
                    if (s4 + 1 < O.ns) dst[1] = v[i].y;
                    if (s4 + 2 < O.ns) dst[2] = v[i].z;
                    if (s4 + 3 < O.ns) dst[3] = v[i].w;
                }
            }
        }
    }
}

// one k-step of MFMAs with the epilogue of block (m0, n0) spread between them
template <bool BF16, int MB, int NB>
__device__ __forceinline__ void bf_mma_epi(f16v (&acc)[MB][NB], const stage<MB, NB> &f,
                                           const f16v &pa, int m0, int n0, const bf_epi<NB> &E,
                                           float *wl, int lane, const bf_out &O)
{
    constexpr int NM = 2 * MB * NB, PER = NM / 8;
    float4 v[4];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++) {
                acc[m][n] = mtype<BF16>::mfma(f.a[m][h], f.b[n], acc[m][n]);
                const int idx = (h * MB + m) * NB + n;
                if ((idx + 1) % PER == 0) {
                    __builtin_amdgcn_sched_barrier(0);
                    bf_epi_part<NB>((idx + 1) / PER - 1, m0, n0, pa, E, wl, lane, O, v);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
}

// last k-step of a tile: hi products into acc, lo products into accp
template <bool BF16, int MB, int NB>
__device__ __forceinline__ void bf_mma_final(f16v (&acc)[MB][NB], f16v (&accp)[MB][NB],
                                             const stage<MB, NB> &f)
{
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) acc[m][n] = mtype<BF16>::mfma(f.a[m][0], f.b[n], acc[m][n]);
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) accp[m][n] = mtype<BF16>::mfma(f.a[m][1], f.b[n], acc[m][n]);
}

template <bool BF16, int MB, int NB, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, 1) void k_beamform_mfma_pt(
    const uint4 *__restrict__ Bf, const uint4 *__restrict__ Af, const int *__restrict__ sexp,
    const uint32_t *__restrict__ wmax, const float4 *__restrict__ N, int ns, int nsb, int nkstep,
    int nrb, int npointing, int nchan, int ntx, int nty, int order, float invw, int nstokes,
    int soffset, int ns_total, float *__restrict__ S)
{
    constexpr int NBLK = MB * NB;
    __shared__ __attribute__((aligned(16))) float stg[WR * WC][32 * 32];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float *wl = stg[w];
    const int ntiles = nchan * ntx * nty;
    const int astr = nkstep * 2048;
    const int vl = lane * 16;
    const int ngroup = nkstep >> 2;
    bf_out O;
    O.S = S;
    O.pstride = (size_t)nstokes * nchan * ns_total;
    O.ststride = (size_t)nchan * ns_total;
    O.ns = ns;
    O.ns_total = ns_total;
    O.soffset = soffset;
    O.npointing = npointing;
    O.nstokes = nstokes;
    O.invw = invw;
    O.vec = ((ns_total | soffset) & 3) == 0;

    int t = blockIdx.x;
    if (t >= ntiles) return;
    bf_tile<NB> cur, nxt;
    bf_tile_at<MB, NB, WR, WC>(cur, t, w, Af, Bf, nchan, ntx, nty, order, nrb, nsb, nkstep);
    f16v acc[MB][NB], accp[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++)
#pragma unroll
            for (int q = 0; q < 16; q++) accp[m][n][q] = 0.f;
    stage<MB, NB> f0, f1, f2, f3;
    // the first tile has no predecessor: its epilogue steps run with no
    // pointing to store (npointing 0) instead of taking another code path
    bf_epi<NB> E;
    E.c = E.rb0 = E.sb0 = 0;
#pragma unroll
    for (int n = 0; n < NB; n++) {
        E.nn[n] = make_float4(0.f, 0.f, 0.f, 0.f);
        E.us[n] = 0.f;
    }
    O.npointing = 0;

#define VM_SB() __builtin_amdgcn_sched_barrier(0)
#define VM_LD(f, T, k) bf_load<MB, NB>(f, T.rA, T.rB, vl, T.aoff, astr, T.boff, (k))
#define VM_MM(f) bf_mma<BF16, MB, NB>(acc, f)
#define VM_MME(f, B) bf_mma_epi<BF16, MB, NB>(acc, f, accp[(B) / NB][(B) % NB], (B) / NB, (B) % NB, E, wl, lane, O)
    VM_LD(f0, cur, 0);
    VM_LD(f1, cur, 1);
    VM_LD(f2, cur, 2);
    for (;;) {
        const int tn = t + (int)gridDim.x;
        const bool more = tn < ntiles;
        bf_tile_at<MB, NB, WR, WC>(nxt, more ? tn : t, w, Af, Bf, nchan, ntx, nty, order, nrb,
                                   nsb, nkstep);
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[m][n][q] = 0.f;
        // groups 0 .. NBLK-2 carry the epilogue of blocks 0 .. NBLK-2
#pragma unroll
        for (int g = 0; g < NBLK - 1; g++) {
            const int ks = 4 * g;
            VM_SB(); VM_LD(f3, cur, ks + 3); VM_MME(f0, g); VM_SB();
            VM_SB(); VM_LD(f0, cur, ks + 4); VM_MM(f1); VM_SB();
            VM_SB(); VM_LD(f1, cur, ks + 5); VM_MM(f2); VM_SB();
            VM_SB(); VM_LD(f2, cur, ks + 6); VM_MM(f3); VM_SB();
        }
        for (int g = NBLK - 1; g < ngroup - 1; g++) {
            const int ks = 4 * g;
            VM_SB(); VM_LD(f3, cur, ks + 3); VM_MM(f0); VM_SB();
            VM_SB(); VM_LD(f0, cur, ks + 4); VM_MM(f1); VM_SB();
            VM_SB(); VM_LD(f1, cur, ks + 5); VM_MM(f2); VM_SB();
            VM_SB(); VM_LD(f2, cur, ks + 6); VM_MM(f3); VM_SB();
        }
        // last group: the last block's epilogue, and the ring moves on to the
        // next tile's first k-steps
        VM_SB(); VM_LD(f3, cur, nkstep - 1); VM_MME(f0, NBLK - 1); VM_SB();
        VM_SB(); VM_LD(f0, nxt, 0); VM_MM(f1); VM_SB();
        VM_SB(); VM_LD(f1, nxt, 1); VM_MM(f2); VM_SB();
        VM_SB(); VM_LD(f2, nxt, 2); bf_mma_final<BF16, MB, NB>(acc, accp, f3); VM_SB();
        bf_epi_load<NB>(E, cur, lane, N, sexp, wmax, ns, nsb);
        O.npointing = npointing;
        if (!more) break;
        cur = nxt;
        t = tn;
    }
#undef VM_MME
#undef VM_MM
#undef VM_LD
#undef VM_SB
    // drain the last tile's epilogue
    float4 v[4];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++)
#pragma unroll
            for (int q = 0; q < 8; q++) bf_epi_part<NB>(q, m, n, accp[m][n], E, wl, lane, O, v);
}

// Kernel variants the autotuner chooses from: register tile per wavefront
// (MB x NB blocks of 32 x 32), wavefront grid per workgroup (WR x WC) and
// register stages.
// pers = 1: the persistent kernel with the overlapped epilogue (4 stages).
struct bf_variant { int mb, nb, wr, wc, nst, epi, occ, pers; };
static const bf_variant kVariants[] = {
    {4, 4, 2, 2, 3, 1, 1, 0}, {4, 4, 1, 4, 3, 1, 1, 0}, {4, 4, 2, 2, 4, 1, 1, 0},
    {4, 2, 2, 2, 2, 1, 2, 0}, {2, 4, 2, 2, 2, 1, 2, 0}, {2, 4, 4, 2, 2, 1, 1, 0},
    {4, 4, 2, 2, 3, 0, 1, 0}, {2, 4, 2, 2, 4, 1, 1, 1}, {2, 4, 4, 1, 4, 1, 1, 1},
    {2, 4, 1, 4, 4, 1, 1, 1},
};
#define N_VARIANTS ((int)(sizeof(kVariants) / sizeof(kVariants[0])))
#define VM_N_ORDERS 5

template <bool BF16>
static void launch_variant(int vi, dim3 grid, dim3 pgrid, hipStream_t st, const uint4 *Bf, const uint4 *Af,
                           const int *sexp, const uint32_t *wmax, const float4 *N, int ns, int nsb,
                           int nkstep, int nrb, int np, int nchan, int ntx, int nty, int order,
                           float invw, int nstokes, int soffset, int ns_total, float *S)
{
#define VM_LAUNCH(MB_, NB_, WR_, WC_, NST_, EPI_, OCC_)                                         \
    hipLaunchKernelGGL((k_beamform_mfma<BF16, MB_, NB_, WR_, WC_, NST_, EPI_, OCC_>), grid,    \
                       dim3(64 * WR_ * WC_), 0, st, Bf, Af, sexp, wmax, N, ns, nsb, nkstep, nrb, \
                       np, nchan, ntx, nty, order, invw, nstokes, soffset, ns_total, S)
#define VM_LAUNCH_PT(MB_, NB_, WR_, WC_)                                                        \
    hipLaunchKernelGGL((k_beamform_mfma_pt<BF16, MB_, NB_, WR_, WC_>), pgrid,                  \
                       dim3(64 * WR_ * WC_), 0, st, Bf, Af, sexp, wmax, N, ns, nsb, nkstep, nrb, \
                       np, nchan, ntx, nty, order, invw, nstokes, soffset, ns_total, S)
    switch (vi) {
    case 0: VM_LAUNCH(4, 4, 2, 2, 3, 1, 1); break;
    case 1: VM_LAUNCH(4, 4, 1, 4, 3, 1, 1); break;
    case 2: VM_LAUNCH(4, 4, 2, 2, 4, 1, 1); break;
    case 3: VM_LAUNCH(4, 2, 2, 2, 2, 1, 2); break;
    case 4: VM_LAUNCH(2, 4, 2, 2, 2, 1, 2); break;
    case 5: VM_LAUNCH(2, 4, 4, 2, 2, 1, 1); break;
    case 6: VM_LAUNCH(4, 4, 2, 2, 3, 0, 1); break;
    case 7: VM_LAUNCH_PT(2, 4, 2, 2); break;
    case 8: VM_LAUNCH_PT(2, 4, 4, 1); break;
    default: VM_LAUNCH_PT(2, 4, 1, 4); break;
    }
#undef VM_LAUNCH_PT
#undef VM_LAUNCH
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

#define ROW_PAD 256 /* GEMM rows are padded to the largest workgroup row tile */

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The beamformer's prepared per-chunk state.
struct mfma_plan {
    int nant, ni, nchan, np, nkstep, nrb, ns, nsb, chunk;
    uint4 *Af, *Bf;
    float4 *N;
    int *sexp;
    uint32_t *wmax;
    float invw;
};

static void mfma_prepare(vcsbeam_context *vm, mfma_plan *pl)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const bool bf16 = vm->beamform_form == VM_BF_MFMA_BF16;
    if (vm->datatype != VM_FLT)
        VM_FATAL("MFMA beamformer needs VM_FLT voltages (forward PFB output); got %d",
                 (int)vm->datatype);
    const int nant = (int)vm->obs_metadata->num_ants, ni = (int)vm->obs_metadata->num_rf_inputs;
    const int nchan = vm->nfine_chan, np = (int)vm->npointing;
    if (ni != 2 * nant) VM_FATAL("MFMA beamformer: %d RF inputs for %d antennas", ni, nant);
    if (ni % (KS / 2) != 0) VM_FATAL("MFMA beamformer: RF inputs (%d) must be a multiple of 8", ni);
    if (nant > 256) VM_FATAL("MFMA beamformer: at most 256 antennas (%d)", nant);
    const int nkstep = 2 * ni / KS;
    const int nrb = ((4 * np + ROW_PAD - 1) / ROW_PAD) * ROW_PAD / RBK;
    const int ns = vm->fine_sample_rate / vm->chunks_per_second;
    const int nsb = (ns + SB - 1) / SB;

    // scratch: Af | Bf | N | sexp | wmax | Jf | jflag | inv
    const size_t af_bytes = align256((size_t)nchan * nrb * nkstep * 128 * sizeof(uint4));
    const size_t bf_bytes = align256((size_t)nchan * nsb * nkstep * 64 * sizeof(uint4));
    const size_t n_bytes = align256((size_t)nchan * ns * sizeof(float4));
    const size_t se_bytes = align256((size_t)nchan * nsb * sizeof(int));
    const size_t wm_bytes = align256((size_t)nchan * sizeof(uint32_t));
    const size_t jf_bytes = align256((size_t)nchan * nant * 4 * sizeof(cf));
    const size_t inv_bytes = align256((size_t)ni * sizeof(uint16_t));
    const size_t need = af_bytes + bf_bytes + n_bytes + se_bytes + wm_bytes + jf_bytes + 256 +
                        inv_bytes;
    if (d->Wm_bytes < need) {
        ensure_buf(&d->d_Wm, &d->Wm_bytes, need);
        d->weights_dirty = 1;
    }
    uint8_t *q = (uint8_t *)d->d_Wm;
    uint4 *Af = (uint4 *)q;                 q += af_bytes;
    uint4 *Bf = (uint4 *)q;                 q += bf_bytes;
    float4 *N = (float4 *)q;                q += n_bytes;
    int *sexp = (int *)q;                   q += se_bytes;
    uint32_t *wmax = (uint32_t *)q;         q += wm_bytes;
    cf *Jf = (cf *)q;                       q += jf_bytes;
    uint32_t *jflag = (uint32_t *)q;        q += 256;
    uint16_t *inv = (uint16_t *)q;

    if (d->weights_dirty) {
        vm_timer_start(d, VM_T_BF_WEIGHTS);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, jflag, (size_t)1, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3((nchan + 255) / 256), dim3(256), 0, st, wmax,
                           (size_t)nchan, 0u);
        hipLaunchKernelGGL(k_inv_pol, dim3((nant + 255) / 256), dim3(256), 0, st, vm->d_polQ_idxs,
                           vm->d_polP_idxs, nant, inv);
        hipLaunchKernelGGL(k_jf, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256), 0, st,
                           (const double2 *)vm->d_J, nant, nchan, Jf, wmax);
        const size_t total = (size_t)nchan * nrb * nkstep * 64;
        const dim3 g((unsigned)((total + 255) / 256));
        if (bf16)
            hipLaunchKernelGGL(k_mfma_weights<true>, g, dim3(256), 0, st, (const double2 *)vm->d_J,
                               (const double2 *)vm->gdelays.d_phi, inv, wmax, np, nant, nchan, nrb,
                               nkstep, Af, jflag);
        else
            hipLaunchKernelGGL(k_mfma_weights<false>, g, dim3(256), 0, st, (const double2 *)vm->d_J,
                               (const double2 *)vm->gdelays.d_phi, inv, wmax, np, nant, nchan, nrb,
                               nkstep, Af, jflag);
        HIPCHECK(hipGetLastError());
        vm_timer_stop(d, VM_T_BF_WEIGHTS);
        d->weights_dirty = 0;
        if (np > 1) {
            uint32_t hflag = 0;
            HIPCHECK(hipMemcpyAsync(&hflag, jflag, 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (hflag)
                VM_FATAL("MFMA beamformer: J^-1 differs between pointings (per-beam primary "
                         "beam); the shared auto-correlation terms need a beam-independent J^-1");
        }
    }
    // auto-correlation terms, block scales and the FP16/BF16 voltage operand
    vm_timer_start(d, VM_T_BF_PREP);
    const dim3 pg((unsigned)nsb, (unsigned)nchan);
#define VM_PREP(B_, A_)                                                                       \
    hipLaunchKernelGGL((k_bf_prep<B_, A_>), pg, dim3(256), 0, st, (const cf *)vm->d_v, ns, ni, \
                       nsb, nkstep, (const cf *)Jf, vm->d_polQ_idxs, vm->d_polP_idxs, nant, Bf,  \
                       sexp, N)
    if (nant <= 64) {
        if (bf16) VM_PREP(true, 8); else VM_PREP(false, 8);
    } else if (nant <= 128) {
        if (bf16) VM_PREP(true, 16); else VM_PREP(false, 16);
    } else {
        if (bf16) VM_PREP(true, 32); else VM_PREP(false, 32);
    }
#undef VM_PREP
    HIPCHECK(hipGetLastError());
    vm_timer_stop(d, VM_T_BF_PREP);
    pl->nant = nant; pl->ni = ni; pl->nchan = nchan; pl->np = np; pl->nkstep = nkstep;
    pl->nrb = nrb; pl->ns = ns; pl->nsb = nsb;
    pl->chunk = vm->chunk_to_load % vm->chunks_per_second;
    pl->Af = Af; pl->Bf = Bf; pl->N = N; pl->sexp = sexp; pl->wmax = wmax;
    pl->invw = (float)(1.0 / (double)vm->num_not_flagged);
}

static bool variant_ok(int vi, const mfma_plan *pl)
{
    if (vi < 0 || vi >= N_VARIANTS) return false;
    const bf_variant &V = kVariants[vi];
    if (pl->nrb % (V.mb * V.wr) != 0) return false;
    if (V.pers && (pl->nkstep % 4 != 0 || pl->nkstep / 4 < V.mb * V.nb)) return false;
    return true;
}

static void mfma_run(vcsbeam_context *vm, const mfma_plan *pl, int vi, int order)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const bf_variant &V = kVariants[vi];
    const int ncol = V.nb * V.wc, nrow = V.mb * V.wr; // fragment blocks per workgroup
    const int ntx = (pl->nsb + ncol - 1) / ncol, nty = pl->nrb / nrow;
    const size_t ntiles = (size_t)pl->nchan * ntx * nty;
    const dim3 grid((unsigned)ntiles);
    // persistent kernel: one workgroup per CU, a multiple of 8 (one per XCD)
    if (!d->ncu) {
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&d->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    size_t g = (size_t)d->ncu;
    if (g >= 8) g &= ~(size_t)7;
    if (g > ntiles) g = ntiles;
    const dim3 pgrid((unsigned)g);
    if (vm->beamform_form == VM_BF_MFMA_BF16)
        launch_variant<true>(vi, grid, pgrid, st, pl->Bf, pl->Af, pl->sexp, pl->wmax, pl->N,
                             pl->ns, pl->nsb, pl->nkstep, pl->nrb, pl->np, pl->nchan, ntx, nty,
                             order, pl->invw, vm->out_nstokes, pl->chunk * pl->ns,
                             vm->fine_sample_rate, (float *)vm->d_S);
    else
        launch_variant<false>(vi, grid, pgrid, st, pl->Bf, pl->Af, pl->sexp, pl->wmax, pl->N,
                              pl->ns, pl->nsb, pl->nkstep, pl->nrb, pl->np, pl->nchan, ntx, nty,
                              order, pl->invw, vm->out_nstokes, pl->chunk * pl->ns,
                              vm->fine_sample_rate, (float *)vm->d_S);
    HIPCHECK(hipGetLastError());
}

int vmBeamformerVariantInfo(int variant, int *rows, int *samples, int *waves)
{
    if (variant < 0 || variant >= N_VARIANTS) return 0;
    const bf_variant &V = kVariants[variant];
    if (rows) *rows = RBK * V.mb * V.wr;
    if (samples) *samples = SB * V.nb * V.wc;
    if (waves) *waves = V.wr * V.wc;
    return N_VARIANTS;
}

void vmSetBeamformerVariant(vcsbeam_context *vm, int variant, int order)
{
    vm_device_state *d = vm_dev(vm);
    if (variant < 0 || variant >= N_VARIANTS) VM_FATAL("no beamformer variant %d", variant);
    if (order < 0 || order >= VM_N_ORDERS) VM_FATAL("no beamformer tile order %d", order);
    d->bf_variant = variant + 1;
    d->bf_order = order;
}

int vmTuneBeamformer(vcsbeam_context *vm, int reps, float *best_ms)
{
    // times every valid (variant, tile order) on the context's current chunk
    // and keeps the fastest; the outputs of the last run are those of the
    // chosen variant
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    mfma_plan pl;
    mfma_prepare(vm, &pl);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    float best = 1e30f;
    int bv = 0, bo = 1;
    for (int vi = 0; vi < N_VARIANTS; vi++) {
        if (!variant_ok(vi, &pl)) continue;
        for (int order = 0; order < VM_N_ORDERS; order++) {
            mfma_run(vm, &pl, vi, order); // warm-up
            HIPCHECK(hipEventRecord(e0, st));
            for (int r = 0; r < reps; r++) mfma_run(vm, &pl, vi, order);
            HIPCHECK(hipEventRecord(e1, st));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (getenv("VCSBEAM_TUNE_VERBOSE"))
                fprintf(stderr,
                        "beamformer variant %d (%dx%d per wave, %dx%d waves%s) order %d: %.3f ms\n",
                        vi, 32 * kVariants[vi].mb, 32 * kVariants[vi].nb, kVariants[vi].wr,
                        kVariants[vi].wc, kVariants[vi].pers ? ", persistent" : "", order, ms);
            if (ms < best) { best = ms; bv = vi; bo = order; }
        }
    }
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    d->bf_variant = bv + 1;
    d->bf_order = bo;
    mfma_run(vm, &pl, bv, bo);
    HIPCHECK(hipStreamSynchronize(st));
    if (best_ms) *best_ms = best;
    return bv;
}

void vmBeamformChunkMFMA(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    mfma_plan pl;
    mfma_prepare(vm, &pl);
    int vi = d->bf_variant > 0 ? d->bf_variant - 1 : 0;
    int order = d->bf_variant > 0 ? d->bf_order : 1;
    if (!variant_ok(vi, &pl)) {
        vi = 0;
        if (!variant_ok(vi, &pl)) VM_FATAL("no beamformer variant fits this problem");
    }
    vm_timer_start(d, VM_T_BF_MFMA);
    mfma_run(vm, &pl, vi, order);
    vm_timer_stop(d, VM_T_BF_MFMA);
    HIPCHECK(hipStreamSynchronize((hipStream_t)d->stream));
}

} // extern "C"
