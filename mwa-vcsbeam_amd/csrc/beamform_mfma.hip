// beamform_mfma.hip -- precision many-beam tied-array beamformer on MFMA.
//
// Reference arithmetic: MWA-vcsbeam src/form_beam.cpp:122-204 and 239-371:
//     e_pa = phi_pa J^-1_pa v_a,  E_p = sum_a e_pa,  N_p = sum_a e_pa e_pa^H,
// then auto-correlation-subtracted Stokes.
//
// Factoring.  When J^-1_pa = J^-1_a does not depend on the pointing (no
// per-pointing primary beam, B = I), e_pa = phi_pa u_a with u_a = J^-1_a v_a:
// the 2x2 Jones product is done once per (antenna, sample) in the prep pass,
// and the beams only differ by the complex scalars phi_pa.  The sum over
// antennas is then, per fine channel c, a complex GEMM
//     E[beam][(s, pol)] = sum_a phi_c[beam][a] u_c[a][(s, pol)]
// (M = beams, K = tiles, N = samples x 2 pols): half the flops of applying
// phi J^-1 as a 2x2 per beam.  N_p = sum_a u_a u_a^H is beam-independent and
// is formed in the prep pass.
//
// It runs as a real GEMM on v_mfma_f32_32x32x16_{f16,bf16}: K interleaves
// (re, im) of u (k' = 2a + {0,1}); each beam owns two output rows (re, im) and
// the two polarisations are separate 32-column blocks, so one lane holds a
// beam's (Ex, Ey) at one sample across two accumulator blocks and forms the
// Stokes parameters in registers.
//
// Precision form: u is converted to FP16 (or BF16) with a power-of-two scale
// per (channel, 32-sample block); the phase weights are split phi = phi_hi +
// phi_lo into two FP16 (BF16) parts and both products accumulate in FP32.
// Phases and Jones inverses are produced in FP64 (geometry.hip) and demoted
// only at the split (phases) or to FP32 for the prep pass (J^-1).
//
// Data flow per chunk:
//   k_bf_prep        FP32 voltages -> u = J^-1 v, N terms, block scales and
//                    the FP16 B operand in MFMA fragment order (one pass).
//   k_beamform_mfma  the GEMM + Stokes epilogue.  Both operands are stored
//                    fragment-native (each 64-lane x 16-byte fragment is one
//                    contiguous kilobyte), so every wavefront loads its own
//                    fragments straight into registers with buffer loads: no
//                    LDS staging, no bank conflicts, no conversion work.
// Weights (k_phase_weights) are rebuilt only when phases or Jones change.
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

#define KS 16   // k' (real) per MFMA step = 8 tiles
#define SB 32   // samples per B fragment block
#define RBK 32  // rows per A fragment block (16 beams)

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

// scale of the phase weights: |Re phi|, |Im phi| <= 1
#define PHASE_EXP 13

// ------------------------------------------------------------ weights -----
// Operand A in fragment order: Af[c][rb][ks][h][lane] (uint4), rb a block of
// 32 GEMM rows (16 beams), ks a step of 16 k' (8 tiles), h = 0 hi part, 1 lo
// part.  Lane l holds row 32 rb + l%32 and k' = 16 ks + 8 (l/32) + j, j =
// 0..7, i.e. tiles a = 8 ks + 4 (l/32) + j/2 with j even the coefficient of
// u.re and j odd of u.im.  Rows: 2 beam + {re, im}: re rows (Re phi, -Im phi),
// im rows (Im phi, Re phi); rows of beams >= npointing are zero.  Weights are
// scaled by 2^PHASE_EXP.  phi: [p][a][c] complex double.
template <bool BF16>
__global__ __launch_bounds__(256) void k_phase_weights(const double2 *__restrict__ phi,
                                                       int npointing, int nant, int nchan,
                                                       int nrb, int nkstep, uint4 *__restrict__ Af)
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
    const int p = row >> 1, im = row & 1;
    const double ws = ldexp(1.0, PHASE_EXP);
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint16_t hb[2] = {0, 0}, lb[2] = {0, 0};
        const int a = 8 * ks + 4 * (lane >> 5) + q;
        if (p < npointing && a < nant) {
            const double2 f = phi[((size_t)p * nant + a) * nchan + c];
            const double vv[2] = {(im ? f.y : f.x) * ws, (im ? f.x : -f.y) * ws};
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
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
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

// flag when J^-1 of any pointing p >= 1 differs from pointing 0's
__global__ void k_jcheck(const double2 *__restrict__ J, size_t per_pointing, size_t total,
                         uint32_t *__restrict__ flag)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x + per_pointing;
    bool d = false;
    if (j < total) {
        const double2 m = J[j], m0 = J[j % per_pointing];
        d = (m.x != m0.x) || (m.y != m0.y);
    }
    if (__any(d) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// ------------------------------------------------------- voltage prep -----
// One workgroup per (32-sample block sb, channel c), one pass over the block:
// lane (sl, g) loads sample 32 sb + sl of both inputs of tiles g, g+8, ...
// (APT tiles at most), forms u = J^-1 v with Jf = J^-1 (float, [c][a][4]),
// the auto-correlation sums N = sum_a (|u_x|^2, |u_y|^2, Re u_x u_y^*,
// Im u_x u_y^*) and the block's largest |component| of u.  The 8 partial
// sums per sample are reduced in LDS and written as N[c][s]; u, scaled by
// 2^es, is placed in LDS in fragment order and written out as B fragments
// Bf[c][sb][ks][pol][lane] (uint4): lane l holds sample 32 sb + l%32 and
// tiles 8 ks + 4 (l/32) + {0..3} as (re, im) pairs of polarisation pol.
// Samples >= ns are zero; tiles >= nant are zero.  sexp[c][sb] = es.
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
    __shared__ uint4 frag[APT * 2 * 64]; // nkstep (<= APT) x 2 pols x 64 lanes
    const int sb = blockIdx.x, c = blockIdx.y;
    const int t = threadIdx.x, sl = t & 31, g = t >> 5;
    const int s = SB * sb + sl;
    const bool ok = s < ns;
    const cf *vc = v + (size_t)c * ni * ns;
    const int nfrag = nkstep * 2 * 64;

    for (int k = t; k < nfrag; k += 256) frag[k] = make_uint4(0u, 0u, 0u, 0u);

    cf ux[APT], uy[APT];
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = g + 8 * j;
        cf q = {0.f, 0.f}, p = {0.f, 0.f};
        if (ok && a < nant) {
            q = vc[(size_t)polQ[a] * ns + s];
            p = vc[(size_t)polP[a] * ns + s];
        }
        ux[j] = uy[j] = cf{0.f, 0.f};
        if (a < nant) {
            const cf *w = Jf + ((size_t)c * nant + a) * 4;
            const cf w00 = w[0], w01 = w[1], w10 = w[2], w11 = w[3];
            ux[j].x = w00.x * q.x - w00.y * q.y + w01.x * p.x - w01.y * p.y;
            ux[j].y = w00.x * q.y + w00.y * q.x + w01.x * p.y + w01.y * p.x;
            uy[j].x = w10.x * q.x - w10.y * q.y + w11.x * p.x - w11.y * p.y;
            uy[j].y = w10.x * q.y + w10.y * q.x + w11.x * p.y + w11.y * p.x;
        }
    }
    float nxx = 0.f, nyy = 0.f, xyr = 0.f, xyi = 0.f, mx = 0.f;
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const cf x = ux[j], y = uy[j];
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(y.x), fabsf(y.y))));
        nxx += x.x * x.x + x.y * x.y;
        nyy += y.x * y.x + y.y * y.y;
        xyr += x.x * y.x + x.y * y.y;
        xyi += x.y * y.x - x.x * y.y;
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

    // tile a, polarisation P, sample sl -> dword (a%4) of lane sl + 32 ((a%8)/4)
    // of fragment (ks = a/8, P)
    uint32_t *fd = (uint32_t *)frag;
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = g + 8 * j;
        if (a >= nant) break;
        const int base = ((a >> 3) * 2) * 256 + (sl + 32 * ((a >> 2) & 1)) * 4 + (a & 3);
        fd[base] = (uint32_t)M::bits(ux[j].x * scale) | ((uint32_t)M::bits(ux[j].y * scale) << 16);
        fd[base + 256] =
            (uint32_t)M::bits(uy[j].x * scale) | ((uint32_t)M::bits(uy[j].y * scale) << 16);
    }
    __syncthreads();
    uint4 *dst = Bf + ((size_t)c * nsb + sb) * nfrag;
    for (int k = t; k < nfrag; k += 256) dst[k] = frag[k];
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
// k-step ks at + 2048 ks (the two polarisation fragments of a k-step are
// adjacent).
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
    for (int n = 0; n < NB; n++) f.b[n] = bload(rB, vl, boff[n] + 2048 * ks);
}

// Scheduling pattern for one k-step: each of the NL fragment loads is
// issued between MFMAs (one load, then NM/NL MFMAs).
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

// Stokes (I, Q, U, V) from a beam's (Ex, Ey) = (xr + i xi, yr + i yi): |E|^2
// terms carry the operand scales squared (unscale2 removes them), the
// auto-correlation terms nn are unscaled.
__device__ __forceinline__ void stokes4(float xr, float xi, float yr, float yi, float unscale2,
                                        float4 nn, float invw, float (&st)[4])
{
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
// (WR MB 32) rows (16 WR MB beams) x (WC NB/2 32) samples of one channel;
// a wavefront's NB column blocks are NB/2 sample blocks x 2 polarisations
// (block n = sample block n/2, polarisation n%2).  NST register stages keep
// NST-1 k-steps of fragments in flight while one is multiplied.
// Tile order: order 0 walks the channels in turn, row tiles fastest.  Orders
// 1-4 are XCD-aware: the 8 workgroups dispatched together (one per XCD) take
// the same tile of 8 different channels, so each XCD walks its own channels;
// within a channel tiles go in bands of gr row tiles, column by column (gr =
// all row tiles for order 1, 1 for order 2, 2 for order 3, 4 for order 4).
// Epilogue: each (32-row block, sample block) pair is passed through a 4 KB
// per-wavefront LDS tile, 8 beams at a time, so that every lane stores four
// consecutive samples of one (beam, Stokes) row with one 16-byte store.
// S: [p][st][c][s_total] (device layout), sample offset soffset.
template <bool BF16, int MB, int NB, int WR, int WC, int NST, int OCC>
__global__ __launch_bounds__(64 * WR * WC, OCC) void k_beamform_mfma(
    const uint4 *__restrict__ Bf, const uint4 *__restrict__ Af, const int *__restrict__ sexp,
    const float4 *__restrict__ N, int ns, int nsb, int nkstep, int nrb, int npointing, int nchan,
    int ntx, int nty, int order, float invw, int nstokes, int soffset, int ns_total,
    float *__restrict__ S)
{
    static_assert(NB % 2 == 0, "column blocks come in polarisation pairs");
    constexpr int NSB = NB / 2; // sample blocks per wavefront
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
    const int sb0 = (tx * WC + wc) * NSB;

    const int astr = nkstep * 2048;              // bytes per 32-row block of A
    const int bstr = nkstep * 2048;              // bytes per sample block of B (2 pols)
    const rsrc_t rA = make_rsrc(Af + (size_t)c * nrb * nkstep * 128, (uint32_t)nrb * astr);
    const rsrc_t rB = make_rsrc(Bf + (size_t)c * nsb * nkstep * 128, (uint32_t)nsb * bstr);
    const int vl = lane * 16;
    const int aoff = rb0 * astr;
    int boff[NB];
#pragma unroll
    for (int n = 0; n < NB; n++) {
        // blocks past the last sample block re-read it (in bounds); their
        // columns are never stored
        boff[n] = min(sb0 + n / 2, nsb - 1) * bstr + 1024 * (n & 1);
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

    // epilogue: in a 32-row block lane (lr, lk) holds rows 8 jg + 4 lk + {0..3}
    // = beams 4 jg + 2 lk + {0, 1} as (re, im), column lr; block n = 2 j holds
    // Ex and n = 2 j + 1 holds Ey of sample block j.  Half h of the block
    // (jg = 2 h + jj) is 8 beams; LDS rows r = st * 8 + (4 jj + 2 b + lk) for
    // beam 4 jg + 2 lk + b: the two lane halves write rows 32 floats apart and
    // each 16-lane read group covers two whole rows.
    __shared__ __attribute__((aligned(16))) float stg[WR * WC][32 * 32];
    float *wl = stg[w];
    const int lr = lane & 31, lk = lane >> 5, qq = lane & 7;
    const size_t pstride = (size_t)nstokes * nchan * ns_total;
    const size_t ststride = (size_t)nchan * ns_total;
    const bool vec = ((ns_total | soffset) & 3) == 0;
#pragma unroll
    for (int j = 0; j < NSB; j++) {
        const int sb = sb0 + j;
        const int s = SB * sb + lr;
        float4 nn = make_float4(0.f, 0.f, 0.f, 0.f);
        if (s < ns) nn = N[(size_t)c * ns + s];
        const float unscale2 =
            ldexpf(1.0f, -2 * (sexp[(size_t)c * nsb + min(sb, nsb - 1)] + PHASE_EXP));
        const int s4 = SB * sb + 4 * qq;
        float *Sc = S + (size_t)c * ns_total + soffset + s4;
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int jg = 2 * h + jj;
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const f16v &ax = acc[m][2 * j], &ay = acc[m][2 * j + 1];
                        float st4[4];
                        stokes4(ax[4 * jg + 2 * b], ax[4 * jg + 2 * b + 1], ay[4 * jg + 2 * b],
                                ay[4 * jg + 2 * b + 1], unscale2, nn, invw, st4);
#pragma unroll
                        for (int st = 0; st < 4; st++)
                            wl[(st * 8 + 4 * jj + 2 * b + lk) * 32 + lr] = st4[st];
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = 8 * i + (lane >> 3);
                    const int st = r >> 3, bbp = r & 7;
                    // bbp = 4 jj + 2 b + lk  ->  beam in the half = 4 jj + 2 lk + b
                    const int bb = 4 * (bbp >> 2) + 2 * (bbp & 1) + ((bbp >> 1) & 1);
                    const int p = 16 * (rb0 + m) + 8 * h + bb;
                    const float4 val = *(const float4 *)(wl + r * 32 + 4 * qq);
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

// Kernel variants the autotuner chooses from: register tile per wavefront
// (MB row blocks of 32 x NB column blocks of 32, NB = 2 x sample blocks),
// wavefront grid per workgroup (WR x WC), register stages and the minimum
// wavefronts per SIMD the registers must allow.
struct bf_variant { int mb, nb, wr, wc, nst, occ; };
static const bf_variant kVariants[] = {
    {4, 4, 2, 2, 3, 1}, {4, 4, 1, 4, 3, 1}, {4, 4, 2, 2, 4, 1}, {2, 4, 2, 2, 3, 2},
    {4, 2, 2, 2, 3, 2}, {2, 4, 4, 1, 3, 2}, {4, 4, 4, 1, 3, 1},
};
#define N_VARIANTS ((int)(sizeof(kVariants) / sizeof(kVariants[0])))
#define VM_N_ORDERS 5

template <bool BF16>
static void launch_variant(int vi, dim3 grid, hipStream_t st, const uint4 *Bf, const uint4 *Af,
                           const int *sexp, const float4 *N, int ns, int nsb, int nkstep, int nrb,
                           int np, int nchan, int ntx, int nty, int order, float invw, int nstokes,
                           int soffset, int ns_total, float *S)
{
#define VM_LAUNCH(MB_, NB_, WR_, WC_, NST_, OCC_)                                             \
    hipLaunchKernelGGL((k_beamform_mfma<BF16, MB_, NB_, WR_, WC_, NST_, OCC_>), grid,        \
                       dim3(64 * WR_ * WC_), 0, st, Bf, Af, sexp, N, ns, nsb, nkstep, nrb, np, \
                       nchan, ntx, nty, order, invw, nstokes, soffset, ns_total, S)
    switch (vi) {
    case 0: VM_LAUNCH(4, 4, 2, 2, 3, 1); break;
    case 1: VM_LAUNCH(4, 4, 1, 4, 3, 1); break;
    case 2: VM_LAUNCH(4, 4, 2, 2, 4, 1); break;
    case 3: VM_LAUNCH(2, 4, 2, 2, 3, 2); break;
    case 4: VM_LAUNCH(4, 2, 2, 2, 3, 2); break;
    case 5: VM_LAUNCH(2, 4, 4, 1, 3, 2); break;
    default: VM_LAUNCH(4, 4, 4, 1, 3, 1); break;
    }
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

#define ROW_PAD 256 /* GEMM rows are padded to a multiple of the usual workgroup row tile */

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The beamformer's prepared per-chunk state.
struct mfma_plan {
    int nant, ni, nchan, np, nkstep, nrb, ns, nsb, chunk;
    uint4 *Af, *Bf;
    float4 *N;
    int *sexp;
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
    if (nant > 256) VM_FATAL("MFMA beamformer: at most 256 antennas (%d)", nant);
    const int nkstep = (nant + 7) / 8; // 8 tiles (16 k') per step; tiles past nant are zero
    const int nrb = ((2 * np + ROW_PAD - 1) / ROW_PAD) * ROW_PAD / RBK;
    const int ns = vm->fine_sample_rate / vm->chunks_per_second;
    const int nsb = (ns + SB - 1) / SB;

    // scratch: Af | Bf | N | sexp | Jf | jflag
    const size_t af_bytes = align256((size_t)nchan * nrb * nkstep * 128 * sizeof(uint4));
    const size_t bf_bytes = align256((size_t)nchan * nsb * nkstep * 128 * sizeof(uint4));
    const size_t n_bytes = align256((size_t)nchan * ns * sizeof(float4));
    const size_t se_bytes = align256((size_t)nchan * nsb * sizeof(int));
    const size_t jf_bytes = align256((size_t)nchan * nant * 4 * sizeof(cf));
    const size_t need = af_bytes + bf_bytes + n_bytes + se_bytes + jf_bytes + 256;
    if (d->Wm_bytes < need) {
        ensure_buf(&d->d_Wm, &d->Wm_bytes, need);
        d->weights_dirty = 1;
    }
    uint8_t *q = (uint8_t *)d->d_Wm;
    uint4 *Af = (uint4 *)q;                 q += af_bytes;
    uint4 *Bf = (uint4 *)q;                 q += bf_bytes;
    float4 *N = (float4 *)q;                q += n_bytes;
    int *sexp = (int *)q;                   q += se_bytes;
    cf *Jf = (cf *)q;                       q += jf_bytes;
    uint32_t *jflag = (uint32_t *)q;

    if (d->weights_dirty) {
        vm_timer_start(d, VM_T_BF_WEIGHTS);
        hipLaunchKernelGGL(k_jf, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256), 0, st,
                           (const double2 *)vm->d_J, nant, nchan, Jf);
        const size_t total = (size_t)nchan * nrb * nkstep * 64;
        const dim3 g((unsigned)((total + 255) / 256));
        if (bf16)
            hipLaunchKernelGGL(k_phase_weights<true>, g, dim3(256), 0, st,
                               (const double2 *)vm->gdelays.d_phi, np, nant, nchan, nrb, nkstep, Af);
        else
            hipLaunchKernelGGL(k_phase_weights<false>, g, dim3(256), 0, st,
                               (const double2 *)vm->gdelays.d_phi, np, nant, nchan, nrb, nkstep, Af);
        if (np > 1) {
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, jflag, (size_t)1, 0u);
            const size_t per = (size_t)nant * nchan * 4, tot = per * np;
            hipLaunchKernelGGL(k_jcheck, dim3((unsigned)((tot - per + 255) / 256)), dim3(256), 0,
                               st, (const double2 *)vm->d_J, per, tot, jflag);
        }
        HIPCHECK(hipGetLastError());
        vm_timer_stop(d, VM_T_BF_WEIGHTS);
        d->weights_dirty = 0;
        if (np > 1) {
            uint32_t hflag = 0;
            HIPCHECK(hipMemcpyAsync(&hflag, jflag, 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (hflag)
                VM_FATAL("MFMA beamformer: J^-1 differs between pointings (per-beam primary "
                         "beam); this form needs a beam-independent J^-1 (use the FP32 form)");
        }
    }
    // u = J^-1 v, auto-correlation terms, block scales and the B operand
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
    pl->Af = Af; pl->Bf = Bf; pl->N = N; pl->sexp = sexp;
    pl->invw = (float)(1.0 / (double)vm->num_not_flagged);
}

static bool variant_ok(int vi, const mfma_plan *pl)
{
    return vi >= 0 && vi < N_VARIANTS && pl->nrb % (kVariants[vi].mb * kVariants[vi].wr) == 0;
}

static void mfma_run(vcsbeam_context *vm, const mfma_plan *pl, int vi, int order)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const bf_variant &V = kVariants[vi];
    const int ncol = (V.nb / 2) * V.wc, nrow = V.mb * V.wr; // sample / row blocks per workgroup
    const int ntx = (pl->nsb + ncol - 1) / ncol, nty = pl->nrb / nrow;
    const dim3 grid((unsigned)((size_t)pl->nchan * ntx * nty));
    if (vm->beamform_form == VM_BF_MFMA_BF16)
        launch_variant<true>(vi, grid, st, pl->Bf, pl->Af, pl->sexp, pl->N, pl->ns, pl->nsb,
                             pl->nkstep, pl->nrb, pl->np, pl->nchan, ntx, nty, order, pl->invw,
                             vm->out_nstokes, pl->chunk * pl->ns, vm->fine_sample_rate,
                             (float *)vm->d_S);
    else
        launch_variant<false>(vi, grid, st, pl->Bf, pl->Af, pl->sexp, pl->N, pl->ns, pl->nsb,
                              pl->nkstep, pl->nrb, pl->np, pl->nchan, ntx, nty, order, pl->invw,
                              vm->out_nstokes, pl->chunk * pl->ns, vm->fine_sample_rate,
                              (float *)vm->d_S);
    HIPCHECK(hipGetLastError());
}

int vmBeamformerVariantInfo(int variant, int *rows, int *samples, int *waves)
{
    if (variant < 0 || variant >= N_VARIANTS) return 0;
    const bf_variant &V = kVariants[variant];
    if (rows) *rows = RBK * V.mb * V.wr;
    if (samples) *samples = SB * (V.nb / 2) * V.wc;
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
                fprintf(stderr, "beamformer variant %d (%d rows x %d samples per wave, %dx%d waves) "
                        "order %d: %.3f ms\n", vi, 32 * kVariants[vi].mb,
                        16 * kVariants[vi].nb, kVariants[vi].wr, kVariants[vi].wc, order, ms);
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
