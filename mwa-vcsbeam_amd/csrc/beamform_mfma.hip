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
// Tiling: workgroup 256 lanes (4 wavefronts) computes 256 rows (64 beams) x
// 128 samples; each wavefront a 128 x 64 block as 4 x 2 MFMA tiles; K steps of
// 16 staged through double-buffered LDS with padded rows (conflict-free
// 16-byte fragment reads), global loads issued two steps ahead.
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
// 16 hi parts then 16 lo parts (64 bytes).  Rows: 4*beam + {Xre, Xim, Yre,
// Yim}; rows of beams >= npointing are zero.  One lane per (c, kstep, row)
// writes its 64-byte row segment; inv[i] = 2*antenna + pol of VCS input i.
// Lanes of beams p >= 1 also compare J^-1 with beam 0's and raise jflag on a
// difference (the shared auto-correlation terms need a beam-independent J^-1).
template <bool BF16>
__global__ void k_mfma_weights(const double2 *__restrict__ J, const double2 *__restrict__ phi,
                               const uint16_t *__restrict__ inv, int npointing, int nant,
                               int nchan, int nrows_pad, int nkstep, uint16_t *__restrict__ Wm,
                               uint32_t *__restrict__ jflag)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)nchan * nkstep * nrows_pad;
    if (j >= total) return;
    const int row = (int)(j % nrows_pad);
    const int ks = (int)((j / nrows_pad) % nkstep);
    const int c = (int)(j / ((size_t)nrows_pad * nkstep));
    const int p = row >> 2, comp = row & 3; // comp: 0 Xre, 1 Xim, 2 Yre, 3 Yim
    uint32_t hi[8], lo[8];
    bool differs = false;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        uint16_t hb[2] = {0, 0}, lb[2] = {0, 0};
        if (p < npointing) {
            int code = inv[8 * ks + q];
            int a = code >> 1, pol = code & 1;           // pol 0: Q input, 1: P input
            size_t pac = ((size_t)p * nant + a) * nchan + c;
            double2 f = phi[pac];
            double2 m = J[pac * 4 + 2 * (comp >> 1) + pol]; // J_x0/J_x1 or J_y0/J_y1
            if (p > 0) {
                double2 m0 = J[((size_t)a * nchan + c) * 4 + 2 * (comp >> 1) + pol];
                differs |= (m.x != m0.x) || (m.y != m0.y);
            }
            double wr = f.x * m.x - f.y * m.y, wi = f.x * m.y + f.y * m.x;
            // Xre/Yre rows: (Re w, -Im w) against (v.re, v.im); Xim/Yim: (Im w, Re w)
            double v0 = (comp & 1) ? wi : wr;
            double v1 = (comp & 1) ? wr : -wi;
            double vv[2] = {v0, v1};
#pragma unroll
            for (int e2 = 0; e2 < 2; e2++) {
                hb[e2] = mtype<BF16>::bits((float)vv[e2]);
                lb[e2] = mtype<BF16>::bits((float)(vv[e2] - (double)mtype<BF16>::back(hb[e2])));
            }
        }
        hi[q] = (uint32_t)hb[0] | ((uint32_t)hb[1] << 16);
        lo[q] = (uint32_t)lb[0] | ((uint32_t)lb[1] << 16);
    }
    uint4 *dst = (uint4 *)(Wm + j * 32);
    dst[0] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    dst[1] = make_uint4(hi[4], hi[5], hi[6], hi[7]);
    dst[2] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    dst[3] = make_uint4(lo[4], lo[5], lo[6], lo[7]);
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

// ------------------------------------------------------------- GEMM -------
// 1-D grid over (channel, tile).  Tiles of a channel are 64*RB rows x NT
// samples.  With `xcd` set the order is XCD-aware: the 8 workgroups
// dispatched together (one per XCD) take the same tile of 8 different
// channels, so each XCD walks its own channels tile by tile (row tiles
// fastest) and keeps the channel's weights and voltage tile in its L2.
// Each LDS stage holds KSUB k-steps of 16 (one barrier per stage).
// v: VM_FLT [c][i][s]; Wm as above; N [c][s] float4; vmax [c] float bits.
// S: [p][st][c][s_total] (device layout), sample offset soffset.
template <bool BF16, int RB, int KSUB>
__global__ __launch_bounds__(256) void k_beamform_mfma(
    const cf *__restrict__ v, int ns, int ni, const uint16_t *__restrict__ Wm, int nrows_pad,
    int nkstep, const float4 *__restrict__ N, const uint32_t *__restrict__ vmax, int npointing,
    int nchan, int nxt, int nyt, int xcd, float invw, int nstokes, int soffset, int ns_total,
    float *__restrict__ S)
{
    typedef mtype<BF16> M;
    constexpr int MTR = 64 * RB;                  // rows per workgroup
    constexpr int AROW = KSUB * 64 + 16;          // LDS bytes per A row (stride 16*KSUB+4 dwords)
    constexpr int BROW = KSUB * 32 + 16;          // LDS bytes per B column
    constexpr int NAU = MTR * KSUB * 4 / 256;     // uint4 of A per lane per stage
    __shared__ __attribute__((aligned(16))) uint8_t ldsA[2][MTR * AROW];
    __shared__ __attribute__((aligned(16))) uint8_t ldsB[2][NT * BROW];

    // ---- tile decode
    const int ntile = nxt * nyt;
    const int L = blockIdx.x;
    int c, tile;
    if (xcd && (nchan & 7) == 0) {
        const int g = L >> 3, x = L & 7;
        c = (g / ntile) * 8 + x;
        tile = g % ntile;
    } else {
        c = L / ntile;
        tile = L % ntile;
    }
    const int yt = tile % nyt, xt = tile / nyt;
    const int s0 = xt * NT, r0 = yt * MTR;

    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;

    // per-channel power-of-two scale: largest |component| -> [2^13, 2^14)
    float vm = __uint_as_float(vmax[c]);
    int ex = 0;
    if (vm > 0.f) frexpf(vm, &ex);
    const float scale = ldexpf(1.0f, 14 - ex);
    const float unscale2 = ldexpf(1.0f, 2 * (ex - 14));

    const cf *vc = v + (size_t)c * ni * ns;
    const uint16_t *Wc = Wm + (size_t)c * nkstep * nrows_pad * 32;
    const int nstage = nkstep / KSUB;

    const int b_col = t & 127, b_grp = t >> 7;
    const bool b_ok = (s0 + b_col) < ns;
    uint4 ra[NAU];
    cf rb[KSUB][4];

    auto gload = [&](int stg) {
#pragma unroll
        for (int r = 0; r < NAU; r++) {
            int u = t + 256 * r;                   // uint4 index within the stage
            int quarter = u & 3, row = (u >> 2) % MTR, sub = (u >> 2) / MTR;
            int ks = stg * KSUB + sub;
            ra[r] = *((const uint4 *)(Wc + ((size_t)ks * nrows_pad + r0 + row) * 32) + quarter);
        }
#pragma unroll
        for (int sub = 0; sub < KSUB; sub++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int i = (stg * KSUB + sub) * (KS / 2) + 4 * b_grp + q;
                rb[sub][q] = b_ok ? vc[(size_t)i * ns + s0 + b_col] : cf{0.f, 0.f};
            }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int r = 0; r < NAU; r++) {
            int u = t + 256 * r;
            int quarter = u & 3, row = (u >> 2) % MTR, sub = (u >> 2) / MTR;
            *(uint4 *)(ldsA[buf] + row * AROW + sub * 64 + quarter * 16) = ra[r];
        }
#pragma unroll
        for (int sub = 0; sub < KSUB; sub++) {
            uint32_t pk[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                pk[q] = (uint32_t)M::bits(rb[sub][q].x * scale) |
                        ((uint32_t)M::bits(rb[sub][q].y * scale) << 16);
            *(uint4 *)(ldsB[buf] + b_col * BROW + sub * 32 + b_grp * 16) =
                make_uint4(pk[0], pk[1], pk[2], pk[3]);
        }
    };

    f16v acc[RB][2];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0.f;

    const int wr = w >> 1, wc = w & 1;
    const int lr = lane & 31, lk = lane >> 5;

    gload(0);
    lstore(0);
    __syncthreads();
    for (int stg = 0; stg < nstage; stg++) {
        const int buf = stg & 1;
        if (stg + 1 < nstage) gload(stg + 1);
#pragma unroll
        for (int sub = 0; sub < KSUB; sub++) {
            typename M::v8 Ah[RB], Al[RB], Bf[2];
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const uint8_t *pa = ldsA[buf] + (32 * RB * wr + 32 * r + lr) * AROW + sub * 64 + 16 * lk;
                Ah[r] = __builtin_bit_cast(typename M::v8, *(const uint4 *)pa);
                Al[r] = __builtin_bit_cast(typename M::v8, *(const uint4 *)(pa + 32));
            }
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const uint8_t *pb = ldsB[buf] + (64 * wc + 32 * cb + lr) * BROW + sub * 32 + 16 * lk;
                Bf[cb] = __builtin_bit_cast(typename M::v8, *(const uint4 *)pb);
            }
            // independent hi products first, then the lo products
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++) acc[r][cb] = M::mfma(Ah[r], Bf[cb], acc[r][cb]);
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++) acc[r][cb] = M::mfma(Al[r], Bf[cb], acc[r][cb]);
        }
        if (stg + 1 < nstage) lstore(buf ^ 1);
        __syncthreads();
    }

    // epilogue: lane holds rows 8*jg + 4*lk + {0..3} = one beam's (Xre, Xim, Yre, Yim)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const int s = s0 + 64 * wc + 32 * cb + lr;
        if (s >= ns) continue;
        const float4 n = N[(size_t)c * ns + s];
        const int so = s + soffset;
#pragma unroll
        for (int r = 0; r < RB; r++)
#pragma unroll
            for (int jg = 0; jg < 4; jg++) {
                const int row = r0 + 32 * RB * wr + 32 * r + 8 * jg + 4 * lk;
                const int p = row >> 2;
                if (p >= npointing) continue;
                float xr = acc[r][cb][4 * jg + 0], xi = acc[r][cb][4 * jg + 1];
                float yr = acc[r][cb][4 * jg + 2], yi = acc[r][cb][4 * jg + 3];
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

// Kernel variants the autotuner chooses from: (row blocks per wavefront,
// k-steps per LDS stage).
struct bf_variant { int rb, ksub; };
static const bf_variant kVariants[] = {{2, 1}, {2, 2}, {2, 4}, {4, 1}, {4, 2}, {1, 2}, {1, 4}};
#define N_VARIANTS ((int)(sizeof(kVariants) / sizeof(kVariants[0])))

template <bool BF16>
static void launch_variant(int vi, dim3 grid, hipStream_t st, const cf *v, int ns, int ni,
                           const uint16_t *Wm, int nrows_pad, int nkstep, const float4 *N,
                           const uint32_t *vmax, int np, int nchan, int nxt, int nyt, int xcd,
                           float invw, int nstokes, int soffset, int ns_total, float *S)
{
#define VM_LAUNCH(RB_, KS_)                                                                       \
    hipLaunchKernelGGL((k_beamform_mfma<BF16, RB_, KS_>), grid, dim3(256), 0, st, v, ns, ni, Wm,  \
                       nrows_pad, nkstep, N, vmax, np, nchan, nxt, nyt, xcd, invw, nstokes,      \
                       soffset, ns_total, S)
    switch (vi) {
    case 0: VM_LAUNCH(2, 1); break;
    case 1: VM_LAUNCH(2, 2); break;
    case 2: VM_LAUNCH(2, 4); break;
    case 3: VM_LAUNCH(4, 1); break;
    case 4: VM_LAUNCH(4, 2); break;
    case 5: VM_LAUNCH(1, 2); break;
    default: VM_LAUNCH(1, 4); break;
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

#define ROW_PAD 256 /* GEMM rows are padded to the largest workgroup tile */

// The beamformer's prepared per-chunk state (weights, N terms, scale).
struct mfma_plan {
    int nant, ni, nchan, np, nkstep, nrows_pad, ns, chunk;
    uint16_t *Wm;
    float4 *N;
    uint32_t *vmax;
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
    if (ni % (KS / 2) != 0) VM_FATAL("MFMA beamformer: RF inputs (%d) must be a multiple of 8", ni);
    if (ni > 65535) VM_FATAL("MFMA beamformer: too many RF inputs (%d)", ni);
    const int nkstep = 2 * ni / KS;
    const int nrows_pad = ((4 * np + ROW_PAD - 1) / ROW_PAD) * ROW_PAD;
    const int ns = vm->fine_sample_rate / vm->chunks_per_second;

    // scratch: Wm | N | vmax | Jf | jflag | inv
    size_t wm_bytes = (size_t)nchan * nkstep * nrows_pad * 32 * sizeof(uint16_t);
    size_t n_bytes = (size_t)nchan * ns * sizeof(float4);
    size_t vmax_bytes = 256 * ((nchan * 4 + 255) / 256);
    size_t jf_bytes = (size_t)nchan * nant * 4 * sizeof(cf);
    size_t inv_bytes = 256 * ((ni * 2 + 255) / 256);
    size_t need = wm_bytes + n_bytes + vmax_bytes + jf_bytes + 256 + inv_bytes;
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
    uint16_t *inv = (uint16_t *)(base + wm_bytes + n_bytes + vmax_bytes + jf_bytes + 256);

    if (d->weights_dirty) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, jflag, (size_t)1, 0u);
        hipLaunchKernelGGL(k_inv_pol, dim3((nant + 255) / 256), dim3(256), 0, st, vm->d_polQ_idxs,
                           vm->d_polP_idxs, nant, inv);
        size_t total = (size_t)nchan * nkstep * nrows_pad;
        dim3 g((unsigned)((total + 255) / 256));
        if (bf16)
            hipLaunchKernelGGL(k_mfma_weights<true>, g, dim3(256), 0, st, (const double2 *)vm->d_J,
                               (const double2 *)vm->gdelays.d_phi, inv, np, nant, nchan, nrows_pad,
                               nkstep, Wm, jflag);
        else
            hipLaunchKernelGGL(k_mfma_weights<false>, g, dim3(256), 0, st, (const double2 *)vm->d_J,
                               (const double2 *)vm->gdelays.d_phi, inv, np, nant, nchan, nrows_pad,
                               nkstep, Wm, jflag);
        hipLaunchKernelGGL(k_jf, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256), 0, st,
                           (const double2 *)vm->d_J, nant, nchan, Jf);
        HIPCHECK(hipGetLastError());
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
    // auto-correlation terms and the per-channel scale of this chunk
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, vmax, (size_t)nchan, 0u);
    vm_timer_start(d, VM_T_NTERMS);
    hipLaunchKernelGGL(k_nterms, dim3((ns + 255) / 256, nchan), dim3(256), 0, st,
                       (const cf *)vm->d_v, ns, ni, (const cf *)Jf, vm->d_polQ_idxs,
                       vm->d_polP_idxs, nant, N, vmax);
    HIPCHECK(hipGetLastError());
    vm_timer_stop(d, VM_T_NTERMS);
    pl->nant = nant; pl->ni = ni; pl->nchan = nchan; pl->np = np; pl->nkstep = nkstep;
    pl->nrows_pad = nrows_pad; pl->ns = ns;
    pl->chunk = vm->chunk_to_load % vm->chunks_per_second;
    pl->Wm = Wm; pl->N = N; pl->vmax = vmax;
    pl->invw = (float)(1.0 / (double)vm->num_not_flagged);
}

static bool variant_ok(int vi, const mfma_plan *pl)
{
    return vi >= 0 && vi < N_VARIANTS && pl->nkstep % kVariants[vi].ksub == 0 &&
           pl->nrows_pad % (64 * kVariants[vi].rb) == 0;
}

static void mfma_run(vcsbeam_context *vm, const mfma_plan *pl, int vi, int xcd)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const int mtr = 64 * kVariants[vi].rb;
    const int nxt = (pl->ns + NT - 1) / NT, nyt = pl->nrows_pad / mtr;
    dim3 grid((unsigned)(pl->nchan * nxt * nyt));
    if (vm->beamform_form == VM_BF_MFMA_BF16)
        launch_variant<true>(vi, grid, st, (const cf *)vm->d_v, pl->ns, pl->ni, pl->Wm,
                             pl->nrows_pad, pl->nkstep, pl->N, pl->vmax, pl->np, pl->nchan, nxt,
                             nyt, xcd, pl->invw, vm->out_nstokes, pl->chunk * pl->ns,
                             vm->fine_sample_rate, (float *)vm->d_S);
    else
        launch_variant<false>(vi, grid, st, (const cf *)vm->d_v, pl->ns, pl->ni, pl->Wm,
                              pl->nrows_pad, pl->nkstep, pl->N, pl->vmax, pl->np, pl->nchan, nxt,
                              nyt, xcd, pl->invw, vm->out_nstokes, pl->chunk * pl->ns,
                              vm->fine_sample_rate, (float *)vm->d_S);
    HIPCHECK(hipGetLastError());
}

void vmSetBeamformerVariant(vcsbeam_context *vm, int variant, int xcd)
{
    vm_device_state *d = vm_dev(vm);
    if (variant < 0 || variant >= N_VARIANTS) VM_FATAL("no beamformer variant %d", variant);
    d->bf_variant = variant + 1;
    d->bf_xcd = xcd != 0;
}

int vmTuneBeamformer(vcsbeam_context *vm, int reps, float *best_ms)
{
    // times every valid (variant, XCD order) on the context's current chunk
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
    int bv = 0, bx = 1;
    for (int vi = 0; vi < N_VARIANTS; vi++) {
        if (!variant_ok(vi, &pl)) continue;
        for (int xcd = 0; xcd < 2; xcd++) {
            mfma_run(vm, &pl, vi, xcd); // warm-up
            HIPCHECK(hipEventRecord(e0, st));
            for (int r = 0; r < reps; r++) mfma_run(vm, &pl, vi, xcd);
            HIPCHECK(hipEventRecord(e1, st));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (getenv("VCSBEAM_TUNE_VERBOSE"))
                fprintf(stderr, "beamformer variant rb=%d ksub=%d xcd=%d: %.3f ms\n",
                        kVariants[vi].rb, kVariants[vi].ksub, xcd, ms);
            if (ms < best) { best = ms; bv = vi; bx = xcd; }
        }
    }
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    d->bf_variant = bv + 1;
    d->bf_xcd = bx;
    mfma_run(vm, &pl, bv, bx);
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
    int xcd = d->bf_variant > 0 ? d->bf_xcd : 1;
    if (!variant_ok(vi, &pl)) {
        vi = 0;
        if (!variant_ok(vi, &pl)) VM_FATAL("no beamformer variant fits this problem");
    }
    vm_timer_start(d, VM_T_BF_MFMA);
    mfma_run(vm, &pl, vi, xcd);
    vm_timer_stop(d, VM_T_BF_MFMA);
    HIPCHECK(hipStreamSynchronize((hipStream_t)d->stream));
}

} // extern "C"
