// pfb.hip -- forward fine PFB for MWAX 8+8-bit voltages on gfx950.
//
// Reference behaviour: MWA-vcsbeam src/pfb.cpp:121-343 (kernels) and
// 378-796 (host): unpack -> integer weighted overlap-add (WOLA) ->
// FPGA rounding (or int->float) -> unnormalised 128-point forward FFT ->
// /K, fftshift, (im,re) swap, VCS reorder, and INT4 or FP64 packing.
//
// MI355X design:
//  * Fused path (default): one persistent kernel per chunk does unpack +
//    WOLA + rounding + FFT + reorder, tile by tile (32 consecutive spectra of
//    one RF input).  The 43-block input window of the next tile is fetched
//    into registers (16-byte loads) while the current tile computes; WOLA
//    accumulates exactly in int32 registers (no atomics, no memset); the FFT
//    runs in registers/LDS (fft128.h); output goes to the device layout
//    VM_FLT = complex float [c][i_vcs][s] so that each wavefront stores two
//    256-byte contiguous runs.
//  * Staged path (PFB_STAGED): the reference's four passes over the
//    [m][i][n] int2/float2 buffer, kept for callers that inspect the
//    intermediate buffer; it shares the FFT code and is bit-identical.
//  * Host formats (legacy 4+4-bit and complex double [s][c'][i]) are produced
//    by an LDS-tiled transpose from VM_FLT.
#include <hip/hip_runtime.h>
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

#define TILE_M 32 // spectra per workgroup
#define ROW 129    // LDS row stride in complex values (128 + 1 pad)

// W128^j table, j = 0..127, built in FP64 on the device at initialisation.
static __device__ cf g_tw128[KCH];

__global__ void k_init_tw128()
{
    int j = threadIdx.x;
    double s, c;
    sincospi(-2.0 * (double)j / KCH, &s, &c);
    g_tw128[j] = cf{(float)c, (float)s};
}

static void ensure_twiddles()
{
    static int done = 0;
    if (done) return;
    hipLaunchKernelGGL(k_init_tw128, dim3(1), dim3(KCH), 0, 0);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    done = 1;
}

// ------------------------------------------------------ rounding (a3) -----
template <bool FPGA>
__device__ __forceinline__ float demote(int X, float post)
{
    if (FPGA) {
        // src/pfb.cpp:225-240: (X > 0 ? X + 0x2000 : X) >> 14; the bias is
        // (median of X, 0, 1) << 13, one v_med3_i32 (the compiler would
        // otherwise compare and select); the sum wraps like the reference's int
        int m;
        asm("v_med3_i32 %0, %1, 0, 1" : "=v"(m) : "v"(X));
        const uint32_t bias = (uint32_t)m << 13;
        const int y = (int)((uint32_t)X + bias) >> 14;
        return (float)y * post;
    } else {
        // src/pfb.cpp:266: (float)X * 2^-14
        return (float)X * (0x1p-14f * post);
    }
}

__device__ __forceinline__ size_t mwax_char2(size_t s, int i, int I)
{
    return ((s / VSPVB) * (size_t)I + i) * VSPVB + (s % VSPVB);
}

// ------------------------------------------------------------ fused -------
typedef short vm_short2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c)
{
    // v_dot2_i32_i16: a.lo*b.lo + a.hi*b.hi + c, exact in int32 here
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(vm_short2, a),
                                  __builtin_bit_cast(vm_short2, b), c, false);
}

__device__ __forceinline__ uint32_t pack16(int lo, int hi)
{
    return (uint32_t)(lo & 0xffff) | ((uint32_t)hi << 16);
}

// Tile = TILE_M consecutive spectra m0.. of one RF input i; block 256.  htr
// points at the chunk's first voltage block (the margin); out is VM_FLT
// [c][i_vcs][nspc].
//
// WOLA: b[m][n] = sum_p h[KP - Kp - n - 1] x[(m + 500 - P)K + pK + n]
// (src/pfb.cpp:149-161).  Lane (col, g) owns samples n = 2col, 2col+1 of
// spectra 8g..8g+7.  Consecutive-block sample pairs (x[q][n], x[q+1][n]) are
// assembled as sign-extended int16 pairs with v_perm_b32 and multiplied by
// tap pairs with v_dot2_i32_i16.  SPLIT: taps wider than int16 are split
// exactly as h = 2^15 hh + hl (0 <= hl < 2^15) and summed as two products;
// int32 wrap-around is identical to the reference's int atomics.
#define PFB_P 12
#define PFB_WROWS 48                  // window rows staged (>= TILE_M + PFB_P - 1)
#define PFB_WIN (PFB_WROWS * 16)      // uint4 of one tile's input window
#define PFB_WPL (PFB_WIN / 256)       // window uint4 per lane (exact)

typedef __amdgpu_buffer_rsrc_t pfb_rsrc;
#define PFB_RSRC_DWORD3 0x00020000 /* raw buffer, 32-bit data format (gfx9 family) */

// tap pairs (h_p, h_{p+1}) for p = 0, 2, ..., P-2 at n = 2col and 2col+1
struct pfb_taps {
    uint32_t l0[PFB_P / 2], l1[PFB_P / 2], h0[PFB_P / 2], h1[PFB_P / 2];
};

template <bool SPLIT>
__device__ __forceinline__ void pfb_load_taps(pfb_taps &T, const int *__restrict__ h, int col)
{
    constexpr int P = PFB_P;
#pragma unroll
    for (int k = 0; k < P / 2; k++) {
        int a0 = h[KCH * P - KCH * (2 * k) - 2 * col - 1];
        int b0 = h[KCH * P - KCH * (2 * k + 1) - 2 * col - 1];
        int a1 = h[KCH * P - KCH * (2 * k) - 2 * col - 2];
        int b1 = h[KCH * P - KCH * (2 * k + 1) - 2 * col - 2];
        if (SPLIT) {
            T.l0[k] = pack16(a0 & 0x7fff, b0 & 0x7fff);
            T.l1[k] = pack16(a1 & 0x7fff, b1 & 0x7fff);
            T.h0[k] = pack16(a0 >> 15, b0 >> 15);
            T.h1[k] = pack16(a1 >> 15, b1 >> 15);
        } else {
            T.l0[k] = pack16(a0, b0);
            T.l1[k] = pack16(a1, b1);
        }
    }
}

// the tile's input window: 128-sample blocks q = m0 + ql + 500 - P, ql <
// PFB_WROWS, 16 bytes per slot, by buffer loads that return zeros past the
// end of the chunk's buffer (rows past the tile's last spectrum + P - 1 only
// feed spectra that are never stored).  No branches: every lane issues the
// same PFB_WPL loads for every tile.
__device__ __forceinline__ void pfb_fetch_window(uint4 (&wr)[PFB_WPL], pfb_rsrc rs, int i, int I,
                                                 int m0, int t)
{
#pragma unroll
    for (int r = 0; r < PFB_WPL; r++) {
        const int j = t + 256 * r;
        const int ql = j >> 4, c4 = j & 15;
        const int q = m0 + ql + 500 - PFB_P;        // 128-sample block index
        const int vb = q / 500;                      // MWAX voltage block
        const int within = (q - vb * 500) * KCH;     // sample offset in the block
        const uint32_t byte = ((uint32_t)(vb * I + i) * (uint32_t)VSPVB + (uint32_t)within) * 2u +
                              16u * (uint32_t)c4;
        wr[r] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)byte, 0, 0));
    }
}

__device__ __forceinline__ void pfb_put_window(const uint4 (&wr)[PFB_WPL], uint32_t *win, int t)
{
#pragma unroll
    for (int r = 0; r < PFB_WPL; r++) ((uint4 *)win)[t + 256 * r] = wr[r];
}

// One output value: complex float, or (H16) complex FP16 packed in a dword
// (each part rounded to nearest even).
template <bool H16> struct pfb_out;
template <> struct pfb_out<false> {
    typedef cf T;
    static __device__ __forceinline__ T make(float a, float b) { return cf{a, b}; }
};
template <> struct pfb_out<true> {
    typedef uint32_t T;
    static __device__ __forceinline__ T make(float a, float b)
    {
        return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)a) |
               ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)b) << 16);
    }
};

// WOLA + rounding + FFT + reorder of one tile whose window is in LDS (win
// aliases F); ends with the tile's stores.  Contains workgroup barriers.
// IMF: output as (im, re) instead of (re, im) (src/pfb.cpp:322-331).
// H16: complex FP16 output (VM_HALF) in 32-sample blocks [c][s/32][i][s%32]
// (the last block padded), so that one block of every input is contiguous.
template <bool FPGA, bool SPLIT, bool IMF, bool H16>
__device__ __forceinline__ void pfb_tile(cf *F, const uint32_t *win, const cf *tw,
                                         const pfb_taps &T, int i, int I, int m0, int nspc, int oi,
                                         typename pfb_out<H16>::T *__restrict__ out, int t)
{
    constexpr int P = PFB_P;
    constexpr int NW = 8 + P - 1; // window blocks per lane
    const int nvalid = min(TILE_M, nspc - m0);
    const int col = t & 63, g = t >> 6;

    // WOLA into registers, walking block pairs (q, q+1): each pair's
    // sign-extended int16 pairs are built once and feed every spectrum j and
    // tap pair k with j + 2k = q (per accumulator the k order is ascending,
    // as in the reference's sum; int32 addition is exact or wraps the same)
    uint32_t w[NW];
#pragma unroll
    for (int q = 0; q < NW; q++) w[q] = win[(8 * g + q) * 64 + col];
    int acc[8][4], bcc[8][4];
#pragma unroll
    for (int j = 0; j < 8; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) acc[j][e] = bcc[j][e] = 0;
#pragma unroll
    for (int q = 0; q < NW - 1; q++) {
        const uint32_t lo = w[q], hi = w[q + 1];
        // bytes of lo: re_n0, im_n0, re_n1, im_n1 (block q); hi: block q+1
        const uint32_t I0 = __builtin_amdgcn_perm(hi, lo, 0x0A050801u);
        const uint32_t I1 = __builtin_amdgcn_perm(hi, lo, 0x0B070903u);
        const uint32_t R0 = __builtin_amdgcn_perm(hi << 8, lo << 8, 0x0A050801u);
        const uint32_t R1 = __builtin_amdgcn_perm(hi << 8, lo << 8, 0x0B070903u);
#pragma unroll
        for (int k = 0; k < P / 2; k++) {
            const int j = q - 2 * k;
            if (j < 0 || j >= 8) continue;
            acc[j][0] = dot2(T.l0[k], R0, acc[j][0]);
            acc[j][1] = dot2(T.l0[k], I0, acc[j][1]);
            acc[j][2] = dot2(T.l1[k], R1, acc[j][2]);
            acc[j][3] = dot2(T.l1[k], I1, acc[j][3]);
            if (SPLIT) {
                bcc[j][0] = dot2(T.h0[k], R0, bcc[j][0]);
                bcc[j][1] = dot2(T.h0[k], I0, bcc[j][1]);
                bcc[j][2] = dot2(T.h1[k], R1, bcc[j][2]);
                bcc[j][3] = dot2(T.h1[k], I1, bcc[j][3]);
            }
        }
    }
    if (SPLIT) {
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                acc[j][e] = (int)((uint32_t)acc[j][e] + ((uint32_t)bcc[j][e] << 15));
    }
    __syncthreads(); // window no longer needed: F may overwrite it
    // the /K of the pack step is applied by the FFT (scaled twiddles): a
    // power of two, so the values are those of scaling here
#pragma unroll
    for (int j = 0; j < 8; j++) {
        cf *dst = &F[(8 * g + j) * ROW + 2 * col];
        dst[0] = cf{demote<FPGA>(acc[j][0], 1.0f), demote<FPGA>(acc[j][1], 1.0f)};
        dst[1] = cf{demote<FPGA>(acc[j][2], 1.0f), demote<FPGA>(acc[j][3], 1.0f)};
    }
    __syncthreads();

    // 32 FFTs, 8 lanes each
    const int f = t & 31, l = t >> 5;
    vmfft::fft_stage1<true>(&F[f * ROW], l, tw);
    cf y[2][8];
    vmfft::fft_stage2(&F[f * ROW], l, y);

    // Every lane issues its 16 stores; in a partial tile the lanes past the
    // last spectrum repeat the store of lane f = nvalid - 1 (same address,
    // same value, taken by a lane shuffle), so the instruction stream never
    // branches around stores.
    // Output channel kp = (k + 64) & 127 (fftshift, src/pfb.cpp:315) of
    // k = l + 8 hh + 16 k2: with l + 8 hh < 16 it is l + 64 + 8 hh + 16 k2,
    // less 128 from k2 = 4 on, so each store is a compile-time multiple of
    // the channel stride I nspc from one per-tile base.
    const int fs = min(f, nvalid - 1);
    const int src = (t & 32) + fs; // lane of (fs, l) in this wavefront
    // channel stride and this lane's first output; a tile's 32 spectra are
    // one block of the H16 layout (m0 is a multiple of TILE_M = 32)
    const size_t nsbp = (size_t)(nspc + TILE_M - 1) / TILE_M;
    const size_t cstride = H16 ? nsbp * I * TILE_M : (size_t)I * nspc;
    typename pfb_out<H16>::T *ob =
        H16 ? out + (((size_t)(l + KCH / 2) * nsbp + m0 / TILE_M) * I + oi) * TILE_M + fs
            : out + ((size_t)(l + KCH / 2) * I + oi) * nspc + m0 + fs;
    if (__builtin_expect(nvalid < TILE_M, 0)) { // the last tile of an input only
#pragma unroll
        for (int hh = 0; hh < 2; hh++)
#pragma unroll
            for (int k2 = 0; k2 < 8; k2++) {
                y[hh][k2].x = __shfl(y[hh][k2].x, src, 64);
                y[hh][k2].y = __shfl(y[hh][k2].y, src, 64);
            }
    }
#pragma unroll
    for (int hh = 0; hh < 2; hh++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
            const long coef = 8 * hh + 16 * k2 - (k2 >= 4 ? KCH : 0);
            const cf v = y[hh][k2];
            ob[coef * (long)cstride] = IMF ? pfb_out<H16>::make(v.x, v.y)
                                           : pfb_out<H16>::make(v.y, v.x);
        }
}

// Fused forward PFB, persistent: workgroup b takes tiles [b per, (b+1) per)
// of the tile list (input-major, spectra blocks in time order), so
// consecutive tiles of one workgroup are consecutive spectra of one input.
// The next tile's window is loaded into registers before this tile's stores
// are issued, so the loads never queue behind the stores and the stores
// drain while the next tile computes; with no branches around vector-memory
// instructions the wait for the window is vmcnt(#stores), not vmcnt(0).
// Taps and twiddles are loaded once per workgroup.  htr_bytes: size of the
// chunk's input buffer (loads past it return zeros).
template <bool FPGA, bool SPLIT, bool IMF, bool H16>
__global__ __launch_bounds__(256) void k_pfb_fused(
    const uint32_t *__restrict__ htr, uint32_t htr_bytes, const int *__restrict__ h,
    const int *__restrict__ out_idx, int I, int nspc, typename pfb_out<H16>::T *__restrict__ out,
    int per)
{
    __shared__ cf F[TILE_M * ROW]; // FFT rows (33 KB); the window aliases it
    __shared__ cf tw[KCH];         // twiddles in LDS: no vector-memory loads in the loop
    uint32_t *win = (uint32_t *)F;
    const int t = threadIdx.x;
    const int nmb = (nspc + TILE_M - 1) / TILE_M;
    const int ntiles = nmb * I;
    const int T0 = blockIdx.x * per;
    const int Tend = min(T0 + per, ntiles);
    if (T0 >= Tend) return;
    const pfb_rsrc rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(htr), (short)0,
                                                          (int)htr_bytes, PFB_RSRC_DWORD3);
    if (t < KCH) tw[t] = cf{g_tw128[t].x * (1.0f / KCH), g_tw128[t].y * (1.0f / KCH)};
    pfb_taps T;
    pfb_load_taps<SPLIT>(T, h, t & 63);
    uint4 wr[PFB_WPL];
    pfb_fetch_window(wr, rs, T0 / nmb, I, (T0 % nmb) * TILE_M, t);
    // first tile peeled, so the loop is entered in the same memory state
    // (one window in flight behind one tile's stores) on every iteration
    __syncthreads();
    pfb_put_window(wr, win, t);
    const int t1 = min(T0 + 1, Tend - 1);
    pfb_fetch_window(wr, rs, t1 / nmb, I, (t1 % nmb) * TILE_M, t);
    __syncthreads();
    pfb_tile<FPGA, SPLIT, IMF, H16>(F, win, tw, T, T0 / nmb, I, (T0 % nmb) * TILE_M, nspc,
                                    out_idx[T0 / nmb], out, t);
    for (int tt = T0 + 1; tt < Tend; tt++) {
        const int i = tt / nmb, m0 = (tt % nmb) * TILE_M;
        __syncthreads(); // the previous tile's FFT is done with F
        pfb_put_window(wr, win, t);
        const int tn = min(tt + 1, Tend - 1); // the last prefetch repeats this tile's
        pfb_fetch_window(wr, rs, tn / nmb, I, (tn % nmb) * TILE_M, t);
        __syncthreads();
        pfb_tile<FPGA, SPLIT, IMF, H16>(F, win, tw, T, i, I, m0, nspc, out_idx[i], out, t);
    }
}

// ------------------------------------------------------------ staged ------
// Pass 1: integer WOLA into int2 b[m][i][n] (src/pfb.cpp:121-179 layout),
// one workgroup per (m, i), one lane per n; no atomics needed.
__global__ __launch_bounds__(128) void k_wola_staged(const int8_t *__restrict__ htr,
                                                     const int *__restrict__ h, int I, int P,
                                                     int2 *__restrict__ b)
{
    const int m = blockIdx.x, i = blockIdx.y, n = threadIdx.x;
    int X = 0, Y = 0;
    for (int p = 0; p < P; p++) {
        int hv = h[KCH * P - KCH * p - n - 1];
        size_t s = (size_t)(m + 500 - P) * KCH + (size_t)p * KCH + n;
        size_t xi = mwax_char2(s, i, I);
        X += hv * (int)htr[2 * xi];
        Y += hv * (int)htr[2 * xi + 1];
    }
    b[((size_t)m * I + i) * KCH + n] = make_int2(X, Y);
}

// Pass 2: rounding / demotion in place (src/pfb.cpp:202-268)
template <bool FPGA>
__global__ void k_round_staged(int *__restrict__ b, size_t n)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float v = demote<FPGA>(b[j], 1.0f);
    b[j] = __float_as_int(v);
}

// Pass 3: in-place FFT of nrows 128-point rows, 32 rows per workgroup
__global__ __launch_bounds__(256) void k_fft_staged(cf *__restrict__ b, size_t nrows)
{
    __shared__ cf F[TILE_M * ROW];
    const int t = threadIdx.x;
    const size_t r0 = (size_t)blockIdx.x * TILE_M;
    for (int j = t; j < TILE_M * KCH; j += 256) {
        int rr = j >> 7, n = j & 127;
        F[rr * ROW + n] = (r0 + rr < nrows) ? b[(r0 + rr) * KCH + n] : cf{0.f, 0.f};
    }
    __syncthreads();
    const int f = t & 31, l = t >> 5;
    vmfft::fft_stage1(&F[f * ROW], l, g_tw128);
    cf y[2][8];
    vmfft::fft_stage2(&F[f * ROW], l, y);
    if (r0 + f < nrows) {
#pragma unroll
        for (int hh = 0; hh < 2; hh++)
#pragma unroll
            for (int k2 = 0; k2 < 8; k2++) b[(r0 + f) * KCH + l + 8 * hh + 16 * k2] = y[hh][k2];
    }
}

// Pass 4: /K, fftshift, swap, reorder into VM_FLT [c][i_vcs][s]
__global__ void k_pack_staged(const cf *__restrict__ b, const int *__restrict__ out_idx, int I,
                              int nspc, int imag_first, cf *__restrict__ out)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)nspc * I * KCH;
    if (j >= total) return;
    int k = (int)(j % KCH);
    int i = (int)((j / KCH) % I);
    int m = (int)(j / ((size_t)KCH * I));
    cf v = b[j];
    v.x *= 1.0f / KCH;
    v.y *= 1.0f / KCH;
    int kp = (k + KCH / 2) & (KCH - 1);
    cf o = imag_first ? v : cf{v.y, v.x};
    out[((size_t)kp * I + out_idx[i]) * nspc + m] = o;
}

// --------------------------------------------------- host-format packing --
__device__ __forceinline__ uint8_t demote4(float x)
{
    // DEMOTE: (int)round(x) (half away from zero), clip [-8,7], two's
    // complement nibble (include/vcsbeam.h.in:453-456)
    int v = (int)roundf(x);
    v = v < -8 ? -8 : (v > 7 ? 7 : v);
    return (uint8_t)(v & 0xf);
}

// VM_FLT [c][i][s] -> legacy [s][c][i] as INT4 (mode 0) or complex double
// (mode 1).  Tile: one channel, 64 inputs x 64 samples through LDS.
__global__ __launch_bounds__(256) void k_flt_to_legacy(const cf *__restrict__ in, int I, int nspc,
                                                       int mode, void *__restrict__ out)
{
    __shared__ cf T[64][65];
    const int c = blockIdx.z, i0 = blockIdx.y * 64, s0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        int i = i0 + r, s = s0 + tx;
        T[r][tx] = (i < I && s < nspc) ? in[((size_t)c * I + i) * nspc + s] : cf{0.f, 0.f};
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        int s = s0 + r, i = i0 + tx;
        if (s >= nspc || i >= I) continue;
        cf v = T[tx][r];
        size_t o = ((size_t)s * KCH + c) * I + i;
        if (mode == 0)
            ((uint8_t *)out)[o] = (uint8_t)((demote4(v.x) << 4) + demote4(v.y));
        else
            ((double2 *)out)[o] = make_double2((double)v.x, (double)v.y);
    }
}

// Persistent launch, contiguous runs: exactly as many workgroups as are
// resident at once (registers and LDS decide how many per CU, queried once
// per kernel instantiation), so no workgroup starts after the first ones
// finish.
template <bool FPGA, bool SPLIT, bool IMF, bool H16>
static void pfb_fused_run(const uint32_t *x, size_t htr_bytes, const int *d_h, const int *d_out_idx,
                          int I, int nspc, void *o, hipStream_t stream)
{
    static int ncu = 0, bpc = 0;
    if (!ncu) {
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &bpc, k_pfb_fused<FPGA, SPLIT, IMF, H16>, 256, 0));
        if (bpc < 1) bpc = 1;
    }
    const int ntiles = ((nspc + TILE_M - 1) / TILE_M) * I;
    const int nwg = min(ntiles, bpc * ncu);
    const int per = (ntiles + nwg - 1) / nwg;
    hipLaunchKernelGGL((k_pfb_fused<FPGA, SPLIT, IMF, H16>), dim3((ntiles + per - 1) / per),
                       dim3(256), 0, stream, x, (uint32_t)htr_bytes, d_h, d_out_idx, I, nspc,
                       (typename pfb_out<H16>::T *)o, per);
}

// =================================================================== host ==
extern "C" {

static hipStream_t vm_stream(vcsbeam_context *vm) { return (hipStream_t)vm_dev(vm)->stream; }

void vmInitForwardPFB(vcsbeam_context *vm, int M, pfb_flags flags)
{
    // src/pfb.cpp:378-497
    ensure_twiddles();
    forward_pfb *fpfb = (forward_pfb *)calloc(1, sizeof(forward_pfb));
    if (!fpfb) VM_FATAL("out of host memory");
    vm->fpfb = fpfb;
    fpfb->flags = flags;
    unsigned int nsamples = (unsigned int)(vm->vcs_metadata->num_samples_per_voltage_block *
                                           vm->vcs_metadata->num_voltage_blocks_per_second);
    fpfb->nspectra = nsamples / M;
    fpfb->M = M;
    fpfb->K = vm->analysis_filter->nchans;
    fpfb->I = (int)vm->obs_metadata->num_rf_inputs;
    fpfb->P = vm->analysis_filter->ncoeffs / fpfb->K;
    if (fpfb->K != KCH || M != KCH)
        VM_FATAL("vmInitForwardPFB: only K = M = %d is supported (the MWAX block offset "
                 "500 = 64000/128 is fixed, src/pfb.cpp:153)", KCH);
    if (fpfb->P < 1 || fpfb->P > 500)
        VM_FATAL("vmInitForwardPFB: invalid number of taps %d", fpfb->P);

    fpfb->i_output_idx = (int *)vm_host_alloc_pinned(fpfb->I * sizeof(int));
    HIPCHECK(hipMalloc((void **)&fpfb->d_i_output_idx, fpfb->I * sizeof(int)));
    for (int i = 0; i < fpfb->I; i++) {
        uint32_t ant = vm->obs_metadata->rf_inputs[i].ant;
        char pol = *(vm->obs_metadata->rf_inputs[i].pol);
        int mwax_idx = 2 * ant + (pol - 'X');
        if (mwax_idx < 0 || mwax_idx >= fpfb->I || vm->obs_metadata->rf_inputs[i].vcs_order >=
                                                       (uint32_t)fpfb->I)
            VM_FATAL("vmInitForwardPFB: rf input %d maps outside the input range", i);
        fpfb->i_output_idx[mwax_idx] = vm->obs_metadata->rf_inputs[i].vcs_order;
    }
    HIPCHECK(hipMemcpy(fpfb->d_i_output_idx, fpfb->i_output_idx, fpfb->I * sizeof(int),
                       hipMemcpyHostToDevice));

    if (vm->chunks_per_second < 1) vm->chunks_per_second = 1;
    while (fpfb->nspectra % vm->chunks_per_second != 0) vm->chunks_per_second++;
    fpfb->nspectra_per_chunk = fpfb->nspectra / vm->chunks_per_second;
    fpfb->chunks_per_second = vm->chunks_per_second;
    if (vm->vcs_metadata->num_voltage_blocks_per_second % vm->chunks_per_second != 0)
        VM_FATAL("vmInitForwardPFB: %u chunks do not divide %lu voltage blocks",
                 vm->chunks_per_second,
                 (unsigned long)vm->vcs_metadata->num_voltage_blocks_per_second);
    int blocks_per_chunk = (int)(vm->vcs_metadata->num_voltage_blocks_per_second / vm->chunks_per_second);
    fpfb->d_htr_size = (size_t)(blocks_per_chunk + 1) * vm->vcs_metadata->voltage_block_size_bytes;

    size_t nsamp_chunk = (size_t)fpfb->nspectra_per_chunk * fpfb->I * fpfb->K;
    int type = flags & PFB_TYPE_MASK;
    if (type & PFB_COMPLEX_INT4) {
        fpfb->vcs_size = (size_t)fpfb->nspectra * fpfb->I * fpfb->K * sizeof(uint8_t);
        fpfb->d_vcs_size = nsamp_chunk * sizeof(uint8_t);
        vm->datatype = VM_INT4;
    } else if (type & PFB_COMPLEX_FLOAT32) {
        fpfb->vcs_size = (size_t)fpfb->nspectra * fpfb->I * fpfb->K * sizeof(gpuFloatComplex);
        fpfb->d_vcs_size = nsamp_chunk * sizeof(gpuFloatComplex);
        vm->datatype = VM_FLT;
    } else if (type & PFB_COMPLEX_FLOAT16) {
        // the fused kernel's values rounded to FP16, VM_FLT's layout
        if (flags & PFB_STAGED)
            VM_FATAL("vmInitForwardPFB: PFB_COMPLEX_FLOAT16 output is written by the fused "
                     "kernel only (not with PFB_STAGED)");
        // 32-sample blocks, the last padded: [c][s/32][i][s%32] per chunk
        const size_t nsbp = ((size_t)fpfb->nspectra_per_chunk + 31) / 32;
        fpfb->d_vcs_size = (size_t)fpfb->K * nsbp * 32 * fpfb->I * 2 * sizeof(uint16_t);
        fpfb->vcs_size = fpfb->d_vcs_size * vm->chunks_per_second;
        vm->datatype = VM_HALF;
    } else {
        // default (PFB_COMPLEX_FLOAT64): host output is complex double legacy
        // layout; on the device the beamformer reads the VM_FLT form
        fpfb->vcs_size = (size_t)fpfb->nspectra * fpfb->I * fpfb->K * sizeof(gpuDoubleComplex);
        fpfb->d_vcs_size = nsamp_chunk * sizeof(gpuFloatComplex);
        vm->datatype = VM_FLT;
    }
    fpfb->vcs_stride = fpfb->vcs_size / vm->chunks_per_second;
    fpfb->htr_stride = (vm->vcs_metadata->num_voltage_blocks_per_second *
                        vm->vcs_metadata->voltage_block_size_bytes) / vm->chunks_per_second;
    fpfb->char2s_per_second = (vm->vcs_metadata->num_voltage_blocks_per_second *
                               vm->vcs_metadata->voltage_block_size_bytes) / 2;
    fpfb->bytes_per_block = vm->vcs_metadata->voltage_block_size_bytes;
    fpfb->ninputs_per_cufft_batch = 32;
    fpfb->cufft_batch_size = fpfb->nspectra_per_chunk * fpfb->ninputs_per_cufft_batch;

    size_t filter_size = (size_t)vm->analysis_filter->ncoeffs * sizeof(int);
    fpfb->filter_coeffs = (int *)vm_host_alloc_pinned(filter_size);
    HIPCHECK(hipMalloc((void **)&fpfb->d_filter_coeffs, filter_size));
    int fits16 = 1;
    for (int i = 0; i < vm->analysis_filter->ncoeffs; i++) {
        fpfb->filter_coeffs[i] = (int)vm->analysis_filter->coeffs[i]; // forcible cast, as reference
        if (fpfb->filter_coeffs[i] > 32767 || fpfb->filter_coeffs[i] < -32768) fits16 = 0;
    }
    fpfb->filter_fits_int16 = fits16;
    HIPCHECK(hipMemcpy(fpfb->d_filter_coeffs, fpfb->filter_coeffs, filter_size,
                       hipMemcpyHostToDevice));

    // src/pfb.cpp:464-474: the device input / output buffers are allocated
    // only with their PFB_MALLOC_DEVICE_* flag; otherwise the caller sets the
    // pointer (a buffer shared by several contexts, for instance) and keeps
    // ownership of it
    if (flags & PFB_MALLOC_HOST_OUTPUT) fpfb->vcs_data = vm_host_alloc_pinned(fpfb->vcs_size);
    if (flags & PFB_MALLOC_DEVICE_INPUT)
        HIPCHECK(hipMalloc((void **)&fpfb->d_htr_data, fpfb->d_htr_size));
    if (flags & PFB_MALLOC_DEVICE_OUTPUT) HIPCHECK(hipMalloc(&fpfb->d_vcs_data, fpfb->d_vcs_size));
    if (vm->datatype == VM_INT4)
        HIPCHECK(hipMalloc(&fpfb->d_flt_out, nsamp_chunk * sizeof(gpuFloatComplex)));
    if (flags & PFB_STAGED) {
        fpfb->weighted_overlap_add_size = nsamp_chunk * sizeof(gpuFloatComplex);
        HIPCHECK(hipMalloc((void **)&fpfb->d_weighted_overlap_add, fpfb->weighted_overlap_add_size));
        if (!fpfb->d_flt_out)
            fpfb->d_flt_out = NULL; // staged pack writes straight into d_vcs_data for VM_FLT
    }
    fpfb->plan = NULL; // the FFT is in-house; no plan object
    vm->fine_sample_rate = fpfb->nspectra;
    vm->nfine_chan = fpfb->K;
    vm_dev(vm);
}

void vmFreeForwardPFB(forward_pfb *fpfb)
{
    if (!fpfb) return;
    HIPCHECK(hipDeviceSynchronize());
    if (fpfb->filter_coeffs) HIPCHECK(hipHostFree(fpfb->filter_coeffs));
    if (fpfb->vcs_data) HIPCHECK(hipHostFree(fpfb->vcs_data));
    if (fpfb->i_output_idx) HIPCHECK(hipHostFree(fpfb->i_output_idx));
    if (fpfb->d_filter_coeffs) HIPCHECK(hipFree(fpfb->d_filter_coeffs));
    if (fpfb->d_htr_data && (fpfb->flags & PFB_MALLOC_DEVICE_INPUT))
        HIPCHECK(hipFree(fpfb->d_htr_data));
    if (fpfb->d_vcs_data && (fpfb->flags & PFB_MALLOC_DEVICE_OUTPUT))
        HIPCHECK(hipFree(fpfb->d_vcs_data));
    if (fpfb->d_i_output_idx) HIPCHECK(hipFree(fpfb->d_i_output_idx));
    if (fpfb->d_weighted_overlap_add) HIPCHECK(hipFree(fpfb->d_weighted_overlap_add));
    if (fpfb->d_flt_out) HIPCHECK(hipFree(fpfb->d_flt_out));
    free(fpfb);
}

void vmUploadForwardPFBChunk(vcsbeam_context *vm)
{
    // src/pfb.cpp:544-561: copies d_htr_size bytes (one block more than the
    // chunk stride, for the overlap)
    logger_start_stopwatch(vm->log, "upload", false);
    int chunk = vm->chunk_to_load % vm->chunks_per_second;
    hipStream_t s = vm_stream(vm);
    if (!vm->fpfb->d_htr_data)
        VM_FATAL("vmUploadForwardPFBChunk: no device input buffer (PFB_MALLOC_DEVICE_INPUT "
                 "not set and d_htr_data not provided)");
    HIPCHECK(hipMemcpyAsync(vm->fpfb->d_htr_data,
                            (char *)vm->v->buffer + (size_t)chunk * vm->fpfb->htr_stride,
                            vm->fpfb->d_htr_size, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "upload");
    if (chunk == (int)vm->chunks_per_second - 1) vm->v->locked = false;
}

// Device buffer that receives the VM_FLT output of the fused / staged pack
static cf *flt_target(vcsbeam_context *vm)
{
    forward_pfb *fpfb = vm->fpfb;
    return (cf *)(vm->datatype == VM_INT4 ? fpfb->d_flt_out : fpfb->d_vcs_data);
}

// Launch of the fused kernel on arbitrary device buffers (used by the chunk
// API and by the benchmark / batched paths).
void vm_pfb_fused_launch(const void *d_htr, size_t htr_bytes, const int *d_h,
                         const int *d_out_idx, int I, int nspc, int P, int fpga, int imag_first,
                         int split, int half, void *d_out, hipStream_t stream)
{
    if (P != PFB_P) VM_FATAL("fused PFB kernel is built for 12 taps (got %d); use PFB_STAGED", P);
    if (htr_bytes >= ((size_t)1 << 31))
        VM_FATAL("fused PFB: chunk input of %zu bytes exceeds 2 GiB; use more chunks", htr_bytes);
    const uint32_t *x = (const uint32_t *)d_htr;
    void *o = d_out;
    const int sel = (half ? 8 : 0) + (fpga ? 4 : 0) + (split ? 2 : 0) + (imag_first ? 1 : 0);
#define VM_PFB_RUN(F_, S_, I_, H_) pfb_fused_run<F_, S_, I_, H_>(x, htr_bytes, d_h, d_out_idx, I, nspc, o, stream)
    switch (sel) {
    case 0: VM_PFB_RUN(false, false, false, false); break;
    case 1: VM_PFB_RUN(false, false, true, false); break;
    case 2: VM_PFB_RUN(false, true, false, false); break;
    case 3: VM_PFB_RUN(false, true, true, false); break;
    case 4: VM_PFB_RUN(true, false, false, false); break;
    case 5: VM_PFB_RUN(true, false, true, false); break;
    case 6: VM_PFB_RUN(true, true, false, false); break;
    case 7: VM_PFB_RUN(true, true, true, false); break;
    case 8: VM_PFB_RUN(false, false, false, true); break;
    case 9: VM_PFB_RUN(false, false, true, true); break;
    case 10: VM_PFB_RUN(false, true, false, true); break;
    case 11: VM_PFB_RUN(false, true, true, true); break;
    case 12: VM_PFB_RUN(true, false, false, true); break;
    case 13: VM_PFB_RUN(true, false, true, true); break;
    case 14: VM_PFB_RUN(true, true, false, true); break;
    default: VM_PFB_RUN(true, true, true, true); break;
    }
#undef VM_PFB_RUN
    HIPCHECK(hipGetLastError());
}

void vmWOLAChunk(vcsbeam_context *vm)
{
    forward_pfb *fpfb = vm->fpfb;
    hipStream_t s = vm_stream(vm);
    if (!fpfb->d_htr_data || !fpfb->d_vcs_data)
        VM_FATAL("vmWOLAChunk: no device %s buffer (PFB_MALLOC_DEVICE_%s not set and none "
                 "provided)", fpfb->d_htr_data ? "output" : "input",
                 fpfb->d_htr_data ? "OUTPUT" : "INPUT");
    logger_start_stopwatch(vm->log, "pfb-wola", false);
    if (fpfb->flags & PFB_STAGED) {
        dim3 grid(fpfb->nspectra_per_chunk, fpfb->I), block(KCH);
        hipLaunchKernelGGL(k_wola_staged, grid, block, 0, s, (const int8_t *)fpfb->d_htr_data,
                           fpfb->d_filter_coeffs, fpfb->I, fpfb->P,
                           (int2 *)fpfb->d_weighted_overlap_add);
        HIPCHECK(hipGetLastError());
    } else {
        // fused: WOLA + rounding + FFT + reorder in one pass
        vm_timer_start(vm_dev(vm), VM_T_PFB);
        vm_pfb_fused_launch(fpfb->d_htr_data, fpfb->d_htr_size, fpfb->d_filter_coeffs,
                            fpfb->d_i_output_idx, fpfb->I,
                            fpfb->nspectra_per_chunk, fpfb->P,
                            (fpfb->flags & PFB_EMULATE_FPGA) != 0,
                            (fpfb->flags & PFB_IMAG_PART_FIRST) != 0, !fpfb->filter_fits_int16,
                            vm->datatype == VM_HALF, flt_target(vm), s);
        vm_timer_stop(vm_dev(vm), VM_T_PFB);
    }
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "pfb-wola");
}

void vmFPGARoundingChunk(vcsbeam_context *vm)
{
    forward_pfb *fpfb = vm->fpfb;
    if (!(fpfb->flags & PFB_STAGED)) return; // fused into vmWOLAChunk
    hipStream_t s = vm_stream(vm);
    logger_start_stopwatch(vm->log, "pfb-round", false);
    size_t n = (size_t)fpfb->nspectra_per_chunk * fpfb->I * KCH * 2;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (fpfb->flags & PFB_EMULATE_FPGA)
        hipLaunchKernelGGL(k_round_staged<true>, grid, block, 0, s,
                           (int *)fpfb->d_weighted_overlap_add, n);
    else
        hipLaunchKernelGGL(k_round_staged<false>, grid, block, 0, s,
                           (int *)fpfb->d_weighted_overlap_add, n);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "pfb-round");
}

void vmFFTChunk(vcsbeam_context *vm)
{
    forward_pfb *fpfb = vm->fpfb;
    if (!(fpfb->flags & PFB_STAGED)) return; // fused into vmWOLAChunk
    hipStream_t s = vm_stream(vm);
    logger_start_stopwatch(vm->log, "pfb-fft", false);
    size_t nrows = (size_t)fpfb->nspectra_per_chunk * fpfb->I;
    hipLaunchKernelGGL(k_fft_staged, dim3((unsigned)((nrows + TILE_M - 1) / TILE_M)), dim3(256), 0,
                       s, (cf *)fpfb->d_weighted_overlap_add, nrows);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "pfb-fft");
}

void vmPackChunk(vcsbeam_context *vm)
{
    forward_pfb *fpfb = vm->fpfb;
    hipStream_t s = vm_stream(vm);
    logger_start_stopwatch(vm->log, "pfb-pack", false);
    int nspc = fpfb->nspectra_per_chunk;
    if (fpfb->flags & PFB_STAGED) {
        size_t total = (size_t)nspc * fpfb->I * KCH;
        hipLaunchKernelGGL(k_pack_staged, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           (const cf *)fpfb->d_weighted_overlap_add, fpfb->d_i_output_idx, fpfb->I,
                           nspc, (fpfb->flags & PFB_IMAG_PART_FIRST) != 0, flt_target(vm));
        HIPCHECK(hipGetLastError());
    }
    if (vm->datatype == VM_INT4) {
        dim3 grid((nspc + 63) / 64, (fpfb->I + 63) / 64, KCH);
        hipLaunchKernelGGL(k_flt_to_legacy, grid, dim3(256), 0, s, (const cf *)fpfb->d_flt_out,
                           fpfb->I, nspc, 0, fpfb->d_vcs_data);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "pfb-pack");
}

void vmDownloadForwardPFBChunk(vcsbeam_context *vm)
{
    // src/pfb.cpp:699-715; the host receives the reference layout/format
    forward_pfb *fpfb = vm->fpfb;
    hipStream_t s = vm_stream(vm);
    logger_start_stopwatch(vm->log, "download", false);
    int chunk = vm->chunk_to_load % vm->chunks_per_second;
    char *dst = (char *)fpfb->vcs_data + (size_t)chunk * fpfb->vcs_stride;
    int type = fpfb->flags & PFB_TYPE_MASK;
    if ((type & PFB_COMPLEX_INT4) || (type & PFB_COMPLEX_FLOAT32) || (type & PFB_COMPLEX_FLOAT16)) {
        HIPCHECK(hipMemcpyAsync(dst, fpfb->d_vcs_data, fpfb->d_vcs_size, hipMemcpyDeviceToHost, s));
    } else {
        vm_device_state *d = vm_dev(vm);
        size_t need = fpfb->vcs_stride;
        if (d->stage_bytes < need) {
            if (d->d_stage) HIPCHECK(hipFree(d->d_stage));
            HIPCHECK(hipMalloc(&d->d_stage, need));
            d->stage_bytes = need;
        }
        int nspc = fpfb->nspectra_per_chunk;
        dim3 grid((nspc + 63) / 64, (fpfb->I + 63) / 64, KCH);
        hipLaunchKernelGGL(k_flt_to_legacy, grid, dim3(256), 0, s, (const cf *)fpfb->d_vcs_data,
                           fpfb->I, nspc, 1, d->d_stage);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(dst, d->d_stage, need, hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "download");
}

void vmExecuteForwardPFB(vcsbeam_context *vm)
{
    // src/pfb.cpp:730-773
    if (vm->v == NULL || vm->v->buffer == NULL)
        VM_FATAL("vmExecuteForwardPFB: Input data buffer have not been set");
    logger_start_stopwatch(vm->log, "pfb", true);
    for (unsigned int chunk = 0; chunk < vm->chunks_per_second; chunk++) {
        vmUploadForwardPFBChunk(vm);
        vmWOLAChunk(vm);
        vmFPGARoundingChunk(vm);
        vmFFTChunk(vm);
        vmPackChunk(vm);
        if (vm->fpfb->vcs_data != NULL) vmDownloadForwardPFBChunk(vm);
        vm->chunk_to_load++;
    }
    logger_stop_stopwatch(vm->log, "pfb");
}

void vmWritePFBOutputToFile(vcsbeam_context *vm)
{
    // src/pfb.cpp:781-796: the legacy recombined file of this second and
    // coarse channel, <obsid>_<gps>_ch<receiver channel>.dat (the name
    // mwalib gives it, src/metadata.c:1245-1262), in vm->datadir (default
    // the working directory)
    logger_start_stopwatch(vm->log, "write", true);
    char filename[1200];
    uint32_t gps = vm->gps_seconds_to_process[vm->current_gps_idx ? vm->current_gps_idx - 1 : 0];
    const int cidx = vm->coarse_chan_idxs_to_process[0];
    if (cidx < 0 || (size_t)cidx >= vm->obs_metadata->num_metafits_coarse_chans)
        VM_FATAL("vmWritePFBOutputToFile: coarse channel index %d of %zu", cidx,
                 vm->obs_metadata->num_metafits_coarse_chans);
    snprintf(filename, sizeof(filename), "%s/%u_%u_ch%03u.dat",
             vm->datadir ? (const char *)vm->datadir : ".", vm->obs_metadata->obs_id, gps,
             vm->obs_metadata->metafits_coarse_chans[cidx].rec_chan_number);
    FILE *f = fopen(filename, "w");
    if (!f) VM_FATAL("vmWritePFBOutputToFile: cannot open %s", filename);
    fwrite(vm->fpfb->vcs_data, vm->fpfb->vcs_size, sizeof(uint8_t), f);
    fclose(f);
    logger_stop_stopwatch(vm->log, "write");
}

} // extern "C"
