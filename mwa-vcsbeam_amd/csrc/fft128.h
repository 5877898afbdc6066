// fft128.h -- in-house 128-point complex FFT building blocks for gfx950.
//
// 128 = 16 x 8 (Cooley-Tukey, decimation in time over the 8-way index):
//   n = 8*n1 + n2 (n1 < 16, n2 < 8),  k = k1 + 16*k2 (k1 < 16, k2 < 8)
//   Y[k1 + 16 k2] = sum_n2 W8^{n2 k2} W128^{n2 k1} sum_n1 W16^{n1 k1} x[8 n1 + n2]
// Stage 1 is a 16-point DFT (radix 4x4) held in one lane's registers, then a
// twiddle by W128^{n2 k1}; stage 2 is an 8-point DFT (radix 4x2) per k1.  The
// two stages exchange data through LDS; eight lanes cooperate on one FFT.
//
// The arithmetic sequence is fixed (no FP contraction beyond the explicit
// fmaf calls) so every kernel that uses these functions produces
// bit-identical results for identical inputs.
#pragma once
#include <hip/hip_runtime.h>

#define VM_FFT_PRAGMA _Pragma("clang fp contract(off)")

namespace vmfft {

struct cf { float x, y; };

__device__ __forceinline__ cf add(cf a, cf b) { VM_FFT_PRAGMA return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf sub(cf a, cf b) { VM_FFT_PRAGMA return {a.x - b.x, a.y - b.y}; }
// a * w
__device__ __forceinline__ cf mul(cf a, cf w)
{
    VM_FFT_PRAGMA
    return {__builtin_fmaf(a.x, w.x, -(a.y * w.y)), __builtin_fmaf(a.x, w.y, a.y * w.x)};
}
// multiply by -i
__device__ __forceinline__ cf mul_mi(cf a) { return {a.y, -a.x}; }

// 4-point forward DFT in place (W4 = -i)
__device__ __forceinline__ void dft4(cf &x0, cf &x1, cf &x2, cf &x3)
{
    cf s02 = add(x0, x2), d02 = sub(x0, x2);
    cf s13 = add(x1, x3), d13 = sub(x1, x3);
    x0 = add(s02, s13);
    x2 = sub(s02, s13);
    cf md = mul_mi(d13);
    x1 = add(d02, md);
    x3 = sub(d02, md);
}

// W16^j = exp(-2 pi i j / 16)
#define VM_C1 0.92387953251128674f
#define VM_S1 0.38268343236508977f
#define VM_R2 0.70710678118654752f

// 16-point forward DFT in place: x[n] -> X[k], both natural order.
__device__ __forceinline__ void dft16(cf x[16])
{
    // n = 4a + b: DFT over a for each b
#pragma unroll
    for (int b = 0; b < 4; b++) dft4(x[b], x[4 + b], x[8 + b], x[12 + b]);
    // now x[4c + b] holds T[b][c]; twiddle by W16^{b c}
    x[4 * 1 + 1] = mul(x[5], cf{VM_C1, -VM_S1});   // b=1 c=1: W^1
    x[4 * 1 + 2] = mul(x[6], cf{VM_R2, -VM_R2});   // b=2 c=1: W^2
    x[4 * 1 + 3] = mul(x[7], cf{VM_S1, -VM_C1});   // b=3 c=1: W^3
    x[4 * 2 + 1] = mul(x[9], cf{VM_R2, -VM_R2});   // b=1 c=2: W^2
    x[4 * 2 + 2] = mul_mi(x[10]);                  // b=2 c=2: W^4 = -i
    x[4 * 2 + 3] = mul(x[11], cf{-VM_R2, -VM_R2}); // b=3 c=2: W^6
    x[4 * 3 + 1] = mul(x[13], cf{VM_S1, -VM_C1});  // b=1 c=3: W^3
    x[4 * 3 + 2] = mul(x[14], cf{-VM_R2, -VM_R2}); // b=2 c=3: W^6
    x[4 * 3 + 3] = mul(x[15], cf{-VM_C1, VM_S1});  // b=3 c=3: W^9
    // DFT over b for each c: X[c + 4d]
#pragma unroll
    for (int c = 0; c < 4; c++) dft4(x[4 * c], x[4 * c + 1], x[4 * c + 2], x[4 * c + 3]);
    // x[4c + d] now holds X[c + 4d]: transpose to natural order
    cf t[16];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int d = 0; d < 4; d++) t[c + 4 * d] = x[4 * c + d];
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = t[k];
}

// 8-point forward DFT in place, natural order in and out.
__device__ __forceinline__ void dft8(cf x[8])
{
    // n = 2a + b: DFT4 over a for b = 0, 1
    dft4(x[0], x[2], x[4], x[6]);
    dft4(x[1], x[3], x[5], x[7]);
    // x[2c + b] = T[b][c]; twiddle T[1][c] by W8^c
    x[3] = mul(x[3], cf{VM_R2, -VM_R2});
    x[5] = mul_mi(x[5]);
    x[7] = mul(x[7], cf{-VM_R2, -VM_R2});
    cf t[8];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        t[c] = add(x[2 * c], x[2 * c + 1]);
        t[c + 4] = sub(x[2 * c], x[2 * c + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = t[k];
}

#define KCH 128 // transform length (the MWA fine-channel count)

// Stage 1 for lane (f, l) of a 256-lane workgroup doing 32 FFTs: reads row f
// at n = 8*n1 + l, 16-point DFT over n1, twiddles by W128^{l k1} (table tw),
// and writes back at index k1*8 + l of the same row.  Contains two
// workgroup barriers, so every lane of the workgroup must call it.
// SCALE: the caller's table carries a factor 1/KCH (exact: a power of two),
// and the k1 = 0 term, which has no twiddle, is scaled here, so the whole
// transform comes out divided by KCH with bit-identical values.
template <bool SCALE = false>
__device__ __forceinline__ void fft_stage1(cf *row, int l, const cf *tw)
{
    cf x[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) x[n1] = row[8 * n1 + l];
    dft16(x);
    if (SCALE) x[0] = cf{x[0].x * (1.0f / KCH), x[0].y * (1.0f / KCH)};
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) x[k1] = mul(x[k1], tw[l * k1]);
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) row[k1 * 8 + l] = x[k1];
    __syncthreads();
}

// Stage 2 for lane (f, l): reads row f at kk*8 + n2 for kk in {l, l+8} and
// returns Y[kk + 16*k2] in y[h][k2] (kk = l + 8*h).
__device__ __forceinline__ void fft_stage2(const cf *row, int l, cf y[2][8])
{
#pragma unroll
    for (int h = 0; h < 2; h++) {
        int kk = l + 8 * h;
#pragma unroll
        for (int n2 = 0; n2 < 8; n2++) y[h][n2] = row[kk * 8 + n2];
        dft8(y[h]);
    }
}

} // namespace vmfft
