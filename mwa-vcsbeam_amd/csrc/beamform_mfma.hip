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
// The single-weight FP16 form uses phi_hi alone: half the MFMA work and
// operand traffic, with the weights' FP16 rounding (relative 2^-11) adding
// about as much Stokes error as the voltages' own FP16 rounding.
// Phases and Jones inverses are produced in FP64 (geometry.hip) and demoted
// only at the split (phases) or to FP32 for the prep pass (J^-1).
// GEMM-side templates take PREC: 0 FP16 split, 1 BF16 split, 2 FP16 single.
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

// global-memory qualifier for pointers the compiler cannot otherwise place
// (loaded from memory rather than passed as kernel arguments): global_*
// rather than flat_* instructions, so stores do not hold LDS waits
#define gmem __attribute__((address_space(1)))
typedef float f4v __attribute__((ext_vector_type(4)));

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

// GEMM precision PREC: operand type (BF16 for 1) and weight parts (1 for 2)
template <int PREC> constexpr bool prec_bf16() { return PREC == 1; }
template <int PREC> constexpr int prec_parts() { return PREC == 2 ? 1 : 2; }
static int form_prec(int form)
{
    return form == VM_BF_MFMA_BF16 ? 1 : form == VM_BF_MFMA_FP16_SINGLE ? 2 : 0;
}

// exponent e with x * 2^e in [2^13, 2^14) for the largest component x > 0
__device__ __forceinline__ int scale_exp(float mx)
{
    int ex = 0;
    if (mx > 0.f) frexpf(mx, &ex);
    return 14 - ex;
}

// scale of the phase weights: |Re phi|, |Im phi| <= 1
#define PHASE_EXP 13
// floats per (channel, pointing) of the per-pointing-beam constants (k_pb_beam)
#define PB_ROW 24

// ------------------------------------------- per-pointing primary beam ----
// With a primary beam B[p][c] that is the same for every tile but differs
// between pointings (vmCalcJDeviceChanB), J^-1_pac = F_pac B_pc^-1 D_ac^-1
// with F_pac = ||D_ac B_pc||_F (src/jones.c:97-157), so for every pointing
//   E_p = B_p^-1 sum_a (phi_pa F_pa) u_a,   u_a = D_a^-1 v_a,
// i.e. the same GEMM with weights phi F and a 2x2 B^-1 in the epilogue, and
//   N_p = B_p^-1 (sum_a F_pa^2 u_a u_a^H) B_p^-H.
// F_pa^2 = tr(P_a Q_p), P_a = D_a^H D_a, Q_p = B_p B_p^H, which for Hermitian
// 2x2 matrices is sum_k g_a[k] h_p[k] with g = (P00, P11, Re P01, Im P01)
// and h = (Q00, Q11, 2 Re Q01, 2 Im Q01); hence
//   sum_a F_pa^2 u_a u_a^H = sum_k h_p[k] T_k,  T_k = sum_a g_a[k] u_a u_a^H,
// and the four T_k are beam-independent sums formed in the prep pass.  The
// epilogue forms X = B^-1 (|E'|^2 - sum_k h_k T_k) B^-H per beam and sample.
// The weights of pointing p are scaled by 2^esp[c][p], chosen so that
// max_a F_pa 2^esp lies in [2^12, 2^13).

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// ||D B||_F^2 of row-major 2x2 complex matrices
__device__ __forceinline__ double frob2_DB(const double2 *D, const double2 *B)
{
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const double2 u = zmul(D[2 * i], B[k]), w = zmul(D[2 * i + 1], B[2 + k]);
            const double re = u.x + w.x, im = u.y + w.y;
            f += re * re + im * im;
        }
    return f;
}

// Per (c, GEMM slot a): D^-1 as float [c][a][4] (zero for a singular D,
// whose J^-1 is zero) and g of P = D^H D as float4 [c][a], for slots a <
// nslot standing for antenna amap[a] (amap NULL: a itself; amap[a] < 0: a
// padding slot, zero).  D: [antenna][c][4] complex double.
__global__ void k_pb_dinv(const double2 *__restrict__ D, int nslot, int nchan,
                          const int *__restrict__ amap, cf *__restrict__ Dinv,
                          float4 *__restrict__ g)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)nslot * nchan) return;
    const int c = (int)(j % nchan), a = (int)(j / nchan);
    const int as = amap ? amap[a] : a;
    cf *o = Dinv + ((size_t)c * nslot + a) * 4;
    if (as < 0) {
        o[0] = o[1] = o[2] = o[3] = cf{0.f, 0.f};
        g[(size_t)c * nslot + a] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const double2 *m = D + ((size_t)as * nchan + c) * 4;
    const double2 d00 = m[0], d01 = m[1], d10 = m[2], d11 = m[3];
    const double2 det = make_double2(d00.x * d11.x - d00.y * d11.y - (d01.x * d10.x - d01.y * d10.y),
                                     d00.x * d11.y + d00.y * d11.x - (d01.x * d10.y + d01.y * d10.x));
    const double dd = det.x * det.x + det.y * det.y;
    if (dd == 0.0) {
        o[0] = o[1] = o[2] = o[3] = cf{0.f, 0.f};
    } else {
        const double2 r = make_double2(det.x / dd, -det.y / dd); // 1 / det
        const double2 e00 = zmul(d11, r), e11 = zmul(d00, r);
        const double2 e01 = zmul(make_double2(-d01.x, -d01.y), r);
        const double2 e10 = zmul(make_double2(-d10.x, -d10.y), r);
        o[0] = cf{(float)e00.x, (float)e00.y};
        o[1] = cf{(float)e01.x, (float)e01.y};
        o[2] = cf{(float)e10.x, (float)e10.y};
        o[3] = cf{(float)e11.x, (float)e11.y};
    }
    // P = D^H D: P00 = |d00|^2 + |d10|^2, P11 = |d01|^2 + |d11|^2,
    // P01 = conj(d00) d01 + conj(d10) d11
    const double p00 = d00.x * d00.x + d00.y * d00.y + d10.x * d10.x + d10.y * d10.y;
    const double p11 = d01.x * d01.x + d01.y * d01.y + d11.x * d11.x + d11.y * d11.y;
    const double p01r = d00.x * d01.x + d00.y * d01.y + d10.x * d11.x + d10.y * d11.y;
    const double p01i = d00.x * d01.y - d00.y * d01.x + d10.x * d11.y - d10.y * d11.x;
    g[(size_t)c * nslot + a] = make_float4((float)p00, (float)p11, (float)p01r, (float)p01i);
}

// Per (c, p): h of Q = B B^H and the products of the entries of B^-1 that
// stokes_pb uses, as bx[c][p][PB_ROW] floats, and the weight exponent
// esp[c][p].  B: [p][c][4], D: [a][c][4].
__global__ void k_pb_beam(const double2 *__restrict__ D, const double2 *__restrict__ B, int np,
                          int nant, int nchan, float *__restrict__ bx, int *__restrict__ esp)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)np * nchan) return;
    const int c = (int)(j % nchan), p = (int)(j / nchan);
    double2 b[4];
#pragma unroll
    for (int q = 0; q < 4; q++) b[q] = B[((size_t)p * nchan + c) * 4 + q];
    const double2 det = make_double2(b[0].x * b[3].x - b[0].y * b[3].y - (b[1].x * b[2].x - b[1].y * b[2].y),
                                     b[0].x * b[3].y + b[0].y * b[3].x - (b[1].x * b[2].y + b[1].y * b[2].x));
    const double dd = det.x * det.x + det.y * det.y;
    const double2 r = dd > 0.0 ? make_double2(det.x / dd, -det.y / dd) : make_double2(0.0, 0.0);
    const double2 inv[4] = {zmul(b[3], r), zmul(make_double2(-b[1].x, -b[1].y), r),
                            zmul(make_double2(-b[2].x, -b[2].y), r), zmul(b[0], r)};
    float *o = bx + ((size_t)c * np + p) * PB_ROW;
    // Q = B B^H: Q00 = |b00|^2 + |b01|^2, Q11 = |b10|^2 + |b11|^2,
    // Q01 = b00 conj(b10) + b01 conj(b11)
    const double q00 = b[0].x * b[0].x + b[0].y * b[0].y + b[1].x * b[1].x + b[1].y * b[1].y;
    const double q11 = b[2].x * b[2].x + b[2].y * b[2].y + b[3].x * b[3].x + b[3].y * b[3].y;
    const double q01r = b[0].x * b[2].x + b[0].y * b[2].y + b[1].x * b[3].x + b[1].y * b[3].y;
    const double q01i = b[0].y * b[2].x - b[0].x * b[2].y + b[1].y * b[3].x - b[1].x * b[3].y;
    o[0] = (float)q00;
    o[1] = (float)q11;
    o[2] = (float)(2.0 * q01r);
    o[3] = (float)(2.0 * q01i);
    // products of the entries of M = B^-1 that the epilogue's M X M^H needs
    // (per pointing, so formed here once in FP64): |m_ij|^2, then m00 m01*,
    // m10 m11*, m00 m10*, m01 m11*, m00 m11*, m01 m10*
#pragma unroll
    for (int q = 0; q < 4; q++) o[4 + q] = (float)(inv[q].x * inv[q].x + inv[q].y * inv[q].y);
    const int pr[6][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {0, 3}, {1, 2}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double2 u = inv[pr[k][0]], w = inv[pr[k][1]];
        o[8 + 2 * k] = (float)(u.x * w.x + u.y * w.y);
        o[9 + 2 * k] = (float)(u.y * w.x - u.x * w.y);
    }
#pragma unroll
    for (int k = 20; k < PB_ROW; k++) o[k] = 0.f;
    double f2 = 0.0;
    for (int a = 0; a < nant; a++) f2 = fmax(f2, frob2_DB(D + ((size_t)a * nchan + c) * 4, b));
    int ex = 0;
    if (f2 > 0.0) frexp(sqrt(f2), &ex);
    esp[(size_t)c * np + p] = 13 - ex;
}

// ------------------------------------------- primary beam per tile group ---
// With a primary beam per tile that takes a few distinct values (one per
// dipole configuration; groups g of tiles sharing B_pg, at most
// VM_MFMA_MAX_GROUPS), J^-1_pa = F_pa B_pg^-1 D_a^-1 for tile a of group g
// (src/jones.c:97-157), so
//   e_p = sum_g B_pg^-1 E_pg,   E_pg = sum_{a in g} (phi_pa F_pa) u_a,
//   N_p = sum_g B_pg^-1 Y_pg B_pg^-H,   Y_pg = sum_k h_pg[k] T_gk,
// with u_a = D_a^-1 v_a and T_gk the prep pass's T_k summed over the tiles
// of group g alone.  The GEMM's K runs over the groups in turn (each group's
// tiles padded to whole k-steps), and between groups the accumulators are
// carried over by R_pg = B_pg B_p,g-1^-1 (a Horner scheme): after the last
// group they hold B_p,G-1 sum_g B_pg^-1 E_pg, so one set of accumulators
// serves every group and the epilogue applies M = B_p,G-1^-1 once.  Each
// group's auto-correlation term goes through S_pg, the linear map
// Y -> B_pg^-1 Y B_pg^-H of Hermitian Y written (Y00, Y11, Re Y01, Im Y01).
#define VM_MFMA_MAX_GROUPS 4
#define GX_G 28                                   // per group: h (4), S (16), R (8)
#define GX_ROW (8 + GX_G * VM_MFMA_MAX_GROUPS)    // M (8) then the groups

// GEMM slots [t[g], t[g + 1]) belong to group g (multiples of 8), n groups;
// rep[g]: an antenna of the group (whose B stands for it)
struct grp_bounds {
    int t[VM_MFMA_MAX_GROUPS + 1];
    int rep[VM_MFMA_MAX_GROUPS];
    int n;
};

__device__ __forceinline__ int slot_group(const grp_bounds &gb, int a)
{
    int g = 0;
#pragma unroll
    for (int k = 1; k < VM_MFMA_MAX_GROUPS; k++) g += (k < gb.n && a >= gb.t[k]) ? 1 : 0;
    return g;
}

// inverse of a 2x2 complex matrix (zero when singular)
__device__ __forceinline__ void inv2x2(const double2 b[4], double2 o[4])
{
    const double2 det = make_double2(b[0].x * b[3].x - b[0].y * b[3].y - (b[1].x * b[2].x - b[1].y * b[2].y),
                                     b[0].x * b[3].y + b[0].y * b[3].x - (b[1].x * b[2].y + b[1].y * b[2].x));
    const double dd = det.x * det.x + det.y * det.y;
    const double2 r = dd > 0.0 ? make_double2(det.x / dd, -det.y / dd) : make_double2(0.0, 0.0);
    o[0] = zmul(b[3], r);
    o[1] = zmul(make_double2(-b[1].x, -b[1].y), r);
    o[2] = zmul(make_double2(-b[2].x, -b[2].y), r);
    o[3] = zmul(b[0], r);
}

// Bg[p][g][c][4] = B[p][rep[g]][4]: the per-tile beam (frequency
// independent) of each group's representative antenna
__global__ void k_bg_build(const double2 *__restrict__ B, int np, int nant, int nchan,
                           grp_bounds gb, double2 *__restrict__ Bg)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)np * gb.n * nchan) return;
    const int p = (int)(j / ((size_t)gb.n * nchan)), g = (int)(j / nchan % gb.n);
    const double2 *src = B + ((size_t)p * nant + gb.rep[g]) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) Bg[j * 4 + q] = src[q];
}

// Per (c, p): gx[c][p][GX_ROW] = M = B_p,G-1^-1 (4 complex), then per group
// g: h of Q = B_pg B_pg^H (as k_pb_beam), S_pg (4 x 4, row-major, mapping
// (Y00, Y11, Re Y01, Im Y01) to the same of M_g Y M_g^H, M_g = B_pg^-1) and
// R_pg = B_pg M_p,g-1 (4 complex; g >= 1); and the weight exponent esp[c][p]
// with max_a F_pa 2^esp in [2^12, 2^13).  Bg: [p][g][c][4]; D: [antenna]
// [c][4]; amap: slot -> antenna (< 0: padding).
__global__ void k_pbg_consts(const double2 *__restrict__ Bg, const double2 *__restrict__ D,
                             const int *__restrict__ amap, grp_bounds gb, int np, int nchan,
                             float *__restrict__ gx, int *__restrict__ esp)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)np * nchan) return;
    const int c = (int)(j % nchan), p = (int)(j / nchan);
    float *o = gx + ((size_t)c * np + p) * GX_ROW;
    double2 Mprev[4];
    for (int g = 0; g < gb.n; g++) {
        double2 b[4], m[4];
#pragma unroll
        for (int q = 0; q < 4; q++) b[q] = Bg[(((size_t)p * gb.n + g) * nchan + c) * 4 + q];
        inv2x2(b, m);
        float *og = o + 8 + GX_G * g;
        const double q00 = b[0].x * b[0].x + b[0].y * b[0].y + b[1].x * b[1].x + b[1].y * b[1].y;
        const double q11 = b[2].x * b[2].x + b[2].y * b[2].y + b[3].x * b[3].x + b[3].y * b[3].y;
        const double q01r = b[0].x * b[2].x + b[0].y * b[2].y + b[1].x * b[3].x + b[1].y * b[3].y;
        const double q01i = b[0].y * b[2].x - b[0].x * b[2].y + b[1].y * b[3].x - b[1].x * b[3].y;
        og[0] = (float)q00;
        og[1] = (float)q11;
        og[2] = (float)(2.0 * q01r);
        og[3] = (float)(2.0 * q01i);
        // S: X00 = |m00|^2 Y00 + |m01|^2 Y11 + 2 Re(m00 m01* Y01), likewise
        // X11 with row 1; X01 = m00 m10* Y00 + m01 m11* Y11 + m00 m11* Y01 +
        // m01 m10* Y01*
        const double2 pp = zmul(m[0], zconj(m[1])), qq = zmul(m[2], zconj(m[3]));
        const double2 aa = zmul(m[0], zconj(m[2])), bb = zmul(m[1], zconj(m[3]));
        const double2 cc = zmul(m[0], zconj(m[3])), dd = zmul(m[1], zconj(m[2]));
        const double S[16] = {
            m[0].x * m[0].x + m[0].y * m[0].y, m[1].x * m[1].x + m[1].y * m[1].y, 2.0 * pp.x, -2.0 * pp.y,
            m[2].x * m[2].x + m[2].y * m[2].y, m[3].x * m[3].x + m[3].y * m[3].y, 2.0 * qq.x, -2.0 * qq.y,
            aa.x, bb.x, cc.x + dd.x, dd.y - cc.y,
            aa.y, bb.y, cc.y + dd.y, cc.x - dd.x};
#pragma unroll
        for (int q = 0; q < 16; q++) og[4 + q] = (float)S[q];
        if (g > 0) {
            double2 r[4];
            r[0] = zadd(zmul(b[0], Mprev[0]), zmul(b[1], Mprev[2]));
            r[1] = zadd(zmul(b[0], Mprev[1]), zmul(b[1], Mprev[3]));
            r[2] = zadd(zmul(b[2], Mprev[0]), zmul(b[3], Mprev[2]));
            r[3] = zadd(zmul(b[2], Mprev[1]), zmul(b[3], Mprev[3]));
#pragma unroll
            for (int q = 0; q < 4; q++) {
                og[20 + 2 * q] = (float)r[q].x;
                og[21 + 2 * q] = (float)r[q].y;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) og[20 + q] = 0.f;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) Mprev[q] = m[q];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        o[2 * q] = (float)Mprev[q].x;
        o[2 * q + 1] = (float)Mprev[q].y;
    }
    for (int q = 8 + GX_G * gb.n; q < GX_ROW; q++) o[q] = 0.f;
    double f2 = 0.0;
    for (int a = 0; a < gb.t[gb.n]; a++) {
        const int as = amap[a];
        if (as < 0) continue;
        const int g = slot_group(gb, a);
        f2 = fmax(f2, frob2_DB(D + ((size_t)as * nchan + c) * 4,
                               Bg + (((size_t)p * gb.n + g) * nchan + c) * 4));
    }
    int ex = 0;
    if (f2 > 0.0) frexp(sqrt(f2), &ex);
    esp[(size_t)c * np + p] = 13 - ex;
}

// ------------------------------------------------------------ weights -----
// Operand A in fragment order: Af[c][rb][ks][h][lane] (uint4), rb a block of
// 32 GEMM rows (16 beams), ks a step of 16 k' (8 tiles), h = 0 hi part, 1 lo
// part.  Lane l holds row 32 rb + l%32 and k' = 16 ks + 8 (l/32) + j, j =
// 0..7, i.e. tiles a = 8 ks + 4 (l/32) + j/2 with j even the coefficient of
// u.re and j odd of u.im.  Rows: 2 beam + {re, im}: re rows (Re phi, -Im phi),
// im rows (Im phi, Re phi); rows of beams >= npointing are zero.  Weights are
// scaled by 2^PHASE_EXP.  phi: [p][a][c] complex double over nant_src
// antennas; GEMM tile a is antenna amap[a] (amap NULL: a itself), nant tiles.
// PB 1 (per-pointing primary beam): the weight is phi F_pa 2^esp[c][p] with
// F_pa = ||D_a B_p||_F (D [a][c][4], B [p][c][4]).  PB 2 (a beam per tile
// group): the same with B = Bg[p][g][c][4] of the slot's group g (gb), and
// slots whose amap is negative (group padding) are zero.
template <bool BF16, int PB>
__global__ __launch_bounds__(256) void k_phase_weights(const double2 *__restrict__ phi,
                                                       int npointing, int nant, int nchan,
                                                       int nrb, int nkstep,
                                                       const int *__restrict__ amap, int nant_src,
                                                       const double2 *__restrict__ D,
                                                       const double2 *__restrict__ B,
                                                       const int *__restrict__ esp,
                                                       uint4 *__restrict__ Af, grp_bounds gb)
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
    double ws = ldexp(1.0, PHASE_EXP);
    if (PB && p < npointing) ws = ldexp(1.0, esp[(size_t)c * npointing + p]);
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint16_t hb[2] = {0, 0}, lb[2] = {0, 0};
        const int a = 8 * ks + 4 * (lane >> 5) + q;
        const int as = (a < nant) ? (amap ? amap[a] : a) : -1;
        if (p < npointing && as >= 0) {
            const double2 f = phi[((size_t)p * nant_src + as) * nchan + c];
            double wf = ws;
            if (PB == 1)
                wf *= sqrt(frob2_DB(D + ((size_t)as * nchan + c) * 4,
                                    B + ((size_t)p * nchan + c) * 4));
            if (PB == 2)
                wf *= sqrt(frob2_DB(D + ((size_t)as * nchan + c) * 4,
                                    B + (((size_t)p * gb.n + slot_group(gb, a)) * nchan + c) * 4));
            const double vv[2] = {(im ? f.y : f.x) * wf, (im ? f.x : -f.y) * wf};
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

// J^-1 of pointing 0 as float [c][a][4], a = 0..nant-1 standing for
// antenna amap[a] (amap NULL: a itself)
__global__ void k_jf(const double2 *__restrict__ J, int nant, int nchan,
                     const int *__restrict__ amap, cf *__restrict__ Jf)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)nant * nchan) return;
    int c = (int)(j % nchan), a = (int)(j / nchan);
    const int as = amap ? amap[a] : a;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double2 m = J[((size_t)as * nchan + c) * 4 + q];
        Jf[((size_t)c * nant + a) * 4 + q] = cf{(float)m.x, (float)m.y};
    }
}

// flag when J^-1 of any pointing p >= 1 differs from pointing 0's; one
// workgroup per (pointing p >= 1, block of 4096 entries), 32-bit indexing,
// and a wavefront sets the flag only while it is still clear
__global__ void k_jcheck(const double2 *__restrict__ J, uint32_t per_pointing,
                         uint32_t *__restrict__ flag)
{
    const uint32_t p = blockIdx.y + 1;
    const double2 *Jp = J + (size_t)p * per_pointing;
    bool d = false;
    for (uint32_t j = blockIdx.x * 4096u + threadIdx.x; j < min(per_pointing, blockIdx.x * 4096u + 4096u);
         j += 256u) {
        const double2 m = Jp[j], m0 = J[j];
        d |= (m.x != m0.x) || (m.y != m0.y);
    }
    if (__any(d) && (threadIdx.x & 63) == 0 && __hip_atomic_load(flag, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT) == 0u)
        atomicOr(flag, 1u);
}

// ------------------------------------------------------- voltage prep -----
// One workgroup of 32 NG lanes per (32-sample block sb, channel c), one pass
// over the block: lane (sl, g) loads sample 32 sb + sl of both inputs of
// its APT tiles (quads of consecutive tiles, prep_ant), forms u = J^-1 v with
// Jf = J^-1 (float, [c][a][4]), the auto-correlation sums N = sum_a
// (|u_x|^2, |u_y|^2, Re u_x u_y^*, Im u_x u_y^*) and the block's largest
// |component| of u.  The NG partial sums per sample are reduced in LDS and
// written as N[c][s]; u, scaled by 2^es, is written straight from registers
// as B fragments Bf[c][sb][ks][pol][lane] (uint4): lane l holds sample
// 32 sb + l%32 and tiles 8 ks + 4 (l/32) + {0..3} as (re, im) pairs of
// polarisation pol, so each thread's quad is one 16-byte slot.
// Samples >= ns are zero; tiles >= nant are zero.  sexp[c][sb] = es.
// v: VM_FLT [c][i][s].  Needs nant <= NG APT.
// PB 1 (per-pointing primary beam): Jf holds D^-1 and, instead of the four N
// terms, N receives T_k = sum_a gw[c][a].k u_a u_a^H (k = 0..3, each as
// (XX, YY, Re XY, Im XY)) as float4 [c][s][4].  PB 2 (a beam per tile
// group): "antennas" are GEMM slots (polQ, polP, Jf and gw given per slot,
// padding slots with zero Jf), and the T_k are summed per group g of gb
// into float4 [c][s][VM_MFMA_MAX_GROUPS][4].
// VIN, the voltages' kind: PREP_VIN_F32 VM_FLT [c][i][s] complex float;
// PREP_VIN_F16 VM_HALF, complex FP16 in 32-sample blocks [c][s/32][i][s%32]
// (a workgroup's block of every input is one contiguous run, staged in LDS:
// read directly in VCS order, its 128-byte runs are slower); PREP_VIN_I4 legacy
// 4+4-bit [s][c][i] (one byte per input-sample, high nibble re, low nibble
// im, two's complement; include/vcsbeam.h.in:997-1003), whose block's 32
// rows of ni bytes are staged in LDS with coalesced loads and each lane
// decodes its inputs from there.
#define PREP_VIN_F32 0
#define PREP_VIN_I4 1
#define PREP_VIN_F16 2
__device__ __forceinline__ cf nibbles_to_cf(uint32_t b)
{
    const int re = (int)((b >> 4) & 0xfu), im = (int)(b & 0xfu);
    return cf{(float)(re >= 8 ? re - 16 : re), (float)(im >= 8 ? im - 16 : im)};
}

#define PREP_I4_MAX_INPUTS 512
#define PREP_F16_MAX_INPUTS 512
// Antenna j of thread group g: groups own quads of consecutive antennas
// (quad g + NG (j / 4), antenna j % 4 in it), so one thread holds the four
// tiles of one 16-byte B-fragment slot and writes it with one store.
template <int NG>
__device__ __forceinline__ int prep_ant(int g, int j)
{
    return 4 * (g + NG * (j >> 2)) + (j & 3);
}

template <bool BF16, int APT, int NG, int PB, int VIN>
__global__ __launch_bounds__(32 * NG) void k_bf_prep(const void *__restrict__ vv, int ns, int ni, int nsb,
                                                 int nkstep, const cf *__restrict__ Jf,
                                                 const float4 *__restrict__ gw,
                                                 const uint32_t *__restrict__ polQ,
                                                 const uint32_t *__restrict__ polP, int nant,
                                                 uint4 *__restrict__ Bf, int *__restrict__ sexp,
                                                 float4 *__restrict__ N, grp_bounds gb)
{
    static_assert(APT % 4 == 0, "antennas per thread come in quads");
    constexpr int NGRP = PB == 2 ? VM_MFMA_MAX_GROUPS : 1;
    constexpr bool I4 = VIN == PREP_VIN_I4;
    constexpr bool H16 = VIN == PREP_VIN_F16;
    __shared__ uint32_t vrow[I4 ? SB * PREP_I4_MAX_INPUTS / 4 : 1];
    const cf *v = (const cf *)vv;
    typedef mtype<BF16> M;
    __shared__ float4 red[NG][SB];
    __shared__ float wmx[NG / 2];
    const int sb = blockIdx.x, c = blockIdx.y;
    const int t = threadIdx.x, sl = t & 31, g = t >> 5;
    const int s = SB * sb + sl;
    const bool ok = s < ns;
    const cf *vc = v + (size_t)c * ni * ns;
    constexpr int NT = 32 * NG; // lanes
    // H16: this block of every input, [c][s/32][i][s%32] complex FP16, is one
    // contiguous run of SB ni dwords, staged in LDS with 16-byte loads (the
    // inputs are then read in VCS order from there; samples past ns are the
    // padding of the last block and are never used)
    __shared__ uint4 stg16[H16 ? SB * PREP_F16_MAX_INPUTS / 4 : 1];
    const uint32_t *vh = (const uint32_t *)stg16 + sl;
    if constexpr (H16) {
        const uint4 *src = (const uint4 *)((const uint32_t *)vv + ((size_t)c * nsb + sb) * ni * SB);
        for (int k = t; k < SB * ni / 4; k += NT) stg16[k] = src[k];
        __syncthreads();
    }
    if constexpr (I4) {
        // rows (s, c) of ni bytes for the block's samples; rows past ns zero
        const uint8_t *vb = (const uint8_t *)vv;
        const int nchan = (int)gridDim.y;
        uint8_t *lb = (uint8_t *)vrow;
        if ((ni & 3) == 0) {
            const int dw = ni >> 2;
            for (int k = t; k < SB * dw; k += NT) {
                const int r = k / dw, q = k % dw, sr = SB * sb + r;
                vrow[r * dw + q] = sr < ns ? ((const uint32_t *)(vb + ((size_t)sr * nchan + c) * ni))[q]
                                           : 0u;
            }
        } else {
            for (int k = t; k < SB * ni; k += NT) {
                const int r = k / ni, q = k % ni, sr = SB * sb + r;
                lb[k] = sr < ns ? vb[((size_t)sr * nchan + c) * ni + q] : (uint8_t)0;
            }
        }
        __syncthreads();
    }

    cf ux[APT], uy[APT];
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const int a = prep_ant<NG>(g, j);
        cf q = {0.f, 0.f}, p = {0.f, 0.f};
        if (ok && a < nant) {
            if constexpr (I4) {
                const uint8_t *lb = (const uint8_t *)vrow + sl * ni;
                q = nibbles_to_cf(lb[polQ[a]]);
                p = nibbles_to_cf(lb[polP[a]]);
            } else if constexpr (H16) {
                const uint32_t hq = vh[polQ[a] * SB], hp = vh[polP[a] * SB];
                q = cf{(float)__builtin_bit_cast(_Float16, (uint16_t)(hq & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(hq >> 16))};
                p = cf{(float)__builtin_bit_cast(_Float16, (uint16_t)(hp & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(hp >> 16))};
            } else {
                q = vc[(size_t)polQ[a] * ns + s];
                p = vc[(size_t)polP[a] * ns + s];
            }
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
    float4 tk[NGRP][4];
#pragma unroll
    for (int q = 0; q < NGRP; q++)
#pragma unroll
        for (int k = 0; k < 4; k++) tk[q][k] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < APT; j++) {
        const cf x = ux[j], y = uy[j];
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(y.x), fabsf(y.y))));
        const float rxx = x.x * x.x + x.y * x.y, ryy = y.x * y.x + y.y * y.y;
        const float rxr = x.x * y.x + x.y * y.y, rxi = x.y * y.x - x.x * y.y;
        if constexpr (PB != 0) {
            const int a = prep_ant<NG>(g, j);
            if (a < nant) {
                const float4 ga = gw[(size_t)c * nant + a];
                const float gk[4] = {ga.x, ga.y, ga.z, ga.w};
                const int ag = PB == 2 ? slot_group(gb, a) : 0;
#pragma unroll
                for (int q = 0; q < NGRP; q++) {
                    const float on = (q == ag) ? 1.f : 0.f; // the slot's group only
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float w = on * gk[k];
                        tk[q][k].x += w * rxx;
                        tk[q][k].y += w * ryy;
                        tk[q][k].z += w * rxr;
                        tk[q][k].w += w * rxi;
                    }
                }
            }
        } else {
            nxx += rxx;
            nyy += ryy;
            xyr += rxr;
            xyi += rxi;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if ((t & 63) == 0) wmx[t >> 6] = mx;
    if constexpr (PB != 0) {
        const int ngrp = PB == 2 ? gb.n : 1;
#pragma unroll
        for (int q = 0; q < NGRP; q++) {
            if (q >= ngrp) break; // uniform over the workgroup
#pragma unroll
            for (int k = 0; k < 4; k++) {
                red[g][sl] = tk[q][k];
                __syncthreads();
                if (g == 0 && ok) {
                    float4 acc = red[0][sl];
#pragma unroll
                    for (int r2 = 1; r2 < NG; r2++) {
                        const float4 r = red[r2][sl];
                        acc.x += r.x; acc.y += r.y; acc.z += r.z; acc.w += r.w;
                    }
                    N[(((size_t)c * ns + s) * NGRP + q) * 4 + k] = acc;
                }
                __syncthreads();
            }
        }
    } else {
        red[g][sl] = make_float4(nxx, nyy, xyr, xyi);
        __syncthreads();
        if (g == 0 && ok) {
            float4 acc = red[0][sl];
#pragma unroll
            for (int k = 1; k < NG; k++) {
                const float4 r = red[k][sl];
                acc.x += r.x; acc.y += r.y; acc.z += r.z; acc.w += r.w;
            }
            N[(size_t)c * ns + s] = acc;
        }
    }
    float bm = wmx[0];
#pragma unroll
    for (int k = 1; k < NG / 2; k++) bm = fmaxf(bm, wmx[k]);
    const int es = scale_exp(bm);
    if (t == 0) sexp[(size_t)c * nsb + sb] = es;
    const float scale = ldexpf(1.0f, es);

    // quad Q = tiles 4Q..4Q+3 of polarisation P at sample sl is the 16-byte
    // slot of lane sl + 32 (Q % 2) in fragment (ks = Q / 2, P); quads past the
    // last k-step are not written (their tiles are past nant and zero)
    uint4 *dst = Bf + ((size_t)c * nsb + sb) * nkstep * 128 + sl;
#pragma unroll
    for (int j4 = 0; j4 < APT / 4; j4++) {
        const int Q = g + NG * j4;
        if ((Q >> 1) >= nkstep) break;
        uint32_t wx[4], wy[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const cf x = ux[4 * j4 + k], y = uy[4 * j4 + k];
            wx[k] = (uint32_t)M::bits(x.x * scale) | ((uint32_t)M::bits(x.y * scale) << 16);
            wy[k] = (uint32_t)M::bits(y.x * scale) | ((uint32_t)M::bits(y.y * scale) << 16);
        }
        uint4 *f = dst + (Q >> 1) * 128 + 32 * (Q & 1);
        f[0] = make_uint4(wx[0], wx[1], wx[2], wx[3]);
        f[64] = make_uint4(wy[0], wy[1], wy[2], wy[3]);
    }
}

// ------------------------------------------------------------- GEMM -------
// Fragments are fetched with buffer loads: the per-channel operand base and
// every wave-uniform offset live in scalar registers, the only vector
// address is the lane's 16-byte slot.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
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

template <int MB, int WP> struct afrag {
    uint4 a[MB][WP]; // hi, lo (WP 2) or hi alone (WP 1)
};
template <int NB> struct bfrag {
    uint4 b[NB];
};

// Operand offsets (bytes): A block m of this wavefront at aoff + m astr,
// k-step ks at + 2048 ks (hi) and + 1024 more (lo); B block n at boff[n],
// k-step ks at + 2048 ks (the two polarisation fragments of a k-step are
// adjacent).
// (valid false: an offset past any weights buffer, so the loads return zeros)
template <int MB, int WP>
__device__ __forceinline__ void load_a(afrag<MB, WP> &f, rsrc_t rA, int vl, int aoff, int astr,
                                       int ks, bool valid)
{
    const int base = valid ? aoff + 2048 * ks : 0x7fff0000;
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int h = 0; h < WP; h++) f.a[m][h] = bload(rA, vl, base + m * astr + 1024 * h);
}

template <int NB>
__device__ __forceinline__ void load_b(bfrag<NB> &f, rsrc_t rB, int vl, const int (&boff)[NB], int ks)
{
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

template <int PREC, int MB, int NB>
__device__ __forceinline__ void mma_ab(f16v (&acc)[MB][NB], const afrag<MB, prec_parts<PREC>()> &A,
                                       const bfrag<NB> &B)
{
    // independent hi products first, then the lo products
#pragma unroll
    for (int h = 0; h < prec_parts<PREC>(); h++)
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++)
                acc[m][n] = mtype<prec_bf16<PREC>()>::mfma(A.a[m][h], B.b[n], acc[m][n]);
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

// Operands of one GEMM problem: a whole chunk in the single-channel call, or
// one coarse channel of a batch.  Af, Bf, sexp and N are indexed by the fine
// channel c as laid out by k_phase_weights / k_bf_prep; S is the blocked
// device layout (VM_S_IDX) of a second of ns_total samples, written from
// sample soffset on.
// Per-pointing primary-beam form: N holds T_k as float4 [c][s][4], bx the
// pointing's constants as float [c][p][PB_ROW] (k_pb_beam) and esp its
// weight exponents.
// Beam per tile group (PB 2): N holds the groups' T_k as float4
// [c][s][VM_MFMA_MAX_GROUPS][4], gx the constants of k_pbg_consts, gb the
// groups' slot ranges.
struct bf_geom {
    const uint4 *Af;
    const uint4 *Bf;
    const int *sexp;
    const float4 *N;
    float *S;
    const float *bx;
    const int *esp;
    const float *gx;
    int nkstep, nrb, npointing, nstokes;
    int soffset, ns_total;
    float invw;
    int pad;
    grp_bounds gb;
};

struct c2f { float x, y; };
__device__ __forceinline__ c2f cmulf(c2f a, c2f b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c2f cconjf(c2f a) { return {a.x, -a.y}; }

// Stokes of one pointing with its own primary beam from the GEMM output E'
// (x, y; scaled, u2 removes the scales squared): X' = E' E'^H u2 -
// sum_k h_k T_k, then X = M X' M^H with M = B^-1, whose (00, 11, 01) are
// (bnXX, bnYY, bnXY) of src/form_beam.cpp:339-355.  R: the pointing's
// constants from k_pb_beam: h (R[0]), |m00|^2 |m01|^2 |m10|^2 |m11|^2 (R[1]),
// m00 m01* and m10 m11* (R[2]), m00 m10* and m01 m11* (R[3]), m00 m11* and
// m01 m10* (R[4]).
__device__ __forceinline__ void stokes_pb(float xr, float xi, float yr, float yi, float u2,
                                          const float4 (&T)[4], const float4 (&R)[5], float invw,
                                          float (&st)[4])
{
    const float h0 = R[0].x, h1 = R[0].y, h2 = R[0].z, h3 = R[0].w;
    const float X00 = (xr * xr + xi * xi) * u2 - (h0 * T[0].x + h1 * T[1].x + h2 * T[2].x + h3 * T[3].x);
    const float X11 = (yr * yr + yi * yi) * u2 - (h0 * T[0].y + h1 * T[1].y + h2 * T[2].y + h3 * T[3].y);
    const c2f X01 = {(xr * yr + xi * yi) * u2 - (h0 * T[0].z + h1 * T[1].z + h2 * T[2].z + h3 * T[3].z),
                     (xi * yr - xr * yi) * u2 - (h0 * T[0].w + h1 * T[1].w + h2 * T[2].w + h3 * T[3].w)};
    const c2f p0 = {R[2].x, R[2].y}, p1 = {R[2].z, R[2].w};
    const c2f q0 = {R[3].x, R[3].y}, q1 = {R[3].z, R[3].w};
    const c2f c0 = {R[4].x, R[4].y}, c1 = {R[4].z, R[4].w};
    // (M X M^H)_00 = |m00|^2 X00 + |m01|^2 X11 + 2 Re(m00 m01* X01*), with X01
    // = X'_01 and X10 = X01*; likewise row 1; (M X M^H)_01 = m00 m10* X00 +
    // m01 m11* X11 + m00 m11* X01 + m01 m10* X01*
    const float Y00 = R[1].x * X00 + R[1].y * X11 + 2.0f * (p0.x * X01.x - p0.y * X01.y);
    const float Y11 = R[1].z * X00 + R[1].w * X11 + 2.0f * (p1.x * X01.x - p1.y * X01.y);
    const c2f r0 = cmulf(X01, c0), r1 = cmulf(cconjf(X01), c1);
    const float Y01r = q0.x * X00 + q1.x * X11 + r0.x + r1.x;
    const float Y01i = q0.y * X00 + q1.y * X11 + r0.y + r1.y;
    st[0] = invw * (Y00 + Y11);
    st[1] = invw * (Y00 - Y11);
    st[2] = 2.0f * invw * Y01r;
    st[3] = -2.0f * invw * Y01i;
}

// Stokes of one beam-sample with a beam per tile group (PB 2) from the GEMM
// output A = (x, y) carried into the last group's frame: e = M A, then
// X = e e^H u2 - sum_g S_g(Y_g), Y_g = sum_k h_g[k] T_gk, and the Stokes
// parameters of X as in stokes4.  gxp: the pointing's constants (M, then
// h, S per group); Tl: the sample's T_gk, 16 floats per group.
__device__ __forceinline__ void stokes_pbg(float xr, float xi, float yr, float yi, float u2,
                                           const float *Tl, const float *gxp, int ngroup,
                                           float invw, float (&st)[4])
{
    const gmem f4v *gq = (const gmem f4v *)gxp;
    const f4v m0 = gq[0], m1 = gq[1]; // (m00, m01), (m10, m11) as (re, im) pairs
    const float exr = m0.x * xr - m0.y * xi + m0.z * yr - m0.w * yi;
    const float exi = m0.x * xi + m0.y * xr + m0.z * yi + m0.w * yr;
    const float eyr = m1.x * xr - m1.y * xi + m1.z * yr - m1.w * yi;
    const float eyi = m1.x * xi + m1.y * xr + m1.z * yi + m1.w * yr;
    float X00 = (exr * exr + exi * exi) * u2, X11 = (eyr * eyr + eyi * eyi) * u2;
    float X01r = (exr * eyr + exi * eyi) * u2, X01i = (exi * eyr - exr * eyi) * u2;
    for (int g = 0; g < ngroup; g++) {
        const gmem f4v *G = gq + 2 + (GX_G / 4) * g;
        const f4v h = G[0];
        const float4 *T = (const float4 *)(Tl + 16 * g);
        const float4 t0 = T[0], t1 = T[1], t2 = T[2], t3 = T[3];
        const float y0 = h.x * t0.x + h.y * t1.x + h.z * t2.x + h.w * t3.x;
        const float y1 = h.x * t0.y + h.y * t1.y + h.z * t2.y + h.w * t3.y;
        const float y2 = h.x * t0.z + h.y * t1.z + h.z * t2.z + h.w * t3.z;
        const float y3 = h.x * t0.w + h.y * t1.w + h.z * t2.w + h.w * t3.w;
        const f4v s0 = G[1], s1 = G[2], s2 = G[3], s3 = G[4];
        X00 -= s0.x * y0 + s0.y * y1 + s0.z * y2 + s0.w * y3;
        X11 -= s1.x * y0 + s1.y * y1 + s1.z * y2 + s1.w * y3;
        X01r -= s2.x * y0 + s2.y * y1 + s2.z * y2 + s2.w * y3;
        X01i -= s3.x * y0 + s3.y * y1 + s3.z * y2 + s3.w * y3;
    }
    st[0] = invw * (X00 + X11);
    st[1] = invw * (X00 - X11);
    st[2] = 2.0f * invw * X01r;
    st[3] = -2.0f * invw * X01i;
}

// Horner step between tile groups (PB 2): every beam's (Ex, Ey) of each
// sample, A <- R_p,g1 A, R from the pointing's constants (clamped to the
// last pointing for padding rows, whose accumulators are zero).  Lane
// (lr, lk) of 32-row block m holds beams 16 (rb0 + m) + 4 jg + 2 lk + b.
template <int MB, int NB>
__device__ __forceinline__ void bf_horner(f16v (&acc)[MB][NB], const float *gxc, int np, int p0,
                                          int lk, int g1)
{
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int jg = 0; jg < 4; jg++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int p = min(p0 + 16 * m + 4 * jg + 2 * lk + b, np - 1);
                const gmem f4v *R =
                    (const gmem f4v *)(gxc + (size_t)p * GX_ROW + 8 + GX_G * g1 + 20);
                const f4v r0 = R[0], r1 = R[1];
#pragma unroll
                for (int j = 0; j < NB / 2; j++) {
                    f16v &ax = acc[m][2 * j], &ay = acc[m][2 * j + 1];
                    const float xr = ax[4 * jg + 2 * b], xi = ax[4 * jg + 2 * b + 1];
                    const float yr = ay[4 * jg + 2 * b], yi = ay[4 * jg + 2 * b + 1];
                    ax[4 * jg + 2 * b] = r0.x * xr - r0.y * xi + r0.z * yr - r0.w * yi;
                    ax[4 * jg + 2 * b + 1] = r0.x * xi + r0.y * xr + r0.z * yi + r0.w * yr;
                    ay[4 * jg + 2 * b] = r1.x * xr - r1.y * xi + r1.z * yr - r1.w * yi;
                    ay[4 * jg + 2 * b + 1] = r1.x * xi + r1.y * xr + r1.z * yi + r1.w * yr;
                }
            }
}

// Each wavefront's staging tile holds 32 rows of 64 samples (68-float rows,
// padded against bank conflicts).
#define BF_WAVE_LDS (32 * 68)
// Per-sample T rows of the tile-group form: 64 floats used, 68 apart
#define BF_TROW 68
// Fragments (1 KB each) of one k-step of a workgroup in the shared-operand
// form: A (WP weight parts) of its WR MB row blocks and B of its WC NB column
// blocks.
template <int MB, int NB, int WR, int WC, int WP>
constexpr int bf_wg_frags() { return WP * WR * MB + WC * NB; }
template <int PREC, int MB, int NB, int WR, int WC, int NST, int PB, bool LDSOP = false>
constexpr int bf_lds_floats()
{
    return WR * WC * BF_WAVE_LDS + (PB == 1 ? WR * WC * MB * 16 * 32 : 0) +
           (PB == 2 ? WR * WC * (NB / 2) * 32 * BF_TROW : 0) +
           (LDSOP ? 2 * bf_wg_frags<MB, NB, WR, WC, prec_parts<PREC>()>() * 256 : 0);
}

// Stokes of one beam-sample from the GEMM outputs (x, y), in either form;
// the per-pointing form reads the pointing's B^-1, h and weight exponent from
// its LDS row pl of wpb and gives zeros for pointings past the last (valid
// false).
template <int PB>
__device__ __forceinline__ void bf_stokes(float xr, float xi, float yr, float yi, float unscale2,
                                          int esb, const float4 &nn, const float4 (&T4)[4],
                                          const float *wpb, int pl, bool valid, float invw,
                                          const float *Tl, const float *gxp, int ngroup,
                                          int espp, float (&st4)[4])
{
    if constexpr (PB == 2) {
        stokes_pbg(xr, xi, yr, yi, ldexpf(1.0f, -2 * (esb + espp)), Tl, gxp, ngroup, invw, st4);
        if (!valid) st4[0] = st4[1] = st4[2] = st4[3] = 0.f;
    } else if constexpr (PB == 1) {
        const float4 *bl = (const float4 *)(wpb + 32 * pl);
        const float4 R[5] = {bl[0], bl[1], bl[2], bl[3], bl[4]};
        const float u2 = ldexpf(1.0f, -2 * (esb + __float_as_int(wpb[32 * pl + PB_ROW])));
        stokes_pb(xr, xi, yr, yi, u2, T4, R, invw, st4);
        if (!valid) st4[0] = st4[1] = st4[2] = st4[3] = 0.f;
    } else {
        stokes4(xr, xi, yr, yi, unscale2, nn, invw, st4);
    }
}

// Epilogue of a wavefront tile of MB row blocks x two sample blocks (column
// blocks Ex, Ey of each), Stokes of its 16 MB pointings x 64 samples.
// ALL: every one of the 16 MB store instructions is issued, as raw-buffer
// stores over the channel's part of S; rows and lanes with nothing to store
// (Stokes past nstokes, pointings past the last, samples past the chunk) get
// an offset past the buffer, and the hardware drops them.  The count of
// vector-memory operations a tile leaves is then fixed (waits can count on
// it); the channel's part of S must be under 2 GB.
template <int PB, int MB, bool ALL = false>
__device__ __forceinline__ void bf_store_runs(const f16v (&acc)[MB][4], const bf_geom &G, int ns,
                                              int c, int rb0, int sb0, int lane,
                                              const float (&us2)[2], const int (&esbv)[2],
                                              const float4 (&nnv)[2],
                                              const float4 (&T4v)[PB == 1 ? 2 : 1][4],
                                              const float *wpb, const float *wT, float *wl)
{
    const int lk = lane >> 5, lr = lane & 31;
    const int nstokes = G.nstokes, npointing = G.npointing;
    const float invw = G.invw;
    const int nblk = VM_S_NBLK(G.ns_total);
    // Both sample blocks of a half block go through one 32-row x 64-sample
    // LDS tile, and 16 lanes store one row: every store instruction writes
    // 4 rows x 256 contiguous bytes (HBM takes such runs markedly faster
    // than 128-byte ones).
    // With the blocked layout of S a wavefront's 64 pointings x nstokes x
    // 64 samples are one contiguous run when the chunk offset is a
    // multiple of 64 (a lane's 4 samples never straddle a block).
    const int q16 = lane & 15;
    const int s4 = SB * sb0 + 4 * q16;
    gmem float *Sc =
        (gmem float *)G.S + VM_S_IDX(0, 0, c, G.soffset + s4, npointing, nstokes, nblk);
    // ALL: the channel's part of S as a buffer, and the lane's byte offset in
    // it of (pointing 0, Stokes 0) of its 4 samples
    const uint32_t chan_bytes = (uint32_t)nblk * npointing * nstokes * VM_SBLK * 4;
    const rsrc_t rS = make_rsrc((const float *)G.S + (size_t)c * nblk * npointing * nstokes * VM_SBLK,
                                chan_bytes);
    const int so = G.soffset + s4;
    const uint32_t lane_off = 4u * ((uint32_t)(so / VM_SBLK) * npointing * nstokes * VM_SBLK +
                                    (uint32_t)(so % VM_SBLK));
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int jj = 0; jj < 2; jj++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const int jg = 2 * h + jj;
                        const f16v &ax = acc[m][2 * j], &ay = acc[m][2 * j + 1];
                        const int pl = 16 * m + 8 * h + 4 * jj + 2 * lk + b;
                        float st4[4];
                        const int pc = min(16 * rb0 + pl, npointing - 1);
                        bf_stokes<PB>(ax[4 * jg + 2 * b], ax[4 * jg + 2 * b + 1],
                                      ay[4 * jg + 2 * b], ay[4 * jg + 2 * b + 1], us2[j],
                                      esbv[j], nnv[j], T4v[PB == 1 ? j : 0], wpb, pl,
                                      16 * rb0 + pl < npointing, invw,
                                      wT + (j * 32 + lr) * BF_TROW,
                                      G.gx + ((size_t)c * npointing + pc) * GX_ROW, G.gb.n,
                                      PB == 2 ? ((const gmem int *)G.esp)[(size_t)c * npointing + pc] : 0,
                                      st4);
#pragma unroll
                        for (int st = 0; st < 4; st++)
                            wl[(st * 8 + 4 * jj + 2 * b + lk) * 68 + 32 * j + lr] = st4[st];
                    }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int r = 4 * i + (lane >> 4);
                const int st = r >> 3, bbp = r & 7;
                // bbp = 4 jj + 2 b + lk  ->  beam in the half = 4 jj + 2 lk + b
                const int bb = 4 * (bbp >> 2) + 2 * (bbp & 1) + ((bbp >> 1) & 1);
                const int p = 16 * (rb0 + m) + 8 * h + bb;
                const float4 val = *(const float4 *)(wl + r * 68 + 4 * q16);
                if constexpr (ALL) {
                    const bool ok = st < nstokes && p < npointing && s4 < ns;
                    const uint32_t off =
                        ok ? lane_off + 4u * (uint32_t)(p * nstokes + st) * VM_SBLK : 0x7ffffff0u;
                    // streamed output (nt), never read back by this kernel
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, val), rS,
                                                           (int)off, 0, 2);
                    continue;
                }
                if (st >= nstokes || p >= npointing || s4 >= ns) continue;
                // streamed output, never read back by this kernel
                __builtin_nontemporal_store(f4v{val.x, val.y, val.z, val.w},
                                            (gmem f4v *)(Sc + ((size_t)p * nstokes + st) * VM_SBLK));
            }
        }
}

// Workgroup barrier that orders LDS only: the fences wait for this
// wavefront's LDS operations (lgkmcnt), not for its global loads, so
// prefetches stay in flight across it.
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// k-loop with workgroup-shared operands (the LDSOP form of bf_tile).  The
// fragments of one k-step of the whole workgroup (bf_wg_frags: A parts of
// its WR MB row blocks, B of its WC NB column blocks) are fetched once per
// workgroup -- wavefront w takes fragments w, w + NW, ... -- two k-steps
// ahead into registers (R0, R1), written to one of two LDS stages, and read
// from there by each wavefront one k-step ahead of its MFMAs (fa0/fb0,
// fa1/fb1).  Each operand byte then crosses a CU's vector-memory path once
// per workgroup instead of once per wavefront.  One barrier per k-step;
// stage s is rewritten only after every wavefront's reads of it completed
// before the previous barrier.  Prefetches past the last k-step re-read it.
template <int MB, int NB, int WR, int WC, int WP, int PER>
__device__ __forceinline__ void wg_gload(uint4 (&R)[PER], rsrc_t rA, rsrc_t rB, int vl,
                                         const int (&goff)[PER], int k)
{
    constexpr int NW = WR * WC, FA = WP * WR * MB;
#pragma unroll
    for (int i = 0; i < PER; i++) R[i] = bload(NW * i < FA ? rA : rB, vl, goff[i] + 2048 * k);
}

template <int PER, int F>
__device__ __forceinline__ void wg_lwrite(uint4 *ops, const uint4 (&R)[PER], int stage, int w,
                                          int lane)
{
    constexpr int NW = F / PER;
#pragma unroll
    for (int i = 0; i < PER; i++) ops[(stage * F + w + NW * i) * 64 + lane] = R[i];
}

template <int MB, int NB, int WP, int F, int FA>
__device__ __forceinline__ void wg_lread(afrag<MB, WP> &fa, bfrag<NB> &fb, const uint4 *ops,
                                         int stage, int wr, int wc, int lane)
{
    const uint4 *o = ops + stage * F * 64 + lane;
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int h = 0; h < WP; h++) fa.a[m][h] = o[(WP * (wr * MB + m) + h) * 64];
#pragma unroll
    for (int n = 0; n < NB; n++) fb.b[n] = o[(FA + wc * NB + n) * 64];
}

template <int PREC, int MB, int NB, int WR, int WC, int PB>
__device__ __forceinline__ void bf_kloop_lds(f16v (&acc)[MB][NB], rsrc_t rA, rsrc_t rB, int vl,
                                             int astr, int bstr, int nkstep, int nsb, int ty,
                                             int tx, int w, int lane, float *stg)
{
    constexpr int WP = prec_parts<PREC>();
    constexpr int NW = WR * WC, FA = WP * WR * MB, F = bf_wg_frags<MB, NB, WR, WC, WP>();
    static_assert(F % NW == 0 && FA % NW == 0, "fragments divide among the wavefronts");
    // two fragment sets and two prefetch sets beside the accumulators: small
    // wavefront tiles only (a 128 x 128 tile's 256 accumulators leave no room)
    static_assert(MB * NB <= 8, "shared-operand k-loop is for 64 x 64 wavefront tiles");
    constexpr int PER = F / NW, NSB = NB / 2;
    static_assert(PB != 2, "the tile-group form uses the register-ring k-loop");
    uint4 *ops = (uint4 *)(stg + WR * WC * BF_WAVE_LDS + (PB == 1 ? WR * WC * MB * 16 * 32 : 0));
    const int wr = w / WC, wc = w % WC;
    const int rbw0 = ty * WR * MB, sbw0 = tx * WC * NSB;
    int goff[PER]; // this wavefront's loader fragments: byte offsets at k-step 0
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int f = w + NW * i;
        if (NW * i < FA) {
            goff[i] = (rbw0 + f / WP) * astr + 1024 * (f % WP);
        } else {
            const int g = f - FA;
            // blocks past the last sample block re-read it (in bounds); their
            // columns are never stored
            goff[i] = min(sbw0 + (g >> 1), nsb - 1) * bstr + 1024 * (g & 1);
        }
    }
    const int kl = nkstep - 1;
    uint4 R0[PER], R1[PER];
    afrag<MB, WP> fa0, fa1;
    bfrag<NB> fb0, fb1;
    wg_gload<MB, NB, WR, WC, WP, PER>(R0, rA, rB, vl, goff, 0);
    wg_gload<MB, NB, WR, WC, WP, PER>(R1, rA, rB, vl, goff, min(1, kl));
    wg_lwrite<PER, F>(ops, R0, 0, w, lane);
    wg_gload<MB, NB, WR, WC, WP, PER>(R0, rA, rB, vl, goff, min(2, kl));
    wg_lwrite<PER, F>(ops, R1, 1, w, lane);
    wg_gload<MB, NB, WR, WC, WP, PER>(R1, rA, rB, vl, goff, min(3, kl));
    lds_sync();
    wg_lread<MB, NB, WP, F, FA>(fa0, fb0, ops, 0, wr, wc, lane);
    lds_sync(); // every wavefront holds k-step 0 before stage 0 is rewritten
    for (int ks = 0; ks < nkstep; ks += 2) {
        // fa0/fb0: k-step ks; stage 1: ks + 1; R0: ks + 2; R1: ks + 3
        wg_lread<MB, NB, WP, F, FA>(fa1, fb1, ops, 1, wr, wc, lane);
        __builtin_amdgcn_sched_barrier(0);
        mma_ab<PREC, MB, NB>(acc, fa0, fb0);
        __builtin_amdgcn_sched_barrier(0);
        wg_lwrite<PER, F>(ops, R0, 0, w, lane);
        wg_gload<MB, NB, WR, WC, WP, PER>(R0, rA, rB, vl, goff, min(ks + 4, kl));
        lds_sync();
        if (ks + 1 < nkstep) {
            // fa1/fb1: k-step ks + 1; stage 0: ks + 2; R1: ks + 3; R0: ks + 4
            wg_lread<MB, NB, WP, F, FA>(fa0, fb0, ops, 0, wr, wc, lane);
            __builtin_amdgcn_sched_barrier(0);
            mma_ab<PREC, MB, NB>(acc, fa1, fb1);
            __builtin_amdgcn_sched_barrier(0);
            wg_lwrite<PER, F>(ops, R1, 1, w, lane);
            wg_gload<MB, NB, WR, WC, WP, PER>(R1, rA, rB, vl, goff, min(ks + 5, kl));
            lds_sync();
        }
    }
}

// Register-ring k-loop over k-steps [k0, k1).  Two register rings: A
// (weights, small and L2-resident) two k-steps deep, B (voltage fragments,
// streamed from HBM) NST deep, so each B fragment is requested NST-1
// k-steps before its MFMAs.  The scheduling barriers pin each step's loads
// ahead of its MFMAs; a super-step of U k-steps returns both rings to slot
// 0, so ring slots are compile-time.  The loop runs whole super-steps with no
// branches: past k1 - 1 the B loads re-read that step and the A loads fall
// outside the weights buffer, which returns zeros, so those MFMAs add
// nothing.
#define VM_SB() __builtin_amdgcn_sched_barrier(0)
template <int PREC, int MB, int NB, int NST>
__device__ __forceinline__ void bf_kloop_ring(f16v (&acc)[MB][NB], rsrc_t rA, rsrc_t rB, int vl,
                                              int aoff, int astr, const int (&boff)[NB], int k0,
                                              int k1)
{
    constexpr int U = (NST % 2 == 0) ? NST : 2 * NST;
    constexpr int WP = prec_parts<PREC>();
    afrag<MB, WP> fa[2];
    bfrag<NB> fb[NST];
    const int kl = k1 - 1;
    load_a<MB, WP>(fa[0], rA, vl, aoff, astr, k0, true);
#pragma unroll
    for (int r = 0; r < NST - 1; r++) load_b<NB>(fb[r], rB, vl, boff, min(k0 + r, kl));
    for (int ks = k0; ks < k1; ks += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            VM_SB();
            load_a<MB, WP>(fa[(u + 1) & 1], rA, vl, aoff, astr, ks + u + 1, ks + u + 1 <= kl);
            load_b<NB>(fb[(u + NST - 1) % NST], rB, vl, boff, min(ks + u + NST - 1, kl));
            mma_ab<PREC, MB, NB>(acc, fa[u & 1], fb[u % NST]);
            interleave_loads<WP * MB + NB, WP * MB * NB>();
            VM_SB();
        }
    }
}
#undef VM_SB

template <int PREC, int MB, int NB, int WR, int WC, int NST, bool VEC, int PB, bool LDSOP>
__device__ __forceinline__ void bf_tile(const bf_geom &G, int ns, int nsb, int nchan, int c,
                                        int tx, int ty, float *stg)
{
    static_assert(NB % 2 == 0, "column blocks come in polarisation pairs");
    constexpr int NSB = NB / 2; // sample blocks per wavefront
    const int nkstep = G.nkstep, nrb = G.nrb;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = w / WC, wc = w % WC;
    const int rb0 = (ty * WR + wr) * MB;
    const int sb0 = (tx * WC + wc) * NSB;

    const int astr = nkstep * 2048;              // bytes per 32-row block of A
    const int bstr = nkstep * 2048;              // bytes per sample block of B (2 pols)
    const rsrc_t rA = make_rsrc(G.Af + (size_t)c * nrb * nkstep * 128, (uint32_t)nrb * astr);
    const rsrc_t rB = make_rsrc(G.Bf + (size_t)c * nsb * nkstep * 128, (uint32_t)nsb * bstr);
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

    // The epilogue's per-sample inputs (auto-correlation terms N, block
    // exponents) are requested before the k-loop: the vector memory counter
    // retires loads and stores in issue order, so a load issued after the
    // first Stokes stores would wait for all of them to complete.
    const int lr = lane & 31;
    float4 nnv[NSB];
    int esbv[NSB];
#pragma unroll
    for (int j = 0; j < NSB; j++) {
        const int sb = sb0 + j;
        const int s = SB * sb + lr;
        nnv[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (PB == 0 && s < ns) {
            const f4v n4 = ((const gmem f4v *)G.N)[(size_t)c * ns + s];
            nnv[j] = make_float4(n4.x, n4.y, n4.z, n4.w);
        }
        // one exponent per sample block: a wave-uniform address, read by a
        // scalar load (its own counter) through the constant address space
        esbv[j] = ((const __attribute__((address_space(4))) int *)G.sexp)[(size_t)c * nsb +
                                                                          min(sb, nsb - 1)];
    }

    if constexpr (LDSOP) {
        bf_kloop_lds<PREC, MB, NB, WR, WC, PB>(acc, rA, rB, vl, astr, bstr, nkstep, nsb, ty, tx, w,
                                               lane, stg);
    } else if constexpr (PB == 2) {
        // the tile groups in turn, carried over by the Horner step
        const int ng = G.gb.n;
        for (int g = 0; g < ng; g++) {
            bf_kloop_ring<PREC, MB, NB, NST>(acc, rA, rB, vl, aoff, astr, boff, G.gb.t[g] / 8,
                                             G.gb.t[g + 1] / 8);
            if (g + 1 < ng)
                bf_horner<MB, NB>(acc, G.gx + (size_t)c * G.npointing * GX_ROW, G.npointing,
                                  16 * rb0, lane >> 5, g + 1);
        }
    } else {
        bf_kloop_ring<PREC, MB, NB, NST>(acc, rA, rB, vl, aoff, astr, boff, 0, nkstep);
    }

    // epilogue: in a 32-row block lane (lr, lk) holds rows 8 jg + 4 lk + {0..3}
    // = beams 4 jg + 2 lk + {0, 1} as (re, im), column lr; block n = 2 j holds
    // Ex and n = 2 j + 1 holds Ey of sample block j.  Half h of the block
    // (jg = 2 h + jj) is 8 beams; LDS rows r = st * 8 + (4 jj + 2 b + lk) for
    // beam 4 jg + 2 lk + b: the two lane halves write rows 32 floats apart and
    // each 16-lane read group covers two whole rows.
    // per-pointing-beam form: its T_k terms (16 floats per sample) would
    // cost registers through the k-loop, so they are requested here, after
    // it and before the first store
    // Its pointings' constants (PB_ROW floats from k_pb_beam) and weight
    // exponents go to a 32-float LDS row per pointing of the wavefront (16 MB
    // pointings).
    float4 T4v[PB == 1 ? NSB : 1][4];
    float *wpb = stg + WR * WC * BF_WAVE_LDS + (PB == 1 ? w * MB * 16 * 32 : 0);
    // tile-group form: each sample's T_gk (VM_MFMA_MAX_GROUPS x 4 float4)
    // staged in this wavefront's LDS rows, the lane halves taking half the
    // groups each
    float *wT = stg + WR * WC * BF_WAVE_LDS + (PB == 2 ? w * NSB * 32 * BF_TROW : 0);
    if constexpr (PB == 2) {
#pragma unroll
        for (int j = 0; j < NSB; j++) {
            const int s = SB * (sb0 + j) + lr;
#pragma unroll
            for (int q = 0; q < 2 * VM_MFMA_MAX_GROUPS; q++) {
                const int gk = (lane >> 5) * 2 * VM_MFMA_MAX_GROUPS + q; // group * 4 + k
                f4v t4 = {0.f, 0.f, 0.f, 0.f};
                if (s < ns && (gk >> 2) < G.gb.n)
                    t4 = ((const gmem f4v *)G.N)[((size_t)c * ns + s) * 4 * VM_MFMA_MAX_GROUPS + gk];
                *(float4 *)(wT + (j * 32 + lr) * BF_TROW + 4 * gk) = make_float4(t4.x, t4.y, t4.z, t4.w);
            }
        }
    }
    if constexpr (PB == 1) {
        for (int pl = lane; pl < MB * 16; pl += 64) {
            const int cp = c * G.npointing + min(16 * rb0 + pl, G.npointing - 1);
            const gmem f4v *bp = (const gmem f4v *)(G.bx + (size_t)cp * PB_ROW);
#pragma unroll
            for (int q = 0; q < PB_ROW / 4; q++) {
                const f4v t = bp[q];
                *(float4 *)(wpb + 32 * pl + 4 * q) = make_float4(t.x, t.y, t.z, t.w);
            }
            wpb[32 * pl + PB_ROW] = __int_as_float(((const gmem int *)G.esp)[cp]);
        }
#pragma unroll
        for (int j = 0; j < NSB; j++) {
            const int s = SB * (sb0 + j) + lr;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                T4v[j][k] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (s < ns) {
                    const f4v t4 = ((const gmem f4v *)G.N)[((size_t)c * ns + s) * 4 + k];
                    T4v[j][k] = make_float4(t4.x, t4.y, t4.z, t4.w);
                }
            }
        }
    }
    float *wl = stg + BF_WAVE_LDS * w;
    const int lk = lane >> 5;
    const int nstokes = G.nstokes, npointing = G.npointing;
    const float invw = G.invw;
    const int nblk = VM_S_NBLK(G.ns_total);
    float us2[NSB];
#pragma unroll
    for (int j = 0; j < NSB; j++) us2[j] = ldexpf(1.0f, -2 * (esbv[j] + PHASE_EXP));
    if constexpr (VEC && NSB == 2) {
        bf_store_runs<PB, MB>(acc, G, ns, c, rb0, sb0, lane, us2, esbv, nnv, T4v, wpb, wT, wl);
        return;
    }
    const int qq = lane & 7;
#pragma unroll
    for (int j = 0; j < NSB; j++) {
        const int sb = sb0 + j;
        const int s4 = SB * sb + 4 * qq;
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
                        const int pl = 16 * m + 8 * h + 4 * jj + 2 * lk + b;
                        float st4[4];
                        const int pc = min(16 * rb0 + pl, npointing - 1);
                        bf_stokes<PB>(ax[4 * jg + 2 * b], ax[4 * jg + 2 * b + 1],
                                      ay[4 * jg + 2 * b], ay[4 * jg + 2 * b + 1], us2[j], esbv[j],
                                      nnv[j], T4v[PB == 1 ? j : 0], wpb, pl,
                                      16 * rb0 + pl < npointing, invw,
                                      wT + (j * 32 + lr) * BF_TROW,
                                      G.gx + ((size_t)c * npointing + pc) * GX_ROW, G.gb.n,
                                      PB == 2 ? ((const gmem int *)G.esp)[(size_t)c * npointing + pc] : 0,
                                      st4);
#pragma unroll
                        for (int st = 0; st < 4; st++)
                            wl[(st * 8 + 4 * jj + 2 * b + lk) * 32 + lr] = st4[st];
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = 8 * i + (lane >> 3);
                    const int st = r >> 3, bbp = r & 7;
                    const int bb = 4 * (bbp >> 2) + 2 * (bbp & 1) + ((bbp >> 1) & 1);
                    const int p = 16 * (rb0 + m) + 8 * h + bb;
                    const float4 val = *(const float4 *)(wl + r * 32 + 4 * qq);
                    if (st >= nstokes || p >= npointing || s4 >= ns) continue;
                    gmem float *S = (gmem float *)G.S;
                    const int so = G.soffset + s4;
                    if constexpr (VEC) {
                        // so is a multiple of 4: the 4 samples share a block
                        __builtin_nontemporal_store(
                            f4v{val.x, val.y, val.z, val.w},
                            (gmem f4v *)(S + VM_S_IDX(p, st, c, so, npointing, nstokes, nblk)));
                    } else {
                        // any offset: each sample finds its own block
                        const float v4[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (s4 + k < ns)
                                S[VM_S_IDX(p, st, c, so + k, npointing, nstokes, nblk)] = v4[k];
                    }
                }
            }
    }
}

// 1-D grid over (channel, tile) of one GEMM problem.
// Tile order: order 0 walks the channels in turn, row tiles fastest.  Orders
// 1-4 are XCD-aware: the 8 workgroups dispatched together (one per XCD) take
// the same tile of 8 different channels, so each XCD walks its own channels;
// within a channel tiles go in bands of gr row tiles, column by column (gr =
// all row tiles for order 1, 1 for order 2, 2 for order 3, 4 for order 4).
template <int PREC, int MB, int NB, int WR, int WC, int NST, int OCC, bool VEC, int PB,
          bool LDSOP>
__global__ __launch_bounds__(64 * WR * WC, OCC) void k_beamform_mfma(bf_geom G, int ns, int nsb,
                                                                     int nchan, int ntx, int nty,
                                                                     int order)
{
    __shared__ __attribute__((aligned(16))) float
        stg[bf_lds_floats<PREC, MB, NB, WR, WC, NST, PB, LDSOP>()];
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
    bf_tile<PREC, MB, NB, WR, WC, NST, VEC, PB, LDSOP>(G, ns, nsb, nchan, c, tx, ty, stg);
}

// Batched form: one persistent launch over the GEMM problems of several
// coarse channels (geoms[g]), all with ns samples in nsb blocks and nchan
// fine channels.  Work items are (g, c, tx, ty0, cnt): cnt row tiles from
// ty0 of column tile tx of fine channel c, encoded as x = g << 16 | c,
// y = tx << 16 | ty0 << 8 | cnt and listed in the order they should start
// (most expensive first).  Workgroups take items from a queue counter that
// must be zero at launch; every workgroup leaves the loop when the counter
// passes nitems, so the grid always drains.  Every channel's ns, ns_total
// and sample offset are multiples of 4 (16-byte Stokes stores).
template <int PREC, int MB, int NB, int WR, int WC, int NST, int OCC>
__global__ __launch_bounds__(64 * WR * WC, OCC) void k_beamform_batch(
    const bf_geom *__restrict__ geoms, const uint2 *__restrict__ items, int nitems,
    uint32_t *__restrict__ queue, int ns, int nsb, int nchan)
{
    __shared__ __attribute__((aligned(16))) float stg[bf_lds_floats<PREC, MB, NB, WR, WC, NST, 0>()];
    __shared__ int slot[2];
    for (int it = 0;; it++) {
        // two slots: a slot is rewritten only after the barrier of the next
        // item, which every wavefront passes after reading it
        if (threadIdx.x == 0) slot[it & 1] = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(slot[it & 1]);
        if (idx >= nitems) break;
        const uint2 e = items[idx];
        const int g = __builtin_amdgcn_readfirstlane((int)(e.x >> 16));
        const int c = __builtin_amdgcn_readfirstlane((int)(e.x & 0xffffu));
        const int tx = __builtin_amdgcn_readfirstlane((int)(e.y >> 16));
        const int ty0 = __builtin_amdgcn_readfirstlane((int)((e.y >> 8) & 0xffu));
        const int cnt = __builtin_amdgcn_readfirstlane((int)(e.y & 0xffu));
        const bf_geom G = geoms[g];
        for (int ty = ty0; ty < ty0 + cnt; ty++)
            bf_tile<PREC, MB, NB, WR, WC, NST, true, 0, false>(G, ns, nsb, nchan, c, tx, ty, stg);
    }
}

// Resident-weights form (pointing-independent J^-1, at most RES_NK k-steps,
// i.e. 128 antennas).  A workgroup of WR wavefronts owns one (fine channel,
// row tile) for the whole chunk: each wavefront holds its MB row blocks of A
// for every k-step in registers, and the workgroup walks the column tiles (64
// samples) with each tile's B fragments and N terms in LDS, double-buffered.
// Tile tx + 1 is copied HBM -> LDS by direct (LDS-DMA) loads issued before
// tile tx's k-loop, so the k-loop reads only LDS and registers.
// Why: a CU's vector-memory path serves loads and stores in order, and while
// the CU's Stokes stores are held up by HBM, a load issued behind them waits
// as well; the register-ring k-loop then stalls under every store phase.
// Here a wavefront's only loads are the DMA of the next tile, issued before
// its stores of the current one, and the MFMAs run under the store stream.
// The waits are explicit: before tile tx the wavefront waits until at most
// as many operations are outstanding as it issued after tile tx's DMA (the
// stores of tile tx - 1, always 16 MB: bf_store_runs with ALL), and the
// barriers order LDS only, so no wait covers the stores.
#define RES_NK 16      // k-steps held in registers
#define RES_BSB 32768  // LDS bytes of one sample block's B fragments (RES_NK k-steps)
template <int WR> constexpr int res_buf_bytes()
{
    // the B fragments or the Stokes staging, whichever is larger
    return WR * BF_WAVE_LDS * 4 > 2 * RES_BSB ? WR * BF_WAVE_LDS * 4 : 2 * RES_BSB;
}

// LDS byte address of a __shared__ array
#define LDS_ADDR(x) ((uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void *)(x))

// One LDS-DMA load: 64 lanes x 16 bytes from the lanes' global addresses g
// to LDS bytes [lds, lds + 1024), lane i at lds + 16 i (M0 holds lds while
// it issues; M0 is restored after).
// Written as asm so that the compiler's wait insertion does not see it: it
// cannot always tell an LDS-DMA load from the LDS accesses around it and
// would then wait for the load (and for every store issued before it) at
// unrelated LDS reads.  The kernel waits for it explicitly.  The instruction
// writes no registers; its address registers and M0 are read when it issues
// (one wait state after the M0 write).
__device__ __forceinline__ void lds_dma16(const void *g, uint32_t lds)
{
    uint32_t m0save;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %2\n\t"
                 "s_nop 0\n\t"
                 "global_load_lds_dwordx4 %1, off\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(m0save)
                 : "v"(g), "s"(lds)
                 : "memory");
}

// s_waitcnt with vmcnt n alone (expcnt and lgkmcnt at their maxima)
#define VM_WAIT_VMCNT(n) \
    __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))

// Workgroup barrier with no memory fence: this wavefront's LDS operations
// complete (lgkmcnt 0), then s_barrier; the empty asm statements keep the
// compiler from moving memory operations across it.  LDS-DMA completion is
// waited for explicitly before it (VM_WAIT_VMCNT).
__device__ __forceinline__ void res_barrier()
{
    __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (0 << 8) | (3 << 14));
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// This wavefront's share of tile tx's DMA into buf: sample block j's B
// fragments (2 nkstep chunks of 1 KB, in their HBM order) at j RES_BSB, the
// chunks taken round-robin by the wavefronts; the last wavefront also copies
// the N terms of the tile's 64 samples to nbuf (samples past the last: the
// last).
// Blocks past the last sample block re-read it; their columns are never
// stored.
template <int WR>
__device__ __forceinline__ void res_dma(uint32_t buf, uint32_t nbuf, const char *Bc,
                                        const float4 *Nc, int tx, int nsb, int bstr, int nkstep,
                                        int ns, int w, int lane)
{
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const char *src = Bc + (size_t)min(2 * tx + j, nsb - 1) * bstr + 16 * lane;
        for (int r = w; r < 2 * nkstep; r += WR)
            lds_dma16(src + 1024 * r, buf + j * RES_BSB + 1024 * r);
    }
    if (w == WR - 1) lds_dma16(Nc + min(64 * tx + lane, ns - 1), nbuf);
}

// One column tile tx from buffer cur (and N terms ncur); starts the DMA of
// tile tx + 1 into the buffers at LDS addresses nxt, nnxt.  FULLK: exactly
// RES_NK k-steps, so the k-loop needs no per-step test and each step's B
// fragments are read one step ahead of its MFMAs.
template <int PREC, int MB, int WR, bool FULLK>
__device__ __forceinline__ void res_tile(char *cur, const float4 *ncur, uint32_t nxt, uint32_t nnxt,
                                         const bf_geom &G,
                                         const afrag<MB, prec_parts<PREC>()> (&A)[RES_NK],
                                         const char *Bc, const float4 *Nc, int c, int tx, int ntx,
                                         int rb0, int ns, int nsb, int bstr, int w, int lane)
{
    const int nkstep = G.nkstep;
    // tile tx's DMA was followed by this wavefront's 16 MB stores of tile
    // tx - 1 and nothing else (none at tile 0, whose DMA the prologue waited for)
    VM_WAIT_VMCNT(16 * MB);
    res_barrier(); // tile tx is in cur; every wavefront is done with nxt
    if (tx + 1 < ntx) res_dma<WR>(nxt, nnxt, Bc, Nc, tx + 1, nsb, bstr, nkstep, ns, w, lane);
    const int lr = lane & 31;
    float4 nnv[2];
    int esbv[2];
    float us2[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        nnv[j] = ncur[32 * j + lr];
        esbv[j] = ((const __attribute__((address_space(4))) int *)G.sexp)[(size_t)c * nsb +
                                                                          min(2 * tx + j, nsb - 1)];
        us2[j] = ldexpf(1.0f, -2 * (esbv[j] + PHASE_EXP));
    }
    f16v acc[MB][4];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[m][n][q] = 0.f;
    const uint4 *cb = (const uint4 *)cur + lane;
    if constexpr (FULLK) {
        bfrag<4> fb[2];
#pragma unroll
        for (int n = 0; n < 4; n++) fb[0].b[n] = cb[((n >> 1) * RES_BSB + 1024 * (n & 1)) / 16];
#pragma unroll
        for (int ks = 0; ks < RES_NK; ks++) {
            __builtin_amdgcn_sched_barrier(0);
            if (ks + 1 < RES_NK) {
#pragma unroll
                for (int n = 0; n < 4; n++)
                    fb[(ks + 1) & 1].b[n] =
                        cb[((n >> 1) * RES_BSB + 2048 * (ks + 1) + 1024 * (n & 1)) / 16];
            }
            mma_ab<PREC, MB, 4>(acc, A[ks], fb[ks & 1]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < RES_NK; ks++) {
            if (ks < nkstep) {
                bfrag<4> fb;
#pragma unroll
                for (int n = 0; n < 4; n++)
                    fb.b[n] = cb[((n >> 1) * RES_BSB + 2048 * ks + 1024 * (n & 1)) / 16];
                mma_ab<PREC, MB, 4>(acc, A[ks], fb);
            }
        }
    }
    res_barrier(); // every wavefront's k-loop reads of cur are done: staging may rewrite it
    const float4 T4v[1][4] = {};
    bf_store_runs<0, MB, true>(acc, G, ns, c, rb0, 2 * tx, lane, us2, esbv, nnv, T4v, nullptr,
                               nullptr, (float *)(cur + w * BF_WAVE_LDS * 4));
}

template <int PREC, int MB, int WR, bool FULLK>
__global__ __launch_bounds__(64 * WR, 1) void k_beamform_res(bf_geom G, int ns, int nsb, int nchan,
                                                             int nty)
{
    __shared__ __attribute__((aligned(16))) char bufA[res_buf_bytes<WR>()];
    __shared__ __attribute__((aligned(16))) char bufB[res_buf_bytes<WR>()];
    __shared__ float4 nA[64], nB[64];
    constexpr int WP = prec_parts<PREC>();
    // XCD-aware: the nty row tiles of a channel go to one XCD (workgroups
    // L, L + 8, ... share an XCD), so their common B tiles meet in its L2
    const int L = blockIdx.x;
    int c, ty;
    if ((nchan & 7) == 0) {
        const int g = L >> 3;
        c = (g / nty) * 8 + (L & 7);
        ty = g % nty;
    } else {
        c = L / nty;
        ty = L % nty;
    }
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nkstep = G.nkstep, nrb = G.nrb;
    const int rb0 = (ty * WR + w) * MB;
    const int astr = nkstep * 2048, bstr = nkstep * 2048;
    const rsrc_t rA = make_rsrc(G.Af + (size_t)c * nrb * nkstep * 128, (uint32_t)nrb * astr);
    const char *Bc = (const char *)(G.Bf + (size_t)c * nsb * nkstep * 128);
    const float4 *Nc = G.N + (size_t)c * ns;
    const int ntx = (nsb + 1) / 2;
    const uint32_t lA = LDS_ADDR(bufA), lB = LDS_ADDR(bufB), lnA = LDS_ADDR(nA), lnB = LDS_ADDR(nB);
    res_dma<WR>(lA, lnA, Bc, Nc, 0, nsb, bstr, nkstep, ns, w, lane);
    afrag<MB, WP> A[RES_NK];
#pragma unroll
    for (int ks = 0; ks < RES_NK; ks++) {
        if (ks < nkstep) {
            load_a<MB, WP>(A[ks], rA, lane * 16, rb0 * astr, astr, ks, true);
        } else {
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int h = 0; h < WP; h++) A[ks].a[m][h] = make_uint4(0, 0, 0, 0);
        }
    }
    VM_WAIT_VMCNT(0); // tile 0's DMA and A
    for (int tx = 0; tx < ntx; tx += 2) {
        res_tile<PREC, MB, WR, FULLK>(bufA, nA, lB, lnB, G, A, Bc, Nc, c, tx, ntx, rb0, ns, nsb,
                                      bstr, w, lane);
        if (tx + 1 < ntx)
            res_tile<PREC, MB, WR, FULLK>(bufB, nB, lA, lnA, G, A, Bc, Nc, c, tx + 1, ntx, rb0, ns,
                                          nsb, bstr, w, lane);
    }
}

// Kernel variants the autotuner chooses from: register tile per wavefront
// (MB row blocks of 32 x NB column blocks of 32, NB = 2 x sample blocks),
// wavefront grid per workgroup (WR x WC), depth of the B register ring (the
// A ring is two deep) and the minimum wavefronts per SIMD the registers must
// allow.
// lds: the k-loop with workgroup-shared operands (bf_kloop_lds); res: the
// resident-weights form (k_beamform_res; one workgroup per channel and row
// tile, tile orders do not apply)
struct bf_variant { int mb, nb, wr, wc, nst, occ, lds, res; };
static const bf_variant kVariants[] = {
    {4, 4, 2, 2, 4, 1}, {4, 4, 1, 4, 4, 1}, {4, 4, 1, 4, 6, 1}, {2, 4, 2, 2, 4, 2},
    {4, 2, 2, 2, 4, 2}, {2, 4, 4, 1, 4, 2}, {4, 4, 4, 1, 4, 1}, {4, 4, 2, 2, 6, 1},
    {2, 4, 1, 4, 4, 2}, {2, 4, 4, 2, 4, 2}, {2, 4, 2, 2, 4, 2, 1}, {2, 4, 8, 1, 4, 2},
    {1, 4, 8, 1, 0, 1, 0, 1},
};
#define N_VARIANTS ((int)(sizeof(kVariants) / sizeof(kVariants[0])))
#define VM_N_ORDERS 5

// variants with the tile-group form (register-ring k-loop, at most four
// wavefronts: its T rows take 17 KB of LDS per wavefront)
static bool variant_has_groups(int vi) { return vi == 0 || vi == 3 || vi == 4 || vi == 5 || vi == 8; }

template <int PREC, int PB>
static void launch_variant(int vi, dim3 grid, hipStream_t st, const bf_geom &G, int ns, int nsb,
                           int nchan, int ntx, int nty, int order)
{
    if ((ns | G.ns_total | G.soffset) & 3) {
        // sample counts not a multiple of 4: the element-wise store form,
        // built for the first variant's shape only (its grid is the caller's
        // when vi = 0)
        if (vi != 0) VM_FATAL("MFMA beamformer: variant %d needs ns %% 4 == 0", vi);
        hipLaunchKernelGGL((k_beamform_mfma<PREC, 4, 4, 2, 2, 4, 1, false, PB, false>), grid,
                           dim3(256),
                           0, st, G, ns, nsb, nchan, ntx, nty, order);
        return;
    }
#define VM_LAUNCH(MB_, NB_, WR_, WC_, NST_, OCC_, LDS_)                                       \
    hipLaunchKernelGGL((k_beamform_mfma<PREC, MB_, NB_, WR_, WC_, NST_, OCC_, true, PB, LDS_>),  \
                       grid, dim3(64 * WR_ * WC_), 0, st, G, ns, nsb, nchan, ntx, nty, order)
    if constexpr (PB == 2) {
        switch (vi) {
        case 0: VM_LAUNCH(4, 4, 2, 2, 4, 1, false); break;
        case 3: VM_LAUNCH(2, 4, 2, 2, 4, 2, false); break;
        case 4: VM_LAUNCH(4, 2, 2, 2, 4, 2, false); break;
        case 5: VM_LAUNCH(2, 4, 4, 1, 4, 2, false); break;
        case 8: VM_LAUNCH(2, 4, 1, 4, 4, 2, false); break;
        default: VM_FATAL("MFMA beamformer: variant %d has no tile-group form", vi);
        }
        return;
    } else {
    switch (vi) {
    case 0: VM_LAUNCH(4, 4, 2, 2, 4, 1, false); break;
    case 1: VM_LAUNCH(4, 4, 1, 4, 4, 1, false); break;
    case 2: VM_LAUNCH(4, 4, 1, 4, 6, 1, false); break;
    case 3: VM_LAUNCH(2, 4, 2, 2, 4, 2, false); break;
    case 4: VM_LAUNCH(4, 2, 2, 2, 4, 2, false); break;
    case 5: VM_LAUNCH(2, 4, 4, 1, 4, 2, false); break;
    case 6: VM_LAUNCH(4, 4, 4, 1, 4, 1, false); break;
    case 7: VM_LAUNCH(4, 4, 2, 2, 6, 1, false); break;
    case 8: VM_LAUNCH(2, 4, 1, 4, 4, 2, false); break;
    case 9: VM_LAUNCH(2, 4, 4, 2, 4, 2, false); break;
    case 10: VM_LAUNCH(2, 4, 2, 2, 4, 2, true); break;
    case 11: VM_LAUNCH(2, 4, 8, 1, 4, 2, false); break;
    case 12:
        if constexpr (PB == 0) {
            if (G.nkstep == RES_NK)
                hipLaunchKernelGGL((k_beamform_res<PREC, 1, 8, true>), dim3((unsigned)(nchan * nty)),
                                   dim3(512), 0, st, G, ns, nsb, nchan, nty);
            else
                hipLaunchKernelGGL((k_beamform_res<PREC, 1, 8, false>), dim3((unsigned)(nchan * nty)),
                                   dim3(512), 0, st, G, ns, nsb, nchan, nty);
        }
        break;
    default: VM_FATAL("MFMA beamformer: no variant %d", vi);
    }
    }
#undef VM_LAUNCH
}

template <int PREC>
static void launch_batch_variant(int vi, unsigned nwg, hipStream_t st, const bf_geom *geoms,
                                 const uint2 *items, int nitems, uint32_t *queue, int ns, int nsb,
                                 int nchan)
{
#define VM_LAUNCH(MB_, NB_, WR_, WC_, NST_, OCC_)                                             \
    hipLaunchKernelGGL((k_beamform_batch<PREC, MB_, NB_, WR_, WC_, NST_, OCC_>), dim3(nwg),  \
                       dim3(64 * WR_ * WC_), 0, st, geoms, items, nitems, queue, ns, nsb, nchan)
    switch (vi) {
    case 0: VM_LAUNCH(4, 4, 2, 2, 4, 1); break;
    case 1: VM_LAUNCH(4, 4, 1, 4, 4, 1); break;
    case 2: VM_LAUNCH(4, 4, 1, 4, 6, 1); break;
    case 3: VM_LAUNCH(2, 4, 2, 2, 4, 2); break;
    case 4: VM_LAUNCH(4, 2, 2, 2, 4, 2); break;
    case 5: VM_LAUNCH(2, 4, 4, 1, 4, 2); break;
    case 6: VM_LAUNCH(4, 4, 4, 1, 4, 1); break;
    case 7: VM_LAUNCH(4, 4, 2, 2, 6, 1); break;
    case 8: VM_LAUNCH(2, 4, 1, 4, 4, 2); break;
    case 9: VM_LAUNCH(2, 4, 4, 2, 4, 2); break;
    default: VM_FATAL("batched beamformer: no persistent kernel for variant %d", vi);
    }
#undef VM_LAUNCH
}

// live[a] = 1 when J^-1 of antenna a (pointing 0, [a][c][4]) is non-zero in
// some channel: the antennas whose terms can be non-zero.  One wavefront per
// antenna.
__global__ void k_ant_live(const double2 *__restrict__ J, int nant, int nchan,
                           uint32_t *__restrict__ live)
{
    const int a = blockIdx.x, t = threadIdx.x;
    bool nz = false;
    for (int k = t; k < nchan * 4; k += 64) {
        const double2 m = J[(size_t)a * nchan * 4 + k];
        nz |= (m.x != 0.0) || (m.y != 0.0);
    }
    const bool any = __any(nz);
    if (t == 0) live[a] = any ? 1u : 0u;
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

// k_bf_prep over (nsb sample blocks) x (nchan channels) for nant <= 256
// antennas or GEMM slots (the template's antennas-per-thread bound is chosen
// from nant).  pbm 1: the per-pointing primary-beam form (Jf = D^-1, N =
// T_k); pbm 2: the tile-group form (per-slot Jf = D^-1 and gw, N = T_gk of
// the groups of gb).
// vin: the voltages' kind in v (PREP_VIN_*).
static void launch_prep(bool bf16, hipStream_t st, const void *v, int ns, int ni, int nsb,
                        int nkstep, int nchan, const cf *Jf, const float4 *gw,
                        const uint32_t *polQ, const uint32_t *polP, int nant, uint4 *Bf, int *sexp,
                        float4 *N, int pbm, const grp_bounds &gb, int vin)
{
    if (nant > 256) VM_FATAL("MFMA beamformer: at most 256 antennas (%d)", nant);
    if (vin == PREP_VIN_I4 && ni > PREP_I4_MAX_INPUTS)
        VM_FATAL("MFMA beamformer: at most %d RF inputs for 4+4-bit input (%d)",
                 PREP_I4_MAX_INPUTS, ni);
    if (vin == PREP_VIN_F16 && ni > PREP_F16_MAX_INPUTS)
        VM_FATAL("MFMA beamformer: at most %d RF inputs for FP16 input (%d)",
                 PREP_F16_MAX_INPUTS, ni);
    const dim3 pg((unsigned)nsb, (unsigned)nchan);
#define VM_PREP1(B_, A_, G_, P_, V_)                                                            \
    hipLaunchKernelGGL((k_bf_prep<B_, A_, G_, P_, V_>), pg, dim3(32 * G_), 0, st, v, ns, ni, nsb, \
                       nkstep, Jf, gw, polQ, polP, nant, Bf, sexp, N, gb)
#define VM_PREP(B_, A_, G_, P_)                                                               \
    do {                                                                                       \
        if (vin == PREP_VIN_I4) VM_PREP1(B_, A_, G_, P_, PREP_VIN_I4);                         \
        else if (vin == PREP_VIN_F16) VM_PREP1(B_, A_, G_, P_, PREP_VIN_F16);                  \
        else VM_PREP1(B_, A_, G_, P_, PREP_VIN_F32);                                           \
    } while (0)
#define VM_PREP2(A_, G_)                                                                      \
    if (pbm == 2) { if (bf16) VM_PREP(true, A_, G_, 2); else VM_PREP(false, A_, G_, 2); }     \
    else if (pbm == 1) { if (bf16) VM_PREP(true, A_, G_, 1); else VM_PREP(false, A_, G_, 1); } \
    else { if (bf16) VM_PREP(true, A_, G_, 0); else VM_PREP(false, A_, G_, 0); }
    // 8 antennas per thread; more antennas take more lanes, not more
    // registers per lane (occupancy hides the loads' latency)
    if (nant <= 64) {
        VM_PREP2(8, 8);
    } else if (nant <= 128) {
        VM_PREP2(8, 16);
    } else {
        VM_PREP2(8, 32);
    }
#undef VM_PREP2
#undef VM_PREP
#undef VM_PREP1
    HIPCHECK(hipGetLastError());
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The beamformer's prepared per-chunk state.
struct mfma_plan {
    int nant, ni, nchan, np, nkstep, nrb, ns, nsb, chunk;
    uint4 *Af, *Bf;
    float4 *N;
    int *sexp;
    float invw;
    int pb;      // 1: per-pointing primary beam; 2: primary beam per tile group
    float *bx;   // pb 1: B^-1 and h per (c, p)
    int *esp;    // pb 1, 2: weight exponents per (c, p)
    float *gx;   // pb 2: the groups' constants per (c, p)
    grp_bounds gb;
    int prec;    // GEMM precision (form_prec)
    int ns_total; // samples per second (the blocked layout of S)
};

// Whether J^-1 (vm->d_J) differs between pointings, decided on the device
// (flag: a device word, cleared here) and read back (synchronises st).
static int j_differs(vcsbeam_context *vm, hipStream_t st, uint32_t *flag)
{
    const int np = (int)vm->npointing;
    const uint32_t per = (uint32_t)vm->obs_metadata->num_ants * vm->nfine_chan * 4;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, flag, (size_t)1, 0u);
    hipLaunchKernelGGL(k_jcheck, dim3((per + 4095) / 4096, (unsigned)(np - 1)), dim3(256), 0, st,
                       (const double2 *)vm->d_J, per, flag);
    HIPCHECK(hipGetLastError());
    uint32_t h = 0;
    HIPCHECK(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return h != 0;
}

// Whether J^-1 may differ between pointings.  When it was computed from D
// and the primary beam d_B (vmCalcJDevice*, or vmCalcJ + vmPushJ), J^-1[p]
// is a fixed function of D and B[p], so it is the same for every pointing
// wherever B is: the check reads B (at most a few MB) instead of J^-1
// (hundreds of MB).  A beam that differs only on flagged tiles (D = 0) is
// reported as differing, and a per-tile beam that differs is then checked
// on J^-1 itself.
static int pointings_differ(vcsbeam_context *vm, hipStream_t st, uint32_t *flag)
{
    vm_device_state *d = vm_dev(vm);
    const int np = (int)vm->npointing;
    if (d->J_shared) return 0; // one J^-1 for every pointing (vmCalcJDevice)
    // d_B must hold the current pointings (a changed pointing count without
    // a new J^-1 falls back to comparing J^-1 itself)
    if (!d->J_fromB || !d->d_B || d->B_np != np) return j_differs(vm, st, flag);
    const uint32_t per =
        (uint32_t)(d->B_chan ? vm->nfine_chan : (int)vm->obs_metadata->num_ants) * 4;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, flag, (size_t)1, 0u);
    hipLaunchKernelGGL(k_jcheck, dim3((per + 4095) / 4096, (unsigned)(np - 1)), dim3(256), 0, st,
                       (const double2 *)d->d_B, per, flag);
    HIPCHECK(hipGetLastError());
    uint32_t h = 0;
    HIPCHECK(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (h && !d->B_chan) return j_differs(vm, st, flag);
    return h != 0;
}

// Tiles grouped by their primary beam pb.B[p][a] (equal in every pointing),
// groups in order of first appearance, each group's antennas ascending and
// padded to a multiple of 8 GEMM slots.  Fills slot_ant (-1: padding) and
// gb; returns the number of groups (more than VM_MFMA_MAX_GROUPS: not
// filled).
static int build_tile_groups(const vcsbeam_context *vm, int *slot_ant, int *nslot, grp_bounds *gb)
{
    const int nant = (int)vm->obs_metadata->num_ants, np = (int)vm->npointing;
    const gpuDoubleComplex *B = vm->pb.B;
    int grp[256], rep[VM_MFMA_MAX_GROUPS + 1], ng = 0;
    for (int a = 0; a < nant; a++) {
        int g = 0;
        for (; g < ng; g++) {
            bool same = true;
            for (int p = 0; p < np && same; p++)
                same = memcmp(&B[PB_IDX(p, a, 0, nant, 4)], &B[PB_IDX(p, rep[g], 0, nant, 4)],
                              4 * sizeof(gpuDoubleComplex)) == 0;
            if (same) break;
        }
        if (g == ng) {
            if (ng == VM_MFMA_MAX_GROUPS) return ng + 1;
            rep[ng++] = a;
        }
        grp[a] = g;
    }
    int n = 0;
    memset(gb, 0, sizeof(*gb));
    for (int g = 0; g < ng; g++) {
        gb->t[g] = n;
        gb->rep[g] = rep[g];
        for (int a = 0; a < nant; a++)
            if (grp[a] == g) slot_ant[n++] = a;
        while (n % 8) slot_ant[n++] = -1;
    }
    gb->t[ng] = n;
    gb->n = ng;
    *nslot = n;
    return ng;
}

static void mfma_prepare(vcsbeam_context *vm, mfma_plan *pl)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t st = (hipStream_t)d->stream;
    const bool bf16 = vm->beamform_form == VM_BF_MFMA_BF16;
    if (vm->datatype != VM_FLT && vm->datatype != VM_HALF && vm->datatype != VM_INT4)
        VM_FATAL("MFMA beamformer needs VM_FLT or VM_HALF (forward PFB output) or VM_INT4 "
                 "(legacy) voltages; got %d", (int)vm->datatype);
    const int vin = vm->datatype == VM_INT4   ? PREP_VIN_I4
                    : vm->datatype == VM_HALF ? PREP_VIN_F16
                                              : PREP_VIN_F32;
    const int nant = (int)vm->obs_metadata->num_ants, ni = (int)vm->obs_metadata->num_rf_inputs;
    const int nchan = vm->nfine_chan, np = (int)vm->npointing;
    if (ni != 2 * nant) VM_FATAL("MFMA beamformer: %d RF inputs for %d antennas", ni, nant);
    if (nant > 256) VM_FATAL("MFMA beamformer: at most 256 antennas (%d)", nant);
    const int nrb = ((2 * np + ROW_PAD - 1) / ROW_PAD) * ROW_PAD / RBK;
    const int ns = vm->fine_sample_rate / vm->chunks_per_second;
    const int nsb = (ns + SB - 1) / SB;

    bool rebuild = d->Wm_gen != d->weights_gen || d->Wm_form != vm->beamform_form;
    int slot_ant[256 + 8 * VM_MFMA_MAX_GROUPS];
    int nslot = 0;
    if (rebuild) {
        // the form: J^-1 shared by every pointing (u = J^-1 v), a primary
        // beam B[p][c] per pointing shared by every tile, or a beam per tile
        // that takes a few distinct values (tile groups)
        int pb = 0;
        if (np > 1) {
            ensure_buf(&d->d_flag, &d->flag_bytes, 256);
            if (pointings_differ(vm, st, (uint32_t *)d->d_flag)) {
                if (d->B_chan && d->d_B) {
                    pb = 1;
                } else if (d->J_fromB && d->d_B && d->B_np == np && vm->pb.B) {
                    grp_bounds gb;
                    const int ng = build_tile_groups(vm, slot_ant, &nslot, &gb);
                    if (ng > VM_MFMA_MAX_GROUPS)
                        VM_FATAL("MFMA beamformer: the tiles have more than %d distinct primary "
                                 "beams; use the FP32 form", VM_MFMA_MAX_GROUPS);
                    pb = 2;
                    memcpy(d->grp_t, gb.t, sizeof(gb.t));
                    memcpy(d->grp_rep, gb.rep, sizeof(gb.rep));
                    d->grp_n = gb.n;
                } else {
                    VM_FATAL("MFMA beamformer: J^-1 differs between pointings and was not "
                             "computed from D and a primary beam (vmCalcJ + vmPushJ, "
                             "vmCalcJDevice, vmCalcJDeviceChanB); use the FP32 form");
                }
            }
        }
        d->Wm_pb = pb;
    }
    const int pb = d->Wm_pb;
    grp_bounds gb = {};
    if (pb == 2) {
        memcpy(gb.t, d->grp_t, sizeof(gb.t));
        memcpy(gb.rep, d->grp_rep, sizeof(gb.rep));
        gb.n = d->grp_n;
    }
    // GEMM slots: the antennas, or (tile groups) the groups' padded runs
    const int nk = pb == 2 ? gb.t[gb.n] : nant;
    const int nkstep = (nk + 7) / 8; // 8 tiles (16 k') per step; slots past nk are zero

    // scratch: Af | Bf | N | sexp | Jf | g | bx (pb 1) or gx (pb 2) | esp
    const int ngrp = pb == 2 ? VM_MFMA_MAX_GROUPS : 1;
    const size_t af_bytes = align256((size_t)nchan * nrb * nkstep * 128 * sizeof(uint4));
    const size_t bf_bytes = align256((size_t)nchan * nsb * nkstep * 128 * sizeof(uint4));
    const size_t n_bytes = align256((size_t)nchan * ns * sizeof(float4) * (pb ? 4 * ngrp : 1));
    const size_t se_bytes = align256((size_t)nchan * nsb * sizeof(int));
    const size_t jf_bytes = align256((size_t)nchan * nk * 4 * sizeof(cf));
    const size_t g_bytes = pb ? align256((size_t)nchan * nk * sizeof(float4)) : 0;
    const size_t bx_bytes = pb == 1   ? align256((size_t)nchan * np * PB_ROW * sizeof(float))
                            : pb == 2 ? align256((size_t)nchan * np * GX_ROW * sizeof(float))
                                      : 0;
    const size_t esp_bytes = pb ? align256((size_t)nchan * np * sizeof(int)) : 0;
    const size_t need = af_bytes + bf_bytes + n_bytes + se_bytes + jf_bytes + g_bytes + bx_bytes +
                        esp_bytes;
    if (d->Wm_bytes < need) {
        ensure_buf(&d->d_Wm, &d->Wm_bytes, need);
        rebuild = true;
    }
    uint8_t *q = (uint8_t *)d->d_Wm;
    uint4 *Af = (uint4 *)q;                      q += af_bytes;
    uint4 *Bf = (uint4 *)q;                      q += bf_bytes;
    float4 *N = (float4 *)q;                     q += n_bytes;
    int *sexp = (int *)q;                        q += se_bytes;
    cf *Jf = (cf *)q;                            q += jf_bytes;
    float4 *gw = pb ? (float4 *)q : NULL;        q += g_bytes;
    float *bx = pb == 1 ? (float *)q : NULL;
    float *gx = pb == 2 ? (float *)q : NULL;     q += bx_bytes;
    int *esp = pb ? (int *)q : NULL;
    // tile groups: slot arrays (antenna, polQ, polP) on the device
    int *d_amap = NULL;
    uint32_t *d_sq = NULL, *d_sp = NULL;
    if (pb == 2) {
        ensure_buf(&d->d_grp, &d->grp_bytes, (size_t)3 * nk * sizeof(int));
        d_amap = (int *)d->d_grp;
        d_sq = (uint32_t *)(d_amap + nk);
        d_sp = d_sq + nk;
    }

    if (rebuild) {
        vm_timer_start(d, VM_T_BF_WEIGHTS);
        const size_t total = (size_t)nchan * nrb * nkstep * 64;
        const dim3 g((unsigned)((total + 255) / 256));
        const double2 *dD = (const double2 *)vm->d_D, *dB = (const double2 *)d->d_B;
        const double2 *phi = (const double2 *)vm->gdelays.d_phi;
        const grp_bounds none = {};
        if (pb == 2) {
            if (nslot != nk) VM_FATAL("MFMA beamformer: tile groups changed without a rebuild");
            uint32_t *hs = (uint32_t *)malloc((size_t)3 * nk * sizeof(uint32_t));
            if (!hs) VM_FATAL("out of host memory");
            for (int a = 0; a < nk; a++) {
                const int as = slot_ant[a];
                hs[a] = (uint32_t)as;
                hs[nk + a] = as >= 0 ? vm->polQ_idxs[as] : 0u;
                hs[2 * nk + a] = as >= 0 ? vm->polP_idxs[as] : 0u;
            }
            HIPCHECK(hipMemcpyAsync(d->d_grp, hs, (size_t)3 * nk * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, st));
            ensure_buf(&d->d_Bg, &d->Bg_bytes, (size_t)np * gb.n * nchan * 4 * sizeof(double2));
            const size_t nbg = (size_t)np * gb.n * nchan;
            hipLaunchKernelGGL(k_bg_build, dim3((unsigned)((nbg + 255) / 256)), dim3(256), 0, st,
                               dB, np, nant, nchan, gb, (double2 *)d->d_Bg);
            hipLaunchKernelGGL(k_pb_dinv, dim3((unsigned)((nk * nchan + 255) / 256)), dim3(256),
                               0, st, dD, nk, nchan, (const int *)d_amap, Jf, gw);
            hipLaunchKernelGGL(k_pbg_consts, dim3((unsigned)((np * nchan + 255) / 256)), dim3(256),
                               0, st, (const double2 *)d->d_Bg, dD, (const int *)d_amap, gb, np,
                               nchan, gx, esp);
            if (bf16)
                hipLaunchKernelGGL((k_phase_weights<true, 2>), g, dim3(256), 0, st, phi, np, nk,
                                   nchan, nrb, nkstep, (const int *)d_amap, nant, dD,
                                   (const double2 *)d->d_Bg, (const int *)esp, Af, gb);
            else
                hipLaunchKernelGGL((k_phase_weights<false, 2>), g, dim3(256), 0, st, phi, np, nk,
                                   nchan, nrb, nkstep, (const int *)d_amap, nant, dD,
                                   (const double2 *)d->d_Bg, (const int *)esp, Af, gb);
            HIPCHECK(hipStreamSynchronize(st)); // the host slot array is released
            free(hs);
        } else if (pb == 1) {
            hipLaunchKernelGGL(k_pb_dinv, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256),
                               0, st, dD, nant, nchan, (const int *)NULL, Jf, gw);
            hipLaunchKernelGGL(k_pb_beam, dim3((unsigned)((np * nchan + 255) / 256)), dim3(256), 0,
                               st, dD, dB, np, nant, nchan, bx, esp);
            if (bf16)
                hipLaunchKernelGGL((k_phase_weights<true, 1>), g, dim3(256), 0, st, phi, np,
                                   nant, nchan, nrb, nkstep, (const int *)NULL, nant, dD, dB,
                                   (const int *)esp, Af, none);
            else
                hipLaunchKernelGGL((k_phase_weights<false, 1>), g, dim3(256), 0, st, phi, np,
                                   nant, nchan, nrb, nkstep, (const int *)NULL, nant, dD, dB,
                                   (const int *)esp, Af, none);
        } else {
            hipLaunchKernelGGL(k_jf, dim3((unsigned)((nant * nchan + 255) / 256)), dim3(256), 0,
                               st, (const double2 *)vm->d_J, nant, nchan, (const int *)NULL, Jf);
            if (bf16)
                hipLaunchKernelGGL((k_phase_weights<true, 0>), g, dim3(256), 0, st, phi, np,
                                   nant, nchan, nrb, nkstep, (const int *)NULL, nant,
                                   (const double2 *)NULL, (const double2 *)NULL,
                                   (const int *)NULL, Af, none);
            else
                hipLaunchKernelGGL((k_phase_weights<false, 0>), g, dim3(256), 0, st, phi, np,
                                   nant, nchan, nrb, nkstep, (const int *)NULL, nant,
                                   (const double2 *)NULL, (const double2 *)NULL,
                                   (const int *)NULL, Af, none);
        }
        HIPCHECK(hipGetLastError());
        vm_timer_stop(d, VM_T_BF_WEIGHTS);
        d->Wm_gen = d->weights_gen;
        d->Wm_form = vm->beamform_form;
    }
    // u = J^-1 v (D^-1 v), auto-correlation terms, block scales and the B operand
    vm_timer_start(d, VM_T_BF_PREP);
    if (pb == 2)
        launch_prep(bf16, st, vm->d_v, ns, ni, nsb, nkstep, nchan, Jf, gw, d_sq, d_sp, nk, Bf,
                    sexp, N, 2, gb, vin);
    else
        launch_prep(bf16, st, vm->d_v, ns, ni, nsb, nkstep, nchan, Jf, gw, vm->d_polQ_idxs,
                    vm->d_polP_idxs, nant, Bf, sexp, N, pb, gb, vin);
    vm_timer_stop(d, VM_T_BF_PREP);
    pl->nant = nant; pl->ni = ni; pl->nchan = nchan; pl->np = np; pl->nkstep = nkstep;
    pl->nrb = nrb; pl->ns = ns; pl->nsb = nsb;
    pl->chunk = vm->chunk_to_load % vm->chunks_per_second;
    pl->Af = Af; pl->Bf = Bf; pl->N = N; pl->sexp = sexp;
    pl->pb = pb; pl->bx = bx; pl->esp = esp; pl->gx = gx; pl->gb = gb;
    pl->invw = (float)(1.0 / (double)vm->num_not_flagged);
    pl->prec = form_prec(vm->beamform_form);
    pl->ns_total = (int)vm->fine_sample_rate;
}

static bool variant_ok(int vi, const mfma_plan *pl)
{
    // sample counts that are not multiples of 4 have only variant 0
    if (pl->ns & 3 && vi != 0) return false;
    if (pl->pb == 2 && !variant_has_groups(vi)) return false;
    // resident weights: pointing-independent J^-1 and A within the registers;
    // each channel's part of S under 2 GB
    if (vi >= 0 && vi < N_VARIANTS && kVariants[vi].res) {
        const size_t chan_bytes = (size_t)VM_S_NBLK(pl->ns_total) * pl->np * 4 * VM_SBLK * 4;
        if (pl->pb != 0 || pl->nkstep > RES_NK || chan_bytes >= ((size_t)1 << 31)) return false;
    }
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
    bf_geom G;
    G.Af = pl->Af; G.Bf = pl->Bf; G.sexp = pl->sexp; G.N = pl->N; G.S = (float *)vm->d_S;
    G.nkstep = pl->nkstep; G.nrb = pl->nrb; G.npointing = pl->np; G.nstokes = vm->out_nstokes;
    G.soffset = pl->chunk * pl->ns; G.ns_total = vm->fine_sample_rate; G.invw = pl->invw;
    G.bx = pl->bx; G.esp = pl->esp; G.gx = pl->gx; G.gb = pl->gb; G.pad = 0;
    const int prec = form_prec(vm->beamform_form);
#define VM_RUN(B_, P_) launch_variant<B_, P_>(vi, grid, st, G, pl->ns, pl->nsb, pl->nchan, ntx, nty, order)
#define VM_RUN_PB(B_)                       \
    do {                                    \
        if (pl->pb == 2) VM_RUN(B_, 2);      \
        else if (pl->pb == 1) VM_RUN(B_, 1); \
        else VM_RUN(B_, 0);                  \
    } while (0)
    if (prec == 1) VM_RUN_PB(1);
    else if (prec == 2) VM_RUN_PB(2);
    else VM_RUN_PB(0);
#undef VM_RUN_PB
#undef VM_RUN
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

const char *vmBeamformerVariantKernel(int variant)
{
    if (variant < 0 || variant >= N_VARIANTS) return NULL;
    return kVariants[variant].res ? "k_beamform_res" : "k_beamform_mfma";
}

int vmBeamformerSelection(vcsbeam_context *vm, int *variant, int *order)
{
    vm_device_state *d = vm_dev(vm);
    const int tuned = d->bf_variant > 0;
    if (variant) *variant = tuned ? d->bf_variant - 1 : 0;
    if (order) *order = tuned ? d->bf_order : 1;
    return tuned;
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
        // (the resident-weights form has one workgroup order)
        for (int order = 0; order < (kVariants[vi].res ? 1 : VM_N_ORDERS); order++) {
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

// ------------------------------------------------------------- batches ---
// Several coarse channels (one context each) beamformed together: per
// channel the fused forward PFB and the prep pass, then ONE persistent GEMM
// launch over every channel's tiles, the whole sequence captured once in a
// hipGraph and replayed each second.  Channels may differ in beam count and
// in flagged tiles: antennas whose J^-1 is zero in every fine channel (flagged
// tiles, D = 0) are dropped from the channel's GEMM (its K), and each channel
// has its own row count, so the work items differ in cost; they are listed
// most expensive first and handed out from a queue counter.

void vm_pfb_fused_launch(const void *d_htr, size_t htr_bytes, const int *d_h,
                         const int *d_out_idx, int I, int nspc, int P, int fpga, int imag_first,
                         int split, int half, void *d_out, hipStream_t stream);

struct vm_batch_chan {
    vcsbeam_context *vm;
    int valid;
    unsigned gen;                 // context weights generation the weights were built from
    int nant_src, nact, nkstep, nrb, np;
    float invw;
    void *mem;
    size_t mem_bytes;
    int *d_amap;
    uint32_t *d_polQ, *d_polP;
    cf *Jf;
    uint4 *Af, *Bf;
    float4 *N;
    int *sexp;
};

struct vm_batch {
    int n, nchan, ns, nsb, bf16, prec, ncu;
    vm_batch_chan *ch;
    void *d_v;                    // forward-PFB output shared by contexts without their own
    bf_geom *d_geoms;
    uint2 *d_items;
    size_t items_cap;
    int nitems, variant, band;
    uint32_t *d_small;            // [0] queue counter, [1] J check flag, [64..319] live antennas
    hipStream_t st;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int layout_changed;
};

static bool batch_variant_ok(int vi)
{
    // the persistent batch kernel has the register-ring k-loop only
    return vi >= 0 && vi < N_VARIANTS && !kVariants[vi].lds && !kVariants[vi].res &&
           ROW_PAD % (RBK * kVariants[vi].mb * kVariants[vi].wr) == 0;
}

static void batch_drop_graph(vm_batch *b)
{
    if (b->exec) HIPCHECK(hipGraphExecDestroy(b->exec));
    if (b->graph) HIPCHECK(hipGraphDestroy(b->graph));
    b->exec = NULL;
    b->graph = NULL;
}

// Work items for variant vi, band row tiles per item (0: a whole column).
static void batch_items(vm_batch *b, int vi, int band)
{
    const bf_variant &V = kVariants[vi];
    const int TR = RBK * V.mb * V.wr, cols = (V.nb / 2) * V.wc;
    const int ntx = (b->nsb + cols - 1) / cols;
    size_t total = 0;
    for (int g = 0; g < b->n; g++) {
        const vm_batch_chan &C = b->ch[g];
        const int nty = (2 * C.np + TR - 1) / TR;
        const int R = band > 0 ? band : nty;
        total += (size_t)b->nchan * ntx * ((nty + R - 1) / R);
    }
    if (total >= ((size_t)1 << 31)) VM_FATAL("batch: %zu work items", total);
    uint2 *h = (uint2 *)malloc(total * sizeof(uint2));
    uint32_t *cost = (uint32_t *)malloc(total * sizeof(uint32_t));
    if (!h || !cost) VM_FATAL("out of host memory");
    size_t k = 0;
    uint32_t maxcost = 0;
    for (int g = 0; g < b->n; g++) {
        const vm_batch_chan &C = b->ch[g];
        const int nty = (2 * C.np + TR - 1) / TR;
        const int R = band > 0 ? band : nty;
        if (nty * TR > C.nrb * RBK) VM_FATAL("batch: row tiles exceed the padded rows");
        if (nty > 255 || R > 255)
            VM_FATAL("batch: %d beams in channel %d need %d row tiles (at most 255)", C.np, g, nty);
        if (b->nchan > 65535 || ntx > 65535) VM_FATAL("batch: too many channels / sample tiles");
        for (int c = 0; c < b->nchan; c++)
            for (int tx = 0; tx < ntx; tx++)
                for (int ty0 = 0; ty0 < nty; ty0 += R) {
                    const int cnt = min(R, nty - ty0);
                    h[k] = make_uint2(((uint32_t)g << 16) | (uint32_t)c,
                                      ((uint32_t)tx << 16) | ((uint32_t)ty0 << 8) | (uint32_t)cnt);
                    cost[k] = (uint32_t)(C.nkstep * cnt);
                    if (cost[k] > maxcost) maxcost = cost[k];
                    k++;
                }
    }
    // most expensive first (a stable counting sort on cost keeps each
    // column's items together)
    uint2 *sorted = (uint2 *)malloc(total * sizeof(uint2));
    size_t *start = (size_t *)calloc((size_t)maxcost + 2, sizeof(size_t));
    if (!sorted || !start) VM_FATAL("out of host memory");
    for (size_t i = 0; i < total; i++) start[maxcost - cost[i] + 1]++;
    for (uint32_t c = 1; c <= maxcost + 1; c++) start[c] += start[c - 1];
    for (size_t i = 0; i < total; i++) sorted[start[maxcost - cost[i]]++] = h[i];
    if (b->items_cap < total) {
        if (b->d_items) HIPCHECK(hipFree(b->d_items));
        HIPCHECK(hipMalloc((void **)&b->d_items, total * sizeof(uint2)));
        b->items_cap = total;
    }
    HIPCHECK(hipMemcpy(b->d_items, sorted, total * sizeof(uint2), hipMemcpyHostToDevice));
    free(h);
    free(cost);
    free(sorted);
    free(start);
    b->nitems = (int)total;
    b->variant = vi;
    b->band = band;
    batch_drop_graph(b);
}

// Bring each channel's weights, antenna list and operand buffers up to date
// with its context (J^-1, phases, pointings); outside any graph.
static void batch_refresh(vm_batch *b)
{
    hipStream_t st = b->st;
    uint32_t *queue_flag = b->d_small + 1, *d_live = b->d_small + 64;
    for (int g = 0; g < b->n; g++) {
        vm_batch_chan &C = b->ch[g];
        vcsbeam_context *vm = C.vm;
        vm_device_state *d = vm_dev(vm);
        HIPCHECK(hipStreamSynchronize((hipStream_t)d->stream));
        const int np = (int)vm->npointing;
        if (C.valid && C.gen == d->weights_gen && C.np == np) continue;
        const int nant = (int)vm->obs_metadata->num_ants, nchan = b->nchan, nsb = b->nsb;
        hipLaunchKernelGGL(k_ant_live, dim3((unsigned)nant), dim3(64), 0, st,
                           (const double2 *)vm->d_J, nant, nchan, d_live);
        HIPCHECK(hipGetLastError());
        uint32_t live[256];
        HIPCHECK(hipMemcpyAsync(live, d_live, nant * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        int amap[256];
        uint32_t pq[256], pp[256];
        int nact = 0;
        for (int a = 0; a < nant; a++)
            if (live[a]) {
                pq[nact] = vm->polQ_idxs[a];
                pp[nact] = vm->polP_idxs[a];
                amap[nact++] = a;
            }
        const int nkstep = max(1, (nact + 7) / 8);
        const int nrb = ((2 * np + ROW_PAD - 1) / ROW_PAD) * ROW_PAD / RBK;
        const size_t af = align256((size_t)nchan * nrb * nkstep * 128 * sizeof(uint4));
        const size_t bfb = align256((size_t)nchan * nsb * nkstep * 128 * sizeof(uint4));
        const size_t nb = align256((size_t)nchan * b->ns * sizeof(float4));
        const size_t se = align256((size_t)nchan * nsb * sizeof(int));
        const size_t jf = align256((size_t)nchan * max(nact, 1) * 4 * sizeof(cf));
        const size_t need = af + bfb + nb + se + jf + 3 * 1024;
        if (C.mem_bytes < need) {
            if (C.mem) HIPCHECK(hipFree(C.mem));
            HIPCHECK(hipMalloc(&C.mem, need));
            C.mem_bytes = need;
            b->layout_changed = 1;
        }
        uint8_t *q = (uint8_t *)C.mem;
        C.Af = (uint4 *)q;     q += af;
        C.Bf = (uint4 *)q;     q += bfb;
        C.N = (float4 *)q;     q += nb;
        C.sexp = (int *)q;     q += se;
        C.Jf = (cf *)q;        q += jf;
        C.d_amap = (int *)q;   q += 1024;
        C.d_polQ = (uint32_t *)q; q += 1024;
        C.d_polP = (uint32_t *)q;
        if (!C.valid || nact != C.nact || nkstep != C.nkstep || nrb != C.nrb || np != C.np)
            b->layout_changed = 1;
        if (nact > 0) {
            HIPCHECK(hipMemcpyAsync(C.d_amap, amap, nact * sizeof(int), hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(C.d_polQ, pq, nact * sizeof(uint32_t), hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(C.d_polP, pp, nact * sizeof(uint32_t), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_jf, dim3((unsigned)((nact * nchan + 255) / 256)), dim3(256), 0, st,
                               (const double2 *)vm->d_J, nact, nchan, (const int *)C.d_amap, C.Jf);
        }
        const size_t total = (size_t)nchan * nrb * nkstep * 64;
        const dim3 gw((unsigned)((total + 255) / 256));
        const grp_bounds none = {};
        if (b->bf16)
            hipLaunchKernelGGL((k_phase_weights<true, 0>), gw, dim3(256), 0, st,
                               (const double2 *)vm->gdelays.d_phi, np, nact, nchan, nrb, nkstep,
                               (const int *)C.d_amap, nant, (const double2 *)NULL,
                               (const double2 *)NULL, (const int *)NULL, C.Af, none);
        else
            hipLaunchKernelGGL((k_phase_weights<false, 0>), gw, dim3(256), 0, st,
                               (const double2 *)vm->gdelays.d_phi, np, nact, nchan, nrb, nkstep,
                               (const int *)C.d_amap, nant, (const double2 *)NULL,
                               (const double2 *)NULL, (const int *)NULL, C.Af, none);
        int hflag = 0;
        if (np > 1) hflag = pointings_differ(vm, st, queue_flag);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(st));
        if (hflag && vm->dev->B_chan)
            hflag = j_differs(vm, st, queue_flag); // a per-pointing beam, but is J^-1 uniform?
        if (hflag)
            VM_FATAL("batch channel %d: J^-1 differs between pointings; the MFMA form needs a "
                     "beam-independent J^-1", g);
        C.nant_src = nant;
        C.nact = nact;
        C.nkstep = nkstep;
        C.nrb = nrb;
        C.np = np;
        C.invw = (float)(1.0 / (double)vm->num_not_flagged);
        C.gen = d->weights_gen;
        C.valid = 1;
    }
    if (b->layout_changed) {
        bf_geom *G = (bf_geom *)calloc(b->n, sizeof(bf_geom));
        if (!G) VM_FATAL("out of host memory");
        for (int g = 0; g < b->n; g++) {
            const vm_batch_chan &C = b->ch[g];
            G[g].Af = C.Af; G[g].Bf = C.Bf; G[g].sexp = C.sexp; G[g].N = C.N;
            G[g].S = (float *)C.vm->d_S;
            G[g].nkstep = C.nkstep; G[g].nrb = C.nrb; G[g].npointing = C.np;
            G[g].nstokes = C.vm->out_nstokes; G[g].soffset = 0;
            G[g].ns_total = C.vm->fine_sample_rate; G[g].invw = C.invw;
        }
        HIPCHECK(hipMemcpy(b->d_geoms, G, b->n * sizeof(bf_geom), hipMemcpyHostToDevice));
        free(G);
        batch_items(b, b->variant, b->band);
        b->layout_changed = 0;
    }
}

static void batch_pfb(vm_batch *b, int g, hipStream_t st)
{
    forward_pfb *f = b->ch[g].vm->fpfb;
    vm_pfb_fused_launch(f->d_htr_data, f->d_htr_size, f->d_filter_coeffs, f->d_i_output_idx, f->I,
                        f->nspectra_per_chunk, f->P, (f->flags & PFB_EMULATE_FPGA) != 0,
                        (f->flags & PFB_IMAG_PART_FIRST) != 0, !f->filter_fits_int16,
                        b->ch[g].vm->datatype == VM_HALF, f->d_vcs_data, st);
}

static void batch_prep(vm_batch *b, int g, hipStream_t st)
{
    const vm_batch_chan &C = b->ch[g];
    const grp_bounds none = {};
    launch_prep(b->bf16, st, (const cf *)C.vm->fpfb->d_vcs_data, b->ns,
                (int)C.vm->obs_metadata->num_rf_inputs, b->nsb, C.nkstep, b->nchan, C.Jf,
                (const float4 *)NULL, C.d_polQ, C.d_polP, C.nact, C.Bf, C.sexp, C.N, 0, none,
                C.vm->datatype == VM_HALF ? PREP_VIN_F16 : PREP_VIN_F32);
}

static void batch_gemm(vm_batch *b, hipStream_t st)
{
    const bf_variant &V = kVariants[b->variant];
    // queue counter cleared by a kernel (no memset node in the graph)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, b->d_small, (size_t)1, 0u);
    const int per_cu = max(1, 4 * V.occ / (V.wr * V.wc));
    const unsigned nwg = (unsigned)min(b->nitems, b->ncu * per_cu);
    if (b->prec == 1)
        launch_batch_variant<1>(b->variant, nwg, st, b->d_geoms, b->d_items, b->nitems,
                                b->d_small, b->ns, b->nsb, b->nchan);
    else if (b->prec == 2)
        launch_batch_variant<2>(b->variant, nwg, st, b->d_geoms, b->d_items, b->nitems,
                                b->d_small, b->ns, b->nsb, b->nchan);
    else
        launch_batch_variant<0>(b->variant, nwg, st, b->d_geoms, b->d_items, b->nitems,
                                b->d_small, b->ns, b->nsb, b->nchan);
    HIPCHECK(hipGetLastError());
}

static void batch_sequence(vm_batch *b, hipStream_t st)
{
    for (int g = 0; g < b->n; g++) {
        batch_pfb(b, g, st);
        batch_prep(b, g, st);
    }
    batch_gemm(b, st);
}

vm_batch *vmBatchCreate(vcsbeam_context **vms, int n)
{
    if (!vms || n < 1 || n > 65535) VM_FATAL("vmBatchCreate: %d contexts (1..65535)", n);
    vm_batch *b = (vm_batch *)calloc(1, sizeof(vm_batch));
    if (!b) VM_FATAL("out of host memory");
    b->ch = (vm_batch_chan *)calloc(n, sizeof(vm_batch_chan));
    if (!b->ch) VM_FATAL("out of host memory");
    b->n = n;
    const vcsbeam_context *v0 = vms[0];
    b->nchan = v0->nfine_chan;
    b->ns = (int)v0->fine_sample_rate;
    if (b->ns & 3) VM_FATAL("vmBatchCreate: %d samples per second (a multiple of 4 is needed)", b->ns);
    b->nsb = (b->ns + SB - 1) / SB;
    b->bf16 = v0->beamform_form == VM_BF_MFMA_BF16;
    b->prec = form_prec(v0->beamform_form);
    size_t vneed = 0;
    for (int g = 0; g < n; g++) {
        vcsbeam_context *vm = vms[g];
        forward_pfb *f = vm ? vm->fpfb : NULL;
        if (!f) VM_FATAL("vmBatchCreate: context %d has no forward PFB", g);
        if ((vm->datatype != VM_FLT && vm->datatype != VM_HALF) || (f->flags & PFB_STAGED) ||
            vm->datatype != v0->datatype)
            VM_FATAL("vmBatchCreate: context %d: the batch runs the fused PFB to VM_FLT or "
                     "VM_HALF, the same in every context", g);
        if (vm->chunks_per_second != 1)
            VM_FATAL("vmBatchCreate: context %d: one chunk per second is required", g);
        if (vm->nfine_chan != b->nchan || (int)vm->fine_sample_rate != b->ns)
            VM_FATAL("vmBatchCreate: context %d differs in channels or samples", g);
        if (vm->beamform_form != v0->beamform_form ||
            (vm->beamform_form != VM_BF_MFMA_FP16 && vm->beamform_form != VM_BF_MFMA_BF16 &&
             vm->beamform_form != VM_BF_MFMA_FP16_SINGLE))
            VM_FATAL("vmBatchCreate: context %d: all contexts need the same MFMA form", g);
        if (!vm->d_S || !vm->d_J || !vm->gdelays.d_phi || !vm->polQ_idxs || !f->d_htr_data)
            VM_FATAL("vmBatchCreate: context %d lacks S, J, phi, pol lists or PFB input", g);
        if (vm->obs_metadata->num_rf_inputs != 2 * vm->obs_metadata->num_ants ||
            vm->obs_metadata->num_ants > 256)
            VM_FATAL("vmBatchCreate: context %d: %u antennas / %u inputs", g,
                     vm->obs_metadata->num_ants, vm->obs_metadata->num_rf_inputs);
        if (!f->d_vcs_data && f->d_vcs_size > vneed) vneed = f->d_vcs_size;
        b->ch[g].vm = vm;
    }
    if (vneed) {
        HIPCHECK(hipMalloc(&b->d_v, vneed));
        for (int g = 0; g < n; g++) {
            vcsbeam_context *vm = vms[g];
            if (vm->fpfb->d_vcs_data) continue;
            vm->fpfb->d_vcs_data = b->d_v;
            vm->d_v = b->d_v;
            vm->d_v_size_bytes = vm->fpfb->d_vcs_size;
        }
    }
    HIPCHECK(hipMalloc((void **)&b->d_geoms, n * sizeof(bf_geom)));
    HIPCHECK(hipMalloc((void **)&b->d_small, 1024 * sizeof(uint32_t)));
    HIPCHECK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    b->variant = 1;
    b->band = 4;
    b->layout_changed = 1;
    return b;
}

void vmBatchBeamformSecond(vm_batch *b, int use_graph)
{
    batch_refresh(b);
    if (!use_graph) {
        batch_sequence(b, b->st);
    } else {
        if (!b->exec) {
            HIPCHECK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
            batch_sequence(b, b->st);
            HIPCHECK(hipStreamEndCapture(b->st, &b->graph));
            HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0));
        }
        HIPCHECK(hipGraphLaunch(b->exec, b->st));
    }
    HIPCHECK(hipStreamSynchronize(b->st));
}

void vmBatchStageTimes(vm_batch *b, float *pfb_ms, float *prep_ms, float *gemm_ms)
{
    batch_refresh(b);
    const int ne = 2 * b->n + 3;
    hipEvent_t *ev = (hipEvent_t *)calloc(ne, sizeof(hipEvent_t));
    if (!ev) VM_FATAL("out of host memory");
    for (int i = 0; i < ne; i++) HIPCHECK(hipEventCreate(&ev[i]));
    HIPCHECK(hipEventRecord(ev[0], b->st));
    for (int g = 0; g < b->n; g++) {
        batch_pfb(b, g, b->st);
        HIPCHECK(hipEventRecord(ev[2 * g + 1], b->st));
        batch_prep(b, g, b->st);
        HIPCHECK(hipEventRecord(ev[2 * g + 2], b->st));
    }
    batch_gemm(b, b->st);
    HIPCHECK(hipEventRecord(ev[2 * b->n + 1], b->st));
    HIPCHECK(hipEventSynchronize(ev[2 * b->n + 1]));
    float tp = 0.f, tq = 0.f, t;
    for (int g = 0; g < b->n; g++) {
        HIPCHECK(hipEventElapsedTime(&t, ev[2 * g], ev[2 * g + 1]));
        tp += t;
        HIPCHECK(hipEventElapsedTime(&t, ev[2 * g + 1], ev[2 * g + 2]));
        tq += t;
    }
    HIPCHECK(hipEventElapsedTime(&t, ev[2 * b->n], ev[2 * b->n + 1]));
    if (pfb_ms) *pfb_ms = tp;
    if (prep_ms) *prep_ms = tq;
    if (gemm_ms) *gemm_ms = t;
    for (int i = 0; i < ne; i++) HIPCHECK(hipEventDestroy(ev[i]));
    free(ev);
}

void vmBatchSetSchedule(vm_batch *b, int variant, int band)
{
    if (!batch_variant_ok(variant)) VM_FATAL("batch: no usable variant %d", variant);
    if (band < 0 || band > 255) VM_FATAL("batch: band %d (0..255)", band);
    b->variant = variant;
    b->band = band;
    if (!b->layout_changed && b->nitems) batch_items(b, variant, band);
}

float vmBatchTune(vm_batch *b, int reps)
{
    // times every usable (variant, band) on the current second and keeps
    // the fastest; S holds the chosen schedule's output afterwards
    batch_refresh(b);
    for (int g = 0; g < b->n; g++) {
        batch_pfb(b, g, b->st);
        batch_prep(b, g, b->st);
    }
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    static const int bands[] = {1, 2, 4, 8, 0};
    float best = 1e30f;
    int bv = b->variant, bb = b->band;
    for (int vi = 0; vi < N_VARIANTS; vi++) {
        if (!batch_variant_ok(vi)) continue;
        for (int bi = 0; bi < (int)(sizeof(bands) / sizeof(bands[0])); bi++) {
            batch_items(b, vi, bands[bi]);
            batch_gemm(b, b->st); // warm-up
            HIPCHECK(hipEventRecord(e0, b->st));
            for (int r = 0; r < reps; r++) batch_gemm(b, b->st);
            HIPCHECK(hipEventRecord(e1, b->st));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (getenv("VCSBEAM_TUNE_VERBOSE"))
                fprintf(stderr, "batch variant %d band %d: %d items, %.3f ms\n", vi, bands[bi],
                        b->nitems, ms);
            if (ms < best) { best = ms; bv = vi; bb = bands[bi]; }
        }
    }
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    batch_items(b, bv, bb);
    batch_gemm(b, b->st);
    HIPCHECK(hipStreamSynchronize(b->st));
    return best;
}

int vmBatchInfo(vm_batch *b, int *variant, int *band, int *nitems)
{
    if (variant) *variant = b->variant;
    if (band) *band = b->band;
    if (nitems) *nitems = b->nitems;
    return b->n;
}

int vmBatchActiveTiles(vm_batch *b, int g)
{
    if (g < 0 || g >= b->n) return -1;
    batch_refresh(b);
    return b->ch[g].nact;
}

void vmBatchDestroy(vm_batch *b)
{
    if (!b) return;
    HIPCHECK(hipStreamSynchronize(b->st));
    batch_drop_graph(b);
    for (int g = 0; g < b->n; g++) {
        vm_batch_chan &C = b->ch[g];
        if (C.mem) HIPCHECK(hipFree(C.mem));
        if (b->d_v && C.vm->fpfb && C.vm->fpfb->d_vcs_data == b->d_v) {
            C.vm->fpfb->d_vcs_data = NULL;
            if (C.vm->d_v == b->d_v) {
                C.vm->d_v = NULL;
                C.vm->d_v_size_bytes = 0;
            }
        }
    }
    if (b->d_v) HIPCHECK(hipFree(b->d_v));
    if (b->d_items) HIPCHECK(hipFree(b->d_items));
    HIPCHECK(hipFree(b->d_geoms));
    HIPCHECK(hipFree(b->d_small));
    HIPCHECK(hipStreamDestroy(b->st));
    free(b->ch);
    free(b);
}

} // extern "C"
