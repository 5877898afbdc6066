// geometry.hip -- FP64 device kernels for the per-second Jones / delay pass.
//
// Reference behaviour: MWA-vcsbeam src/geometry.c:40-135 (calc_geometric_delays,
// vmCalcPhi) and src/jones.c:97-157 (vmCalcJ), which run single-threaded on
// the host once per second.  Here both run on the GPU in FP64:
//  * phi = exp(i 2 pi f dt): the cycle count f*dt (up to ~1e4 cycles for
//    3 km baselines at 300 MHz) is split exactly with an FMA into its rounded
//    value and residual, the integer part is removed, and sincospi takes the
//    reduced argument, so the phase is accurate to ~1e-16 cycles instead of
//    losing ~4 digits to the 2*pi*f*dt product.
//  * J^-1 = F (D B)^-1 with F the Frobenius norm of D B, as the reference's
//    norm2x2 + inv2x2S sequence.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define HIPCHECK(x)                                                               \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) VM_FATAL("HIP: %s (%s)", hipGetErrorString(e_), #x); \
    } while (0)

// pos[a] = (E, N, H - Href, L); unit[p] = (uE, uN, uH); phi[p][a][c]
__global__ void k_calc_phi(const double4 *__restrict__ pos, const double *__restrict__ unit,
                           const double *__restrict__ freq, int npointing, int nant, int nchan,
                           double2 *__restrict__ phi)
{
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)npointing * nant * nchan;
    if (j >= total) return;
    int c = (int)(j % nchan);
    int a = (int)((j / nchan) % nant);
    int p = (int)(j / ((size_t)nchan * nant));
    double4 q = pos[a];
    const double *u = unit + 3 * p;
    double w = q.x * u[0] + q.y * u[1] + q.z * u[2];      // src/geometry.c:82-84
    double dt = (w - q.w) / SPEED_OF_LIGHT_IN_VACUUM_M_PER_S; // line 91
    double f = freq[c];
    double cyc = dt * f;
    double err = fma(dt, f, -cyc);                        // exact residual of dt*f
    double fr = (cyc - rint(cyc)) + err;                  // reduced cycles, |fr| <= 0.5+
    double s, co;
    sincospi(2.0 * fr, &s, &co);
    phi[j] = make_double2(co, s);
}

__device__ __forceinline__ double2 cm(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 ca(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// J^-1 of element j = (p, a, c) into out[0..3]
__device__ __forceinline__ void calc_J_one(const double2 *__restrict__ D,
                                           const double2 *__restrict__ B, int nant, int nchan,
                                           int bchan, size_t j, double2 *out)
{
    int c = (int)(j % nchan);
    int a = (int)((j / nchan) % nant);
    int p = (int)(j / ((size_t)nchan * nant));
    const double2 *M1 = D + ((size_t)a * nchan + c) * 4;
    const double2 *M2 = B + ((size_t)p * (bchan ? nchan : nant) + (bchan ? c : a)) * 4;
    double2 Ji[4];
    Ji[0] = ca(cm(M1[0], M2[0]), cm(M1[1], M2[2]));
    Ji[1] = ca(cm(M1[0], M2[1]), cm(M1[1], M2[3]));
    Ji[2] = ca(cm(M1[2], M2[0]), cm(M1[3], M2[2]));
    Ji[3] = ca(cm(M1[2], M2[1]), cm(M1[3], M2[3]));
    double F = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) F += Ji[q].x * Ji[q].x + Ji[q].y * Ji[q].y;
    F = sqrt(F);
    if (F == 0.0) {
#pragma unroll
        for (int q = 0; q < 4; q++) out[q] = make_double2(0.0, 0.0);
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) Ji[q] = make_double2(Ji[q].x / F, Ji[q].y / F);
    double2 det = make_double2(Ji[0].x * Ji[3].x - Ji[0].y * Ji[3].y - (Ji[1].x * Ji[2].x - Ji[1].y * Ji[2].y),
                               Ji[0].x * Ji[3].y + Ji[0].y * Ji[3].x - (Ji[1].x * Ji[2].y + Ji[1].y * Ji[2].x));
    double sc = 1.0 / (det.x * det.x + det.y * det.y);
    double2 inv = make_double2(sc * det.x, -sc * det.y), ninv = make_double2(-inv.x, -inv.y);
    out[0] = cm(inv, Ji[3]);
    out[1] = cm(ninv, Ji[1]);
    out[2] = cm(ninv, Ji[2]);
    out[3] = cm(inv, Ji[0]);
}

// D[a][c][4], B[p][a][4] -> J[p][a][c][4]; with bchan, B is [p][c][4] (a
// per-fine-channel beam shared by all tiles).
// block 256.  Each lane's 2x2 result (64 bytes) goes through LDS so that
// the workgroup's 16 KB of output (consecutive j) is stored with 16-byte
// lanes side by side.
__global__ __launch_bounds__(256) void k_calc_J(const double2 *__restrict__ D,
                                                const double2 *__restrict__ B, int npointing,
                                                int nant, int nchan, int bchan,
                                                double2 *__restrict__ J)
{
    __shared__ double2 sh[256 * 4];
    const size_t j0 = (size_t)blockIdx.x * 256;
    const size_t j = j0 + threadIdx.x;
    const size_t total = (size_t)npointing * nant * nchan;
    calc_J_one(D, B, nant, nchan, bchan, j < total ? j : total - 1, &sh[threadIdx.x * 4]);
    __syncthreads();
    const size_t nout = ((total < j0 + 256 ? total : j0 + 256) - j0) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const size_t e = threadIdx.x + 256 * q;
        if (e < nout) J[j0 * 4 + e] = sh[e];
    }
}

extern "C" {

static void ensure(void **ptr, size_t *cap, size_t need)
{
    if (*cap >= need && *ptr) return;
    if (*ptr) HIPCHECK(hipFree(*ptr));
    HIPCHECK(hipMalloc(ptr, need));
    *cap = need;
}

void vm_calc_phi_launch(const void *d_pos, const double *d_unit, const double *d_freq,
                        int npointing, int nant, int nchan, void *d_phi, hipStream_t s)
{
    size_t total = (size_t)npointing * nant * nchan;
    hipLaunchKernelGGL(k_calc_phi, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (const double4 *)d_pos, d_unit, d_freq, npointing, nant, nchan,
                       (double2 *)d_phi);
    HIPCHECK(hipGetLastError());
}

void vm_calc_J_launch(const void *d_D, const void *d_B, int npointing, int nant, int nchan,
                      int bchan, void *d_J, hipStream_t s)
{
    size_t total = (size_t)npointing * nant * nchan;
    hipLaunchKernelGGL(k_calc_J, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (const double2 *)d_D, (const double2 *)d_B, npointing, nant, nchan, bchan,
                       (double2 *)d_J);
    HIPCHECK(hipGetLastError());
}

void vmCalcPhiDevice(vcsbeam_context *vm, beam_geom *bg)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    geometric_delays *g = &vm->gdelays;
    int nant = (int)g->nant, nchan = (int)g->nchan, np = (int)g->npointings;
    if (!d->d_pos) {
        // per-antenna geometry from the X-pol RF inputs (src/geometry.c:66-80)
        double *pos = (double *)calloc((size_t)nant * 4, sizeof(double));
        for (size_t i = 0; i < vm->obs_metadata->num_rf_inputs; i++) {
            Rfinput *rf = &vm->obs_metadata->rf_inputs[i];
            if (*(rf->pol) == 'Y') continue;
            pos[4 * rf->ant + 0] = rf->east_m;
            pos[4 * rf->ant + 1] = rf->north_m;
            pos[4 * rf->ant + 2] = rf->height_m - MWA_ALTITUDE_METRES;
            pos[4 * rf->ant + 3] = rf->electrical_length_m;
        }
        HIPCHECK(hipMalloc((void **)&d->d_pos, (size_t)nant * 4 * sizeof(double)));
        HIPCHECK(hipMemcpy(d->d_pos, pos, (size_t)nant * 4 * sizeof(double), hipMemcpyHostToDevice));
        free(pos);
        HIPCHECK(hipMalloc((void **)&d->d_freqs, (size_t)nchan * sizeof(double)));
        HIPCHECK(hipMemcpy(d->d_freqs, g->chan_freqs_hz, (size_t)nchan * sizeof(double),
                           hipMemcpyHostToDevice));
    }
    double *unit = (double *)malloc((size_t)np * 3 * sizeof(double));
    for (int p = 0; p < np; p++) {
        unit[3 * p + 0] = bg[p].unit_E;
        unit[3 * p + 1] = bg[p].unit_N;
        unit[3 * p + 2] = bg[p].unit_H;
    }
    ensure((void **)&d->d_unit, &d->unit_bytes, (size_t)np * 3 * sizeof(double));
    HIPCHECK(hipMemcpyAsync(d->d_unit, unit, (size_t)np * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    vm_timer_start(d, VM_T_GEOM);
    vm_calc_phi_launch(d->d_pos, d->d_unit, d->d_freqs, np, nant, nchan, g->d_phi, s);
    vm_timer_stop(d, VM_T_GEOM);
    HIPCHECK(hipStreamSynchronize(s));
    free(unit);
    d->weights_gen++;
}

// *flag |= 1 where block p (p = 1 + blockIdx.y) of `per` values differs
// from block 0 (an exact comparison; NaN counts as differing)
__global__ void k_differs_from_first(const double2 *__restrict__ x, uint32_t per,
                                     uint32_t *__restrict__ flag)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool dif = false;
    if (j < per) {
        const double2 a = x[j], b = x[(size_t)(blockIdx.y + 1) * per + j];
        dif = !(a.x == b.x && a.y == b.y);
    }
    if (__any(dif) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// blocks 1..n-1 of `per` values = block 0
__global__ void k_copy_first_block(double2 *__restrict__ x, uint32_t per)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < per) x[(size_t)(blockIdx.y + 1) * per + j] = x[j];
}

void vm_expand_J(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    if (!d->J_shared) return;
    const int np = (int)vm->npointing;
    const uint32_t per = (uint32_t)vm->obs_metadata->num_ants * vm->nfine_chan * 4;
    if (np > 1)
        hipLaunchKernelGGL(k_copy_first_block, dim3((per + 255) / 256, (unsigned)(np - 1)),
                           dim3(256), 0, (hipStream_t)d->stream, (double2 *)vm->d_J, per);
    HIPCHECK(hipGetLastError());
    d->J_shared = 0;
}

void vmExpandJDevice(vcsbeam_context *vm)
{
    vm_expand_J(vm);
    HIPCHECK(hipStreamSynchronize((hipStream_t)vm_dev(vm)->stream));
}

void vmCalcJDevice(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    int nant = (int)vm->obs_metadata->num_ants, nchan = vm->nfine_chan, np = (int)vm->npointing;
    size_t bbytes = (size_t)np * nant * 4 * sizeof(gpuDoubleComplex);
    ensure(&d->d_B, &d->B_bytes, bbytes);
    HIPCHECK(hipMemcpyAsync(d->d_B, vm->pb.B, bbytes, hipMemcpyHostToDevice, s));
    // every pointing with the same beam (B = I, or one beam for all) gives
    // every pointing the same J^-1: compute pointing 0's only and share it
    // (vm_expand_J writes the others when something reads them)
    uint32_t differs = 1;
    if (np > 1) {
        ensure(&d->d_flag, &d->flag_bytes, 256);
        HIPCHECK(hipMemsetAsync(d->d_flag, 0, 4, s));
        const uint32_t per = (uint32_t)nant * 4;
        hipLaunchKernelGGL(k_differs_from_first, dim3((per + 255) / 256, (unsigned)(np - 1)),
                           dim3(256), 0, s, (const double2 *)d->d_B, per, (uint32_t *)d->d_flag);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(&differs, d->d_flag, 4, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
    }
    const int shared = np > 1 && !differs;
    vm_timer_start(d, VM_T_GEOM);
    vm_calc_J_launch(vm->d_D, d->d_B, shared ? 1 : np, nant, nchan, 0, vm->d_J, s);
    vm_timer_stop(d, VM_T_GEOM);
    d->J_shared = shared;
    d->B_chan = 0;
    d->J_fromB = 1;
    d->B_np = np;
    HIPCHECK(hipStreamSynchronize(s));
    d->weights_gen++;
}

void vmCalcJDeviceChanB(vcsbeam_context *vm, const gpuDoubleComplex *B)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    int nant = (int)vm->obs_metadata->num_ants, nchan = vm->nfine_chan, np = (int)vm->npointing;
    size_t bbytes = (size_t)np * nchan * 4 * sizeof(gpuDoubleComplex);
    ensure(&d->d_B, &d->B_bytes, bbytes);
    HIPCHECK(hipMemcpyAsync(d->d_B, B, bbytes, hipMemcpyHostToDevice, s));
    vm_calc_J_launch(vm->d_D, d->d_B, np, nant, nchan, 1, vm->d_J, s);
    d->J_shared = 0;
    d->B_chan = 1;
    d->J_fromB = 1;
    d->B_np = np;
    HIPCHECK(hipStreamSynchronize(s));
    d->weights_gen++;
}

} // extern "C"
