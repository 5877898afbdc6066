// vm_device.hip -- device-side context management: the vmMalloc*/vmFree*
// family, host<->device pushes, streams, statistics buffers and context
// teardown.
//
// Reference: MWA-vcsbeam src/metadata.c:448-1076 (allocation, push, streams),
// src/metadata.c:323-377 (destroy_vcsbeam_context), src/form_beam.cpp:679-804.
//
// Every context kernel runs on one HIP stream per context (vm->dev->stream);
// entry points whose outputs the host reads synchronise before returning.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define HIPCHECK(x)                                                          \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) VM_FATAL("HIP: %s (%s)", hipGetErrorString(e_), #x); \
    } while (0)

extern "C" {

void *vm_host_alloc_pinned(size_t bytes)
{
    void *p = NULL;
    HIPCHECK(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
    return p;
}

void vm_host_free_pinned(void *p)
{
    if (p) HIPCHECK(hipHostFree(p));
}

static void *dmalloc(size_t bytes)
{
    void *p = NULL;
    HIPCHECK(hipMalloc(&p, bytes ? bytes : 1));
    return p;
}

static void dfree(void *p)
{
    if (p) HIPCHECK(hipFree(p));
}

vm_device_state *vm_dev(vcsbeam_context *vm)
{
    if (vm->dev) return vm->dev;
    vm_device_state *d = (vm_device_state *)calloc(1, sizeof(vm_device_state));
    if (!d) VM_FATAL("out of host memory");
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    d->stream = (void *)s;
    d->weights_gen++;
    vm->dev = d;
    return d;
}

void vm_dev_free(vcsbeam_context *vm)
{
    vm_device_state *d = vm->dev;
    if (!d) return;
    HIPCHECK(hipStreamSynchronize((hipStream_t)d->stream));
    dfree(d->d_W);
    dfree(d->d_Wm);
    dfree(d->d_vscale);
    dfree(d->d_stage);
    dfree(d->d_pos);
    dfree(d->d_freqs);
    dfree(d->d_B);
    dfree(d->d_grp);
    dfree(d->d_Bg);
    dfree(d->d_gx);
    dfree(d->d_unit);
    dfree(d->d_ipfb_out);
    dfree(d->d_ipfb_jobs);
    dfree(d->d_nterm);
    dfree(d->d_flag);
    vm_timer_free(d);
    HIPCHECK(hipStreamDestroy((hipStream_t)d->stream));
    free(d);
    vm->dev = NULL;
}

gpuStream_t vmGetStream(vcsbeam_context *vm) { return (gpuStream_t)vm_dev(vm)->stream; }

void vmSynchronize(vcsbeam_context *vm)
{
    HIPCHECK(hipStreamSynchronize((hipStream_t)vm_dev(vm)->stream));
}

/* ---------------------------------------------------------- host mallocs */
void vmMallocJVHost(vcsbeam_context *vm)
{
    /* J^-1 v is never materialised (fused into the beamformer); the size is
     * still recorded as in src/metadata.c:464-473 */
    vm->Jv_size_bytes = (uintptr_t)vm->npointing * vm->obs_metadata->num_ants *
                        vm->nfine_chan * vm->fine_sample_rate * sizeof(gpuDoubleComplex);
    vm->Jv_P = vm->Jv_Q = NULL;
}
void vmFreeJVHost(vcsbeam_context *vm) { (void)vm; }

void vmMallocEHost(vcsbeam_context *vm)
{
    vm->e_size_bytes = (uintptr_t)vm->npointing * vm->fine_sample_rate * vm->nfine_chan *
                       vm->obs_metadata->num_ant_pols * sizeof(gpuDoubleComplex);
    vm->e = (gpuDoubleComplex *)vm_host_alloc_pinned(vm->e_size_bytes);
}
void vmFreeEHost(vcsbeam_context *vm) { vm_host_free_pinned(vm->e); vm->e = NULL; }

void vmMallocSHost(vcsbeam_context *vm)
{
    vm->S_size_bytes = (uintptr_t)vm->npointing * vm->nfine_chan * vm->out_nstokes *
                       vm->fine_sample_rate * sizeof(float);
    vm->S = vm_host_alloc_pinned(vm->S_size_bytes);
}
void vmFreeSHost(vcsbeam_context *vm) { vm_host_free_pinned(vm->S); vm->S = NULL; }

void vmMallocJHost(vcsbeam_context *vm)
{
    vm->J_size_bytes = (uintptr_t)vm->npointing * vm->obs_metadata->num_ants * vm->nfine_chan *
                       vm->obs_metadata->num_visibility_pols * sizeof(gpuDoubleComplex);
    vm->J = (gpuDoubleComplex *)vm_host_alloc_pinned(vm->J_size_bytes);
    memset(vm->J, 0, vm->J_size_bytes);
}
void vmFreeJHost(vcsbeam_context *vm) { vm_host_free_pinned(vm->J); vm->J = NULL; }

void vmMallocDHost(vcsbeam_context *vm)
{
    vm->D_size_bytes = (uintptr_t)vm->obs_metadata->num_ants * vm->nfine_chan *
                       vm->obs_metadata->num_visibility_pols * sizeof(gpuDoubleComplex);
    vm->D = (gpuDoubleComplex *)vm_host_alloc_pinned(vm->D_size_bytes);
    memset(vm->D, 0, vm->D_size_bytes);
}
void vmFreeDHost(vcsbeam_context *vm) { vm_host_free_pinned(vm->D); vm->D = NULL; }

void vmMallocPQIdxsHost(vcsbeam_context *vm)
{
    vm->pol_idxs_size_bytes = vm->obs_metadata->num_ants * sizeof(uint32_t);
    vm->polP_idxs = (uint32_t *)vm_host_alloc_pinned(vm->pol_idxs_size_bytes);
    vm->polQ_idxs = (uint32_t *)vm_host_alloc_pinned(vm->pol_idxs_size_bytes);
}
void vmFreePQIdxsHost(vcsbeam_context *vm)
{
    vm_host_free_pinned(vm->polP_idxs);
    vm_host_free_pinned(vm->polQ_idxs);
    vm->polP_idxs = vm->polQ_idxs = NULL;
}

/* -------------------------------------------------------- device mallocs */
void vmMallocVDevice(vcsbeam_context *vm)
{
    /* src/metadata.c:647-660: legacy input gets its own chunk buffer; MWAX
     * input aliases the forward PFB's output */
    if (vm->obs_metadata->mwa_version == VCSLegacyRecombined) {
        vm->d_v_size_bytes = vm->bytes_per_second / vm->chunks_per_second;
        vm->d_v = dmalloc(vm->d_v_size_bytes);
    } else {
        if (!vm->fpfb) VM_FATAL("vmMallocVDevice: forward PFB not initialised");
        vm->d_v = vm->fpfb->d_vcs_data;
        vm->d_v_size_bytes = vm->fpfb->d_vcs_size;
    }
}
void vmFreeVDevice(vcsbeam_context *vm)
{
    if (vm->obs_metadata->mwa_version == VCSLegacyRecombined) dfree(vm->d_v);
    vm->d_v = NULL;
}

void vmMallocJVDevice(vcsbeam_context *vm)
{
    vm->d_Jv_size_bytes = (uintptr_t)vm->npointing * vm->obs_metadata->num_ants *
                          vm->nfine_chan * vm->fine_sample_rate *
                          sizeof(gpuDoubleComplex) / vm->chunks_per_second;
    vm->d_Jv_P = vm->d_Jv_Q = NULL; /* fused: never allocated */
}
void vmFreeJVDevice(vcsbeam_context *vm) { (void)vm; }

void vmMallocEDevice(vcsbeam_context *vm)
{
    /* device layout [p][c][s][pol] complex float (half the reference's bytes) */
    vm->d_e_size_bytes = (uintptr_t)vm->npointing * vm->fine_sample_rate * vm->nfine_chan *
                         vm->obs_metadata->num_ant_pols * sizeof(gpuFloatComplex);
    vm->d_e = (gpuDoubleComplex *)dmalloc(vm->d_e_size_bytes);
}
void vmFreeEDevice(vcsbeam_context *vm) { dfree(vm->d_e); vm->d_e = NULL; }

void vmMallocSDevice(vcsbeam_context *vm)
{
    /* blocked layout, the last sample block padded to VM_SBLK */
    vm->d_S_size_bytes = (uintptr_t)vm->npointing * vm->nfine_chan * vm->out_nstokes *
                         VM_S_NBLK(vm->fine_sample_rate) * VM_SBLK * sizeof(float);
    vm->d_S = dmalloc(vm->d_S_size_bytes);
}
void vmFreeSDevice(vcsbeam_context *vm) { dfree(vm->d_S); vm->d_S = NULL; }

void vmMallocJDevice(vcsbeam_context *vm)
{
    vm->d_J_size_bytes = (uintptr_t)vm->npointing * vm->obs_metadata->num_ants * vm->nfine_chan *
                         vm->obs_metadata->num_visibility_pols * sizeof(gpuDoubleComplex);
    vm->d_J = (gpuDoubleComplex *)dmalloc(vm->d_J_size_bytes);
    vm_dev(vm)->J_fromB = 0; /* not computed from d_B yet */
}
void vmFreeJDevice(vcsbeam_context *vm) { dfree(vm->d_J); vm->d_J = NULL; }

void vmMallocDDevice(vcsbeam_context *vm)
{
    vm->d_D_size_bytes = (uintptr_t)vm->obs_metadata->num_ants * vm->nfine_chan *
                         vm->obs_metadata->num_visibility_pols * sizeof(gpuDoubleComplex);
    vm->d_D = (gpuDoubleComplex *)dmalloc(vm->d_D_size_bytes);
}
void vmFreeDDevice(vcsbeam_context *vm) { dfree(vm->d_D); vm->d_D = NULL; }

void vmMallocPQIdxsDevice(vcsbeam_context *vm)
{
    vm->d_pol_idxs_size_bytes = vm->obs_metadata->num_ants * sizeof(uint32_t);
    vm->d_polP_idxs = (uint32_t *)dmalloc(vm->d_pol_idxs_size_bytes);
    vm->d_polQ_idxs = (uint32_t *)dmalloc(vm->d_pol_idxs_size_bytes);
}
void vmFreePQIdxsDevice(vcsbeam_context *vm)
{
    dfree(vm->d_polP_idxs);
    dfree(vm->d_polQ_idxs);
    vm->d_polP_idxs = vm->d_polQ_idxs = NULL;
}

/* ------------------------------------------------------------ set-up ----- */
void vmSetMaxGPUMem(vcsbeam_context *vm, int nchunks)
{
    /* src/metadata.c:892-906 */
    vm->chunks_per_second = nchunks;
    if (nchunks <= 0 || vm->sample_rate % vm->chunks_per_second != 0) {
        logger_timed_message(vm->log, "Bad number of chunks requested: must divide "
                                      "number of samples per second of data.");
        VM_FATAL("bad number of chunks (%d)", nchunks);
    }
    vm->d_v_size_bytes = vm->bytes_per_second / vm->chunks_per_second;
    vm->d_Jv_size_bytes = vm->Jv_size_bytes / vm->chunks_per_second;
}

void vmPushChunk(vcsbeam_context *vm)
{
    /* src/metadata.c:913-923 (legacy input) */
    logger_start_stopwatch(vm->log, "upload", false);
    int chunk = vm->chunk_to_load % vm->chunks_per_second;
    const char *src = (const char *)vm->v->buffer + (size_t)chunk * vm->d_v_size_bytes;
    hipStream_t s = (hipStream_t)vm_dev(vm)->stream;
    HIPCHECK(hipMemcpyAsync(vm->d_v, src, vm->d_v_size_bytes, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    logger_stop_stopwatch(vm->log, "upload");
}

void vmPushJ(vcsbeam_context *vm)
{
    vm_device_state *d = vm_dev(vm);
    hipStream_t s = (hipStream_t)d->stream;
    HIPCHECK(hipMemcpyAsync(vm->d_J, vm->J, vm->J_size_bytes, hipMemcpyHostToDevice, s));
    d->J_fromB = 0;
    d->J_shared = 0;
    if (d->J_host_fromB && vm->pb.B && vm->D && vm->d_D &&
        vm->pb.npointings == vm->npointing && vm->pb.nant == vm->obs_metadata->num_ants) {
        /* J^-1 came from vmCalcJ, i.e. from D and the primary beam pb.B:
         * hand both to the device as well, so that the precision form can
         * factor J^-1 = F B^-1 D^-1 when the beam differs between tiles */
        const size_t bbytes = (size_t)vm->npointing * vm->pb.nant * 4 * sizeof(gpuDoubleComplex);
        if (d->B_bytes < bbytes) {
            dfree(d->d_B);
            d->d_B = dmalloc(bbytes);
            d->B_bytes = bbytes;
        }
        HIPCHECK(hipMemcpyAsync(d->d_B, vm->pb.B, bbytes, hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(vm->d_D, vm->D, vm->D_size_bytes, hipMemcpyHostToDevice, s));
        d->B_chan = 0;
        d->B_np = (int)vm->npointing;
        d->J_fromB = 1;
    }
    d->J_host_fromB = 0;
    HIPCHECK(hipStreamSynchronize(s));
    d->weights_gen++;
}

void vmPushD(vcsbeam_context *vm)
{
    hipStream_t s = (hipStream_t)vm_dev(vm)->stream;
    HIPCHECK(hipMemcpyAsync(vm->d_D, vm->D, vm->D_size_bytes, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
}

void vmPushPolIdxLists(vcsbeam_context *vm)
{
    hipStream_t s = (hipStream_t)vm_dev(vm)->stream;
    HIPCHECK(hipMemcpyAsync(vm->d_polQ_idxs, vm->polQ_idxs, vm->pol_idxs_size_bytes,
                            hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(vm->d_polP_idxs, vm->polP_idxs, vm->pol_idxs_size_bytes,
                            hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    vm->dev->weights_gen++;
}

void vmCreateGeometricDelays(vcsbeam_context *vm)
{
    /* src/geometry.c:142-171 */
    geometric_delays *g = &vm->gdelays;
    g->npointings = vm->npointing;
    g->nant = vm->obs_metadata->num_ants;
    g->nchan = vm->nfine_chan;
    g->obs_metadata = vm->obs_metadata;
    g->chan_freqs_hz = (double *)malloc(sizeof(double) * vm->nfine_chan);
    int cc = vm->coarse_chan_idxs_to_process[0];
    CoarseChannel C = vm->obs_metadata->metafits_coarse_chans[cc];
    uint32_t fine_chan_width = vm->obs_metadata->coarse_chan_width_hz / vm->nfine_chan;
    for (int c = 0; c < vm->nfine_chan; c++)
        g->chan_freqs_hz[c] = (double)C.chan_start_hz + (double)fine_chan_width * c;
    size_t size = g->npointings * g->nant * g->nchan * sizeof(gpuDoubleComplex);
    g->phi = (gpuDoubleComplex *)vm_host_alloc_pinned(size);
    memset(g->phi, 0, size);
    g->d_phi = (gpuDoubleComplex *)dmalloc(size);
}

void free_geometric_delays(geometric_delays *g)
{
    free(g->chan_freqs_hz);
    vm_host_free_pinned(g->phi);
    dfree(g->d_phi);
    g->chan_freqs_hz = NULL;
    g->phi = g->d_phi = NULL;
}

void vmPushPhi(vcsbeam_context *vm)
{
    geometric_delays *g = &vm->gdelays;
    size_t size = g->npointings * g->nant * g->nchan * sizeof(gpuDoubleComplex);
    hipStream_t s = (hipStream_t)vm_dev(vm)->stream;
    HIPCHECK(hipMemcpyAsync(g->d_phi, g->phi, size, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    vm->dev->weights_gen++;
}

void vmCreatePrimaryBeam(vcsbeam_context *vm)
{
    /* src/primary_beam.c:200-229 allocates B [p][a][npol*npol] and takes
     * the dipole delays and amplitudes from the metadata, at the
     * coarse-channel centre frequency.  B starts as the identity: vmCalcB
     * replaces it from the dipole configurations (when the metadata has
     * dipole delays), or the caller may overwrite it */
    primary_beam *pb = &vm->pb;
    pb->npointings = vm->npointing;
    pb->nant = vm->obs_metadata->num_ants;
    pb->npol = vm->obs_metadata->num_visibility_pols;
    pb->obs_metadata = vm->obs_metadata;
    pb->beam = NULL;
    pb->delays = NULL;
    pb->amps = NULL;
    if (vm->obs_metadata->num_delays > 0 && vm->obs_metadata->num_rf_inputs > 0 &&
        vm->obs_metadata->rf_inputs[0].dipole_delays)
        create_delays_amps_from_metafits(vm->obs_metadata, &pb->delays, &pb->amps);
    pb->freq_hz = (vm->obs_metadata->metafits_coarse_chans && vm->coarse_chan_idx >= 0 &&
                   (size_t)vm->coarse_chan_idx < vm->obs_metadata->num_metafits_coarse_chans)
                      ? vm->obs_metadata->metafits_coarse_chans[vm->coarse_chan_idx].chan_centre_hz
                      : 0;
    size_t n = pb->npointings * pb->nant * pb->npol;
    /* pinned, as it is copied to the device every second (vmCalcJDevice) */
    pb->B = (gpuDoubleComplex *)vm_host_alloc_pinned(n * sizeof(gpuDoubleComplex));
    memset(pb->B, 0, n * sizeof(gpuDoubleComplex));
    for (size_t j = 0; j < pb->npointings * pb->nant; j++) {
        pb->B[4 * j + 0].x = 1.0;
        pb->B[4 * j + 3].x = 1.0;
    }
}

void free_primary_beam(primary_beam *pb)
{
    /* src/primary_beam.c (free_primary_beam): releases B and the dipole
     * delays and amplitudes */
    if (!pb) return;
    vm_host_free_pinned(pb->B);
    pb->B = NULL;
    if (pb->amps) free_delays_amps(pb->obs_metadata, pb->delays, pb->amps);
    pb->delays = NULL;
    pb->amps = NULL;
}

void vmCreateCudaStreams(vcsbeam_context *vm)
{
    /* src/metadata.c:1023-1032.  The streams are created for callers that
     * use them; this library issues all work on vm->dev->stream. */
    vm->streams = (gpuStream_t *)malloc(sizeof(gpuStream_t) * (vm->npointing ? vm->npointing : 1));
    for (unsigned int p = 0; p < vm->npointing; p++) {
        hipStream_t s;
        HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        vm->streams[p] = (gpuStream_t)s;
    }
}

void vmDestroyCudaStreams(vcsbeam_context *vm)
{
    if (!vm->streams) return;
    for (unsigned int p = 0; p < vm->npointing; p++)
        HIPCHECK(hipStreamDestroy((hipStream_t)vm->streams[p]));
    free(vm->streams);
    vm->streams = NULL;
}

void vmCreateStatistics(vcsbeam_context *vm, mpi_psrfits *mpfs)
{
    /* src/metadata.c:1061-1076; the PSRFITS sub-integration size is
     * nchan * nstokes * nsamples bytes */
    size_t nchan = vm->nfine_chan;
    vm->offsets_size = vm->npointing * nchan * vm->out_nstokes * sizeof(float);
    vm->scales_size = vm->npointing * nchan * vm->out_nstokes * sizeof(float);
    size_t per = (mpfs && mpfs[0].coarse_chan_pf.sub.bytes_per_subint > 0)
                     ? (size_t)mpfs[0].coarse_chan_pf.sub.bytes_per_subint
                     : nchan * vm->out_nstokes * (size_t)vm->fine_sample_rate;
    vm->Cscaled_size = vm->npointing * per;
    vm->d_offsets = (float *)dmalloc(vm->offsets_size);
    vm->d_scales = (float *)dmalloc(vm->scales_size);
    vm->d_Cscaled = (uint8_t *)dmalloc(vm->Cscaled_size);
    vm->offsets = (float *)vm_host_alloc_pinned(vm->offsets_size);
    vm->scales = (float *)vm_host_alloc_pinned(vm->scales_size);
    vm->Cscaled = (uint8_t *)vm_host_alloc_pinned(vm->Cscaled_size);
}

void vmDestroyStatistics(vcsbeam_context *vm)
{
    vm_host_free_pinned(vm->offsets);
    vm_host_free_pinned(vm->scales);
    vm_host_free_pinned(vm->Cscaled);
    dfree(vm->d_offsets);
    dfree(vm->d_scales);
    dfree(vm->d_Cscaled);
    vm->offsets = vm->scales = vm->d_offsets = vm->d_scales = NULL;
    vm->Cscaled = vm->d_Cscaled = NULL;
}

float *create_pinned_data_buffer(size_t size)
{
    float *p = (float *)vm_host_alloc_pinned(size);
    memset(p, 0, size);
    return p;
}

gpuDoubleComplex *create_data_buffer_fine(int npointing, int nsamples, int nchan, int npol)
{
    size_t n = (size_t)npointing * nsamples * nchan * npol * sizeof(gpuDoubleComplex);
    gpuDoubleComplex *p = (gpuDoubleComplex *)malloc(n);
    if (!p) VM_FATAL("out of host memory");
    memset(p, 0, n);
    return p;
}

void allocate_input_output_arrays(void **data, void **d_data, size_t size)
{
    *data = vm_host_alloc_pinned(size);
    *d_data = dmalloc(size);
}

void free_input_output_arrays(void *data, void *d_data)
{
    vm_host_free_pinned(data);
    dfree(d_data);
}

/* ---------------------------------------------- device-memory helpers ---- */
void *vmDeviceMalloc(size_t bytes) { return dmalloc(bytes); }
void vmDeviceFree(void *p) { dfree(p); }
void vmMemcpyHtoD(void *dst, const void *src, size_t bytes)
{
    HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
void vmMemcpyDtoH(void *dst, const void *src, size_t bytes)
{
    HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
void vmMemcpy2DDtoH(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height)
{
    HIPCHECK(hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost));
}
void vmMemsetDevice(void *dst, int value, size_t bytes)
{
    /* complete before returning: later work may run on a non-blocking stream */
    HIPCHECK(hipMemset(dst, value, bytes));
    HIPCHECK(hipDeviceSynchronize());
}
void vmDeviceSynchronize(void) { HIPCHECK(hipDeviceSynchronize()); }
int vmDeviceCount(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
void *vmEventCreate(void)
{
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return (void *)e;
}
void vmEventRecord(void *ev, gpuStream_t stream)
{
    HIPCHECK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
}
float vmEventElapsedMs(void *start, void *stop)
{
    float ms = 0.f;
    HIPCHECK(hipEventSynchronize((hipEvent_t)stop));
    HIPCHECK(hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop));
    return ms;
}
void vmEventDestroy(void *ev) { HIPCHECK(hipEventDestroy((hipEvent_t)ev)); }

void destroy_vcsbeam_context(vcsbeam_context *vm)
{
    /* src/metadata.c:323-377: frees what the context itself owns */
    if (!vm) return;
    if (vm->fpfb) vmFreeForwardPFB(vm->fpfb);
    vm->fpfb = NULL;
    if (vm->v) vmFreeVHost(vm);
    free_pfb_filter(vm->analysis_filter);
    free_pfb_filter(vm->synth_filter);
    free(vm->gps_seconds_to_process);
    free(vm->coarse_chan_idxs_to_process);
    free(vm->cal_coarse_chan_idxs_to_process);
    free(vm->ras_hours);
    free(vm->decs_degs);
    vm_free_obs_contexts(vm);
    free(vm->datadir);
    free_calibration(&vm->cal);
    if (vm->vf) {
        for (unsigned p = 0; p < vm->npointing; p++) {
            free(vm->vf[p].b_scales);
            free(vm->vf[p].b_offsets);
        }
        free(vm->vf);
    }
    vm_dev_free(vm);
    destroy_logger(vm->log);
    free(vm);
}

} // extern "C"

/* ------------------------------------------------------ kernel timing ---- */
extern "C" {

static const char *vm_timer_names[VM_T_NSLOT] = {
    "pfb",      "beamform_fp32",    "beamform_mfma", "beamform_prep", "ipfb",
    "geometry", "beamform_weights", "incoh",         "renormalise",   "vdif"};

/* timing state of the calls that take no context (their own streams) */
static vm_device_state g_dev_global;
vm_device_state *vm_dev_global(void) { return &g_dev_global; }

// The timers record events and do not wait for them, so timing adds no
// host synchronisation to the timed calls: each slot keeps a ring of
// VM_T_RING (start, stop) pairs, resolved when the times are read (or when
// the ring is full).
static void vm_timer_flush_slot(vm_device_state *d, int slot)
{
    for (int k = 0; k < d->t_pend[slot]; k++) {
        HIPCHECK(hipEventSynchronize((hipEvent_t)d->ev[slot][k][1]));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, (hipEvent_t)d->ev[slot][k][0],
                                     (hipEvent_t)d->ev[slot][k][1]));
        d->t_ms[slot] += ms;
        d->t_n[slot] += 1;
    }
    d->t_pend[slot] = 0;
}

void vm_timer_flush(vm_device_state *d)
{
    for (int i = 0; i < VM_T_NSLOT; i++) vm_timer_flush_slot(d, i);
}

void vm_timer_free(vm_device_state *d)
{
    for (int i = 0; i < VM_T_NSLOT; i++)
        for (int k = 0; k < VM_T_RING; k++)
            for (int e = 0; e < 2; e++)
                if (d->ev[i][k][e]) {
                    HIPCHECK(hipEventDestroy((hipEvent_t)d->ev[i][k][e]));
                    d->ev[i][k][e] = NULL;
                }
}

void vm_timer_start_on(vm_device_state *d, int slot, void *stream)
{
    if (!d->timing) return;
    if (d->t_pend[slot] == VM_T_RING) vm_timer_flush_slot(d, slot);
    void **e = d->ev[slot][d->t_pend[slot]];
    if (!e[0]) {
        HIPCHECK(hipEventCreate((hipEvent_t *)&e[0]));
        HIPCHECK(hipEventCreate((hipEvent_t *)&e[1]));
    }
    HIPCHECK(hipEventRecord((hipEvent_t)e[0], (hipStream_t)stream));
    d->t_open[slot] = 1;
}

void vm_timer_start(vm_device_state *d, int slot) { vm_timer_start_on(d, slot, d->stream); }

void vm_timer_stop(vm_device_state *d, int slot) { vm_timer_stop_on(d, slot, d->stream); }

void vm_timer_stop_on(vm_device_state *d, int slot, void *stream)
{
    if (!d->timing || !d->t_open[slot]) return;
    HIPCHECK(hipEventRecord((hipEvent_t)d->ev[slot][d->t_pend[slot]][1], (hipStream_t)stream));
    d->t_pend[slot]++;
    d->t_open[slot] = 0;
}

void vmSetKernelTiming(vcsbeam_context *vm, int on)
{
    vm_device_state *d = vm ? vm_dev(vm) : vm_dev_global();
    vm_timer_flush(d);
    for (int i = 0; i < VM_T_NSLOT; i++) d->t_open[i] = 0;
    d->timing = on;
    if (!on) return; /* keep the counts for vmKernelTimeMs */
    for (int i = 0; i < VM_T_NSLOT; i++) {
        d->t_ms[i] = 0.0;
        d->t_n[i] = 0;
    }
}

double vmKernelTimeMs(vcsbeam_context *vm, const char *kernel, long *count)
{
    vm_device_state *d = vm ? vm_dev(vm) : vm_dev_global();
    vm_timer_flush(d);
    for (int i = 0; i < VM_T_NSLOT; i++)
        if (strcmp(kernel, vm_timer_names[i]) == 0) {
            if (count) *count = d->t_n[i];
            return d->t_n[i] ? d->t_ms[i] / d->t_n[i] : 0.0;
        }
    if (count) *count = 0;
    return -1.0;
}

} // extern "C"

/* ------------------------------------------------- plain streaming copy ---- */
__global__ void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" void vmDeviceCopy(void *dst, const void *src, size_t bytes)
{
    /* bytes must be a multiple of 16; a plain HBM-bound copy (used as the
     * reference kernel when calibrating memory counters) */
    if (bytes % 16) VM_FATAL("vmDeviceCopy: size %zu is not a multiple of 16", bytes);
    size_t n = bytes / 16;
    if (!n) return;
    // up to 65536 workgroups of 256 lanes, a few 16-byte elements per lane:
    // the fastest plain copy measured on MI355X (5.1 TB/s read + write, where
    // 4096 workgroups with one element in flight per lane reach 4.4-4.6)
    const size_t nwg = (n + 255) / 256;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)(nwg < 65536 ? nwg : 65536)), dim3(256), 0, 0,
                       (const uint4 *)src, (uint4 *)dst, n);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
}

/* ------------------------------------------------ dense MFMA issue rate ---- */
typedef _Float16 mp_h8 __attribute__((ext_vector_type(8)));
typedef float mp_f16 __attribute__((ext_vector_type(16)));
#define MP_CHAINS 4 /* independent accumulations per wavefront */
#define MP_ITERS 4096

/* MP_CHAINS chains of v_mfma_f32_32x32x16_f16 per wavefront on operands held
 * in registers: the dense FP16 MFMA rate with no memory traffic.  The sums
 * are stored only where they equal a value they cannot reach (operands of
 * at most 0.008, so |sum| < 1), which keeps the work from being removed. */
__global__ __launch_bounds__(256) void k_mfma_peak(float *out)
{
    mp_h8 a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (_Float16)(0.001f * (float)((threadIdx.x + i) & 7));
        b[i] = (_Float16)(0.001f * (float)((threadIdx.x * 3 + i) & 7));
    }
    mp_f16 acc[MP_CHAINS];
#pragma unroll
    for (int c = 0; c < MP_CHAINS; c++)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[c][e] = 0.f;
    for (int it = 0; it < MP_ITERS; it++) {
#pragma unroll
        for (int c = 0; c < MP_CHAINS; c++)
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[c], 0, 0, 0);
    }
    float sum = 0.f;
#pragma unroll
    for (int c = 0; c < MP_CHAINS; c++)
#pragma unroll
        for (int e = 0; e < 16; e++) sum += acc[c][e];
    if (sum == 1.0e6f) out[threadIdx.x] = sum;
}

extern "C" double vmMFMAPeakTFLOPS(void)
{
    /* dense FP16 MFMA rate of this device, measured: 8 workgroups of 4
     * wavefronts per CU, ten timed launches after five warm-up launches (the
     * clock rises under a sustained matrix load) */
    int dev = 0, ncu = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    float *out = nullptr;
    HIPCHECK(hipMalloc(&out, 256 * sizeof(float)));
    const int nwg = 8 * ncu, reps = 10;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k_mfma_peak, dim3(nwg), dim3(256), 0, 0, out);
    HIPCHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_mfma_peak, dim3(nwg), dim3(256), 0, 0, out);
    HIPCHECK(hipEventRecord(e1, 0));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    HIPCHECK(hipFree(out));
    const double flops = (double)reps * nwg * 4 /* wavefronts */ * MP_ITERS * MP_CHAINS *
                         (32.0 * 32.0 * 16.0 * 2.0);
    return flops / (ms * 1e-3) / 1e12;
}

/* ---------------------------------------------- seeded synthetic voltages ---- */
/* Counter-based generator: element idx of stream `seed` is the SplitMix64
 * finaliser of seed * 0xD1B54A32D192ED03 + (idx + 1) * 0x9E3779B97F4A7C15;
 * uniforms take its top 53 bits; a Gaussian pair is Box-Muller on uniforms
 * 2 pair and 2 pair + 1 (r = sqrt(-2 ln(1 - u1)), angle 2 pi u2). */
__device__ __forceinline__ uint64_t synth_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double synth_unif(uint64_t seed, uint64_t idx)
{
    const uint64_t z = synth_mix(seed * 0xD1B54A32D192ED03ULL + (idx + 1) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ int8_t synth_q8(double x)
{
    double r = round(x);
    r = r > 127.0 ? 127.0 : (r < -127.0 ? -127.0 : r);
    return (int8_t)r;
}

// out: [nblocks][ninput][spb] char2; block b draws pair ((first + b) I + i) spb + t
__global__ void k_synth_mwax(uint64_t seed, uint64_t first_block, int nblocks, int ninput, int spb,
                             double sigma, int zero_blocks, uint16_t *__restrict__ out)
{
    const size_t per_block = (size_t)ninput * spb, n = per_block * nblocks;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const size_t b = j / per_block;
        uint16_t v = 0;
        if ((long)b >= zero_blocks) {
            const uint64_t pair = (first_block + b) * per_block + (j - b * per_block);
            const double u1 = synth_unif(seed, 2 * pair), u2 = synth_unif(seed, 2 * pair + 1);
            const double r = sqrt(-2.0 * log(1.0 - u1)), t = 2.0 * 3.141592653589793 * u2;
            const int8_t re = synth_q8(sigma * (r * cos(t))), im = synth_q8(sigma * (r * sin(t)));
            v = (uint16_t)(uint8_t)re | ((uint16_t)(uint8_t)im << 8);
        }
        out[j] = v;
    }
}

extern "C" void vmSynthMWAXDevice(uint64_t seed, uint64_t first_block, int nblocks, int ninput,
                                  int spb, double sigma, int zero_blocks, void *d_out)
{
    if (nblocks < 0 || ninput < 0 || spb < 0)
        VM_FATAL("vmSynthMWAXDevice: negative size (%d blocks, %d inputs, %d samples)", nblocks,
                 ninput, spb);
    hipLaunchKernelGGL(k_synth_mwax, dim3(8192), dim3(256), 0, 0, seed, first_block, nblocks,
                       ninput, spb, sigma, zero_blocks, (uint16_t *)d_out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
}
