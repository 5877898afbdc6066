/*
 * vm_internal.h -- library-private declarations shared by the host C module
 * and the HIP translation units.  Not part of the C-ABI.
 */
#ifndef VM_INTERNAL_H
#define VM_INTERNAL_H

#include "vcsbeam.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fatal error: print and exit, like the reference's gpuAssert / GPU_CHECK_ERROR
 * (src/pfb.cpp:19-33, src/gpu_macros.h:27-35). */
void vm_fatal(const char *file, int line, const char *fmt, ...);
#define VM_FATAL(...) vm_fatal(__FILE__, __LINE__, __VA_ARGS__)

/* releases the metadata vmLoadObsMetafits / vmLoadCalMetafits loaded and the
 * file reader vmGetVoltageMetadata set up (obsfiles.c) */
void vm_free_obs_contexts(vcsbeam_context *vm);

/* Devices a process may use (per-device tables of library state). */
#define VM_MAX_DEVICES 64

/* Device Stokes layout d_S: [c][s / VM_SBLK][p][st][s % VM_SBLK] float, the
   second's samples cut into blocks of VM_SBLK, the last block padded.  A
   tile of pointings x VM_SBLK samples is then one contiguous run of memory
   (the beamformers write whole runs; HBM takes them at near its sequential
   rate).  VM_S_IDX gives the element of (p, st, c, s) for np pointings, nst
   Stokes and nblk = VM_S_NBLK(samples per second) blocks. */
#define VM_SBLK 64
#define VM_S_NBLK(ns) (((ns) + VM_SBLK - 1) / VM_SBLK)
#define VM_S_IDX(p, st, c, s, np, nst, nblk)                                              \
    (((((size_t)(c) * (size_t)(nblk) + (size_t)((s) / VM_SBLK)) * (size_t)(np) + (size_t)(p)) * \
          (size_t)(nst) +                                                                  \
      (size_t)(st)) *                                                                      \
         VM_SBLK +                                                                         \
     (size_t)((s) % VM_SBLK))

/* Device-side state hung off vcsbeam_context.dev (created lazily). */
#define VM_T_RING 32 /* timed launches per slot between two reads */
struct vm_device_state {
    void *stream;             /* hipStream_t on which all context work runs */
    /* FP32 beamformer weights W = phi * J^-1, complex float
       [p][c][a][2][2] */
    void *d_W;
    size_t W_bytes;
    /* bumped whenever J, D, phi or the pol index lists change; each set of
       derived weights remembers the generation it was built from */
    unsigned weights_gen;
    unsigned W_gen;  /* FP32 weights d_W */
    unsigned Wm_gen; /* MFMA weights in d_Wm */
    int Wm_form;     /* beamform_form (FP16 / BF16) the MFMA weights are in */
    int Wm_pb;       /* MFMA weights in the per-pointing primary-beam form */
    void *d_flag;    /* small device flag buffer */
    size_t flag_bytes;
    /* MFMA weights: [p-block][c][row][k] packed half/bf16 planes */
    void *d_Wm;
    size_t Wm_bytes;
    /* MFMA: per-channel scale applied to voltages before conversion */
    float *d_vscale;
    /* staging for host-layout pulls */
    void *d_stage;
    size_t stage_bytes;
    /* tile positions (E, N, H, L) per antenna for device phi */
    double *d_pos;
    double *d_freqs;
    /* primary beam on device, complex double: [p][a][2][2] after
       vmCalcJDevice (B_chan 0), [p][c][2][2] after vmCalcJDeviceChanB
       (B_chan 1: the same beam for every tile, per fine channel) */
    void *d_B;
    size_t B_bytes;
    int B_chan;
    /* d_J was last computed on the device from D and d_B (vmCalcJDevice*),
       so it is the same for every pointing wherever d_B is; 0 after vmPushJ
       or a new d_J.  B_np: the pointings d_B holds (B_chan 0: [B_np][a][4],
       1: [B_np][c][4]). */
    int J_fromB;
    int B_np;
    /* d_J holds pointing 0's J^-1 only, which every pointing shares (every
       pointing had the same beam when vmCalcJDevice computed it); expanded
       to all pointings by vm_expand_J before anything reads d_J per pointing */
    int J_shared;
    /* J^-1 on the host was last computed by vmCalcJ from D and pb.B (set
       there, taken by the next vmPushJ, which then also hands D and B to the
       device) */
    int J_host_fromB;
    /* MFMA tile-group form: GEMM slot arrays (antenna, polQ, polP per
       slot), the groups' beams [p][g][c][4] complex double and per-(c, p)
       constants; the groups' slot bounds and representative antennas */
    void *d_grp;
    size_t grp_bytes;
    void *d_Bg;
    size_t Bg_bytes;
    void *d_gx;
    size_t gx_bytes;
    int grp_t[5];
    int grp_rep[4];
    int grp_n;
    /* unit vectors per pointing (uE, uN, uH) */
    double *d_unit;
    size_t unit_bytes;
    /* inverse-PFB device buffers for the device-resident path */
    void *d_ipfb_out;
    size_t ipfb_out_bytes;
    void *d_ipfb_jobs; /* job descriptors of vmInvertPFBDeviceVDIFBatch */
    size_t ipfb_jobs_bytes;
    /* per-channel N-term scratch for the MFMA form (floats) */
    void *d_nterm;
    size_t nterm_bytes;
    /* MFMA beamformer variant chosen by vmTuneBeamformer (0: not tuned) */
    int bf_variant;
    int bf_order;
    /* compute units of the device (persistent launches) */
    int ncu;
    /* per-kernel event timing (vmSetKernelTiming): per slot a ring of
       VM_T_RING (start, stop) event pairs, recorded without waiting and
       resolved into t_ms / t_n when read (vm_timer_flush) */
    int timing;
    void *ev[16][VM_T_RING][2];
    int t_pend[16];  /* pairs recorded and not yet resolved */
    int t_open[16];  /* a start without its stop yet */
    double t_ms[16];
    long t_n[16];
};

/* kernel timing slots */
enum {
    VM_T_PFB,
    VM_T_BF_FP32,
    VM_T_BF_MFMA,
    VM_T_BF_PREP,
    VM_T_IPFB,
    VM_T_GEOM,
    VM_T_BF_WEIGHTS,
    VM_T_INCOH,
    VM_T_RENORM,
    VM_T_VDIF,
    VM_T_NSLOT
};
/* bracket a launch on the context stream with events when timing is on */
void vm_timer_start(vm_device_state *d, int slot);
void vm_timer_stop(vm_device_state *d, int slot);
/* the same on another stream (calls without a context use vm_dev_global()) */
void vm_timer_start_on(vm_device_state *d, int slot, void *stream);
void vm_timer_stop_on(vm_device_state *d, int slot, void *stream);
/* resolve the recorded pairs into t_ms / t_n (waits for their stops) */
void vm_timer_flush(vm_device_state *d);
/* destroy the timing events */
void vm_timer_free(vm_device_state *d);
vm_device_state *vm_dev_global(void);

/* Host helpers implemented in vm_host.c */
vm_device_state *vm_dev(vcsbeam_context *vm); /* create on first use */
void vm_dev_free(vcsbeam_context *vm);
/* d_J for every pointing (expands a shared J^-1, see J_shared); enqueued on
   the context stream */
void vm_expand_J(vcsbeam_context *vm);

/* Pinned host memory (hipHostMalloc), implemented in vm_device.hip */
void *vm_host_alloc_pinned(size_t bytes);
void vm_host_free_pinned(void *p);

#ifdef __cplusplus
}
#endif
#endif
