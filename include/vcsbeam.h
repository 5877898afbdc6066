/*
 * vcsbeam.h -- C-ABI of the MI355X-native voltage-beamforming path.
 *
 * Drop-in for the part of VCSBeam's public header (include/vcsbeam.h.in in
 * MWA-vcsbeam) that the voltage-beamforming path uses: the context, forward
 * fine PFB, tied-array / incoherent beamformer, inverse PFB and VDIF
 * requantisation entry points keep their names, argument meaning and
 * "exit on error" behaviour.  Citations are file:line in the reference tree.
 *
 * Differences from the reference header, all deliberate:
 *  - no CUDA/HIP, mwalib, PAL, psrfits, vdifio, hyperbeam or MPI headers are
 *    included; gpuDoubleComplex / gpuFloatComplex are plain structs with the
 *    layout of hipDoubleComplex / hipFloatComplex;
 *  - the mwalib metadata types are replaced by minimal structs carrying only
 *    the fields the path reads (same field names);
 *  - device buffers behind the context (d_v, d_S, d_e, ...) use the layouts
 *    documented at each field; host buffers keep the reference layouts;
 *  - MI355X extensions (precision MFMA beamformer and its autotuner, device
 *    FP64 Jones/phase, primary-beam surrogate, device-resident inverse PFB
 *    and VDIF requantisation, device memory/event helpers, kernel timing) are
 *    appended at the end of the header and of vcsbeam_context.
 */
#ifndef VCSBEAM_H
#define VCSBEAM_H

#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

typedef struct { double x, y; } gpuDoubleComplex; /* == hipDoubleComplex */
typedef struct { float x, y; } gpuFloatComplex;   /* == hipFloatComplex */
typedef struct { int8_t x, y; } vm_char2;         /* == HIP char2 */
typedef void *gpuStream_t;                        /* hipStream_t */
typedef void *gpufftHandle;                       /* opaque plan of the in-house FFT */

#define ERROR_MESSAGE_LEN 1024
#define MAX_COMMAND_LENGTH 1024

#define SPEED_OF_LIGHT_IN_VACUUM_M_PER_S 299792458.0
#define MWA_LATITUDE_RADIANS (-0.4660608448386394)   /* -26.703319 deg */
#define MWA_LONGITUDE_RADIANS (2.0362898668561042)   /* 116.67081524 deg */
#define MWA_ALTITUDE_METRES 377.827

#define PIBY2 (0.5 * M_PI)
#define D2R (M_PI / 180.0)
#define R2D (180.0 / M_PI)

/* ---------------------------------------- mwalib replacement (minimal) -- */
/* Field names follow mwalib's MetafitsMetadata / VoltageMetadata / Rfinput /
 * CoarseChannel; only the fields the beamforming path reads are present. */

typedef enum MWAVersion {
    CorrOldLegacy = 1,
    CorrLegacy = 2,
    CorrMWAXv2 = 3,
    VCSLegacyRecombined = 4,
    VCSMWAXv2 = 5
} MWAVersion;

typedef struct Rfinput {
    uint32_t input;
    uint32_t ant;
    uint32_t tile_id;
    char *tile_name;
    char *pol;                  /* "X" or "Y" */
    double electrical_length_m;
    double north_m;
    double east_m;
    double height_m;
    uint32_t vcs_order;
    uint32_t subfile_order;
    bool flagged;
    /* mwalib's per-input analog beamformer delays: num_delays (16) steps of
     * 435 ps, 32 marking a dead dipole; NULL when the metadata has none */
    uint32_t *dipole_delays;
} Rfinput;

typedef struct CoarseChannel {
    uint32_t corr_chan_number;
    uint32_t rec_chan_number;
    uint32_t gpubox_number;
    uint32_t chan_width_hz;
    uint32_t chan_start_hz;
    uint32_t chan_centre_hz;
    uint32_t chan_end_hz;
} CoarseChannel;

typedef struct MetafitsMetadata {
    uint32_t obs_id;
    MWAVersion mwa_version;
    double sched_start_mjd;
    uint64_t sched_start_gps_time_ms;
    size_t num_ants;
    size_t num_rf_inputs;
    Rfinput *rf_inputs;
    size_t num_ant_pols;
    size_t num_visibility_pols;
    size_t num_metafits_coarse_chans;
    CoarseChannel *metafits_coarse_chans;
    uint32_t coarse_chan_width_hz;
    size_t num_volt_fine_chans_per_coarse;
    double ra_tile_pointing_deg;
    double dec_tile_pointing_deg;
    char *project_id;
    size_t num_corr_fine_chans_per_coarse; /* calibration solutions per coarse channel */
    uint32_t corr_fine_chan_width_hz;
    uint64_t good_time_gps_ms;      /* first good data (metafits GOODTIME) */
    size_t num_metafits_timesteps;  /* 1 s timesteps from the scheduled start */
    size_t num_delays;              /* dipoles per tile (16), the length of dipole_delays */
} MetafitsMetadata;

typedef struct VoltageMetadata {
    MWAVersion mwa_version;
    uint64_t num_samples_per_voltage_block;
    uint64_t num_voltage_blocks_per_second;
    uint64_t voltage_block_size_bytes;
} VoltageMetadata;

typedef struct MetafitsContext MetafitsContext; /* opaque (mwalib) */
typedef struct VoltageContext VoltageContext;   /* opaque (mwalib) */

/* ------------------------------------------------------------ performance */
/* include/vcsbeam.h.in:100-155 */
#define PERFORMANCE_MAX_NUM_STOPWATCHES 16
#define PERFORMANCE_MAX_START_STOP 4096
#define PERFORMANCE_NO_STOPWATCH_FOUND -1
#define PERFORMANCE_NO_MPI -1

typedef struct logger_stopwatch_t {
    char *name;
    char *description;
    double values[PERFORMANCE_MAX_START_STOP];
    int nstart_stops;
    bool running;
    double total, total_sq;
} logger_stopwatch;

typedef struct logger_t {
    double begintime;
    FILE *fout;
    logger_stopwatch stopwatches[PERFORMANCE_MAX_NUM_STOPWATCHES];
    int nstopwatches;
    int world_rank;
} logger;

logger *create_logger(FILE *fout, int world_rank);
void destroy_logger(logger *log);
void logger_add_stopwatch(logger *log, const char *stopwatch_name, const char *description);
void logger_start_stopwatch(logger *log, const char *stopwatch_name, bool print_description);
void logger_stop_stopwatch(logger *log, const char *stopwatch_name);
void logger_timed_message(logger *log, const char *message);
void logger_message(logger *log, const char *message);
void logger_stopwatch_report_stats(logger *log, const char *stopwatch_name);
void logger_report_all_stats(logger *log);

/* ------------------------------------------------------------ calibration */
/* include/vcsbeam.h.in:169-202.  D reaches vm->D either from the loaders
 * below (Offringa binary, RTS text) or directly from the caller. */
#define CAL_NONE 0
#define CAL_RTS 1
#define CAL_OFFRINGA 2
#define OFFRINGA_HEADER_SIZE_BYTES 48
#define JONES_SIZE_BYTES 64 /* 4 complex doubles */
#define NDBL_PER_JONES 8
#define BANDPASS_ROWS_PER_ANT 8
#define CAL_BUFSIZE 4096
typedef struct calibration_t {
    char *metafits;
    char *caldir;
    int cal_type;
    char *ref_ant;
    bool keep_cross_terms;
    double phase_offset;
    double phase_slope;
    bool use_bandpass;
    bool custom_pq_correction;
    char *flags_file;
    int nflags;
    char **flagged_tilenames;
} calibration;

/* include/vcsbeam.h.in:205-217 */
void read_dijones_file(gpuDoubleComplex **Dd, gpuDoubleComplex *A, double *amp, uintptr_t nant,
                       char *fname);
void read_bandpass_file(gpuDoubleComplex ***Jm, gpuDoubleComplex ***Jf,
                        MetafitsMetadata *cal_metadata, char *filename);
void remove_reference_phase(gpuDoubleComplex *J, gpuDoubleComplex *Jref);
void zero_PQ_and_QP(gpuDoubleComplex *J);
/* reads <VCSBEAM_RUNTIME_DIR>/pq_phase_correction.txt */
void parse_calibration_correction_file(uint32_t gpstime, calibration *cal);
bool tilename_is_flagged(char *tilename, calibration *cal);
void init_calibration(calibration *cal);
void free_calibration(calibration *cal);

/* ------------------------------------------------------------------ jones */
#define NSTOKES 4

/* Index macros, identical to include/vcsbeam.h.in:234-268 */
#define v_IDX(s, c, i, nc, ni) ((s) * (ni) * (nc) + (c) * (ni) + (i))
#define VSPVB 64000
#define vMWAX_IDX(s, i, ni) ((((s) / VSPVB) * (ni) + (i)) * VSPVB + ((s) % VSPVB))
#define D_IDX(a, c, p1, p2, nc, npol) \
    ((a) * (npol) * (npol) * (nc) + (c) * (npol) * (npol) + (p1) * (npol) + (p2))
#define J_IDX(p, a, c, p1, p2, nant, nc, npol)                                   \
    ((p) * (nant) * (npol) * (npol) * (nc) + (a) * (npol) * (npol) * (nc) +     \
     (c) * (npol) * (npol) + (p1) * (npol) + (p2))
#define Jv_IDX(p, s, c, a, ns, nc, na) \
    ((p) * (na) * (nc) * (ns) + (s) * (na) * (nc) + (c) * (na) + (a))
#define B_IDX(p, s, c, pol, ns, nc, npol) \
    ((p) * (npol) * (nc) * (ns) + (s) * (npol) * (nc) + (c) * (npol) + (pol))
#define C_IDX(p, s, st, c, ns, nst, nc) \
    ((p) * ((nc) * (nst) * (ns)) + (s) * ((nc) * (nst)) + (st) * (nc) + (c))
#define I_IDX(s, c, nc) ((s) * (nc) + (c))
#define PHI_IDX(p, a, c, na, nc) ((p) * (nc) * (na) + (a) * (nc) + (c))
#define PB_IDX(p, a, pol, na, npol) ((p) * (na) * (npol) + (a) * (npol) + (pol))

/* 2x2 complex matrix operations, src/jones.c:201-578 */
void cp2x2(gpuDoubleComplex *Min, gpuDoubleComplex *Mout);
void inv2x2(gpuDoubleComplex *Min, gpuDoubleComplex *Mout);
void inv2x2d(double *Min, double *Mout);
void inv2x2S(gpuDoubleComplex *Min, gpuDoubleComplex *Mout);
void mult2x2d(gpuDoubleComplex *M1, gpuDoubleComplex *M2, gpuDoubleComplex *Mout);
void mult2x2d_RxC(double *M1, gpuDoubleComplex *M2, gpuDoubleComplex *Mout);
void mult2x2d_CxR(gpuDoubleComplex *M1, double *M2, gpuDoubleComplex *Mout);
void conj2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout);
double norm2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout);
void reverse2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout);
void swaprows2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout);
void swapcols2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout);
bool is2x2zero(gpuDoubleComplex *M);
void calc_hermitian(gpuDoubleComplex *M, gpuDoubleComplex *H);
void calc_coherency_matrix(gpuDoubleComplex *M, gpuDoubleComplex *c);
void fprintf_complex_matrix(FILE *fout, gpuDoubleComplex *M);

/* ----------------------------------------------------------------- filter */
typedef enum filter_type_t { ANALYSIS_FILTER, SYNTHESIS_FILTER } filter_type;

typedef struct pfb_filter_t {
    double *coeffs;
    int ncoeffs;
    int ntaps;
    int nchans;
    gpuDoubleComplex *twiddles;
    filter_type type;
} pfb_filter;

gpuDoubleComplex *roots_of_unity(int N);

/* --------------------------------------------------------------- metadata */
typedef enum vcsbeam_datatype_t {
    VM_INT4,  /* legacy 4+4-bit [s][c][i] (reference) */
    VM_DBL,   /* complex double [s][c][i] (reference) */
    VM_FLT,   /* MI355X: complex float [c][i][s], i in VCS order */
    VM_HALF   /* MI355X: complex FP16 (re, im halves of a dword, each rounded
                 to nearest even) in 32-sample blocks [c][s/32][i][s%32], the
                 last block padded; the MFMA forms' input */
} vcsbeam_datatype;

typedef enum vm_error_t {
    VM_SUCCESS,
    VM_END_OF_DATA,
    VM_READ_BUFFER_NOT_SET,
    VM_READ_BUFFER_LOCKED
} vm_error;

typedef struct geometric_delays_t {
    gpuDoubleComplex *phi;   /* host [p][a][c] (PHI_IDX) */
    gpuDoubleComplex *d_phi; /* device, same layout */
    uintptr_t npointings;
    uintptr_t nant;
    uintptr_t nchan;
    double *chan_freqs_hz;
    MetafitsMetadata *obs_metadata;
} geometric_delays;

typedef struct primary_beam_t {
    gpuDoubleComplex *B;     /* host [p][a][2][2] (PB_IDX with npol=4) */
    void *beam;              /* FEEBeam* in the reference (hyperbeam) */
    uint32_t **delays;
    double **amps;
    uintptr_t npointings;
    uintptr_t nant;
    uintptr_t npol;
    uint32_t freq_hz;
    MetafitsMetadata *obs_metadata;
} primary_beam;

/* VDIF frame header: 8 little-endian 32-bit words, the 32-byte header of
 * the VDIF 1.0 specification (vdifio's vdif_header, used at
 * include/vcsbeam.h.in:637 and src/beam_vdif.c). */
#define VDIF_HEADER_BYTES 32
typedef struct vdif_header_t {
    uint32_t w[8];
} vdif_header;

/* VDIF output description of one pointing (include/vcsbeam.h.in:1107-1146),
 * filled by vmPopulateVDIFHeader. */
struct vdifinfo {
    int frame_length;         /* bytes per frame, header included (544) */
    int frame_rate;           /* frames per second */
    size_t samples_per_frame; /* time samples per frame (128) */
    int dataarraylength;      /* payload bytes per frame (512) */
    int bits;                 /* bits per sample (8) */
    int nchan;                /* channels per frame: one per polarisation (2) */
    int chan_width;           /* channel width, Hz */
    int sample_rate;          /* sample rate, Hz */
    int npol;
    int iscomplex;
    int threadid;
    char stationid[3];
    char exp_name[17];
    char scan_name[17];
    size_t block_size;        /* bytes of one second of frames, headers included */
    size_t sizeof_buffer;     /* payload values of one second */
    size_t sizeof_beam;       /* payload values of one frame */
    float *b_scales;
    float *b_offsets;
    char telescope[24];
    char source[24];
    char obs_mode[8];
    double fctr;              /* centre frequency, MHz */
    char ra_str[16];
    char dec_str[16];
    double BW;                /* bandwidth, MHz */
    double MJD_start;
    double sec_offset;
    double MJD_epoch;
    char date_obs[24];
    char basefilename[1024];
    int got_scales;
};

typedef struct beam_geom_t {
    double mean_ra;
    double mean_dec;
    double az;
    double el;
    double lmst;
    double fracmjd;
    double intmjd;
    double unit_N;
    double unit_E;
    double unit_H;
} beam_geom;

/* include/vcsbeam.h.in:909-920: pointing geometry (J2000 RA in hours, Dec in
 * degrees, UTC MJD) and sexagesimal strings */
void calc_beam_geom(double ras_hours, double decs_degs, double mjd, beam_geom *bg);
void dec2hms(char *out, double in, int sflag);
void mjd2lst(double mjd, double *lst);
double parse_dec(char *dec_ddmmss);
double parse_ra(char *ra_hhmmss);
/* No counterpart in the reference (PAL provides these there): the models
 * calc_beam_geom uses.  Greenwich mean sidereal time (rad) of a UT MJD;
 * IAU 1980 nutation in longitude and obliquity and the mean obliquity (rad);
 * J2000 mean place to apparent place (stages: 1 precession, 2 nutation,
 * 4 aberration, 7 all); hour angle and declination to azimuth (from North
 * through East) and elevation at latitude phi. */
double vmGMST(double mjd);
void vmNutation(double mjd, double *dpsi, double *deps, double *eps0);
void vmMeanToApparent(double ra, double dec, double mjd, int stages, double *ra_ap, double *dec_ap);
void vmHaDecToAzEl(double ha, double dec, double phi, double *az, double *el);
/* include/vcsbeam.h.in:1201-1205: P = [-cos chi, -sin chi; sin chi, -cos chi]
 * for the parallactic angle chi of (az, za) at latitude lat (radians) */
void parallactic_angle_correction(double *P, double lat, double az, double za);

/* ---------------------------------------------------------- primary beam */
/* include/vcsbeam.h.in:1181-1183 */
#define NCONFIGS           138
#define DEAD_CONFIG        (NCONFIGS - 1)
#define MANY_DEAD_DIPOLES  -1
/* Seconds per analog beamformer delay step (the unit of dipole_delays) */
#define MWA_DELAY_STEP_S   435e-12
#define MWA_DEAD_DIPOLE_DELAY 32

/* include/vcsbeam.h.in:1195-1199 (src/primary_beam.c:250-297): per RF input
 * the dipole delays (pointers into the metadata) and amplitudes (1, or 0
 * for a dipole whose delay is 32); free_delays_amps releases them. */
void create_delays_amps_from_metafits(MetafitsMetadata *metafits_metadata, uint32_t ***delays,
                                      double ***amps);
void free_delays_amps(MetafitsMetadata *metafits_metadata, uint32_t **delays, double **amps);
/* include/vcsbeam.h.in:1207 (src/primary_beam.c:355-414): index of a
 * live/dead dipole configuration of 16 amplitudes (0 = dead): 0 for none
 * dead, 1 + d for one, 16 d1 - (d1 + 2)(d1 + 1)/2 + d2 + 17 for two
 * (d1 < d2), 137 for all sixteen, MANY_DEAD_DIPOLES otherwise. */
int hash_dipole_config(double *amps);
/* include/vcsbeam.h.in:1215 (src/primary_beam.c:479-532): the beam Jones
 * J of (az, za, freq_hz) for the given dipole delays and amplitudes (the
 * analytic model; `beam` is unused and may be NULL), optionally rotated
 * J <- P J, and the response IQUV[4] to pure Stokes I, Q, U and V skies:
 * I = Re(C00 + C11)/2, Q = Re(C00 - C11)/2, U = Re C01, V = -Im C01 for the
 * coherency C = J S J^H of each sky S. */
void calc_normalised_beam_response(void *beam, double az, double za, double freq_hz,
                                   uint32_t *delays, double *amps, double *IQUV,
                                   gpuDoubleComplex *J, bool apply_pa_correction);

/* PSRFITS search-mode output: the header and sub-integration fields that
 * populate_psrfits_header / populate_spliced_psrfits_header set
 * (src/beam_psrfits.c:39-466; psrfits_utils' struct psrfits, which is not
 * part of the reference tree), and this library's own writer state.  The
 * first four sub fields are the arrays vmSendSToFits fills
 * (src/form_beam.cpp:666-671). */
typedef struct vm_psrfits_sub_t {
    float *dat_offsets;       /* [npol][nchan] */
    float *dat_scales;        /* [npol][nchan] */
    uint8_t *data;            /* [nsblk][npol][nchan] */
    int bytes_per_subint;
    float *dat_freqs;         /* [nchan], MHz */
    float *dat_weights;       /* [nchan] */
    double tsubint, offs, lst, ra, dec, glon, glat; /* s, s, s, deg x 4 */
    float feed_ang, pos_ang, par_ang, tel_az, tel_zen; /* deg */
} vm_psrfits_sub;
typedef struct vm_psrfits_hdr_t {
    char project_id[24], obs_mode[8], observer[24], telescope[24], frontend[24], backend[24];
    char source[24], date_obs[64], poln_type[8], track_mode[16], cal_mode[8], feed_mode[8];
    char ra_str[16], dec_str[16];
    double scanlen, dt, fctr, BW, df, ra2000, dec2000, azimuth, zenith_ang, beam_FWHM;
    double start_lst, MJD_epoch, fd_sang, fd_xyph, chan_dm;
    int npol, nchan, nbits, nsblk, summed_polns, rcvr_polns, fd_hand, be_phase;
    int start_day, start_sec, scan_number;
} vm_psrfits_hdr;
typedef struct vm_psrfits_t {
    vm_psrfits_sub sub;
    vm_psrfits_hdr hdr;
    char basefilename[1024];
    int rows_per_file; /* sub-integrations (seconds) per file */
    int filenum;       /* files opened so far; files are <basefilename>_%04d.fits */
    int tot_rows;      /* sub-integrations written */
    int rows_in_file;
    void *fptr;        /* the open file (FILE *) */
    long naxis2_pos;   /* byte offset of the SUBINT NAXIS2 card in it */
    long data_start;   /* byte offset of the SUBINT rows in it */
} vm_psrfits;
typedef struct mpi_psrfits_t {
    vm_psrfits coarse_chan_pf;
    int writer_id;
    vm_psrfits spliced_pf;
} mpi_psrfits;

typedef struct host_buffer_t {
    void *buffer;
    size_t buffer_size;
    void *read_ptr;
    size_t read_size;
    void *copy_from_ptr;
    void *copy_to_ptr;
    size_t copy_size;
    bool locked;
} host_buffer;

typedef struct device_buffer_t {
    void *buffer;
    size_t buffer_size;
} device_buffer;

/* ------------------------------------------------------------ forward PFB */
/* include/vcsbeam.h.in:449-522 */
#define CLIP(x, max) ((x) < -(max) ? -(max) : (x) > (max)-1 ? (max)-1 : (x))
#define INT_TO_UINT8(x) ((x) < 0 ? CLIP(x, 8) + 0x10 : CLIP(x, 8))
#define DEMOTE(x) (INT_TO_UINT8((int)round(x)))
#define PACK_NIBBLES(r, i) ((DEMOTE(r) << 4) + DEMOTE(i))

typedef enum pfb_flags_t {
    PFB_MALLOC_HOST_INPUT = 0x01,
    PFB_MALLOC_HOST_OUTPUT = 0x02,
    PFB_MALLOC_DEVICE_INPUT = 0x04,
    PFB_MALLOC_DEVICE_OUTPUT = 0x08,
    PFB_MALLOC_ALL = 0x0F,
    PFB_TYPE_MASK = 0xF0,
    PFB_COMPLEX_INT4 = 0x10,
    PFB_COMPLEX_FLOAT64 = 0x20,
    PFB_COMPLEX_FLOAT32 = 0x40, /* MI355X: device output VM_FLT [c][i][s] */
    PFB_COMPLEX_FLOAT16 = 0x80, /* MI355X: device output VM_HALF [c][s/32][i][s%32]
                                   (fused kernel only), read by the MFMA forms */
    PFB_IMAG_PART_FIRST = 0x100,
    PFB_EMULATE_FPGA = 0x200,
    PFB_STAGED = 0x400,         /* MI355X: run WOLA/round/FFT/pack as the
                                   reference's four separate passes */
    PFB_SMART = 0x21D,
    PFB_FULL_PRECISION = 0x2D
} pfb_flags;

typedef struct forward_pfb_t {
    vm_char2 *d_htr_data;
    void *vcs_data;              /* host output, reference layout/format */
    void *d_vcs_data;            /* device output: VM_FLT [c][i][s] complex
                                    float, VM_HALF [c][s/32][i][s%32] complex
                                    FP16, or VM_INT4 legacy [s][c][i] */
    size_t d_htr_size;
    size_t htr_stride;
    size_t vcs_size;
    size_t d_vcs_size;
    size_t vcs_stride;
    size_t char2s_per_second;
    size_t bytes_per_block;
    size_t chunks_per_second;
    size_t chunk;
    int ninputs_per_cufft_batch;
    int cufft_batch_size;
    gpuFloatComplex *d_weighted_overlap_add; /* staged mode: int2 / float2
                                                [m][i][n] as in the reference */
    size_t weighted_overlap_add_size;
    int *filter_coeffs;
    int *d_filter_coeffs;
    int nspectra;
    int nspectra_per_chunk;
    int M;
    int K;
    int I;
    int P;
    int *i_output_idx;
    int *d_i_output_idx;
    pfb_flags flags;
    gpufftHandle plan;
    /* MI355X additions */
    void *d_flt_out;             /* complex float [c][i][s] fused-kernel output
                                    when d_vcs_data holds another format */
    int filter_fits_int16;       /* 1 -> packed dot2 WOLA path */
} forward_pfb;

/* Metafits subset and voltage files, replacing the mwalib contexts
 * (src/metadata.c:1297-1424, 930-993).  vmLoadMetafits fills md (and vmd when
 * not NULL) from a metafits file: primary keys GPSTIME, MJD, PROJECT, MODE,
 * CHANNELS, BANDWDTH, NCHANS, FINECHAN, RA, DEC, NINPUTS and the TILEDATA
 * table; rf_inputs come sorted by subfile order (antenna, then X before Y).
 * Returns 0; malformed files end the process.  vmFreeMetafits releases what
 * it allocated.  vmReadSecondFromFiles is a vm_read_second_fn over MWAX
 * subfiles (<obsid>_<subobs>_<rec>.sub) or legacy recombined files
 * (<obsid>_<gps>_ch<rec>.dat) in dir; user points to a vm_voltage_files. */
typedef struct vm_voltage_files_t {
    const char *dir;
    const MetafitsMetadata *md;
    const VoltageMetadata *vmd;
} vm_voltage_files;
int vmLoadMetafits(const char *path, MetafitsMetadata *md, VoltageMetadata *vmd);
void vmFreeMetafits(MetafitsMetadata *md);
int vmReadSecondFromFiles(void *user, uint64_t gps_second, int coarse_chan_idx, void *dst,
                          size_t nbytes);

/* ---------------------------------------------------------------- context */
typedef struct vm_device_state vm_device_state; /* library-private */

/* Reader callback replacing mwalib_voltage_context_read_second: fill
 * `nbytes` bytes at `dst` with one second of voltages for gps_second and
 * coarse channel coarse_chan_idx; return 0 on success. */
typedef int (*vm_read_second_fn)(void *user, uint64_t gps_second,
                                 int coarse_chan_idx, void *dst, size_t nbytes);

typedef enum vm_beamform_form_t {
    VM_BF_FP32 = 0,          /* plain FP32 drop-in (default) */
    VM_BF_MFMA_FP16 = 1,     /* FP16 voltages, split hi+lo FP16 weights, MFMA */
    VM_BF_MFMA_BF16 = 2,     /* BF16 voltages, split hi+lo BF16 weights, MFMA */
    VM_BF_MFMA_FP16_SINGLE = 3 /* FP16 voltages, one FP16 part per weight, MFMA:
                                  half the MFMA work of VM_BF_MFMA_FP16 */
} vm_beamform_form;

typedef struct vcsbeam_context_t {
    bool use_mpi;
    int mpi_size;
    int mpi_rank;
    int ncoarse_chans;
    int coarse_chan_idx;
    int writer;

    MetafitsContext *obs_context;
    MetafitsMetadata *obs_metadata;
    VoltageContext *vcs_context;
    VoltageMetadata *vcs_metadata;
    MetafitsContext *cal_context;
    MetafitsMetadata *cal_metadata;
    MetafitsContext *obs_context_legacy;
    MetafitsMetadata *obs_metadata_legacy;

    char *datadir;
    char **filenames;
    int nfiles;
    unsigned int nfiletimes;
    unsigned int seconds_per_file;

    calibration cal;

    double *ras_hours;
    double *decs_degs;

    pfb_filter *analysis_filter;
    pfb_filter *synth_filter;

    forward_pfb *fpfb;

    host_buffer *v;
    void *d_v;                 /* device voltages (see vcsbeam_datatype) */
    uintptr_t v_size_bytes;
    uintptr_t d_v_size_bytes;

    void *S, *d_S;             /* host S: [p][s][st][c] float (C_IDX);
                                  device d_S: [c][s/64][p][st][s%64] float,
                                  the last 64-sample block padded */
    uintptr_t S_size_bytes;
    uintptr_t d_S_size_bytes;

    gpuDoubleComplex *e, *d_e; /* host e: [p][s][c][pol] complex double;
                                  device d_e: [p][c][s][pol] complex float */
    uintptr_t e_size_bytes;
    uintptr_t d_e_size_bytes;

    gpuDoubleComplex *J, *d_J; /* [p][a][c][2][2] complex double, both */
    uintptr_t J_size_bytes;
    uintptr_t d_J_size_bytes;

    gpuDoubleComplex *Jv_P, *d_Jv_P; /* unused: J^-1 v is fused into the */
    gpuDoubleComplex *Jv_Q, *d_Jv_Q; /* beamformer (kept for layout)      */
    uintptr_t Jv_size_bytes;
    uintptr_t d_Jv_size_bytes;

    gpuDoubleComplex *D, *d_D; /* [a][c][2][2] complex double, both */
    uintptr_t D_size_bytes;
    uintptr_t d_D_size_bytes;

    uint32_t *polP_idxs, *d_polP_idxs;
    uint32_t *polQ_idxs, *d_polQ_idxs;
    uintptr_t pol_idxs_size_bytes;
    uintptr_t d_pol_idxs_size_bytes;

    geometric_delays gdelays;
    primary_beam pb;

    unsigned int npointing;

    uintptr_t max_gpu_mem_bytes;
    uint32_t chunks_per_second;

    int num_coarse_chans_to_process;
    int *coarse_chan_idxs_to_process;
    int *cal_coarse_chan_idxs_to_process;

    int num_gps_seconds_to_process;
    uint32_t *gps_seconds_to_process;

    bool output_fine_channels;
    bool output_coarse_channels;

    size_t current_gps_idx;

    bool do_forward_pfb;
    bool do_inverse_pfb;

    int sample_rate;
    int fine_sample_rate;
    int bytes_per_second;
    int num_not_flagged;
    int nchan;
    int nfine_chan;

    int out_nstokes;

    vdif_header vhdr;          /* VDIF frame header of the next frame (32 bytes) */
    struct vdifinfo *vf;

    float *d_offsets, *offsets;
    float *d_scales, *scales;
    uint8_t *d_Cscaled, *Cscaled;

    size_t offsets_size;
    size_t scales_size;
    size_t Cscaled_size;

    int chunk_to_load;
    vcsbeam_datatype datatype;

    gpuStream_t *streams;

    logger *log;
    char log_message[MAX_COMMAND_LENGTH];
    char error_message[ERROR_MESSAGE_LEN];

    /* ---- MI355X extensions ---- */
    vm_device_state *dev;        /* device-side state owned by the library */
    vm_beamform_form beamform_form;
    vm_read_second_fn read_second;
    void *read_second_user;
    float *bf_err;               /* optional host [p][st] norm-wise errors
                                    filled by vmBeamformErrorReport */
} vcsbeam_context;

#define NO_ANTENNA_FOUND -1

/* -------------------------------------------- context and buffer set-up -- */
/* include/vcsbeam.h.in:667-856 */
void vmCheckError(vm_error err);
vcsbeam_context *vmInit(bool use_mpi);
void destroy_vcsbeam_context(vcsbeam_context *vm);
void vmSetOutputChannelisation(vcsbeam_context *vm, bool out_fine, bool out_coarse);

void vmMallocVHost(vcsbeam_context *vm);
void vmMallocVDevice(vcsbeam_context *vm);
void vmFreeVHost(vcsbeam_context *vm);
void vmFreeVDevice(vcsbeam_context *vm);
void vmMallocJVHost(vcsbeam_context *vm);
void vmMallocJVDevice(vcsbeam_context *vm);
void vmFreeJVHost(vcsbeam_context *vm);
void vmFreeJVDevice(vcsbeam_context *vm);
void vmMallocEHost(vcsbeam_context *vm);
void vmMallocEDevice(vcsbeam_context *vm);
void vmFreeEHost(vcsbeam_context *vm);
void vmFreeEDevice(vcsbeam_context *vm);
void vmMallocSHost(vcsbeam_context *vm);
void vmMallocSDevice(vcsbeam_context *vm);
void vmFreeSHost(vcsbeam_context *vm);
void vmFreeSDevice(vcsbeam_context *vm);
void vmMallocJHost(vcsbeam_context *vm);
void vmMallocJDevice(vcsbeam_context *vm);
void vmFreeJHost(vcsbeam_context *vm);
void vmFreeJDevice(vcsbeam_context *vm);
void vmMallocDHost(vcsbeam_context *vm);
void vmMallocDDevice(vcsbeam_context *vm);
void vmFreeDHost(vcsbeam_context *vm);
void vmFreeDDevice(vcsbeam_context *vm);
void vmMallocPQIdxsHost(vcsbeam_context *vm);
void vmMallocPQIdxsDevice(vcsbeam_context *vm);
void vmFreePQIdxsHost(vcsbeam_context *vm);
void vmFreePQIdxsDevice(vcsbeam_context *vm);

void vmSetMaxGPUMem(vcsbeam_context *vm, int nchunks);
void vmPushChunk(vcsbeam_context *vm);
void vmPushJ(vcsbeam_context *vm);
void vmCreatePrimaryBeam(vcsbeam_context *vm);
/* include/vcsbeam.h.in:1213; called by app/make_mwa_tied_array_beam.c:375 */
void free_primary_beam(primary_beam *pb);
void vmCreateGeometricDelays(vcsbeam_context *vm);
void vmCreateCudaStreams(vcsbeam_context *vm);
void vmDestroyCudaStreams(vcsbeam_context *vm);
void vmCreateStatistics(vcsbeam_context *vm, mpi_psrfits *mpfs);
void vmDestroyStatistics(vcsbeam_context *vm);
void vmSetPolIdxLists(vcsbeam_context *vm);
/* include/vcsbeam.h.in:1209-1211 (src/primary_beam.c:79-191): B[p][a] for
 * every pointing and tile, one beam evaluation per dipole configuration
 * (hash_dipole_config) and pointing, times the parallactic-angle rotation
 * on the right (B = J P).  The beam is this library's analytic 4 x 4
 * dipole-array model at the coarse-channel centre (the reference's FEE
 * model, hyperbeam, is not part of this library): dipole amplitudes from
 * create_delays_amps_from_metafits, analog delays dipole_delays x 435 ps,
 * normalised to the all-live array at zenith.  Each configuration keeps
 * its own beam (the reference's cache aliases one buffer, so a tile whose
 * configuration was seen before receives the last beam evaluated); tiles
 * with more than two dead dipoles are evaluated with their own
 * amplitudes and also rotated; a tile whose sixteen dipoles are dead gets
 * B = 0 (the reference leaves it unset).  Without dipole delays in the
 * metadata B stays the identity. */
void vmCalcB(vcsbeam_context *vm, beam_geom *beam_geom_vals);
void vmCalcJ(vcsbeam_context *vm);
/* include/vcsbeam.h.in:789 (src/jones.c:170-188): the beam geometry of
 * every pointing at the middle of the second about to be beamformed
 * (MJD = sched_start_mjd + (t + first GPS second - obs_id + 0.5) / 86400,
 * t = chunk_to_load / chunks_per_second), then vmCalcPhi, vmCalcB and
 * vmCalcJ.  Call it after vmReadNextSecond and before vmBeamformSecond. */
void vmCalcJonesAndDelays(vcsbeam_context *vm, double *ras_hours, double *decs_degs,
                          beam_geom *beam_geom_vals);
void vmSetNumPointings(vcsbeam_context *vm, unsigned int npointings);
void vmSetNumNotFlaggedRFInputs(vcsbeam_context *vm);
/* include/vcsbeam.h.in (src/metadata.c:1656-1705): "RA Dec" string pairs
 * from a file into vm->ras_hours / vm->decs_degs, setting the pointing count */
void vmParsePointingFile(vcsbeam_context *vm, const char *filename);
/* include/vcsbeam.h.in:689-703, 801-819: the observation and calibration by
 * file name, with this library's metafits reader and voltage-file reader in
 * place of mwalib.  vmLoadObsMetafits / vmLoadCalMetafits load into
 * vm->obs_metadata / vm->cal_metadata (the context owns them);
 * vmBindObsData then binds channels ("+n" index, "-n" from the end, or a
 * receiver number) and seconds ("+n" from the first good second, "-n" from
 * the end, or a GPS second) and registers vmReadSecondFromFiles over datadir
 * (vmGetVoltageMetadata).  destroy_vcsbeam_context releases all of it. */
void vmLoadObsMetafits(vcsbeam_context *vm, char *filename);
void vmLoadCalMetafits(vcsbeam_context *vm, char *filename);
void vmGetVoltageMetadata(vcsbeam_context *vm);
void vmBindObsData(vcsbeam_context *vm, char *first_coarse_chan_str, int num_coarse_chans_to_process,
                   int coarse_chan_idx_offset, char *first_gps_second_str,
                   int num_gps_seconds_to_process, int gps_second_offset, char *datadir);
void vmBindCalibrationData(vcsbeam_context *vm, char *caldir, int cal_type, bool use_bandpass,
                           char *flags_file);
void vmReadCalibration(vcsbeam_context *vm);
void vmPrintTitle(vcsbeam_context *vm, const char *title);
void vmGetVoltFilename(vcsbeam_context *vm, unsigned int coarse_chan_idx, uint64_t gps_second,
                       char *filename);
void vmGetLegacyVoltFilename(vcsbeam_context *vm, unsigned int coarse_chan_idx, uint64_t gps_second,
                             char *filename);
long unsigned int get_relative_gps(MetafitsMetadata *obs_metadata, long int relative_begin);
long unsigned int parse_begin_string(MetafitsMetadata *obs_metadata, char *begin_str);
uintptr_t parse_coarse_chan_string(MetafitsMetadata *obs_metadata, char *begin_coarse_chan_str);
Rfinput *find_matching_rf_input(MetafitsMetadata *metadata, Rfinput *rfinput);
/* include/vcsbeam.h.in:794-809: calibration into vm->D (host, before vmPushD /
 * vmCalcJ).  The Offringa loader reads vm->cal.caldir as a file, the RTS loader
 * as the directory of DI_JonesMatrices_node%03d.dat and
 * BandpassCalibration_node%03d.dat; both use vm->cal_metadata. */
void vmParseFlaggedTilenamesFile(char *filename, calibration *cal);
void vmSetCustomTileFlags(vcsbeam_context *vm);
void vmLoadRTSSolution(vcsbeam_context *vm);
void vmLoadOffringaSolution(vcsbeam_context *vm);
void vmApplyCalibrationCorrections(vcsbeam_context *vm);

void vmLoadFilter(vcsbeam_context *vm, char *filtername, filter_type type, int nchans);
void free_pfb_filter(pfb_filter *filter);
void vmScaleFilterCoeffs(vcsbeam_context *vm, filter_type type, double scale_factor);

host_buffer *vmInitReadBuffer(size_t read_size, size_t margin_size);
vm_error vmReadBufferCopyMargin(host_buffer *rb);
void vmFreeReadBuffer(host_buffer *rb);
vm_error vmReadNextSecond(vcsbeam_context *vm);
void vmReportPerformanceStats(vcsbeam_context *vm);

/* ------------------------------------------------------------- geometry */
/* include/vcsbeam.h.in:878-922 */
void vmCalcPhi(vcsbeam_context *vm, beam_geom *beam_geom_vals);
void calc_geometric_delays(beam_geom *beam_geom_vals, uint32_t freq_hz,
                           MetafitsMetadata *obs_metadata, gpuDoubleComplex *phi);
void free_geometric_delays(geometric_delays *gdelays);
void vmPushPhi(vcsbeam_context *vm);
double calc_array_factor(MetafitsMetadata *obs_metadata, uint32_t freq_hz,
                         beam_geom *bg1, beam_geom *bg2);

/* ------------------------------------------------------------ forward PFB */
/* include/vcsbeam.h.in:924-935 */
void vmInitForwardPFB(vcsbeam_context *vm, int M, pfb_flags flags);
void vmUploadForwardPFBChunk(vcsbeam_context *vm);
void vmWOLAChunk(vcsbeam_context *vm);
void vmFPGARoundingChunk(vcsbeam_context *vm);
void vmFFTChunk(vcsbeam_context *vm);
void vmPackChunk(vcsbeam_context *vm);
void vmDownloadForwardPFBChunk(vcsbeam_context *vm);
void vmExecuteForwardPFB(vcsbeam_context *vm);
void vmWritePFBOutputToFile(vcsbeam_context *vm);
void vmFreeForwardPFB(forward_pfb *fpfb);

/* ------------------------------------------------------------ inverse PFB */
/* include/vcsbeam.h.in:952-979 */
struct gpu_ipfb_arrays {
    int ntaps;
    int in_size;
    int ft_size;
    int out_size;
    float *in_real, *in_imag;   /* unused by this implementation */
    float *ft_real, *ft_imag;   /* host copies of the filter (polyphase) */
    float *d_in_real, *d_in_imag; /* d_in_real: device complex staging */
    float *d_ft_real, *d_ft_imag; /* d_ft_real: device polyphase filter */
    float *d_out;
};

void cu_invert_pfb(gpuDoubleComplex *data_buffer_fine, int file_no, int npointing,
                   int nsamples, int nchan, int npol, int sizeof_buffer,
                   struct gpu_ipfb_arrays *g, float *data_buffer_uvdif);
void cu_load_ipfb_filter(pfb_filter *filter, struct gpu_ipfb_arrays *g);
void malloc_ipfb(struct gpu_ipfb_arrays *g, pfb_filter *filter, int nsamples,
                 int npol, int npointing);
void free_ipfb(struct gpu_ipfb_arrays *g);

/* -------------------------------------------------------------- form beam */
/* include/vcsbeam.h.in:997-1048 */
#define REAL_NIBBLE_TO_UINT8(X) (((X) >> 4) & 0xf)
#define IMAG_NIBBLE_TO_UINT8(X) ((X)&0xf)
#define UINT8_TO_INT(X) ((X) >= 0x8 ? (signed int)(X)-0x10 : (signed int)(X))

void cu_form_incoh_beam(uint8_t *data, uint8_t *d_data, size_t data_size,
                        float *d_incoh, unsigned int nsample, int nchan, int ninput,
                        float *offsets, float *d_offsets, float *scales,
                        float *d_scales, uint8_t *Iscaled, uint8_t *d_Iscaled,
                        size_t Iscaled_size);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the device
 * part of cu_form_incoh_beam on device buffers only, enqueued on `stream`
 * (NULL: the null stream) with no synchronisation, no copies and no
 * allocation.  d_data [nsample][nchan][ninput] 4+4-bit (data_size bytes);
 * writes d_incoh [nsample][nchan], d_offsets and d_scales [nchan] and
 * d_Iscaled [nsample][nchan] (Iscaled_size bytes, at least nsample x nchan).
 * The renormalisation's partial extremes go to a fixed device array of
 * VM_INCOH_SCRATCH_PAIRS (min, max) pairs, 512 KiB, which each GPU the
 * library runs on holds until the process ends; it limits nchan to
 * VM_INCOH_SCRATCH_PAIRS.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race.
 * Returns 0, or -1 with a message on stderr when it rejects the sizes. */
#define VM_INCOH_SCRATCH_PAIRS 65536
int vmFormIncohBeamNoAlloc(const uint8_t *d_data, size_t data_size, float *d_incoh,
                           unsigned int nsample, int nchan, int ninput, float *d_offsets,
                           float *d_scales, uint8_t *d_Iscaled, size_t Iscaled_size,
                           gpuStream_t stream);
void vmApplyJChunk(vcsbeam_context *vm);
void vmBeamformChunk(vcsbeam_context *vm);
void vmBeamformSecond(vcsbeam_context *vm);
void vmPullE(vcsbeam_context *vm);
void vmPullS(vcsbeam_context *vm);
void prepare_data_buffer_fine(gpuDoubleComplex *data_buffer_fine, vcsbeam_context *vm,
                              uintptr_t timestep_idx);
void vmSendSToFits(vcsbeam_context *vm, mpi_psrfits *mpfs);
/* The one-process form of the coarse-channel splice (vmSendSToFits on each
 * rank + gather_splice_psrfits, src/form_beam.cpp:653-673 and
 * src/beam_psrfits.c:608-642): vms[g] holds coarse channel g, in ascending
 * frequency, all with the same pointings, Stokes, fine channels and samples,
 * and each with vmCreateStatistics done.  Each context's Stokes are
 * renormalised on the device straight into one spliced buffer, downloaded
 * once into pf[p].sub (data [s][st][g*nfine + c], offsets and scales
 * [st][g*nfine + c]) for every pointing p.  The spliced device buffer is
 * kept by the library between calls (grown as needed, on the current
 * device); concurrent calls from several host threads race on it. */
void vmSpliceSToFits(vcsbeam_context **vms, int ncoarse, vm_psrfits *pf);
/* src/beam_psrfits.c:39-258 (spliced: vm is the lowest channel's context and
 * vm->ncoarse_chans the number of channels spliced), 281-466 (one channel),
 * 475-488.  The arrays are allocated here; free_psrfits releases them. */
void populate_spliced_psrfits_header(vcsbeam_context *vm, vm_psrfits *pf, int max_sec_per_file,
                                     int outpol, beam_geom *beam_geom_vals, char *basename,
                                     bool is_coherent);
void populate_psrfits_header(vcsbeam_context *vm, vm_psrfits *pf, int max_sec_per_file,
                             int outpol, beam_geom *beam_geom_vals, char *incoh_basename,
                             bool is_coherent);
void free_psrfits(vm_psrfits *pf);
/* No counterpart in the reference (psrfits_utils writes there): appends
 * pf->sub as the next SUBINT row of <basefilename>_%04d.fits, starting a new
 * file every rows_per_file rows, then advances offs; vmClosePsrfits finishes
 * the open file.  Returns 0 on success. */
int vmWritePsrfitsSubint(vm_psrfits *pf);
void vmClosePsrfits(vm_psrfits *pf);
/* J2000 RA/Dec (rad) to galactic longitude and latitude (rad) */
void vmEqToGal(double ra, double dec, double *glon, double *glat);
float *create_pinned_data_buffer(size_t size);
void vmPushPolIdxLists(vcsbeam_context *vm);
gpuDoubleComplex *create_data_buffer_fine(int npointing, int nsamples, int nchan, int npol);
void allocate_input_output_arrays(void **data, void **d_data, size_t size);
void free_input_output_arrays(void *data, void *d_data);

/* ------------------------------------------------------------------- VDIF */
/* src/beam_vdif.c:29-38, 254-260 */
void float2int8_trunc(float *f, int n, float min, float max, int8_t *i);
void to_offset_binary(int8_t *i, int n);

/* ======================================================================
 * MI355X extensions -- no counterpart in the reference header.
 * ====================================================================== */

/* Bind metadata supplied by the caller (replaces vmLoadObsMetafits +
 * vmBindObsData's mwalib glue, src/metadata.c:166-233): sets the shorthand
 * sizes and allocates the read buffer exactly as vmBindObsData does. */
void vmBindObsMetadata(vcsbeam_context *vm, MetafitsMetadata *obs_metadata,
                       VoltageMetadata *vcs_metadata, int coarse_chan_idx,
                       uint32_t first_gps_second, int num_gps_seconds);

/* vmReadNextSecond with the second taken from memory instead of a file. */
vm_error vmReadNextSecondFrom(vcsbeam_context *vm, const void *src, size_t nbytes);

/* In-memory filter (replaces reading RUNTIME_DIR/<name>.dat). */
void vmSetFilter(vcsbeam_context *vm, const double *coeffs, int ncoeffs,
                 filter_type type, int nchans);

/* Device FP64: phi from tile positions and pointing unit vectors, with
 * argument reduction in cycles; result in gdelays.d_phi (and not on host). */
void vmCalcPhiDevice(vcsbeam_context *vm, beam_geom *beam_geom_vals);
/* Device FP64: J^-1 = F (D B)^-1 from d_D and pb.B; result in d_J.
 * d_J is written by this call, vmCalcJDeviceChanB or vmPushJ only: the MFMA
 * form learns from them whether J^-1 can differ between pointings.  When
 * every pointing has the same beam, J^-1 is the same for every pointing and
 * only pointing 0's is written; the library writes the others before it
 * reads J^-1 per pointing (the FP32 form), and a caller that reads d_J
 * itself calls vmExpandJDevice first. */
void vmCalcJDevice(vcsbeam_context *vm);
/* Extension (no counterpart in the reference): every pointing's J^-1 in
 * d_J after vmCalcJDevice shared pointing 0's; synchronous. */
void vmExpandJDevice(vcsbeam_context *vm);
/* Upload vm->D to d_D. */
void vmPushD(vcsbeam_context *vm);
/* Device FP64 J^-1 = F (D B)^-1 with a per-fine-channel primary beam shared
 * by all tiles: B is host [npointing][nfine_chan][4] (e.g. from
 * vmSurrogateJones); result in d_J. */
void vmCalcJDeviceChanB(vcsbeam_context *vm, const gpuDoubleComplex *B);

/* Primary-beam Jones for many directions (stands in for the FEE beam the
 * reference evaluates per pointing with hyperbeam, src/primary_beam.c:79-191,
 * 479-532): a project-defined analytic dipole-array beam (4x4 dual-pol short
 * dipoles, 1.1 m grid, 0.278 m over ground, seeded 1 % complex gain and
 * delay jitter per dipole) and a BF16 MLP surrogate (16 -> 256 -> 256 -> 256
 * -> 8, GELU) trained on the GPU by this library's forward/backward kernels,
 * with the analytic beam as fallback and as the measure of its error. */
typedef struct vm_beam_surrogate vm_beam_surrogate;
/* Seeded dipole jitter (gain: complex std dev, delay: +-seconds) and an
 * untrained network (seeded He-uniform weights). */
vm_beam_surrogate *vmSurrogateCreate(uint64_t seed, double gain_jitter, double delay_jitter_s);
void vmSurrogateDestroy(vm_beam_surrogate *s);
/* Adam training on nsamples seeded samples (za <= za_max_deg, 170-230 MHz);
 * returns the last mean-squared loss; loss_curve (optional) receives the
 * mean loss of every 100 steps. */
double vmSurrogateTrain(vm_beam_surrogate *s, long nsamples, int steps, int batch, float lr,
                        double za_max_deg, uint64_t data_seed, float *loss_curve);
/* One forward/backward pass on host features X [B][16] and targets T [B][8]:
 * gradients of the mean-squared loss (packed layer by layer, W [out][in]
 * then b) and the loss. */
double vmSurrogateGrad(vm_beam_surrogate *s, const float *X, const float *T, int B, float *dW,
                       float *db);
/* Get (set = 0, weights as bf16-rounded fp32) or set (set = 1) the packed
 * weights and biases. */
void vmSurrogateParams(vm_beam_surrogate *s, float *W, float *b, int set);
/* The dipole gains [2][16] (re, im) and delays [2][16] (s): 96 doubles. */
void vmSurrogateConfig(const vm_beam_surrogate *s, double *cfg96);
/* Jones B[p][c][4] (host, complex) for directions (az, za)[np] and
 * frequencies freq[nf] (Hz): mode 0 analytic beam, 1 nominal (jitter-free)
 * beam, 2 surrogate, 3 surrogate with fallback to the analytic beam where its
 * norm-wise error exceeds 1e-3.  err (optional, [np][nf]): the surrogate's
 * error against the analytic beam.  Returns the number of fallbacks. */
long vmSurrogateJones(vm_beam_surrogate *s, int mode, const double *az, const double *za,
                      int np, const double *freq, int nf, double *B, float *err);

/* Precision MFMA beamformer for one chunk (vm->beamform_form selects FP16 or
 * BF16); vmBeamformChunk dispatches here when beamform_form != VM_BF_FP32. */
void vmBeamformChunkMFMA(vcsbeam_context *vm);

/* Autotuner of the MFMA beamformer: times every kernel variant (register
 * tile per wavefront, wavefront grid per workgroup) in every tile order on
 * the context's current chunk with HIP events, keeps the fastest for later
 * calls and returns its index (and its time in *best_ms when not NULL).
 * vmSetBeamformerVariant fixes one; order 0 walks channels in turn with row
 * tiles fastest; orders 1-4 give each XCD its own channels and walk a
 * channel's tiles in bands of all (1), one (2), two (3) or four (4) row
 * tiles, column by column. */
int vmTuneBeamformer(vcsbeam_context *vm, int reps, float *best_ms);
void vmSetBeamformerVariant(vcsbeam_context *vm, int variant, int order);
/* The variant and tile order the MFMA beamformer runs with; returns 1 when
 * they were set by vmTuneBeamformer / vmSetBeamformerVariant, 0 for the
 * defaults. */
int vmBeamformerSelection(vcsbeam_context *vm, int *variant, int *order);
/* Number of MFMA beamformer variants and the tile shape of one: rows and
 * samples per workgroup, wavefronts per workgroup. Returns 0 if no such. */
int vmBeamformerVariantInfo(int variant, int *rows, int *samples, int *waves);
/* Name of the GPU kernel a variant launches ("k_beamform_mfma", or
 * "k_beamform_res" for the resident-weights form), for profiles; NULL if no
 * such variant. */
const char *vmBeamformerVariantKernel(int variant);

/* Batched beamforming of several coarse channels (no counterpart in the
 * reference, which runs one process per coarse channel, app/
 * make_mwa_tied_array_beam.c:250-311 once per rank).  Each context holds one
 * coarse channel-second: forward PFB initialised (VM_FLT, not staged, one
 * chunk per second), MFMA form (all the same), S, J^-1 (pointing 0 used),
 * phases and pol lists on the device, and the second uploaded
 * (vmUploadForwardPFBChunk).  Contexts whose forward PFB was initialised
 * without PFB_MALLOC_DEVICE_OUTPUT share one PFB output buffer owned by the
 * batch.  Channels may differ in beam count and flagged tiles: tiles whose
 * J^-1 is zero in every fine channel (D = 0) are left out of that channel's
 * GEMM.  vmBatchBeamformSecond runs, per channel, the fused PFB and the prep
 * pass, then ONE persistent MFMA launch over all channels' work items; with
 * use_graph the sequence is captured in a hipGraph on first use and replayed
 * (re-captured when beams, flags or the schedule change).  Results are in
 * each context's d_S on return.  Destroy the batch before the contexts. */
typedef struct vm_batch vm_batch;
vm_batch *vmBatchCreate(vcsbeam_context **vms, int n);
void vmBatchBeamformSecond(vm_batch *b, int use_graph);
/* Autotuner: times every usable GEMM variant with bands of 1, 2, 4, 8 or all
 * row tiles per work item on the current second; keeps the fastest and
 * returns its time (ms).  vmBatchSetSchedule fixes one (band 0: all). */
float vmBatchTune(vm_batch *b, int reps);
void vmBatchSetSchedule(vm_batch *b, int variant, int band);
/* Event-timed eager run: summed PFB, prep and GEMM times (ms). */
void vmBatchStageTimes(vm_batch *b, float *pfb_ms, float *prep_ms, float *gemm_ms);
/* Number of channels; the schedule in use and its work-item count. */
int vmBatchInfo(vm_batch *b, int *variant, int *band, int *nitems);
/* Tiles of channel g that enter its GEMM (non-zero J^-1). */
int vmBatchActiveTiles(vm_batch *b, int g);
void vmBatchDestroy(vm_batch *b);

/* Seeded MWAX 8+8-bit voltages generated in place on the device: the same
 * counter-based generator (SplitMix64, Box-Muller, sigma, round, clip to
 * +-127) and layout [nblocks][ninput][spb] (re, im) as the host generator of
 * the synthetic workloads, for inputs too large to build on the host. */
void vmSynthMWAXDevice(uint64_t seed, uint64_t first_block, int nblocks, int ninput, int spb,
                       double sigma, int zero_blocks, void *d_out);

/* Device-resident inverse PFB of pointing 0 straight from d_e (no host
 * round trip), one fused kernel that also keeps the last P fine samples for
 * the next second; output complex float [n][pol] on the device (g->d_out),
 * copied to data_buffer_vdif when it is not NULL.  Seconds are expected in
 * order (file_no 0, 1, ...), as cu_invert_pfb expects them. */
void vmInvertPFBDevice(vcsbeam_context *vm, struct gpu_ipfb_arrays *g,
                       int file_no, float *data_buffer_vdif);
/* The same, and the output's 8-bit VDIF samples (as vmRequantVDIFDevice
 * makes them from g->d_out) written by the same kernel into d_vdif
 * (device, n x pol x 2 bytes). */
void vmInvertPFBDeviceVDIF(vcsbeam_context *vm, struct gpu_ipfb_arrays *g, int file_no,
                           uint8_t *d_vdif);
/* vmInvertPFBDeviceVDIF for the nctx contexts of several coarse channels of
 * the same second (equal sample counts, taps and polarisations) in one
 * launch, one grid row per context, on vm[0]'s stream after each context's
 * own stream has drained.  Each context's g[k]->d_out, d_vdif[k] and history
 * end as its own vmInvertPFBDeviceVDIF call would leave them; the kernel
 * time is recorded on vm[0].  No counterpart in the reference, which
 * inverts one coarse channel per process (src/pfb.cpp:940-1102): for
 * callers that hold several channels on one device. */
void vmInvertPFBDeviceVDIFBatch(vcsbeam_context **vm, struct gpu_ipfb_arrays **g, int nctx,
                                int file_no, uint8_t **d_vdif);
/* Passes per workgroup of the device inverse PFB with 12 taps and two
 * polarisations (1 to 64; 0, the default, chooses from the channel and
 * sample counts).  A workgroup walks a run of passes along one channel's
 * samples: the first gives 21 output samples, every later one 32.  The
 * outputs do not depend on it.  Process-wide.  No counterpart in the
 * reference (its inverse PFB has no such tiling, src/pfb.cpp:842-907): for
 * tests and tuning. */
void vmSetInversePFBPasses(int npass);

/* Device VDIF requantisation of n floats (float2int8_trunc + offset binary). */
void vmRequantVDIFDevice(const float *d_in, uint8_t *d_out, size_t n,
                         gpuStream_t stream);

/* VDIF output (src/beam_vdif.c).
 * vmPopulateVDIFHeader fills vm->vf (one struct vdifinfo per pointing: 8-bit
 * complex samples, 2 channels = 2 polarisations, 128 samples per frame,
 * fine_sample_rate frames per second, station "mw", file base name
 * <project>_<obs id>_<RA>_<Dec>_ch<rec chan>) and vm->vhdr (epoch from
 * beam_geom_vals->intmjd, seconds from intmjd and fracmjd, frame 0)
 * (src/beam_vdif.c:150-246).
 * vdif_write_second frames one second of data_buffer_vdif (sizeof_buffer
 * floats [sample][pol][re, im]) on the GPU: each frame is the current
 * header and sizeof_beam values clamped to [-126, 127], rounded half to even
 * and stored as offset binary; the header advances by one frame per frame,
 * as vdifio's nextVDIFHeader does.  It appends the frames to
 * <basefilename>.vdif and writes the DSPSR ASCII header <basefilename>.hdr
 * (vdif_write_data) (src/beam_vdif.c:52-143).  data_buffer_vdif is not
 * modified (the reference clamps it in place).
 * float2int8_trunc and to_offset_binary are the reference's host helpers
 * (src/beam_vdif.c:29-38, 254-260). */
void vmPopulateVDIFHeader(vcsbeam_context *vm, beam_geom *beam_geom_vals, double mjd_start,
                          double sec_offset);
/* vdif_write_second frames on the GPU through device buffers the library
 * keeps between calls (grown as needed, on the current device); concurrent
 * calls from several host threads race on them. */
void vdif_write_second(struct vdifinfo *vf, vdif_header *vhdr, float *data_buffer_vdif);
void vdif_write_data(struct vdifinfo *vf, int8_t *output);
void float2int8_trunc(float *f, int n, float min, float max, int8_t *i);
void to_offset_binary(int8_t *i, int n);
/* Device VDIF frames (no counterpart in the reference, which frames on the
 * host): nframes frames of payload_bytes requantised values each, from the
 * device floats d_in, headers from *hdr onwards (advanced by nframes on
 * return) at frame_rate frames per second, into d_out (nframes x (32 +
 * payload_bytes) bytes).  payload_bytes is a multiple of 4; asynchronous on
 * stream. */
void vmVDIFFramesDevice(const float *d_in, size_t nframes, int payload_bytes, vdif_header *hdr,
                        int frame_rate, uint8_t *d_out, gpuStream_t stream);

/* Synchronise the context's device work. */
void vmSynchronize(vcsbeam_context *vm);
/* The HIP stream all context kernels are issued on. */
gpuStream_t vmGetStream(vcsbeam_context *vm);

/* Device memory, copies and event timing for callers without a GPU runtime
 * binding of their own (exit on error like every other entry point). */
void *vmDeviceMalloc(size_t bytes);
void vmDeviceFree(void *p);
void vmMemcpyHtoD(void *dst, const void *src, size_t bytes);
void vmMemcpyDtoH(void *dst, const void *src, size_t bytes);
/* height rows of width bytes, rows spitch apart on the device, dpitch on the host */
void vmMemcpy2DDtoH(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height);
void vmMemsetDevice(void *dst, int value, size_t bytes);
void vmDeviceSynchronize(void);
void vmDeviceCopy(void *dst, const void *src, size_t bytes); /* plain copy kernel */
/* The device's dense FP16 MFMA rate in TFLOP/s, measured with chains of
 * 32x32x16 FP16 matrix instructions on operands held in registers (no memory
 * traffic): a measured peak beside the published one.  Synchronous. */
double vmMFMAPeakTFLOPS(void);
int vmDeviceCount(void);
void *vmEventCreate(void);
void vmEventRecord(void *ev, gpuStream_t stream);
float vmEventElapsedMs(void *start, void *stop);
void vmEventDestroy(void *ev);

/* Per-kernel timing with HIP events on the context stream: when on, each
 * launch of the named kernels ("pfb", "beamform_fp32", "beamform_mfma",
 * "beamform_prep", "beamform_weights", "ipfb", "geometry", "renormalise") is
 * bracketed by events; vmKernelTimeMs returns the average milliseconds per
 * launch since timing was switched on.  With vm NULL they apply to the calls
 * that take no context: "incoh" (cu_form_incoh_beam's device part: the
 * sum and the renormalisation) and "vdif" (vmRequantVDIFDevice).  The
 * events are recorded without waiting for them (timing adds no host
 * synchronisation to the timed calls); vmKernelTimeMs and vmSetKernelTiming
 * wait for the recorded ones. */
void vmSetKernelTiming(vcsbeam_context *vm, int on);
double vmKernelTimeMs(vcsbeam_context *vm, const char *kernel, long *count);

/* Library build/version string. */
const char *vmLibraryVersion(void);

/* Struct layout probe for foreign-function mirrors: writes sizeof() and the
 * offsets of the fields listed in the implementation into out[0..n). */
int vmStructLayout(const char *name, size_t *out, int n);

#ifdef __cplusplus
}
#endif
#endif /* VCSBEAM_H */
