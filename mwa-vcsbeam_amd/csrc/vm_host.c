/*
 * vm_host.c -- host-side C module of the voltage-beamforming library:
 * stopwatch logger, 2x2 Jones algebra, filter loading, the overlapping read
 * buffer, context set-up, and the host FP64 Jones / phase calculations of
 * the reference's per-second pass.
 *
 * Reference: MWA-vcsbeam src/performance.c, src/jones.c, src/filter.c,
 * src/buffer.c, src/metadata.c, src/geometry.c, src/beam_vdif.c.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "vm_internal.h"

#ifndef VCSBEAM_RUNTIME_DIR
#define VCSBEAM_RUNTIME_DIR "."
#endif

void vm_fatal(const char *file, int line, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "vcsbeam error (%s:%d): ", file, line);
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(EXIT_FAILURE);
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) VM_FATAL("out of host memory (%zu bytes)", n);
    return p;
}

/* ============================================================ logger ==== */
/* src/performance.c:26-425 (names and semantics of the stopwatches) */

static double now_seconds(void)
{
    struct timespec t;
    clock_gettime(CLOCK_REALTIME, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

logger *create_logger(FILE *fout, int world_rank)
{
    logger *log = (logger *)xmalloc(sizeof(logger));
    memset(log, 0, sizeof(logger));
    log->begintime = now_seconds();
    log->fout = fout;
    log->world_rank = world_rank;
    return log;
}

void destroy_logger(logger *log)
{
    if (!log) return;
    for (int i = 0; i < log->nstopwatches; i++) {
        free(log->stopwatches[i].name);
        free(log->stopwatches[i].description);
    }
    free(log);
}

static int find_stopwatch(logger *log, const char *name)
{
    for (int i = 0; i < log->nstopwatches; i++)
        if (strcmp(log->stopwatches[i].name, name) == 0) return i;
    return PERFORMANCE_NO_STOPWATCH_FOUND;
}

void logger_add_stopwatch(logger *log, const char *name, const char *description)
{
    if (!log || log->nstopwatches >= PERFORMANCE_MAX_NUM_STOPWATCHES) return;
    if (find_stopwatch(log, name) != PERFORMANCE_NO_STOPWATCH_FOUND) return;
    logger_stopwatch *sw = &log->stopwatches[log->nstopwatches++];
    memset(sw, 0, sizeof(*sw));
    sw->name = strdup(name);
    sw->description = description ? strdup(description) : NULL;
}

void logger_start_stopwatch(logger *log, const char *name, bool print_description)
{
    if (!log) return;
    int i = find_stopwatch(log, name);
    if (i < 0) return;
    logger_stopwatch *sw = &log->stopwatches[i];
    if (print_description && sw->description && log->fout)
        logger_timed_message(log, sw->description);
    if (sw->nstart_stops < PERFORMANCE_MAX_START_STOP)
        sw->values[sw->nstart_stops] = now_seconds();
    sw->running = true;
}

void logger_stop_stopwatch(logger *log, const char *name)
{
    if (!log) return;
    int i = find_stopwatch(log, name);
    if (i < 0) return;
    logger_stopwatch *sw = &log->stopwatches[i];
    if (!sw->running || sw->nstart_stops >= PERFORMANCE_MAX_START_STOP) return;
    double dt = now_seconds() - sw->values[sw->nstart_stops];
    sw->values[sw->nstart_stops++] = dt;
    sw->total += dt;
    sw->total_sq += dt * dt;
    sw->running = false;
}

void logger_timed_message(logger *log, const char *message)
{
    if (!log || !log->fout) return;
    double t = now_seconds() - log->begintime;
    if (log->world_rank == PERFORMANCE_NO_MPI)
        fprintf(log->fout, "[%f]  %s\n", t, message);
    else
        fprintf(log->fout, "[%f] [MPI:%2d]  %s\n", t, log->world_rank, message);
}

void logger_message(logger *log, const char *message)
{
    if (!log || !log->fout) return;
    fprintf(log->fout, "%s\n", message);
}

void logger_stopwatch_report_stats(logger *log, const char *name)
{
    if (!log || !log->fout) return;
    int i = find_stopwatch(log, name);
    if (i < 0) return;
    logger_stopwatch *sw = &log->stopwatches[i];
    int n = sw->nstart_stops;
    double mean = n ? sw->total / n : 0.0;
    double var = n ? sw->total_sq / n - mean * mean : 0.0;
    fprintf(log->fout, "Total time taken for %-10s: %f s (mean %f +/- %f s over %d)\n",
            sw->name, sw->total, mean, sqrt(var > 0 ? var : 0), n);
}

void logger_report_all_stats(logger *log)
{
    if (!log) return;
    for (int i = 0; i < log->nstopwatches; i++)
        logger_stopwatch_report_stats(log, log->stopwatches[i].name);
}

/* ============================================================= jones ==== */
/* src/jones.c:201-578.  Same operations, own code. */

static inline gpuDoubleComplex C(double x, double y) { gpuDoubleComplex r = {x, y}; return r; }
static inline gpuDoubleComplex Cmul(gpuDoubleComplex a, gpuDoubleComplex b)
{ return C(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
static inline gpuDoubleComplex Cadd(gpuDoubleComplex a, gpuDoubleComplex b) { return C(a.x + b.x, a.y + b.y); }
static inline gpuDoubleComplex Csub(gpuDoubleComplex a, gpuDoubleComplex b) { return C(a.x - b.x, a.y - b.y); }
static inline gpuDoubleComplex Cconj(gpuDoubleComplex a) { return C(a.x, -a.y); }
static inline gpuDoubleComplex Crecip(gpuDoubleComplex z)
{
    double s = 1.0 / (z.x * z.x + z.y * z.y); /* src/jones.c:215-219 */
    return C(s * z.x, -s * z.y);
}

void cp2x2(gpuDoubleComplex *Min, gpuDoubleComplex *Mout)
{
    for (int i = 0; i < 4; i++) Mout[i] = Min[i];
}

void inv2x2(gpuDoubleComplex *Min, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex m00 = Min[0], m01 = Min[1], m10 = Min[2], m11 = Min[3];
    gpuDoubleComplex inv = Crecip(Csub(Cmul(m00, m11), Cmul(m01, m10)));
    gpuDoubleComplex ninv = C(-inv.x, -inv.y);
    Mout[0] = Cmul(inv, m11);
    Mout[1] = Cmul(ninv, m01);
    Mout[2] = Cmul(ninv, m10);
    Mout[3] = Cmul(inv, m00);
}

void inv2x2d(double *Min, double *Mout)
{
    double m00 = Min[0], m01 = Min[1], m10 = Min[2], m11 = Min[3];
    double inv = 1.0 / (m00 * m11 - m01 * m10);
    Mout[0] = inv * m11;
    Mout[1] = -inv * m01;
    Mout[2] = -inv * m10;
    Mout[3] = inv * m00;
}

void inv2x2S(gpuDoubleComplex *Min, gpuDoubleComplex *Mout)
{
    /* Min and Mout must not alias (as in the reference, src/jones.c:294-306) */
    gpuDoubleComplex inv = Crecip(Csub(Cmul(Min[0], Min[3]), Cmul(Min[1], Min[2])));
    gpuDoubleComplex ninv = C(-inv.x, -inv.y);
    Mout[0] = Cmul(inv, Min[3]);
    Mout[1] = Cmul(ninv, Min[1]);
    Mout[2] = Cmul(ninv, Min[2]);
    Mout[3] = Cmul(inv, Min[0]);
}

void mult2x2d(gpuDoubleComplex *M1, gpuDoubleComplex *M2, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex r0 = Cadd(Cmul(M1[0], M2[0]), Cmul(M1[1], M2[2]));
    gpuDoubleComplex r1 = Cadd(Cmul(M1[0], M2[1]), Cmul(M1[1], M2[3]));
    gpuDoubleComplex r2 = Cadd(Cmul(M1[2], M2[0]), Cmul(M1[3], M2[2]));
    gpuDoubleComplex r3 = Cadd(Cmul(M1[2], M2[1]), Cmul(M1[3], M2[3]));
    Mout[0] = r0; Mout[1] = r1; Mout[2] = r2; Mout[3] = r3;
}

void mult2x2d_RxC(double *M1, gpuDoubleComplex *M2, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex r[4];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            gpuDoubleComplex a = M2[j], b = M2[2 + j];
            r[2 * i + j] = C(M1[2 * i] * a.x + M1[2 * i + 1] * b.x,
                             M1[2 * i] * a.y + M1[2 * i + 1] * b.y);
        }
    for (int i = 0; i < 4; i++) Mout[i] = r[i];
}

void mult2x2d_CxR(gpuDoubleComplex *M1, double *M2, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex r[4];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            gpuDoubleComplex a = M1[2 * i], b = M1[2 * i + 1];
            r[2 * i + j] = C(a.x * M2[j] + b.x * M2[2 + j], a.y * M2[j] + b.y * M2[2 + j]);
        }
    for (int i = 0; i < 4; i++) Mout[i] = r[i];
}

void conj2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout)
{
    for (int i = 0; i < 4; i++) Mout[i] = Cconj(M[i]);
}

double norm2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout)
{
    /* Frobenius norm, src/jones.c:404-424 */
    double F = 0.0;
    for (int i = 0; i < 4; i++) F += M[i].x * M[i].x + M[i].y * M[i].y;
    F = sqrt(F);
    for (int i = 0; i < 4; i++)
        Mout[i] = (F == 0.0) ? C(0.0, 0.0) : C(M[i].x / F, M[i].y / F);
    return F;
}

void reverse2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex t[4] = {M[3], M[2], M[1], M[0]};
    for (int i = 0; i < 4; i++) Mout[i] = t[i];
}

void swaprows2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex t[4] = {M[2], M[3], M[0], M[1]};
    for (int i = 0; i < 4; i++) Mout[i] = t[i];
}

void swapcols2x2(gpuDoubleComplex *M, gpuDoubleComplex *Mout)
{
    gpuDoubleComplex t[4] = {M[1], M[0], M[3], M[2]};
    for (int i = 0; i < 4; i++) Mout[i] = t[i];
}

bool is2x2zero(gpuDoubleComplex *M)
{
    for (int i = 0; i < 4; i++)
        if (M[i].x != 0.0 || M[i].y != 0.0) return false;
    return true;
}

void calc_hermitian(gpuDoubleComplex *M, gpuDoubleComplex *H)
{
    gpuDoubleComplex t[4] = {Cconj(M[0]), Cconj(M[2]), Cconj(M[1]), Cconj(M[3])};
    for (int i = 0; i < 4; i++) H[i] = t[i];
}

void calc_coherency_matrix(gpuDoubleComplex *M, gpuDoubleComplex *c)
{
    gpuDoubleComplex H[4];
    calc_hermitian(M, H);
    mult2x2d(M, H, c);
}

void fprintf_complex_matrix(FILE *fout, gpuDoubleComplex *M)
{
    fprintf(fout, "[ %lf%+lf*i, %lf%+lf*i; %lf%+lf*i, %lf%+lf*i ]\n", M[0].x, M[0].y,
            M[1].x, M[1].y, M[2].x, M[2].y, M[3].x, M[3].y);
}

/* vmCalcJ on the host, src/jones.c:97-157 */
void vmCalcJ(vcsbeam_context *vm)
{
    int nant = (int)vm->obs_metadata->num_ants;
    int nchan = vm->nfine_chan;
    int npol = (int)vm->obs_metadata->num_ant_pols;
    for (unsigned int p = 0; p < vm->npointing; p++)
        for (int ch = 0; ch < nchan; ch++)
            for (int ant = 0; ant < nant; ant++) {
                gpuDoubleComplex Ji[4];
                mult2x2d(&vm->D[D_IDX(ant, ch, 0, 0, nchan, npol)],
                         &vm->pb.B[PB_IDX(p, ant, 0, nant, npol * npol)], Ji);
                double F = norm2x2(Ji, Ji);
                gpuDoubleComplex *out = &vm->J[J_IDX(p, ant, ch, 0, 0, nant, nchan, npol)];
                if (F != 0.0) inv2x2S(Ji, out);
                else for (int q = 0; q < 4; q++) out[q] = C(0.0, 0.0);
            }
    /* the next vmPushJ hands D and B to the device as well */
    if (vm->dev) vm->dev->J_host_fromB = 1;
}

/* src/jones.c:170-188: geometry at the middle of the current second, then
 * phases, primary beam and Jones inverses */
void vmCalcJonesAndDelays(vcsbeam_context *vm, double *ras_hours, double *decs_degs,
                          beam_geom *beam_geom_vals)
{
    logger_start_stopwatch(vm->log, "delay", true);
    const uintptr_t t = (uintptr_t)(vm->chunk_to_load / vm->chunks_per_second);
    const double sec_offset =
        (double)t + (double)vm->gps_seconds_to_process[0] - (double)vm->obs_metadata->obs_id;
    const double mjd = vm->obs_metadata->sched_start_mjd + (sec_offset + 0.5) / 86400.0;
    for (unsigned int p = 0; p < vm->npointing; p++)
        calc_beam_geom(ras_hours[p], decs_degs[p], mjd, &beam_geom_vals[p]);
    vmCalcPhi(vm, beam_geom_vals);
    vmCalcB(vm, beam_geom_vals);
    vmCalcJ(vm);
    logger_stop_stopwatch(vm->log, "delay");
}

/* ====================================================== primary beam ==== */
/* src/primary_beam.c:250-297: delays point into the metadata; amplitudes
 * are 0 for a dead dipole (delay 32) and 1 otherwise */
void create_delays_amps_from_metafits(MetafitsMetadata *md, uint32_t ***delays, double ***amps)
{
    const size_t ni = md->num_rf_inputs, nd = md->num_delays;
    *delays = (uint32_t **)xmalloc(sizeof(uint32_t *) * (ni ? ni : 1));
    *amps = (double **)xmalloc(sizeof(double *) * (ni ? ni : 1));
    for (size_t i = 0; i < ni; i++) {
        (*delays)[i] = md->rf_inputs[i].dipole_delays;
        (*amps)[i] = (double *)xmalloc(sizeof(double) * (nd ? nd : 1));
        for (size_t d = 0; d < nd; d++)
            (*amps)[i][d] = md->rf_inputs[i].dipole_delays &&
                                    md->rf_inputs[i].dipole_delays[d] == MWA_DEAD_DIPOLE_DELAY
                                ? 0.0
                                : 1.0;
    }
}

void free_delays_amps(MetafitsMetadata *md, uint32_t **delays, double **amps)
{
    if (amps)
        for (size_t i = 0; i < md->num_rf_inputs; i++) free(amps[i]);
    free(amps);
    free(delays);
}

/* src/primary_beam.c:355-414.  Configurations are numbered by how many of
 * the 16 dipoles are dead: none (0), one at d (1 + d), two at d1 < d2
 * (17 onwards, pairs in lexicographic order: the pairs before d1 number
 * 15 + 14 + ... = 16 d1 - d1 (d1 + 1) / 2, so the index is
 * 17 + 16 d1 - d1 (d1 + 1) / 2 + (d2 - d1 - 1)), all sixteen (137). */
int hash_dipole_config(double *amps)
{
    int dead[16], ndead = 0;
    for (int i = 0; i < 16; i++)
        if (amps[i] == 0.0) dead[ndead++] = i;
    switch (ndead) {
    case 0: return 0;
    case 1: return 1 + dead[0];
    case 2: return 17 + 16 * dead[0] - dead[0] * (dead[0] + 1) / 2 + (dead[1] - dead[0] - 1);
    case 16: return NCONFIGS - 1;
    default: return MANY_DEAD_DIPOLES;
    }
}

static void check_rf_inputs(const MetafitsMetadata *md);

/* vmSetPolIdxLists, src/jones.c:40-64 */
void vmSetPolIdxLists(vcsbeam_context *vm)
{
    check_rf_inputs(vm->obs_metadata);
    for (size_t i = 0; i < vm->obs_metadata->num_rf_inputs; i++) {
        Rfinput *rf = &vm->obs_metadata->rf_inputs[i];
        if (*(rf->pol) == 'X') vm->polQ_idxs[rf->ant] = rf->vcs_order;
        else vm->polP_idxs[rf->ant] = rf->vcs_order;
    }
}

/* ============================================================ filter ==== */
/* src/filter.c:21-177 */

gpuDoubleComplex *roots_of_unity(int N)
{
    gpuDoubleComplex *r = (gpuDoubleComplex *)xmalloc(sizeof(gpuDoubleComplex) * N);
    for (int n = 0; n < N; n++) {
        double th = 2.0 * M_PI * (double)n / (double)N; /* e^{+2 pi i n/N} */
        r[n] = C(cos(th), sin(th));
    }
    return r;
}

void free_pfb_filter(pfb_filter *filter)
{
    if (!filter) return;
    free(filter->coeffs);
    free(filter->twiddles);
    free(filter);
}

static pfb_filter **filter_slot(vcsbeam_context *vm, filter_type type)
{
    return type == ANALYSIS_FILTER ? &vm->analysis_filter : &vm->synth_filter;
}

void vmSetFilter(vcsbeam_context *vm, const double *coeffs, int ncoeffs,
                 filter_type type, int nchans)
{
    if (ncoeffs <= 0) VM_FATAL("vmSetFilter: no coefficients");
    pfb_filter **slot = filter_slot(vm, type);
    if (*slot) free_pfb_filter(*slot);
    pfb_filter *f = (pfb_filter *)xmalloc(sizeof(pfb_filter));
    f->nchans = nchans;
    f->ncoeffs = ncoeffs;
    f->type = type;
    f->coeffs = (double *)xmalloc(sizeof(double) * ncoeffs);
    memcpy(f->coeffs, coeffs, sizeof(double) * ncoeffs);
    if (ncoeffs % nchans != 0)
        fprintf(stderr, "warning: vmLoadFilter: number of channels (%d) does not "
                        "divide evenly into the number of coefficients (%d)\n",
                nchans, ncoeffs);
    f->ntaps = ncoeffs / nchans;
    f->twiddles = roots_of_unity(nchans);
    *slot = f;
}

void vmLoadFilter(vcsbeam_context *vm, char *filtername, filter_type type, int nchans)
{
    if (filtername == NULL) VM_FATAL("vmLoadFilter: arguments must not be NULL");
    const char *dir = getenv("VCSBEAM_RUNTIME_DIR");
    if (!dir) dir = VCSBEAM_RUNTIME_DIR;
    char path[2048];
    snprintf(path, sizeof(path), "%s/%s.dat", dir, filtername);
    FILE *f = fopen(path, "r");
    if (!f) VM_FATAL("vmLoadFilter: could not find filter '%s'", filtername);
    size_t cap = 2048, n = 0;
    double *buf = (double *)xmalloc(sizeof(double) * cap), d;
    while (fscanf(f, "%lf", &d) == 1) {
        if (n == cap) {
            cap *= 2;
            buf = (double *)realloc(buf, sizeof(double) * cap);
            if (!buf) VM_FATAL("out of memory");
        }
        buf[n++] = d;
    }
    fclose(f);
    if (n == 0) VM_FATAL("vmLoadFilter: No coefficients found in '%s'", filtername);
    vmSetFilter(vm, buf, (int)n, type, nchans);
    free(buf);
}

void vmScaleFilterCoeffs(vcsbeam_context *vm, filter_type type, double scale_factor)
{
    pfb_filter *f = *filter_slot(vm, type);
    if (!f) return;
    for (int i = 0; i < f->ncoeffs; i++) f->coeffs[i] *= scale_factor;
}

/* ======================================================== read buffer ==== */
/* src/buffer.c:52-124.  The pinned allocation is done by the device module
 * (vm_host_alloc_pinned) so that this file stays free of HIP. */

host_buffer *vmInitReadBuffer(size_t read_size, size_t copy_size)
{
    if (read_size == 0) return NULL;
    host_buffer *rb = (host_buffer *)xmalloc(sizeof(host_buffer));
    rb->buffer_size = read_size + copy_size;
    rb->buffer = vm_host_alloc_pinned(rb->buffer_size);
    /* The reference leaves the margin uninitialised on the first second
     * (src/buffer.c:63).  Here it is defined as zeros: the first
     * vmReadBufferCopyMargin copies the tail of the read region into the
     * margin, so that tail is zeroed too (recycled pinned memory would
     * otherwise leak an earlier buffer's contents into the first spectra). */
    memset(rb->buffer, 0, copy_size);
    if (copy_size) memset((char *)rb->buffer + read_size, 0, copy_size);
    rb->read_size = read_size;
    rb->copy_size = copy_size;
    rb->read_ptr = (char *)rb->buffer + copy_size;
    rb->copy_from_ptr = (char *)rb->buffer + read_size;
    rb->copy_to_ptr = rb->buffer;
    rb->locked = false;
    return rb;
}

void vmFreeReadBuffer(host_buffer *rb)
{
    if (!rb) return;
    vm_host_free_pinned(rb->buffer);
    free(rb);
}

vm_error vmReadBufferCopyMargin(host_buffer *rb)
{
    if (rb == NULL || rb->buffer == NULL) return VM_READ_BUFFER_NOT_SET;
    if (rb->copy_from_ptr == NULL || rb->copy_to_ptr == NULL || rb->copy_size == 0)
        return VM_SUCCESS;
    memcpy(rb->copy_to_ptr, rb->copy_from_ptr, rb->copy_size);
    return VM_SUCCESS;
}

/* ============================================================ context ==== */

void vmCheckError(vm_error err)
{
    /* src/metadata.c:1740-1759 */
    switch (err) {
    case VM_SUCCESS: return;
    case VM_END_OF_DATA: VM_FATAL("no more data to read"); break;
    case VM_READ_BUFFER_NOT_SET: VM_FATAL("read buffer has not been set"); break;
    case VM_READ_BUFFER_LOCKED: VM_FATAL("read buffer is locked"); break;
    default: VM_FATAL("unknown error code %d", (int)err);
    }
}

vcsbeam_context *vmInit(bool use_mpi)
{
    /* src/metadata.c:26-147.  MPI is not part of this library: a caller that
     * wants one rank per coarse channel sets mpi_rank / coarse_chan_idx. */
    vcsbeam_context *vm = (vcsbeam_context *)xmalloc(sizeof(vcsbeam_context));
    memset(vm, 0, sizeof(vcsbeam_context));
    vm->use_mpi = use_mpi;
    vm->mpi_size = 1;
    vm->mpi_rank = PERFORMANCE_NO_MPI;
    vm->writer = 0;
    vm->ncoarse_chans = vm->mpi_size;
    vm->coarse_chan_idx = 0;
    vm->chunk_to_load = 0;
    vm->datatype = VM_INT4;
    vm->chunks_per_second = 1;
    vm->out_nstokes = 1;
    vm->beamform_form = VM_BF_FP32;
    vm->log = create_logger(NULL, vm->mpi_rank);
    static const char *sw[][2] = {
        {"read", "Reading in data"},
        {"upload", "Uploading the data to the device"},
        {"pfb", "Performing the PFB"},
        {"pfb-wola", "Weighted overlap-add"},
        {"pfb-round", "FPGA rounding and demotion"},
        {"pfb-fft", "Performing FFT"},
        {"pfb-pack", "Packing the data into the recombined format"},
        {"delay", "Calculating geometric and cable delays"},
        {"calc", "Calculating tied-array beam"},
        {"ipfb", "Inverting the PFB"},
        {"download", "Downloading the data to the host"},
        {"splice", "Splicing coarse channels together"},
        {"write", "Writing out data to file"}};
    for (size_t i = 0; i < sizeof(sw) / sizeof(sw[0]); i++)
        logger_add_stopwatch(vm->log, sw[i][0], sw[i][1]);
    return vm;
}

void vmSetOutputChannelisation(vcsbeam_context *vm, bool out_fine, bool out_coarse)
{
    /* src/metadata.c:410-430 */
    vm->output_fine_channels = out_fine;
    vm->output_coarse_channels = out_coarse;
    if (!out_fine && !out_coarse) {
        vm->do_forward_pfb = false;
        vm->do_inverse_pfb = false;
        return;
    }
    vm->do_inverse_pfb = vm->output_coarse_channels;
    vm->do_forward_pfb = (vm->obs_metadata->mwa_version == VCSMWAXv2);
}

void vmMallocVHost(vcsbeam_context *vm)
{
    /* src/metadata.c:448-454 */
    if (vm->obs_metadata->mwa_version == VCSLegacyRecombined)
        vm->v = vmInitReadBuffer(vm->bytes_per_second, 0);
    else if (vm->obs_metadata->mwa_version == VCSMWAXv2)
        vm->v = vmInitReadBuffer(vm->bytes_per_second,
                                 vm->vcs_metadata->voltage_block_size_bytes);
    vm->v_size_bytes = vm->bytes_per_second;
}

void vmFreeVHost(vcsbeam_context *vm)
{
    vmFreeReadBuffer(vm->v);
    vm->v = NULL;
}

void vmSetNumNotFlaggedRFInputs(vcsbeam_context *vm)
{
    /* src/metadata.c:1554-1561: counts RF inputs, not tiles */
    vm->num_not_flagged = 0;
    for (size_t i = 0; i < vm->obs_metadata->num_rf_inputs; i++)
        if (!vm->obs_metadata->rf_inputs[i].flagged) vm->num_not_flagged++;
}

void vmSetNumPointings(vcsbeam_context *vm, unsigned int npointings)
{
    /* src/metadata.c:1111-1119 */
    size_t npol = vm->obs_metadata->num_ant_pols;
    vm->npointing = npointings;
    vm->S_size_bytes = (size_t)vm->npointing * vm->nchan * vm->out_nstokes *
                       vm->sample_rate * sizeof(float);
    vm->d_S_size_bytes = vm->S_size_bytes;
    vm->e_size_bytes = (size_t)vm->npointing * vm->sample_rate * vm->nchan * npol *
                       sizeof(gpuDoubleComplex);
    vm->d_e_size_bytes = vm->e_size_bytes;
}

/* Every index the kernels take from the RF inputs must stay inside the
 * buffers: antennas < num_ants, pols X or Y with each (antenna, pol) once,
 * and vcs_order a permutation of 0 .. num_rf_inputs - 1. */
static void check_rf_inputs(const MetafitsMetadata *md)
{
    const size_t n = md->num_rf_inputs;
    if (!md->rf_inputs || n == 0 || n != 2 * md->num_ants)
        VM_FATAL("metadata: %zu RF inputs for %zu antennas", n, md->num_ants);
    uint8_t *seen_order = (uint8_t *)calloc(n, 1), *seen_input = (uint8_t *)calloc(n, 1);
    if (!seen_order || !seen_input) VM_FATAL("out of host memory");
    for (size_t i = 0; i < n; i++) {
        const Rfinput *rf = &md->rf_inputs[i];
        if (!rf->pol || (rf->pol[0] != 'X' && rf->pol[0] != 'Y'))
            VM_FATAL("metadata: RF input %zu has no X/Y polarisation", i);
        if (rf->ant >= md->num_ants)
            VM_FATAL("metadata: RF input %zu has antenna %u of %zu", i, rf->ant, md->num_ants);
        if (rf->vcs_order >= n || seen_order[rf->vcs_order]++)
            VM_FATAL("metadata: vcs_order %u of RF input %zu is outside 0..%zu or repeated",
                     rf->vcs_order, i, n - 1);
        const size_t slot = 2 * (size_t)rf->ant + (rf->pol[0] == 'Y');
        if (seen_input[slot]++)
            VM_FATAL("metadata: antenna %u polarisation %c appears twice", rf->ant, rf->pol[0]);
    }
    free(seen_order);
    free(seen_input);
}

/* the channel, second and size set-up shared by vmBindObsMetadata and
 * vmBindObsData (src/metadata.c:177-232) */
static void bind_common(vcsbeam_context *vm, int first_coarse_chan_idx, int num_coarse_chans,
                        uint32_t first_gps_second, int num_gps_seconds)
{
    MetafitsMetadata *md = vm->obs_metadata;
    VoltageMetadata *vcs_metadata = vm->vcs_metadata;
    if (num_coarse_chans < 1 || first_coarse_chan_idx < 0 ||
        (size_t)(first_coarse_chan_idx + num_coarse_chans) > md->num_metafits_coarse_chans)
        VM_FATAL("bind: coarse channels %d..%d of %zu", first_coarse_chan_idx,
                 first_coarse_chan_idx + num_coarse_chans - 1, md->num_metafits_coarse_chans);
    vm->num_gps_seconds_to_process = num_gps_seconds;
    vm->num_coarse_chans_to_process = num_coarse_chans;
    free(vm->gps_seconds_to_process);
    free(vm->coarse_chan_idxs_to_process);
    free(vm->cal_coarse_chan_idxs_to_process);
    vm->gps_seconds_to_process = (uint32_t *)xmalloc(sizeof(uint32_t) * num_gps_seconds);
    for (int g = 0; g < num_gps_seconds; g++) vm->gps_seconds_to_process[g] = first_gps_second + g;
    vm->coarse_chan_idxs_to_process = (int *)xmalloc(sizeof(int) * num_coarse_chans);
    vm->cal_coarse_chan_idxs_to_process = (int *)xmalloc(sizeof(int) * num_coarse_chans);
    for (int c = 0; c < num_coarse_chans; c++) {
        vm->coarse_chan_idxs_to_process[c] = first_coarse_chan_idx + c;
        vm->cal_coarse_chan_idxs_to_process[c] = c;
    }
    vm->coarse_chan_idx = first_coarse_chan_idx;
    vmSetOutputChannelisation(vm, md->mwa_version == VCSLegacyRecombined,
                              md->mwa_version == VCSMWAXv2);
    vm->sample_rate = (int)(vcs_metadata->num_samples_per_voltage_block *
                            vcs_metadata->num_voltage_blocks_per_second);
    vm->bytes_per_second = (int)(vcs_metadata->num_voltage_blocks_per_second *
                                 vcs_metadata->voltage_block_size_bytes);
    vm->nchan = (int)md->num_volt_fine_chans_per_coarse;
    if (md->mwa_version == VCSLegacyRecombined) {
        vm->fine_sample_rate = vm->sample_rate;
        vm->nfine_chan = vm->nchan;
        vm->datatype = VM_INT4;
    } else {
        vm->fine_sample_rate = 0;
        vm->nfine_chan = 0;
    }
    vmSetNumNotFlaggedRFInputs(vm);
    vmMallocVHost(vm);
}

void vmBindObsMetadata(vcsbeam_context *vm, MetafitsMetadata *obs_metadata,
                       VoltageMetadata *vcs_metadata, int coarse_chan_idx,
                       uint32_t first_gps_second, int num_gps_seconds)
{
    /* vmBindObsData, src/metadata.c:166-233, with the mwalib look-ups replaced
     * by caller-supplied metadata */
    check_rf_inputs(obs_metadata);
    vm->obs_metadata = obs_metadata;
    vm->vcs_metadata = vcs_metadata;
    bind_common(vm, coarse_chan_idx, 1, first_gps_second, num_gps_seconds);
}

void vmBindObsData(vcsbeam_context *vm, char *first_coarse_chan_str, int num_coarse_chans_to_process,
                   int coarse_chan_idx_offset, char *first_gps_second_str,
                   int num_gps_seconds_to_process, int gps_second_offset, char *datadir)
{
    /* src/metadata.c:166-233: the observation from vmLoadObsMetafits, the
     * voltages from the MWAX subfiles or legacy files in datadir */
    if (!vm->obs_metadata) VM_FATAL("vmBindObsData: no metafits loaded (vmLoadObsMetafits)");
    check_rf_inputs(vm->obs_metadata);
    const uint32_t first_gps =
        (uint32_t)(parse_begin_string(vm->obs_metadata, first_gps_second_str) + gps_second_offset);
    const int first_cc =
        (int)parse_coarse_chan_string(vm->obs_metadata, first_coarse_chan_str) + coarse_chan_idx_offset;
    free(vm->datadir);
    vm->datadir = strdup(datadir ? datadir : ".");
    if (!vm->datadir) VM_FATAL("out of host memory");
    vmGetVoltageMetadata(vm);
    bind_common(vm, first_cc, num_coarse_chans_to_process, first_gps, num_gps_seconds_to_process);
}

vm_error vmReadNextSecond(vcsbeam_context *vm)
{
    /* src/metadata.c:925-993 with the mwalib read replaced by the registered
     * reader callback */
    if (vm->v == NULL) return VM_READ_BUFFER_NOT_SET;
    if (vm->v->locked) return VM_READ_BUFFER_LOCKED;
    if (vm->current_gps_idx >= (size_t)vm->num_gps_seconds_to_process) return VM_END_OF_DATA;
    vm->v->locked = true;
    logger_start_stopwatch(vm->log, "read", true);
    vmReadBufferCopyMargin(vm->v);
    if (!vm->read_second) VM_FATAL("vmReadNextSecond: no voltage reader registered");
    if (vm->read_second(vm->read_second_user, vm->gps_seconds_to_process[vm->current_gps_idx],
                        vm->coarse_chan_idxs_to_process[0], vm->v->read_ptr,
                        vm->v->read_size) != 0)
        VM_FATAL("vmReadNextSecond: voltage reader failed");
    logger_stop_stopwatch(vm->log, "read");
    vm->current_gps_idx++;
    return VM_SUCCESS;
}

vm_error vmReadNextSecondFrom(vcsbeam_context *vm, const void *src, size_t nbytes)
{
    if (vm->v == NULL) return VM_READ_BUFFER_NOT_SET;
    if (vm->v->locked) return VM_READ_BUFFER_LOCKED;
    if (vm->current_gps_idx >= (size_t)vm->num_gps_seconds_to_process) return VM_END_OF_DATA;
    if (nbytes != vm->v->read_size)
        VM_FATAL("vmReadNextSecondFrom: %zu bytes given, %zu expected", nbytes, vm->v->read_size);
    vm->v->locked = true;
    vmReadBufferCopyMargin(vm->v);
    memcpy(vm->v->read_ptr, src, nbytes);
    vm->current_gps_idx++;
    return VM_SUCCESS;
}

void vmReportPerformanceStats(vcsbeam_context *vm) { logger_report_all_stats(vm->log); }

/* ========================================================== geometry ==== */
/* src/geometry.c:40-171, 440-467 (host FP64) */

void calc_geometric_delays(beam_geom *bg, uint32_t freq_hz,
                           MetafitsMetadata *obs_metadata, gpuDoubleComplex *phi)
{
    const double Href = MWA_ALTITUDE_METRES;
    for (size_t i = 0; i < obs_metadata->num_rf_inputs; i++) {
        Rfinput *rf = &obs_metadata->rf_inputs[i];
        if (*(rf->pol) == 'Y') continue;
        double w = rf->east_m * bg->unit_E + rf->north_m * bg->unit_N +
                   (rf->height_m - Href) * bg->unit_H;
        double dt = (w - rf->electrical_length_m) / SPEED_OF_LIGHT_IN_VACUUM_M_PER_S;
        double phase = 2.0 * M_PI * dt * freq_hz;
        phi[rf->ant] = C(cos(phase), sin(phase));
    }
}

void vmCalcPhi(vcsbeam_context *vm, beam_geom *beam_geom_vals)
{
    geometric_delays *g = &vm->gdelays;
    gpuDoubleComplex *tmp = (gpuDoubleComplex *)xmalloc(sizeof(gpuDoubleComplex) * g->nant);
    for (uintptr_t p = 0; p < g->npointings; p++)
        for (uintptr_t c = 0; c < g->nchan; c++) {
            calc_geometric_delays(&beam_geom_vals[p], (uint32_t)g->chan_freqs_hz[c],
                                  g->obs_metadata, tmp);
            for (uintptr_t a = 0; a < g->nant; a++)
                g->phi[PHI_IDX(p, a, c, g->nant, g->nchan)] = tmp[a];
        }
    free(tmp);
}

double calc_array_factor(MetafitsMetadata *obs_metadata, uint32_t freq_hz,
                         beam_geom *bg1, beam_geom *bg2)
{
    size_t nant = obs_metadata->num_ants;
    gpuDoubleComplex *w = (gpuDoubleComplex *)xmalloc(sizeof(gpuDoubleComplex) * nant);
    gpuDoubleComplex *psi = (gpuDoubleComplex *)xmalloc(sizeof(gpuDoubleComplex) * nant);
    calc_geometric_delays(bg1, freq_hz, obs_metadata, w);
    calc_geometric_delays(bg2, freq_hz, obs_metadata, psi);
    gpuDoubleComplex acc = C(0.0, 0.0);
    for (size_t a = 0; a < nant; a++) acc = Cadd(acc, Cmul(Cconj(w[a]), psi[a]));
    free(w);
    free(psi);
    double af = hypot(acc.x, acc.y) / nant;
    return af * af;
}

/* ============================================================== VDIF ==== */
/* src/beam_vdif.c:29-38, 254-260 */
void float2int8_trunc(float *f, int n, float min, float max, int8_t *i)
{
    for (int j = 0; j < n; j++) {
        f[j] = (f[j] > max) ? max : f[j];
        f[j] = (f[j] < min) ? min : f[j];
        i[j] = (int8_t)rint(f[j]);
    }
}

void to_offset_binary(int8_t *i, int n)
{
    for (int j = 0; j < n; j++) i[j] = (int8_t)(i[j] ^ 0x80);
}

/* ============================================================ layout ==== */
#define OFF(T, f) offsetof(T, f)
int vmStructLayout(const char *name, size_t *out, int n)
{
    size_t tmp[64];
    int k = 0;
#define PUT(x) tmp[k++] = (x)
    if (strcmp(name, "vcsbeam_context") == 0) {
        PUT(sizeof(vcsbeam_context));
        PUT(OFF(vcsbeam_context, obs_metadata));
        PUT(OFF(vcsbeam_context, cal));
        PUT(OFF(vcsbeam_context, fpfb));
        PUT(OFF(vcsbeam_context, v));
        PUT(OFF(vcsbeam_context, d_v));
        PUT(OFF(vcsbeam_context, S));
        PUT(OFF(vcsbeam_context, e));
        PUT(OFF(vcsbeam_context, J));
        PUT(OFF(vcsbeam_context, D));
        PUT(OFF(vcsbeam_context, polP_idxs));
        PUT(OFF(vcsbeam_context, gdelays));
        PUT(OFF(vcsbeam_context, pb));
        PUT(OFF(vcsbeam_context, npointing));
        PUT(OFF(vcsbeam_context, chunks_per_second));
        PUT(OFF(vcsbeam_context, gps_seconds_to_process));
        PUT(OFF(vcsbeam_context, do_inverse_pfb));
        PUT(OFF(vcsbeam_context, sample_rate));
        PUT(OFF(vcsbeam_context, out_nstokes));
        PUT(OFF(vcsbeam_context, vhdr));
        PUT(OFF(vcsbeam_context, d_offsets));
        PUT(OFF(vcsbeam_context, offsets_size));
        PUT(OFF(vcsbeam_context, chunk_to_load));
        PUT(OFF(vcsbeam_context, streams));
        PUT(OFF(vcsbeam_context, log_message));
        PUT(OFF(vcsbeam_context, dev));
        PUT(OFF(vcsbeam_context, beamform_form));
        PUT(OFF(vcsbeam_context, bf_err));
    } else if (strcmp(name, "forward_pfb") == 0) {
        PUT(sizeof(forward_pfb));
        PUT(OFF(forward_pfb, vcs_data));
        PUT(OFF(forward_pfb, d_vcs_data));
        PUT(OFF(forward_pfb, vcs_size));
        PUT(OFF(forward_pfb, d_weighted_overlap_add));
        PUT(OFF(forward_pfb, nspectra));
        PUT(OFF(forward_pfb, P));
        PUT(OFF(forward_pfb, flags));
        PUT(OFF(forward_pfb, d_flt_out));
        PUT(OFF(forward_pfb, filter_fits_int16));
    } else if (strcmp(name, "MetafitsMetadata") == 0) {
        PUT(sizeof(MetafitsMetadata));
        PUT(OFF(MetafitsMetadata, sched_start_mjd));
        PUT(OFF(MetafitsMetadata, rf_inputs));
        PUT(OFF(MetafitsMetadata, metafits_coarse_chans));
        PUT(OFF(MetafitsMetadata, coarse_chan_width_hz));
        PUT(OFF(MetafitsMetadata, project_id));
        PUT(OFF(MetafitsMetadata, num_corr_fine_chans_per_coarse));
        PUT(OFF(MetafitsMetadata, corr_fine_chan_width_hz));
        PUT(OFF(MetafitsMetadata, good_time_gps_ms));
        PUT(OFF(MetafitsMetadata, num_metafits_timesteps));
        PUT(OFF(MetafitsMetadata, num_delays));
    } else if (strcmp(name, "Rfinput") == 0) {
        PUT(sizeof(Rfinput));
        PUT(OFF(Rfinput, pol));
        PUT(OFF(Rfinput, electrical_length_m));
        PUT(OFF(Rfinput, vcs_order));
        PUT(OFF(Rfinput, flagged));
        PUT(OFF(Rfinput, dipole_delays));
    } else if (strcmp(name, "gpu_ipfb_arrays") == 0) {
        PUT(sizeof(struct gpu_ipfb_arrays));
        PUT(OFF(struct gpu_ipfb_arrays, in_real));
        PUT(OFF(struct gpu_ipfb_arrays, d_out));
    } else if (strcmp(name, "vm_psrfits") == 0) {
        PUT(sizeof(vm_psrfits));
        PUT(OFF(vm_psrfits, hdr));
        PUT(OFF(vm_psrfits, basefilename));
        PUT(OFF(vm_psrfits, fptr));
        PUT(OFF(vm_psrfits, data_start));
    } else if (strcmp(name, "vdifinfo") == 0) {
        PUT(sizeof(struct vdifinfo));
        PUT(OFF(struct vdifinfo, samples_per_frame));
        PUT(OFF(struct vdifinfo, stationid));
        PUT(OFF(struct vdifinfo, block_size));
        PUT(OFF(struct vdifinfo, b_scales));
        PUT(OFF(struct vdifinfo, fctr));
        PUT(OFF(struct vdifinfo, ra_str));
        PUT(OFF(struct vdifinfo, BW));
        PUT(OFF(struct vdifinfo, date_obs));
        PUT(OFF(struct vdifinfo, basefilename));
        PUT(OFF(struct vdifinfo, got_scales));
    } else {
        return -1;
    }
#undef PUT
    for (int j = 0; j < k && j < n; j++) out[j] = tmp[j];
    return k;
}

const char *vmLibraryVersion(void)
{
    return "vcsbeam-mi355x 0.1 (gfx950)";
}
