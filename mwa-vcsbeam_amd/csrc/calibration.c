/*
 * calibration.c -- calibration solutions into vm->D: the Offringa binary and
 * RTS text loaders, the corrections applied to D before the beamformer
 * (reference-tile phase, cross terms, PQ phase slope), the PQ correction
 * table and per-tile flag lists.
 *
 * Reference: MWA-vcsbeam src/calibration.c (vmLoadRTSSolution 51-208,
 * read_dijones_file 227-260, read_bandpass_file 281-409,
 * vmLoadOffringaSolution 426-582, remove_reference_phase 617-637,
 * zero_PQ_and_QP 650-654, parse_calibration_correction_file 663-735,
 * vmApplyCalibrationCorrections 747-882, vmParseFlaggedTilenamesFile
 * 890-919, tilename_is_flagged 930-942, vmSetCustomTileFlags 952-1013,
 * init_calibration / free_calibration 1021-1051), src/metadata.c:1575-1594
 * (find_matching_rf_input).
 *
 * All of this is per-second set-up on nant x nchan 2x2 matrices (tens of
 * kilobytes): host code, as in the reference.  D reaches the device through
 * vmPushD / vmPushJ.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#ifndef VCSBEAM_RUNTIME_DIR
#define VCSBEAM_RUNTIME_DIR "."
#endif

static gpuDoubleComplex cdiv(gpuDoubleComplex a, gpuDoubleComplex b)
{
    const double d = b.x * b.x + b.y * b.y;
    gpuDoubleComplex r = {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
    return r;
}

static gpuDoubleComplex cmul(gpuDoubleComplex a, gpuDoubleComplex b)
{
    gpuDoubleComplex r = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
    return r;
}

/* the RF input of md with rf's tile id and polarisation, or NULL
 * (src/metadata.c:1575-1594) */
Rfinput *find_matching_rf_input(MetafitsMetadata *md, Rfinput *rf)
{
    for (size_t i = 0; i < md->num_rf_inputs; i++)
        if (md->rf_inputs[i].tile_id == rf->tile_id && md->rf_inputs[i].pol[0] == rf->pol[0])
            return &md->rf_inputs[i];
    return NULL;
}

/* antenna number of the tile called name, or -1 */
static long ant_by_tile_name(const MetafitsMetadata *md, const char *name)
{
    for (size_t i = 0; i < md->num_rf_inputs; i++)
        if (md->rf_inputs[i].tile_name && strcmp(md->rf_inputs[i].tile_name, name) == 0)
            return (long)md->rf_inputs[i].ant;
    return -1;
}

static void need_D(vcsbeam_context *vm, size_t nant, size_t nchan)
{
    if (!vm->D) VM_FATAL("calibration: vm->D is not allocated (vmMallocDHost)");
    if (vm->D_size_bytes < nant * nchan * 4 * sizeof(gpuDoubleComplex))
        VM_FATAL("calibration: vm->D holds %zu bytes, %zu needed", vm->D_size_bytes,
                 nant * nchan * 4 * sizeof(gpuDoubleComplex));
}

/* ------------------------------------------------------------- Offringa -- */
void vmLoadOffringaSolution(vcsbeam_context *vm)
{
    /* Binary layout: 16 bytes of intro and type words, then uint32 interval,
     * antenna, channel and polarisation counts, two doubles (start and end
     * time): 48 bytes.  Then [interval][antenna][channel] 2x2 complex
     * doubles (64 bytes each).  Only the first interval is used.  A matrix
     * whose first element is NaN becomes zero; each solution channel covers
     * nfine_chan / nchan consecutive fine channels. */
    MetafitsMetadata *cm = vm->cal_metadata;
    if (!cm || !vm->cal.caldir || !vm->cal_coarse_chan_idxs_to_process)
        VM_FATAL("vmLoadOffringaSolution: calibration metadata, file or channel not set");
    const int cidx = vm->cal_coarse_chan_idxs_to_process[0];
    FILE *fp = fopen(vm->cal.caldir, "rb");
    if (!fp) VM_FATAL("vmLoadOffringaSolution: cannot open %s", vm->cal.caldir);
    uint32_t counts[4];
    if (fseek(fp, 16, SEEK_SET) != 0 || fread(counts, sizeof(uint32_t), 4, fp) != 4)
        VM_FATAL("vmLoadOffringaSolution: %s: short header", vm->cal.caldir);
    const uint32_t nintervals = counts[0], file_nant = counts[1], file_nchan = counts[2];
    const size_t nant = cm->num_ants, vcs_nchan = (size_t)vm->nfine_chan;
    size_t nchan = cm->num_corr_fine_chans_per_coarse;
    size_t nChan = nchan * (size_t)(vm->mpi_size > 0 ? vm->mpi_size : 1);
    if (nintervals > 1)
        fprintf(stdout, "Warning: %s holds %u time intervals; using interval 0 only\n",
                vm->cal.caldir, nintervals);
    if (file_nant != nant)
        VM_FATAL("vmLoadOffringaSolution: %s has %u antennas, the calibration metadata %zu",
                 vm->cal.caldir, file_nant, nant);
    if (file_nchan != nChan) {
        fprintf(stdout, "Warning: %s has %u channels where %zu were expected; using the "
                        "file's count\n", vm->cal.caldir, file_nchan, nChan);
        nChan = file_nchan;
        nchan = nChan / (size_t)(vm->mpi_size > 0 ? vm->mpi_size : 1);
    }
    if (cidx >= (int)cm->num_metafits_coarse_chans)
        VM_FATAL("vmLoadOffringaSolution: coarse channel index %d is past the %zu channels "
                 "of %s", cidx, cm->num_metafits_coarse_chans, vm->cal.caldir);
    if (nchan == 0 || vcs_nchan % nchan)
        VM_FATAL("vmLoadOffringaSolution: %zu solution channels per coarse channel do not divide "
                 "%zu fine channels", nchan, vcs_nchan);
    const size_t interp = vcs_nchan / nchan;
    need_D(vm, nant, vcs_nchan);
    gpuDoubleComplex Dread[4];
    for (size_t i = 0; i < cm->num_rf_inputs; i++) {
        const Rfinput *rf = &cm->rf_inputs[i];
        if (rf->pol[0] == 'Y') continue;
        const size_t ant = rf->ant;
        for (size_t ch = 0; ch < nchan; ch++) {
            const size_t Ch = ch + (size_t)cidx * nchan;
            const long pos = (long)(48 + (ant * nChan + Ch) * 64);
            if (fseek(fp, pos, SEEK_SET) != 0 || fread(Dread, 1, 64, fp) != 64)
                VM_FATAL("vmLoadOffringaSolution: %s: short read at byte %ld", vm->cal.caldir, pos);
            if (isnan(Dread[0].x)) memset(Dread, 0, sizeof(Dread));
            for (size_t v = ch * interp; v < (ch + 1) * interp; v++)
                cp2x2(Dread, &vm->D[D_IDX(ant, v, 0, 0, vcs_nchan, 2)]);
        }
    }
    fclose(fp);
}

/* ------------------------------------------------------------------ RTS -- */
void read_dijones_file(gpuDoubleComplex **Dd, gpuDoubleComplex *A, double *amp, uintptr_t nant,
                       char *fname)
{
    /* An amplitude, the alignment matrix A (8 doubles) and one 2x2 complex
     * matrix per antenna (8 doubles each), separated by whitespace or
     * commas. */
    FILE *fp = fopen(fname, "r");
    if (!fp) VM_FATAL("read_dijones_file: cannot open %s", fname);
    double val = 0.0, *a = (double *)A, J[8];
    if (fscanf(fp, "%lf", &val) != 1) VM_FATAL("read_dijones_file: %s: no amplitude", fname);
    if (amp) *amp = val;
    for (int i = 0; i < 8; i++)
        if (fscanf(fp, "%lf,", &a[i]) != 1) VM_FATAL("read_dijones_file: %s: short alignment matrix", fname);
    for (uintptr_t ant = 0; ant < nant; ant++) {
        for (int i = 0; i < 8; i++)
            if (fscanf(fp, "%lf,", &J[i]) != 1)
                VM_FATAL("read_dijones_file: %s: short matrix for antenna %lu", fname,
                         (unsigned long)ant);
        cp2x2((gpuDoubleComplex *)J, Dd[ant]);
    }
    fclose(fp);
}

void read_bandpass_file(gpuDoubleComplex ***Jm, gpuDoubleComplex ***Jf,
                        MetafitsMetadata *cal_metadata, char *filename)
{
    /* First line: the unflagged channels' frequencies in MHz (channel index
     * = round(f / corr_fine_chan_width)).  Then eight rows per unflagged
     * antenna, each starting with the 1-based antenna number: measured and
     * fitted rows alternately for PP, PQ, QP, QQ, each holding (amplitude,
     * phase) pairs per listed channel. */
    const int chan_width = (int)cal_metadata->corr_fine_chan_width_hz;
    const int nchan = (int)cal_metadata->num_corr_fine_chans_per_coarse;
    const int nant = (int)cal_metadata->num_ants;
    FILE *f = fopen(filename, "r");
    if (!f) VM_FATAL("read_bandpass_file: cannot open %s", filename);
    char line[4096];
    if (!fgets(line, sizeof(line), f)) VM_FATAL("read_bandpass_file: %s is empty", filename);
    int *chan_idxs = (int *)malloc(sizeof(int) * (size_t)(nchan > 0 ? nchan : 1));
    if (!chan_idxs) VM_FATAL("out of host memory");
    int count = 0, pos = 0;
    double mhz;
    for (char *p = line; sscanf(p, "%lf,%n", &mhz, &pos) == 1; p += pos) {
        if (++count > nchan)
            VM_FATAL("read_bandpass_file: %s lists more than %d channels", filename, nchan);
        const int idx = (int)roundf((float)(mhz * 1e6 / (double)chan_width));
        if (idx < 0 || idx >= nchan) VM_FATAL("Bandpass file %s: channel at %f MHz out of range", filename, mhz);
        chan_idxs[count - 1] = idx;
    }
    for (int ant = 0; ant < nant; ant++) {
        int cur;
        if (fscanf(f, "%d,", &cur) == EOF) break;
        cur--;
        if (ant < cur) ant = cur;
        if (ant >= nant) VM_FATAL("Bandpass file %s: antenna %d out of range", filename, cur + 1);
        for (int row = 0; row < 8; row++) {
            if (row > 0) {
                if (fscanf(f, "%d,", &cur) != 1 || cur - 1 != ant)
                    VM_FATAL("read_bandpass_file: %s: antenna %d has %d of its 8 rows",
                             filename, ant + 1, row);
            }
            gpuDoubleComplex ***J = (row % 2 == 0) ? Jm : Jf;
            if (!J) {
                if (!fgets(line, sizeof(line), f)) break;
                continue;
            }
            const int pol = row / 2;
            for (int ci = 0; ci < count; ci++) {
                double a, ph;
                if (fscanf(f, "%lf,%lf,", &a, &ph) != 2)
                    VM_FATAL("read_bandpass_file: %s: short row for antenna %d", filename, ant + 1);
                J[ant][chan_idxs[ci]][pol].x = a * cos(ph);
                J[ant][chan_idxs[ci]][pol].y = a * sin(ph);
            }
        }
    }
    free(chan_idxs);
    fclose(f);
}

void vmLoadRTSSolution(vcsbeam_context *vm)
{
    /* D[ant][ch] = reverse(Dd[a] Db[a][ch / interp] A^-1), a = the
     * calibration observation's input matching the tile's X input, halved;
     * Db is the fitted bandpass (identity when use_bandpass is off: the
     * matrices start at zero and only Dd is copied). */
    MetafitsMetadata *cm = vm->cal_metadata, *om = vm->obs_metadata;
    if (!cm || !om || !vm->cal.caldir) VM_FATAL("vmLoadRTSSolution: metadata or directory not set");
    const int cidx = vm->coarse_chan_idx;
    const unsigned long node = cm->metafits_coarse_chans[cidx].corr_chan_number + 1;
    char dij[4200], bp[4200];
    snprintf(dij, sizeof(dij), "%s/DI_JonesMatrices_node%03lu.dat", vm->cal.caldir, node);
    snprintf(bp, sizeof(bp), "%s/BandpassCalibration_node%03lu.dat", vm->cal.caldir, node);
    const size_t nant = cm->num_ants, nchan = cm->num_corr_fine_chans_per_coarse;
    const size_t vcs_nchan = (size_t)vm->nfine_chan;
    if (nchan == 0 || vcs_nchan % nchan)
        VM_FATAL("vmLoadRTSSolution: %zu calibration channels do not divide %zu fine channels",
                 nchan, vcs_nchan);
    const size_t interp = vcs_nchan / nchan;
    need_D(vm, om->num_ants, vcs_nchan);
    gpuDoubleComplex *dd = (gpuDoubleComplex *)calloc(nant * 4, sizeof(gpuDoubleComplex));
    gpuDoubleComplex *db = (gpuDoubleComplex *)calloc(nant * nchan * 4, sizeof(gpuDoubleComplex));
    gpuDoubleComplex **Dd = (gpuDoubleComplex **)calloc(nant, sizeof(*Dd));
    gpuDoubleComplex ***Db = (gpuDoubleComplex ***)calloc(nant, sizeof(*Db));
    gpuDoubleComplex **Dbrows = (gpuDoubleComplex **)calloc(nant * nchan, sizeof(*Dbrows));
    if (!dd || !db || !Dd || !Db || !Dbrows) VM_FATAL("out of host memory");
    for (size_t a = 0; a < nant; a++) {
        Dd[a] = dd + 4 * a;
        Db[a] = Dbrows + a * nchan;
        for (size_t c = 0; c < nchan; c++) Db[a][c] = db + 4 * (a * nchan + c);
    }
    gpuDoubleComplex A[4], Ainv[4];
    read_dijones_file(Dd, A, NULL, nant, dij);
    inv2x2(A, Ainv);
    if (vm->cal.use_bandpass) read_bandpass_file(NULL, Db, cm, bp);
    for (size_t i = 0; i < om->num_rf_inputs; i++) {
        Rfinput *orf = &om->rf_inputs[i];
        if (orf->pol[0] == 'Y') continue;
        Rfinput *crf = find_matching_rf_input(cm, orf);
        if (!crf) VM_FATAL("vmLoadRTSSolution: tile %u has no input in the calibration metadata",
                           orf->tile_id);
        const size_t di = crf->input / 2;
        if (di >= nant) VM_FATAL("vmLoadRTSSolution: calibration input %u out of range", crf->input);
        for (size_t ch = 0; ch < vcs_nchan; ch++) {
            gpuDoubleComplex *D = &vm->D[D_IDX(orf->ant, ch, 0, 0, vcs_nchan, 2)];
            if (vm->cal.use_bandpass)
                mult2x2d(Dd[di], Db[di][ch / interp], D);
            else
                cp2x2(Dd[di], D);
            mult2x2d(D, Ainv, D);
            reverse2x2(D, D);
        }
    }
    free(Dbrows);
    free(Db);
    free(Dd);
    free(db);
    free(dd);
}

/* ---------------------------------------------------------- corrections -- */
void remove_reference_phase(gpuDoubleComplex *J, gpuDoubleComplex *Jref)
{
    /* J[k] / (Jref[k] / |Jref[k]|) elementwise; elements whose reference is
     * zero are left alone */
    for (int k = 0; k < 4; k++) {
        const double s = 1.0 / hypot(Jref[k].x, Jref[k].y);
        if (!isfinite(s)) continue;
        gpuDoubleComplex u = {s * Jref[k].x, s * Jref[k].y};
        J[k] = cdiv(J[k], u);
    }
}

void zero_PQ_and_QP(gpuDoubleComplex *J)
{
    J[1].x = J[1].y = 0.0;
    J[2].x = J[2].y = 0.0;
}

static const char *vm_runtime_dir(void)
{
    const char *dir = getenv("VCSBEAM_RUNTIME_DIR");
    return dir ? dir : VCSBEAM_RUNTIME_DIR;
}

void parse_calibration_correction_file(uint32_t gpstime, calibration *cal)
{
    /* <runtime dir>/pq_phase_correction.txt: lines "from to slope offset
     * tile ..." (GPS seconds, rad/Hz, rad, reference tile); lines not
     * starting with a digit are comments.  The first range holding gpstime
     * wins.  The reference tile is installed unless the caller set one
     * (none found: "NONE"); slope and offset unless custom_pq_correction. */
    char path[4200], line[4096], tile[4096];
    snprintf(path, sizeof(path), "%s/pq_phase_correction.txt", vm_runtime_dir());
    FILE *f = fopen(path, "r");
    if (!f) VM_FATAL("parse_calibration_correction_file: cannot open %s", path);
    unsigned from, to;
    double slope = 0.0, offset = 0.0, s, o;
    int found = 0;
    while (fgets(line, sizeof(line), f)) {
        if (!(line[0] >= '0' && line[0] <= '9')) continue;
        if (sscanf(line, "%u %u %lf %lf %4095s", &from, &to, &s, &o, tile) != 5)
            VM_FATAL("parse_calibration_correction_file: cannot parse line \"%s\"", line);
        if (from <= gpstime && gpstime <= to) {
            slope = s;
            offset = o;
            found = 1;
            break;
        }
    }
    fclose(f);
    if (!cal->ref_ant) cal->ref_ant = strdup(found ? tile : "NONE");
    if (!cal->custom_pq_correction && found) {
        cal->phase_slope = slope;
        cal->phase_offset = offset;
    }
}

void vmApplyCalibrationCorrections(vcsbeam_context *vm)
{
    /* Per fine channel: divide every antenna's D by the reference tile's
     * element phases (unless ref_ant is "NONE"), zero the cross terms
     * (unless keep_cross_terms), and multiply the P column by
     * e^{i (slope f + offset)} with f = chan_start + ch * width (Hz). */
    MetafitsMetadata *om = vm->obs_metadata;
    if (!om) VM_FATAL("vmApplyCalibrationCorrections: no observation metadata");
    const int cidx = vm->coarse_chan_idx;
    int apply_ref = 1;
    const int apply_zero = !vm->cal.keep_cross_terms;
    const int apply_slope = (vm->cal.phase_offset != 0.0 || vm->cal.phase_slope != 0.0);
    long aref = -1;
    if (!vm->cal.ref_ant || strcmp(vm->cal.ref_ant, "NONE") == 0) {
        apply_ref = 0;
    } else {
        aref = ant_by_tile_name(om, vm->cal.ref_ant);
        if (aref < 0)
            VM_FATAL("vmApplyCalibrationCorrections: reference tile %s is not in the "
                     "observation", vm->cal.ref_ant);
    }
    if (!apply_ref && !apply_zero && !apply_slope) return;
    const size_t nant = om->num_ants, nchan = (size_t)vm->nfine_chan;
    need_D(vm, nant, nchan);
    const long f0 = (long)om->metafits_coarse_chans[cidx].chan_start_hz;
    const int width = (int)(om->coarse_chan_width_hz / (uint32_t)vm->nfine_chan);
    gpuDoubleComplex Dref[4], z = {1.0, 0.0};
    for (size_t ch = 0; ch < nchan; ch++) {
        if (apply_ref) cp2x2(&vm->D[D_IDX((size_t)aref, ch, 0, 0, nchan, 2)], Dref);
        if (apply_slope) {
            const long fch = f0 + (long)ch * width;
            const double phi = vm->cal.phase_slope * (double)fch + vm->cal.phase_offset;
            z.x = cos(phi);
            z.y = sin(phi);
        }
        for (size_t a = 0; a < nant; a++) {
            gpuDoubleComplex *D = &vm->D[D_IDX(a, ch, 0, 0, nchan, 2)];
            if (apply_ref) remove_reference_phase(D, Dref);
            if (apply_zero) zero_PQ_and_QP(D);
            if (apply_slope) {
                D[0] = cmul(D[0], z);
                D[2] = cmul(D[2], z);
            }
        }
    }
}

/* ---------------------------------------------------------------- flags -- */
void vmParseFlaggedTilenamesFile(char *filename, calibration *cal)
{
    /* whitespace-separated tile names */
    FILE *f = fopen(filename, "r");
    if (!f) VM_FATAL("vmParseFlaggedTilenamesFile: cannot open %s", filename);
    char name[MAX_COMMAND_LENGTH];
    int n = 0;
    while (fscanf(f, "%1023s", name) == 1) n++;
    rewind(f);
    char **names = (char **)calloc((size_t)(n > 0 ? n : 1), sizeof(char *));
    if (!names) VM_FATAL("out of host memory");
    int k = 0;
    while (k < n && fscanf(f, "%1023s", name) == 1) names[k++] = strdup(name);
    fclose(f);
    cal->nflags = k;
    cal->flagged_tilenames = names;
}

bool tilename_is_flagged(char *tilename, calibration *cal)
{
    for (int i = 0; i < cal->nflags; i++)
        if (strcmp(tilename, cal->flagged_tilenames[i]) == 0) return true;
    return false;
}

void vmSetCustomTileFlags(vcsbeam_context *vm)
{
    /* Zero D for every fine channel of each listed tile found in the
     * observation (unknown names: a warning), then recount the unflagged
     * inputs.  D holds nfine_chan channels per tile (vmMallocDHost); the
     * reference strides it by num_volt_fine_chans_per_coarse, which is 1 for
     * MWAX data and then zeroes the wrong elements. */
    if (!vm->cal.flags_file) {
        vmSetNumNotFlaggedRFInputs(vm);
        return;
    }
    vmParseFlaggedTilenamesFile(vm->cal.flags_file, &vm->cal);
    MetafitsMetadata *om = vm->obs_metadata;
    const size_t nchan = (size_t)vm->nfine_chan;
    need_D(vm, om->num_ants, nchan);
    for (int i = 0; i < vm->cal.nflags; i++) {
        const long a = ant_by_tile_name(om, vm->cal.flagged_tilenames[i]);
        if (a < 0) {
            snprintf(vm->log_message, sizeof(vm->log_message),
                     "Warning: %s names tile %s, which the observation does not have\n",
                     vm->cal.flags_file, vm->cal.flagged_tilenames[i]);
            if (vm->log) logger_message(vm->log, vm->log_message);
            continue;
        }
        for (size_t ch = 0; ch < nchan; ch++)
            memset(&vm->D[D_IDX((size_t)a, ch, 0, 0, nchan, 2)], 0, 4 * sizeof(gpuDoubleComplex));
    }
    vmSetNumNotFlaggedRFInputs(vm);
}

void init_calibration(calibration *cal)
{
    cal->caldir = NULL;
    cal->ref_ant = NULL;
    cal->flags_file = NULL;
    cal->nflags = 0;
    cal->flagged_tilenames = NULL;
}

void free_calibration(calibration *cal)
{
    free(cal->caldir);
    free(cal->ref_ant);
    free(cal->flags_file);
    for (int i = 0; i < cal->nflags; i++) free(cal->flagged_tilenames[i]);
    free(cal->flagged_tilenames);
    cal->caldir = cal->ref_ant = cal->flags_file = NULL;
    cal->flagged_tilenames = NULL;
    cal->nflags = 0;
}
