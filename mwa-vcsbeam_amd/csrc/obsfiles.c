/*
 * obsfiles.c -- observation metadata and voltage files without mwalib: the
 * subset of a metafits file that the beamforming path reads, and a reader
 * callback that fills one second of voltages from MWAX subfiles or legacy
 * recombined files.
 *
 * Reference: MWA-vcsbeam src/metadata.c:166-233 (vmBindObsData), 930-993
 * (vmReadNextSecond -> mwalib_voltage_context_read_second), 1297-1424 (the
 * mwalib metafits / voltage context glue); include/vcsbeam.h.in:238-240.
 * mwalib itself is not part of the reference tree.  What is restated here:
 *   - FITS (NASA/IAU FITS standard 4.0): 2880-byte records of 80-character
 *     header cards, long strings continued with CONTINUE, binary tables
 *     (XTENSION = 'BINTABLE') with big-endian fields described by TTYPEn /
 *     TFORMn;
 *   - the metafits primary keys GPSTIME, MJD, PROJECT, MODE, CHANNELS,
 *     BANDWDTH, NCHANS, FINECHAN, RA, DEC, NINPUTS and the TILEDATA table
 *     columns Input, Antenna, Tile, TileName, Pol, Flag, Length ("EL_<m>"),
 *     North, East, Height, and Delays (16 dipole delay steps) when present;
 *   - MWAX VCS subfiles <obsid>_<subobs>_<rec>.sub: a 4096-byte ASCII header
 *     ("KEY value" lines; HDR_SIZE and NINPUTS are checked), one delay block
 *     the size of a data block, then 160 data blocks of 50 ms (8 s),
 *     [input][64000 samples][re, im] int8 each;
 *   - legacy recombined files <obsid>_<gps>_ch<rec>.dat: one second each,
 *     [10000 samples][128 fine channels][inputs] 4+4-bit.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define FITS_BLOCK 2880

/* ------------------------------------------------------------ FITS I/O -- */
typedef struct {
    char *cards; /* ncards x 80 characters */
    int ncards;
    long data_bytes; /* size of the HDU's data, before padding */
} fits_hdu;

static int fits_read_header(FILE *f, fits_hdu *h)
{
    /* reads 2880-byte records until the END card; 0 on success */
    h->cards = NULL;
    h->ncards = 0;
    char rec[FITS_BLOCK];
    for (;;) {
        if (fread(rec, 1, FITS_BLOCK, f) != FITS_BLOCK) return -1;
        char *grown = (char *)realloc(h->cards, (size_t)(h->ncards + 36) * 80);
        if (!grown) VM_FATAL("out of host memory");
        h->cards = grown;
        for (int c = 0; c < 36; c++) {
            memcpy(h->cards + (size_t)h->ncards * 80, rec + c * 80, 80);
            h->ncards++;
            if (strncmp(rec + c * 80, "END     ", 8) == 0) return 0;
        }
    }
}

/* the value field of the first card named key (NULL if none) */
static const char *fits_card(const fits_hdu *h, const char *key)
{
    const size_t n = strlen(key);
    for (int c = 0; c < h->ncards; c++) {
        const char *k = h->cards + (size_t)c * 80;
        if (strncmp(k, key, n) == 0 && (k[n] == ' ' || k[n] == '=') && k[8] == '=') return k + 10;
    }
    return NULL;
}

static int fits_long(const fits_hdu *h, const char *key, long *v)
{
    const char *p = fits_card(h, key);
    if (!p) return 0;
    char buf[71];
    memcpy(buf, p, 70);
    buf[70] = 0;
    *v = strtol(buf, NULL, 10);
    return 1;
}

static int fits_double(const fits_hdu *h, const char *key, double *v)
{
    const char *p = fits_card(h, key);
    if (!p) return 0;
    char buf[71];
    memcpy(buf, p, 70);
    buf[70] = 0;
    for (char *q = buf; *q; q++)
        if (*q == 'D' || *q == 'd') *q = 'E'; /* Fortran exponents */
    *v = strtod(buf, NULL);
    return 1;
}

/* a quoted string value, with '' as a quote and CONTINUE cards ('&' at the
 * end of a part) appended; trailing blanks dropped; malloc'd (NULL if none) */
static char *fits_string(const fits_hdu *h, const char *key)
{
    const size_t n = strlen(key);
    int c0 = -1;
    for (int c = 0; c < h->ncards && c0 < 0; c++) {
        const char *k = h->cards + (size_t)c * 80;
        if (strncmp(k, key, n) == 0 && (k[n] == ' ' || k[n] == '=') && k[8] == '=') c0 = c;
    }
    if (c0 < 0) return NULL;
    size_t cap = 256, len = 0;
    char *out = (char *)malloc(cap);
    if (!out) VM_FATAL("out of host memory");
    for (int c = c0; c < h->ncards; c++) {
        const char *card = h->cards + (size_t)c * 80;
        const char *p = card + ((c == c0) ? 10 : 8);
        const char *end = card + 80;
        while (p < end && *p == ' ') p++;
        if (p >= end || *p != '\'') break;
        p++;
        size_t part0 = len;
        while (p < end) {
            if (*p == '\'') {
                if (p + 1 < end && p[1] == '\'') {
                    p += 2;
                    if (len + 1 >= cap) out = (char *)realloc(out, cap *= 2);
                    if (!out) VM_FATAL("out of host memory");
                    out[len++] = '\'';
                    continue;
                }
                break;
            }
            if (len + 1 >= cap) out = (char *)realloc(out, cap *= 2);
            if (!out) VM_FATAL("out of host memory");
            out[len++] = *p++;
        }
        while (len > part0 && out[len - 1] == ' ') len--;
        const int more = (len > part0 && out[len - 1] == '&' && c + 1 < h->ncards &&
                          strncmp(h->cards + (size_t)(c + 1) * 80, "CONTINUE", 8) == 0);
        if (!more) break;
        len--; /* drop the '&' */
    }
    out[len] = 0;
    while (len > 0 && out[len - 1] == ' ') out[--len] = 0;
    return out;
}

static void fits_data_size(fits_hdu *h)
{
    long bitpix = 8, naxis = 0, pcount = 0, gcount = 1, prod = 1;
    fits_long(h, "BITPIX", &bitpix);
    fits_long(h, "NAXIS", &naxis);
    fits_long(h, "PCOUNT", &pcount);
    fits_long(h, "GCOUNT", &gcount);
    for (long a = 1; a <= naxis; a++) {
        char k[32];
        long v = 0;
        snprintf(k, sizeof(k), "NAXIS%ld", a);
        fits_long(h, k, &v);
        prod *= v;
    }
    if (naxis == 0) prod = 0;
    h->data_bytes = labs(bitpix) / 8 * gcount * (pcount + prod);
}

static long fits_padded(long n) { return (n + FITS_BLOCK - 1) / FITS_BLOCK * FITS_BLOCK; }

/* bytes of one element of a TFORM type code */
static int tform_size(char t)
{
    switch (t) {
    case 'L': case 'B': case 'A': return 1;
    case 'I': return 2;
    case 'J': case 'E': return 4;
    case 'K': case 'D': case 'C': case 'P': return 8;
    case 'M': case 'Q': return 16;
    default: return -1;
    }
}

typedef struct {
    int offset, count;
    char type;
} fits_col;

static int fits_column(const fits_hdu *h, const char *name, fits_col *col)
{
    long nf = 0;
    fits_long(h, "TFIELDS", &nf);
    int offset = 0;
    for (long i = 1; i <= nf; i++) {
        char k[32];
        snprintf(k, sizeof(k), "TFORM%ld", i);
        char *form = fits_string(h, k);
        if (!form) VM_FATAL("metafits: TILEDATA has no %s", k);
        char *p = form;
        while (*p == ' ') p++;
        int count = isdigit((unsigned char)*p) ? (int)strtol(p, &p, 10) : 1;
        const char type = *p;
        const int sz = tform_size(type);
        free(form);
        if (sz < 0 && type != 'X') VM_FATAL("metafits: unsupported column type '%c'", type);
        snprintf(k, sizeof(k), "TTYPE%ld", i);
        char *tt = fits_string(h, k);
        const int match = tt && strcasecmp(tt, name) == 0;
        free(tt);
        if (match) {
            col->offset = offset;
            col->count = count;
            col->type = type;
            return 1;
        }
        offset += (type == 'X') ? (count + 7) / 8 : count * sz;
    }
    return 0;
}

static double be_num(const unsigned char *p, char type)
{
    /* big-endian field to double */
    uint64_t u = 0;
    const int sz = tform_size(type);
    for (int i = 0; i < sz; i++) u = (u << 8) | p[i];
    switch (type) {
    case 'B': return (double)(uint8_t)u;
    case 'I': return (double)(int16_t)(uint16_t)u;
    case 'J': return (double)(int32_t)(uint32_t)u;
    case 'K': return (double)(int64_t)u;
    case 'E': { uint32_t w = (uint32_t)u; float f; memcpy(&f, &w, 4); return f; }
    case 'D': { double d; memcpy(&d, &u, 8); return d; }
    case 'L': return p[0] == 'T' ? 1.0 : 0.0;
    default: VM_FATAL("metafits: column type '%c' is not numeric", type);
    }
    return 0.0;
}

/* ------------------------------------------------------------ metafits -- */
static void set_coarse_chans(MetafitsMetadata *md, const char *list)
{
    /* receiver channel numbers, comma-separated; kept in ascending order
     * (corr_chan_number = rank) */
    int n = 0, cap = 32;
    int *rec = (int *)malloc(sizeof(int) * (size_t)cap);
    if (!rec) VM_FATAL("out of host memory");
    for (const char *p = list; *p;) {
        while (*p && !isdigit((unsigned char)*p)) p++;
        if (!*p) break;
        if (n == cap) rec = (int *)realloc(rec, sizeof(int) * (size_t)(cap *= 2));
        rec[n++] = (int)strtol(p, (char **)&p, 10);
    }
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0 && rec[j - 1] > rec[j]; j--) {
            const int t = rec[j];
            rec[j] = rec[j - 1];
            rec[j - 1] = t;
        }
    md->num_metafits_coarse_chans = (size_t)n;
    md->metafits_coarse_chans = (CoarseChannel *)calloc((size_t)(n > 0 ? n : 1), sizeof(CoarseChannel));
    if (!md->metafits_coarse_chans) VM_FATAL("out of host memory");
    const uint32_t w = md->coarse_chan_width_hz;
    for (int i = 0; i < n; i++) {
        CoarseChannel *c = &md->metafits_coarse_chans[i];
        c->corr_chan_number = (uint32_t)i;
        c->rec_chan_number = (uint32_t)rec[i];
        c->gpubox_number = (md->mwa_version == VCSMWAXv2) ? (uint32_t)rec[i] : (uint32_t)(i + 1);
        c->chan_width_hz = w;
        c->chan_centre_hz = (uint32_t)rec[i] * w;
        c->chan_start_hz = c->chan_centre_hz - w / 2;
        c->chan_end_hz = c->chan_centre_hz + w / 2;
    }
    free(rec);
}

static int cmp_subfile_order(const void *a, const void *b)
{
    const Rfinput *x = (const Rfinput *)a, *y = (const Rfinput *)b;
    return (x->subfile_order > y->subfile_order) - (x->subfile_order < y->subfile_order);
}

int vmLoadMetafits(const char *path, MetafitsMetadata *md, VoltageMetadata *vmd)
{
    memset(md, 0, sizeof(*md));
    FILE *f = fopen(path, "rb");
    if (!f) VM_FATAL("vmLoadMetafits: cannot open %s", path);
    fits_hdu ph;
    if (fits_read_header(f, &ph) != 0 || strncmp(ph.cards, "SIMPLE  =", 9) != 0)
        VM_FATAL("vmLoadMetafits: %s is not a FITS file", path);
    fits_data_size(&ph);
    long gps = 0, nchans = 24, ninputs = -1, exposure = 0;
    double mjd = NAN, bw = 30.72, finechan = 10.0, goodtime = NAN;
    if (!fits_long(&ph, "GPSTIME", &gps)) VM_FATAL("vmLoadMetafits: %s has no GPSTIME", path);
    fits_double(&ph, "MJD", &mjd);
    fits_long(&ph, "NCHANS", &nchans);
    fits_double(&ph, "BANDWDTH", &bw);
    fits_double(&ph, "FINECHAN", &finechan);
    fits_long(&ph, "NINPUTS", &ninputs);
    fits_long(&ph, "EXPOSURE", &exposure);
    fits_double(&ph, "GOODTIME", &goodtime);
    fits_double(&ph, "RA", &md->ra_tile_pointing_deg);
    fits_double(&ph, "DEC", &md->dec_tile_pointing_deg);
    char *mode = fits_string(&ph, "MODE");
    char *chans = fits_string(&ph, "CHANNELS");
    md->project_id = fits_string(&ph, "PROJECT");
    if (!md->project_id) md->project_id = strdup("");
    if (!mode || !chans) VM_FATAL("vmLoadMetafits: %s lacks MODE or CHANNELS", path);
    if (strcmp(mode, "MWAX_VCS") == 0)
        md->mwa_version = VCSMWAXv2;
    else if (strcmp(mode, "VOLTAGE_START") == 0 || strcmp(mode, "VOLTAGE_BUFFER") == 0)
        md->mwa_version = VCSLegacyRecombined;
    else
        VM_FATAL("vmLoadMetafits: MODE '%s' is not a voltage observation", mode);
    md->obs_id = (uint32_t)gps;
    md->sched_start_gps_time_ms = (uint64_t)gps * 1000;
    /* without an MJD card: GPS epoch MJD 44244, less the 18 leap seconds
     * GPS has gained on UTC since 2017 */
    md->sched_start_mjd = isnan(mjd) ? 44244.0 + ((double)gps - 18.0) / 86400.0 : mjd;
    /* GOODTIME is a UNIX time: GPS = UNIX - 315964800 + 18 leap seconds */
    md->good_time_gps_ms = isnan(goodtime) ? (uint64_t)gps * 1000
                                           : (uint64_t)llround((goodtime - 315964800.0 + 18.0) * 1e3);
    md->num_metafits_timesteps = exposure > 0 ? (size_t)exposure : 0;
    md->coarse_chan_width_hz = (uint32_t)llround(bw * 1e6 / (double)(nchans > 0 ? nchans : 24));
    md->corr_fine_chan_width_hz = (uint32_t)llround(finechan * 1e3);
    md->num_corr_fine_chans_per_coarse =
        md->corr_fine_chan_width_hz ? md->coarse_chan_width_hz / md->corr_fine_chan_width_hz : 0;
    md->num_volt_fine_chans_per_coarse = (md->mwa_version == VCSMWAXv2) ? 1 : 128;
    md->num_ant_pols = 2;
    md->num_visibility_pols = 4;
    set_coarse_chans(md, chans);
    free(mode);
    free(chans);

    /* find TILEDATA */
    if (fseek(f, fits_padded(ph.data_bytes), SEEK_CUR) != 0) VM_FATAL("vmLoadMetafits: %s truncated", path);
    free(ph.cards);
    fits_hdu th;
    for (;;) {
        if (fits_read_header(f, &th) != 0) VM_FATAL("vmLoadMetafits: %s has no TILEDATA table", path);
        fits_data_size(&th);
        char *ext = fits_string(&th, "EXTNAME");
        const int found = ext && strcmp(ext, "TILEDATA") == 0;
        free(ext);
        if (found) break;
        if (fseek(f, fits_padded(th.data_bytes), SEEK_CUR) != 0) VM_FATAL("vmLoadMetafits: %s truncated", path);
        free(th.cards);
    }
    long rowb = 0, nrows = 0;
    fits_long(&th, "NAXIS1", &rowb);
    fits_long(&th, "NAXIS2", &nrows);
    if (ninputs >= 0 && ninputs != nrows)
        VM_FATAL("vmLoadMetafits: NINPUTS = %ld but TILEDATA has %ld rows", ninputs, nrows);
    if (nrows <= 0 || nrows % 2) VM_FATAL("vmLoadMetafits: TILEDATA has %ld rows", nrows);
    unsigned char *tab = (unsigned char *)malloc((size_t)(rowb * nrows));
    if (!tab || fread(tab, 1, (size_t)(rowb * nrows), f) != (size_t)(rowb * nrows))
        VM_FATAL("vmLoadMetafits: %s: short TILEDATA", path);
    fclose(f);
    fits_col cin, cant, ctile, cname, cpol, cflag, clen, cn, ce, ch;
    const char *need[] = {"Input", "Antenna", "Tile", "TileName", "Pol", "Flag", "Length", "North", "East", "Height"};
    fits_col *cols[] = {&cin, &cant, &ctile, &cname, &cpol, &cflag, &clen, &cn, &ce, &ch};
    for (int k = 0; k < 10; k++)
        if (!fits_column(&th, need[k], cols[k])) VM_FATAL("vmLoadMetafits: TILEDATA has no column %s", need[k]);
    /* optional: the analog beamformer's dipole delays (16 integers per
     * input, 32 = dead dipole) */
    fits_col cdel;
    const int have_delays = fits_column(&th, "Delays", &cdel);
    if (have_delays && (cdel.count < 1 || tform_size(cdel.type) < 1 || cdel.type == 'A'))
        VM_FATAL("vmLoadMetafits: TILEDATA column Delays has type %d%c", cdel.count, cdel.type);
    free(th.cards);
    md->num_delays = have_delays ? (size_t)cdel.count : 0;
    md->num_rf_inputs = (size_t)nrows;
    md->num_ants = (size_t)nrows / 2;
    md->rf_inputs = (Rfinput *)calloc((size_t)nrows, sizeof(Rfinput));
    if (!md->rf_inputs) VM_FATAL("out of host memory");
    for (long r = 0; r < nrows; r++) {
        const unsigned char *row = tab + r * rowb;
        Rfinput *rf = &md->rf_inputs[r];
        rf->input = (uint32_t)be_num(row + cin.offset, cin.type);
        rf->ant = (uint32_t)be_num(row + cant.offset, cant.type);
        rf->tile_id = (uint32_t)be_num(row + ctile.offset, ctile.type);
        rf->flagged = be_num(row + cflag.offset, cflag.type) != 0.0;
        rf->north_m = be_num(row + cn.offset, cn.type);
        rf->east_m = be_num(row + ce.offset, ce.type);
        rf->height_m = be_num(row + ch.offset, ch.type);
        char s[64];
        int m = cname.count < 63 ? cname.count : 63;
        memcpy(s, row + cname.offset, (size_t)m);
        s[m] = 0;
        while (m > 0 && (s[m - 1] == ' ' || s[m - 1] == 0)) s[--m] = 0;
        rf->tile_name = strdup(s);
        rf->pol = strdup(row[cpol.offset] == 'Y' ? "Y" : "X");
        m = clen.count < 63 ? clen.count : 63;
        memcpy(s, row + clen.offset, (size_t)m);
        s[m] = 0;
        const char *lp = s;
        while (*lp == ' ') lp++;
        if (strncmp(lp, "EL_", 3) == 0) lp += 3;
        rf->electrical_length_m = strtod(lp, NULL);
        if (have_delays) {
            rf->dipole_delays = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cdel.count);
            if (!rf->dipole_delays) VM_FATAL("out of host memory");
            const int sz = tform_size(cdel.type);
            for (int d = 0; d < cdel.count; d++)
                rf->dipole_delays[d] = (uint32_t)be_num(row + cdel.offset + d * sz, cdel.type);
        }
        /* MWAX subfiles hold inputs by antenna then polarisation; legacy
         * recombined data by the PFB's interleave of each board's 64 inputs
         * (whole boards only: with a partial board the order is the input) */
        if (rf->input >= (uint32_t)nrows || rf->ant >= (uint32_t)(nrows / 2))
            VM_FATAL("vmLoadMetafits: TILEDATA row %ld has input %u, antenna %u for %ld inputs",
                     r, rf->input, rf->ant, nrows);
        rf->subfile_order = rf->ant * 2 + (rf->pol[0] == 'Y');
        rf->vcs_order = (nrows % 64 == 0)
                            ? (rf->input & ~63u) | ((rf->input & 15u) << 2) | ((rf->input >> 4) & 3u)
                            : rf->input;
    }
    free(tab);
    qsort(md->rf_inputs, (size_t)nrows, sizeof(Rfinput), cmp_subfile_order);
    if (vmd) {
        memset(vmd, 0, sizeof(*vmd));
        vmd->mwa_version = md->mwa_version;
        if (md->mwa_version == VCSMWAXv2) {
            vmd->num_samples_per_voltage_block = 64000;
            vmd->num_voltage_blocks_per_second = 20;
            vmd->voltage_block_size_bytes = 64000ull * 2 * (uint64_t)nrows;
        } else {
            vmd->num_samples_per_voltage_block = 10000;
            vmd->num_voltage_blocks_per_second = 1;
            vmd->voltage_block_size_bytes = 10000ull * 128 * (uint64_t)nrows;
        }
    }
    return 0;
}

void vmFreeMetafits(MetafitsMetadata *md)
{
    for (size_t i = 0; md->rf_inputs && i < md->num_rf_inputs; i++) {
        free(md->rf_inputs[i].tile_name);
        free(md->rf_inputs[i].pol);
        free(md->rf_inputs[i].dipole_delays);
    }
    free(md->rf_inputs);
    free(md->metafits_coarse_chans);
    free(md->project_id);
    memset(md, 0, sizeof(*md));
}

/* ------------------------------------------------------- voltage files -- */
static FILE *open_either(char *path, size_t cap, const char *dir, const char *fmt3,
                         const char *fmt, unsigned long a, unsigned long b, unsigned long c)
{
    snprintf(path, cap, fmt3, dir, a, b, c);
    FILE *f = fopen(path, "rb");
    if (f) return f;
    snprintf(path, cap, fmt, dir, a, b, c);
    return fopen(path, "rb");
}

int vmReadSecondFromFiles(void *user, uint64_t gps_second, int coarse_chan_idx, void *dst,
                          size_t nbytes)
{
    const vm_voltage_files *vf = (const vm_voltage_files *)user;
    const MetafitsMetadata *md = vf->md;
    const VoltageMetadata *vmd = vf->vmd;
    if (coarse_chan_idx < 0 || (size_t)coarse_chan_idx >= md->num_metafits_coarse_chans) return -1;
    const unsigned long rec = md->metafits_coarse_chans[coarse_chan_idx].rec_chan_number;
    const size_t bsz = vmd->voltage_block_size_bytes;
    const size_t want = bsz * vmd->num_voltage_blocks_per_second;
    if (nbytes != want) {
        fprintf(stderr, "vmReadSecondFromFiles: %zu bytes asked, one second is %zu\n", nbytes, want);
        return -1;
    }
    if (gps_second < md->obs_id) return -1;
    char path[2200];
    FILE *f;
    long off;
    if (md->mwa_version == VCSMWAXv2) {
        const uint64_t sub = md->obs_id + (gps_second - md->obs_id) / 8 * 8;
        f = open_either(path, sizeof(path), vf->dir, "%s/%lu_%lu_%03lu.sub", "%s/%lu_%lu_%lu.sub",
                        (unsigned long)md->obs_id, (unsigned long)sub, rec);
        if (!f) {
            fprintf(stderr, "vmReadSecondFromFiles: no subfile for GPS %lu, channel %lu in %s\n",
                    (unsigned long)gps_second, rec, vf->dir);
            return -1;
        }
        char hdr[4097];
        if (fread(hdr, 1, 4096, f) != 4096) {
            fclose(f);
            return -1;
        }
        hdr[4096] = 0;
        long hdr_size = 4096, nin = -1;
        for (char *line = hdr; line && *line; line = strchr(line, '\n') ? strchr(line, '\n') + 1 : NULL) {
            if (strncmp(line, "HDR_SIZE ", 9) == 0) hdr_size = strtol(line + 9, NULL, 10);
            if (strncmp(line, "NINPUTS ", 8) == 0) nin = strtol(line + 8, NULL, 10);
        }
        if (nin >= 0 && (size_t)nin != md->num_rf_inputs) {
            fprintf(stderr, "vmReadSecondFromFiles: %s holds %ld inputs, the metadata %zu\n", path,
                    nin, md->num_rf_inputs);
            fclose(f);
            return -1;
        }
        off = hdr_size + (long)bsz * (long)(1 + (gps_second - sub) * vmd->num_voltage_blocks_per_second);
    } else {
        f = open_either(path, sizeof(path), vf->dir, "%s/%lu_%lu_ch%03lu.dat", "%s/%lu_%lu_ch%lu.dat",
                        (unsigned long)md->obs_id, (unsigned long)gps_second, rec);
        if (!f) {
            fprintf(stderr, "vmReadSecondFromFiles: no file for GPS %lu, channel %lu in %s\n",
                    (unsigned long)gps_second, rec, vf->dir);
            return -1;
        }
        off = 0;
    }
    const int ok = fseek(f, off, SEEK_SET) == 0 && fread(dst, 1, nbytes, f) == nbytes;
    fclose(f);
    if (!ok) fprintf(stderr, "vmReadSecondFromFiles: %s: short read at byte %ld\n", path, off);
    return ok ? 0 : -1;
}

/* ------------------------------------------ context-level observation -- */
/* vm->obs_context / vm->cal_context (mwalib's contexts in the reference)
 * hold the metadata loaded here; vm->vcs_context holds the file reader's
 * state.  destroy_vcsbeam_context releases all three. */
typedef struct {
    MetafitsMetadata md;
    VoltageMetadata vmd;
} vm_metafits_ctx;

static MetafitsMetadata *load_into(void **ctx_slot, const char *filename)
{
    if (*ctx_slot) {
        vmFreeMetafits(&((vm_metafits_ctx *)*ctx_slot)->md);
        free(*ctx_slot);
    }
    vm_metafits_ctx *c = (vm_metafits_ctx *)calloc(1, sizeof(vm_metafits_ctx));
    if (!c) VM_FATAL("out of host memory");
    vmLoadMetafits(filename, &c->md, &c->vmd);
    *ctx_slot = c;
    return &c->md;
}

void vmLoadObsMetafits(vcsbeam_context *vm, char *filename)
{
    /* src/metadata.c:1297-1330 */
    vm->obs_metadata = load_into((void **)&vm->obs_context, filename);
}

void vmLoadCalMetafits(vcsbeam_context *vm, char *filename)
{
    /* src/metadata.c:1339-1363 */
    vm->cal_metadata = load_into((void **)&vm->cal_context, filename);
}

void vmGetVoltageMetadata(vcsbeam_context *vm)
{
    /* src/metadata.c:1373-1424: the voltage description of the loaded
     * observation and the reader over the files in vm->datadir */
    vm_metafits_ctx *c = (vm_metafits_ctx *)vm->obs_context;
    if (!c) VM_FATAL("vmGetVoltageMetadata: no metafits loaded (vmLoadObsMetafits)");
    vm->vcs_metadata = &c->vmd;
    vm_voltage_files *f = (vm_voltage_files *)vm->vcs_context;
    if (!f) {
        f = (vm_voltage_files *)calloc(1, sizeof(vm_voltage_files));
        if (!f) VM_FATAL("out of host memory");
        vm->vcs_context = (VoltageContext *)f;
    }
    f->dir = vm->datadir ? vm->datadir : ".";
    f->md = &c->md;
    f->vmd = &c->vmd;
    vm->read_second = vmReadSecondFromFiles;
    vm->read_second_user = f;
}

void vm_free_obs_contexts(vcsbeam_context *vm)
{
    void **slots[2] = {(void **)&vm->obs_context, (void **)&vm->cal_context};
    for (int k = 0; k < 2; k++)
        if (*slots[k]) {
            vm_metafits_ctx *c = (vm_metafits_ctx *)*slots[k];
            if (vm->obs_metadata == &c->md) vm->obs_metadata = NULL;
            if (vm->cal_metadata == &c->md) vm->cal_metadata = NULL;
            if (vm->vcs_metadata == &c->vmd) vm->vcs_metadata = NULL;
            vmFreeMetafits(&c->md);
            free(c);
            *slots[k] = NULL;
        }
    if (vm->vcs_context) {
        if (vm->read_second_user == (void *)vm->vcs_context) {
            vm->read_second = NULL;
            vm->read_second_user = NULL;
        }
        free(vm->vcs_context);
        vm->vcs_context = NULL;
    }
}

long unsigned int get_relative_gps(MetafitsMetadata *obs_metadata, long int relative_begin)
{
    /* src/metadata.c:1448-1455: from the first good second, or (negative)
     * counting back from the end of the 1 s timesteps */
    if (relative_begin >= 0) return obs_metadata->good_time_gps_ms / 1000 + (unsigned long)relative_begin;
    const long idx = (long)obs_metadata->num_metafits_timesteps + relative_begin;
    if (idx < 0) VM_FATAL("get_relative_gps: %ld is before the first timestep", relative_begin);
    return obs_metadata->sched_start_gps_time_ms / 1000 + (unsigned long)idx;
}

long unsigned int parse_begin_string(MetafitsMetadata *obs_metadata, char *begin_str)
{
    /* src/metadata.c:1473-1489 */
    if (!begin_str) VM_FATAL("parse_begin_string: no string");
    if (begin_str[0] == '+' || begin_str[0] == '-') return get_relative_gps(obs_metadata, atol(begin_str));
    if (!(begin_str[0] >= '0' && begin_str[0] <= '9'))
        VM_FATAL("parse_begin_string: cannot parse '%s' as a valid gps time", begin_str);
    return (unsigned long)atol(begin_str);
}

uintptr_t parse_coarse_chan_string(MetafitsMetadata *obs_metadata, char *begin_coarse_chan_str)
{
    /* src/metadata.c:1509-1545: "+n" an index, "-n" from the end, else a
     * receiver channel number */
    if (!begin_coarse_chan_str) VM_FATAL("parse_coarse_chan_string: begin_coarse_chan_str cannot be NULL");
    if (begin_coarse_chan_str[0] == '+') return (uintptr_t)atoi(begin_coarse_chan_str);
    if (begin_coarse_chan_str[0] == '-')
        return (uintptr_t)((long)obs_metadata->num_metafits_coarse_chans + atoi(begin_coarse_chan_str));
    if (!(begin_coarse_chan_str[0] >= '0' && begin_coarse_chan_str[0] <= '9'))
        VM_FATAL("parse_coarse_chan_string: cannot parse '%s' as a valid coarse channel",
                 begin_coarse_chan_str);
    const uint32_t rec = (uint32_t)atoi(begin_coarse_chan_str);
    for (size_t i = 0; i < obs_metadata->num_metafits_coarse_chans; i++)
        if (obs_metadata->metafits_coarse_chans[i].rec_chan_number == rec) return i;
    VM_FATAL("parse_coarse_chan_string: no coarse channel with receiver number %u", rec);
    return 0;
}

void vmGetVoltFilename(vcsbeam_context *vm, unsigned int coarse_chan_idx, uint64_t gps_second,
                       char *filename)
{
    /* the file holding gps_second of the coarse channel: the MWAX subfile
     * <obsid>_<subobs>_<rec>.sub or the legacy <obsid>_<gps>_ch<rec>.dat */
    const MetafitsMetadata *md = vm->obs_metadata;
    if (coarse_chan_idx >= md->num_metafits_coarse_chans)
        VM_FATAL("vmGetVoltFilename: coarse channel index %u", coarse_chan_idx);
    const unsigned rec = md->metafits_coarse_chans[coarse_chan_idx].rec_chan_number;
    if (md->mwa_version == VCSMWAXv2) {
        const uint64_t sub = md->obs_id + (gps_second - md->obs_id) / 8 * 8;
        sprintf(filename, "%u_%lu_%03u.sub", md->obs_id, (unsigned long)sub, rec);
    } else {
        vmGetLegacyVoltFilename(vm, coarse_chan_idx, gps_second, filename);
    }
}

void vmGetLegacyVoltFilename(vcsbeam_context *vm, unsigned int coarse_chan_idx, uint64_t gps_second,
                             char *filename)
{
    /* src/metadata.c:1245-1262 */
    const MetafitsMetadata *md = vm->obs_metadata;
    if (coarse_chan_idx >= md->num_metafits_coarse_chans)
        VM_FATAL("vmGetLegacyVoltFilename: coarse channel index %u", coarse_chan_idx);
    sprintf(filename, "%u_%lu_ch%03u.dat", md->obs_id, (unsigned long)gps_second,
            md->metafits_coarse_chans[coarse_chan_idx].rec_chan_number);
}

void vmBindCalibrationData(vcsbeam_context *vm, char *caldir, int cal_type, bool use_bandpass,
                           char *flags_file)
{
    /* src/metadata.c:247-264 */
    vm->cal.cal_type = cal_type;
    vm->cal.use_bandpass = use_bandpass;
    free(vm->cal.caldir);
    vm->cal.caldir = caldir ? strdup(caldir) : NULL;
    free(vm->cal.flags_file);
    vm->cal.flags_file = flags_file ? strdup(flags_file) : NULL;
}

void vmReadCalibration(vcsbeam_context *vm)
{
    /* src/metadata.c:275-286 */
    if (vm->cal.cal_type == CAL_RTS)
        vmLoadRTSSolution(vm);
    else if (vm->cal.cal_type == CAL_OFFRINGA)
        vmLoadOffringaSolution(vm);
    vmSetCustomTileFlags(vm);
}

void vmPrintTitle(vcsbeam_context *vm, const char *title)
{
    /* src/metadata.c:1733-1741 */
    snprintf(vm->log_message, sizeof(vm->log_message), "------- VCSBeam (%s): %s -------",
             vmLibraryVersion(), title);
    if (vm->log) logger_message(vm->log, vm->log_message);
}
