/*
 * psrfits.c -- PSRFITS search-mode output: header population for one coarse
 * channel or a spliced band, and a writer of SUBINT rows.
 *
 * Reference: MWA-vcsbeam src/beam_psrfits.c:39-258
 * (populate_spliced_psrfits_header), 281-466 (populate_psrfits_header),
 * 475-488 (free_psrfits).  The reference writes through psrfits_utils and
 * cfitsio, neither of which is part of its tree; the file written here
 * follows the PSRFITS definition (ATNF, version 6): a primary HDU of
 * observation cards and a SUBINT binary table with one row per
 * sub-integration (TSUBINT, OFFS_SUB, LST_SUB, RA_SUB, DEC_SUB, GLON_SUB,
 * GLAT_SUB, FD_ANG, POS_ANG, PAR_ANG, TEL_AZ, TEL_ZEN, DAT_FREQ, DAT_WTS,
 * DAT_OFFS, DAT_SCL, DATA), in the FITS 4.0 record format.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "vm_internal.h"

#define FITS_BLOCK 2880

void vmEqToGal(double ra, double dec, double *glon, double *glat)
{
    /* rotation from J2000 equatorial to galactic axes (north galactic pole at
     * RA 192.85948 deg, Dec +27.12825 deg; l = 0 towards RA 266.40500 deg,
     * Dec -28.93617 deg): rows are the galactic x, y, z axes */
    static const double A[3][3] = {{-0.0548755604162154, -0.8734370902348850, -0.4838350155487132},
                                   {+0.4941094278755837, -0.4448296299600112, +0.7469822444972189},
                                   {-0.8676661490190047, -0.1980763734312015, +0.4559837761750669}};
    const double v[3] = {cos(dec) * cos(ra), cos(dec) * sin(ra), sin(dec)};
    double g[3];
    for (int i = 0; i < 3; i++) g[i] = A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2];
    double l = atan2(g[1], g[0]);
    *glon = l < 0.0 ? l + 2.0 * M_PI : l;
    *glat = asin(fmax(-1.0, fmin(1.0, g[2])));
}

static void utc_of_mjd(double mjd, char *out, size_t n)
{
    const time_t t = (time_t)llround((mjd - 40587.0) * 86400.0);
    struct tm ts;
    gmtime_r(&t, &ts);
    strftime(out, n, "%Y-%m-%dT%H:%M:%S", &ts);
}

/* fields common to the single-channel and spliced headers */
static void common_header(vcsbeam_context *vm, vm_psrfits *pf, int max_sec_per_file, int outpol,
                          bool is_coherent)
{
    if (!(outpol == 1 || outpol == 4))
        fprintf(stderr, "warning: populate_psrfits_header: unusual number of output pols = %d\n",
                outpol);
    memset(pf, 0, sizeof(*pf));
    MetafitsMetadata *md = vm->obs_metadata;
    vm_psrfits_hdr *h = &pf->hdr;
    snprintf(h->project_id, sizeof(h->project_id), "%s", md->project_id ? md->project_id : "");
    strcpy(h->obs_mode, "SEARCH");
    strcpy(h->observer, "MWA User");
    strcpy(h->telescope, "MWA");
    strcpy(h->frontend, "MWA-RECVR");
    snprintf(h->source, sizeof(h->source), "%u", md->obs_id);
    snprintf(h->backend, sizeof(h->backend), "%s", vmLibraryVersion());
    utc_of_mjd(md->sched_start_mjd, h->date_obs, sizeof(h->date_obs));
    strcpy(h->poln_type, "LIN");
    strcpy(h->track_mode, "TRACK");
    strcpy(h->cal_mode, "OFF");
    strcpy(h->feed_mode, "FA");
    h->scanlen = 1.0;
    h->dt = 1.0 / vm->fine_sample_rate;
    h->npol = outpol;
    h->scan_number = 1;
    h->rcvr_polns = 2;
    h->summed_polns = is_coherent ? 0 : 1;
    h->df = (double)(md->coarse_chan_width_hz / (uint32_t)vm->nfine_chan) / 1e6;
    h->nbits = 8;
    h->nsblk = vm->fine_sample_rate;
    h->fd_hand = 1;
    h->fd_sang = 45.0;
    pf->rows_per_file = max_sec_per_file;
    pf->sub.tsubint = roundf((float)(h->nsblk * h->dt));
    pf->sub.offs = 0.5 * pf->sub.tsubint;
}

static void alloc_arrays(vm_psrfits *pf)
{
    const vm_psrfits_hdr *h = &pf->hdr;
    pf->sub.bytes_per_subint = h->nbits * h->nchan * h->npol * h->nsblk / 8;
    pf->sub.dat_freqs = (float *)malloc(sizeof(float) * (size_t)h->nchan);
    pf->sub.dat_weights = (float *)malloc(sizeof(float) * (size_t)h->nchan);
    pf->sub.dat_offsets = (float *)calloc((size_t)h->nchan * h->npol, sizeof(float));
    pf->sub.dat_scales = (float *)malloc(sizeof(float) * (size_t)h->nchan * h->npol);
    pf->sub.data = (uint8_t *)calloc((size_t)pf->sub.bytes_per_subint, 1);
    if (!pf->sub.dat_freqs || !pf->sub.dat_weights || !pf->sub.dat_offsets || !pf->sub.dat_scales ||
        !pf->sub.data)
        VM_FATAL("out of host memory");
    for (int i = 0; i < h->nchan * h->npol; i++) pf->sub.dat_scales[i] = 1.0f;
    for (int i = 0; i < h->nchan; i++) pf->sub.dat_weights[i] = 1.0f;
}

static void pointing(vcsbeam_context *vm, vm_psrfits *pf, beam_geom *bg, bool is_coherent)
{
    vm_psrfits_hdr *h = &pf->hdr;
    if (is_coherent) {
        h->ra2000 = bg->mean_ra * R2D;
        h->dec2000 = bg->mean_dec * R2D;
    } else {
        h->ra2000 = vm->obs_metadata->ra_tile_pointing_deg;
        h->dec2000 = vm->obs_metadata->dec_tile_pointing_deg;
    }
    dec2hms(h->ra_str, h->ra2000 / 15.0, 0);
    dec2hms(h->dec_str, h->dec2000, 1);
    h->azimuth = bg->az * R2D;
    h->zenith_ang = 90.0 - bg->el * R2D;
    h->beam_FWHM = 0.25;
    h->start_lst = bg->lmst * 60.0 * 60.0;
    h->start_sec = (int)roundf((float)(bg->fracmjd * 86400.0));
    h->start_day = (int)bg->intmjd;
    h->MJD_epoch = bg->intmjd + bg->fracmjd;
    pf->sub.lst = h->start_lst;
    pf->sub.ra = h->ra2000;
    pf->sub.dec = h->dec2000;
    vmEqToGal(h->ra2000 * D2R, h->dec2000 * D2R, &pf->sub.glon, &pf->sub.glat);
    pf->sub.glon *= R2D;
    pf->sub.glat *= R2D;
    pf->sub.tel_az = (float)h->azimuth;
    pf->sub.tel_zen = (float)h->zenith_ang;
}

void populate_psrfits_header(vcsbeam_context *vm, vm_psrfits *pf, int max_sec_per_file,
                             int outpol, beam_geom *beam_geom_vals, char *incoh_basename,
                             bool is_coherent)
{
    common_header(vm, pf, max_sec_per_file, outpol, is_coherent);
    const CoarseChannel *cc = &vm->obs_metadata->metafits_coarse_chans[vm->coarse_chan_idxs_to_process[0]];
    vm_psrfits_hdr *h = &pf->hdr;
    h->fctr = cc->chan_centre_hz / 1e6;
    h->BW = vm->obs_metadata->coarse_chan_width_hz / 1e6;
    h->nchan = vm->nfine_chan;
    alloc_arrays(pf);
    const double f0 = h->fctr - 0.5 * h->BW + 0.5 * h->df;
    for (int i = 0; i < h->nchan; i++) pf->sub.dat_freqs[i] = (float)(f0 + i * h->df);
    if (beam_geom_vals) {
        pointing(vm, pf, beam_geom_vals, is_coherent);
        if (is_coherent)
            snprintf(pf->basefilename, sizeof(pf->basefilename), "%s_%s_%s_%s_ch%03lu", h->project_id,
                     h->source, h->ra_str, h->dec_str, (unsigned long)cc->rec_chan_number);
        else if (incoh_basename)
            snprintf(pf->basefilename, sizeof(pf->basefilename), "%s", incoh_basename);
        else
            snprintf(pf->basefilename, sizeof(pf->basefilename), "%s_%s_incoh_ch%03lu", h->project_id,
                     h->source, (unsigned long)cc->rec_chan_number);
    }
}

void populate_spliced_psrfits_header(vcsbeam_context *vm, vm_psrfits *pf, int max_sec_per_file,
                                     int outpol, beam_geom *beam_geom_vals, char *basename,
                                     bool is_coherent)
{
    common_header(vm, pf, max_sec_per_file, outpol, is_coherent);
    MetafitsMetadata *md = vm->obs_metadata;
    const int first = vm->coarse_chan_idxs_to_process[0];
    const int nco = vm->ncoarse_chans > 0 ? vm->ncoarse_chans : 1;
    const int last = first + nco - 1;
    if (last >= (int)md->num_metafits_coarse_chans)
        VM_FATAL("populate_spliced_psrfits_header: channels %d..%d of %zu", first, last,
                 md->num_metafits_coarse_chans);
    const CoarseChannel *c0 = &md->metafits_coarse_chans[first], *c1 = &md->metafits_coarse_chans[last];
    vm_psrfits_hdr *h = &pf->hdr;
    h->fctr = 0.5 * ((double)c0->chan_centre_hz + (double)c1->chan_centre_hz) / 1e6;
    h->BW = ((double)c1->chan_end_hz - (double)c0->chan_start_hz) / 1e6;
    h->nchan = vm->nfine_chan * nco;
    alloc_arrays(pf);
    const uint32_t w = md->coarse_chan_width_hz / (uint32_t)vm->nfine_chan;
    for (int i = 0; i < h->nchan; i++) pf->sub.dat_freqs[i] = (float)((c0->chan_start_hz + (double)i * w) * 1e-6);
    if (beam_geom_vals) {
        pointing(vm, pf, beam_geom_vals, is_coherent);
        if (basename) {
            snprintf(pf->basefilename, sizeof(pf->basefilename), "%s", basename);
        } else {
            char chans[32];
            if (first == last)
                snprintf(chans, sizeof(chans), "%03lu", (unsigned long)c0->rec_chan_number);
            else
                snprintf(chans, sizeof(chans), "%03lu-%03lu", (unsigned long)c0->rec_chan_number,
                         (unsigned long)c1->rec_chan_number);
            if (is_coherent)
                snprintf(pf->basefilename, sizeof(pf->basefilename), "%s_%s_%s_%s_ch%s", h->project_id,
                         h->source, h->ra_str, h->dec_str, chans);
            else
                snprintf(pf->basefilename, sizeof(pf->basefilename), "%s_%s_incoh_ch%s", h->project_id,
                         h->source, chans);
        }
    }
}

void free_psrfits(vm_psrfits *pf)
{
    vmClosePsrfits(pf);
    free(pf->sub.dat_freqs);
    free(pf->sub.dat_weights);
    free(pf->sub.dat_offsets);
    free(pf->sub.dat_scales);
    free(pf->sub.data);
    pf->sub.dat_freqs = pf->sub.dat_weights = pf->sub.dat_offsets = pf->sub.dat_scales = NULL;
    pf->sub.data = NULL;
}

/* ------------------------------------------------------------- writer -- */
typedef struct {
    char *buf;
    size_t n, cap;
} cardbuf;

static void put_card(cardbuf *b, const char *card)
{
    if (b->n + 80 > b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64 * 80;
        b->buf = (char *)realloc(b->buf, b->cap);
        if (!b->buf) VM_FATAL("out of host memory");
    }
    char c[81];
    snprintf(c, sizeof(c), "%-80s", card);
    memcpy(b->buf + b->n, c, 80);
    b->n += 80;
}

static void card_str(cardbuf *b, const char *k, const char *v)
{
    char c[96], q[72];
    size_t j = 0;
    for (const char *p = v; *p && j < 66; p++) {
        if (*p == '\'') q[j++] = '\'';
        q[j++] = *p;
    }
    q[j] = 0;
    char val[80];
    snprintf(val, sizeof(val), "'%-8s'", q);
    snprintf(c, sizeof(c), "%-8.8s= %-20s", k, val);
    put_card(b, c);
}

static void card_int(cardbuf *b, const char *k, long v)
{
    char c[96];
    snprintf(c, sizeof(c), "%-8.8s= %20ld", k, v);
    put_card(b, c);
}

static void card_dbl(cardbuf *b, const char *k, double v)
{
    char c[96];
    snprintf(c, sizeof(c), "%-8.8s= %20.14G", k, v);
    put_card(b, c);
}

static void card_log(cardbuf *b, const char *k, int v)
{
    char c[96];
    snprintf(c, sizeof(c), "%-8.8s= %20s", k, v ? "T" : "F");
    put_card(b, c);
}

static void pad_cards(cardbuf *b)
{
    put_card(b, "END");
    while (b->n % FITS_BLOCK) put_card(b, "");
}

static void be_put(uint8_t *dst, const void *src, int size, size_t n)
{
    /* n big-endian elements of size bytes */
    const uint8_t *s = (const uint8_t *)src;
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < size; k++) dst[i * size + k] = s[i * size + size - 1 - k];
}

static int open_next_file(vm_psrfits *pf)
{
    const vm_psrfits_hdr *h = &pf->hdr;
    pf->filenum++;
    char path[1100];
    snprintf(path, sizeof(path), "%s_%04d.fits", pf->basefilename, pf->filenum);
    FILE *f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "vmWritePsrfitsSubint: cannot open %s\n", path);
        return -1;
    }
    char now[32];
    time_t t = time(NULL);
    struct tm ts;
    gmtime_r(&t, &ts);
    strftime(now, sizeof(now), "%Y-%m-%dT%H:%M:%S", &ts);
    cardbuf b = {0};
    card_log(&b, "SIMPLE", 1);
    card_int(&b, "BITPIX", 8);
    card_int(&b, "NAXIS", 0);
    card_log(&b, "EXTEND", 1);
    card_str(&b, "HDRVER", "6.1");
    card_str(&b, "FITSTYPE", "PSRFITS");
    card_str(&b, "DATE", now);
    card_str(&b, "OBSERVER", h->observer);
    card_str(&b, "PROJID", h->project_id);
    card_str(&b, "TELESCOP", h->telescope);
    card_dbl(&b, "ANT_X", -2559454.08);
    card_dbl(&b, "ANT_Y", 5095372.14);
    card_dbl(&b, "ANT_Z", -2849057.18);
    card_str(&b, "FRONTEND", h->frontend);
    card_int(&b, "NRCVR", h->rcvr_polns);
    card_str(&b, "FD_POLN", h->poln_type);
    card_int(&b, "FD_HAND", h->fd_hand);
    card_dbl(&b, "FD_SANG", h->fd_sang);
    card_dbl(&b, "FD_XYPH", h->fd_xyph);
    card_str(&b, "BACKEND", h->backend);
    card_str(&b, "BECONFIG", "N/A");
    card_int(&b, "BE_PHASE", h->be_phase);
    card_str(&b, "OBS_MODE", h->obs_mode);
    card_str(&b, "DATE-OBS", h->date_obs);
    card_dbl(&b, "OBSFREQ", h->fctr);
    card_dbl(&b, "OBSBW", h->BW);
    card_int(&b, "OBSNCHAN", h->nchan);
    card_dbl(&b, "CHAN_DM", h->chan_dm);
    card_str(&b, "SRC_NAME", h->source);
    card_str(&b, "COORD_MD", "J2000");
    card_dbl(&b, "EQUINOX", 2000.0);
    card_str(&b, "RA", h->ra_str);
    card_str(&b, "DEC", h->dec_str);
    card_dbl(&b, "BMAJ", h->beam_FWHM);
    card_dbl(&b, "BMIN", h->beam_FWHM);
    card_str(&b, "STT_CRD1", h->ra_str);
    card_str(&b, "STT_CRD2", h->dec_str);
    card_str(&b, "TRK_MODE", h->track_mode);
    card_str(&b, "STP_CRD1", h->ra_str);
    card_str(&b, "STP_CRD2", h->dec_str);
    card_dbl(&b, "SCANLEN", h->scanlen * pf->rows_per_file);
    card_str(&b, "FD_MODE", h->feed_mode);
    card_str(&b, "CAL_MODE", h->cal_mode);
    card_int(&b, "STT_IMJD", h->start_day);
    card_int(&b, "STT_SMJD", h->start_sec + pf->tot_rows * (int)lround(pf->sub.tsubint));
    card_dbl(&b, "STT_OFFS", 0.0);
    card_dbl(&b, "STT_LST", h->start_lst);
    pad_cards(&b);
    const long nc = h->nchan, np = h->npol;
    const long rowb = 7 * 8 + 5 * 4 + nc * 8 + nc * 4 + 2 * nc * np * 4 + (long)pf->sub.bytes_per_subint;
    const size_t prim = b.n;
    card_str(&b, "XTENSION", "BINTABLE");
    card_int(&b, "BITPIX", 8);
    card_int(&b, "NAXIS", 2);
    card_int(&b, "NAXIS1", rowb);
    const size_t naxis2_card = b.n;
    card_int(&b, "NAXIS2", 0);
    card_int(&b, "PCOUNT", 0);
    card_int(&b, "GCOUNT", 1);
    card_int(&b, "TFIELDS", 17);
    card_str(&b, "EXTNAME", "SUBINT");
    struct { const char *name, *unit; char form[24]; } cols[17] = {
        {"TSUBINT", "s", "1D"}, {"OFFS_SUB", "s", "1D"}, {"LST_SUB", "s", "1D"},
        {"RA_SUB", "deg", "1D"}, {"DEC_SUB", "deg", "1D"}, {"GLON_SUB", "deg", "1D"},
        {"GLAT_SUB", "deg", "1D"}, {"FD_ANG", "deg", "1E"}, {"POS_ANG", "deg", "1E"},
        {"PAR_ANG", "deg", "1E"}, {"TEL_AZ", "deg", "1E"}, {"TEL_ZEN", "deg", "1E"},
        {"DAT_FREQ", "MHz", ""}, {"DAT_WTS", "", ""}, {"DAT_OFFS", "", ""}, {"DAT_SCL", "", ""},
        {"DATA", "Jy", ""}};
    snprintf(cols[12].form, 24, "%ldD", nc);
    snprintf(cols[13].form, 24, "%ldE", nc);
    snprintf(cols[14].form, 24, "%ldE", nc * np);
    snprintf(cols[15].form, 24, "%ldE", nc * np);
    snprintf(cols[16].form, 24, "%ldB", (long)pf->sub.bytes_per_subint);
    for (int i = 0; i < 17; i++) {
        char k[16];
        snprintf(k, sizeof(k), "TTYPE%d", i + 1);
        card_str(&b, k, cols[i].name);
        snprintf(k, sizeof(k), "TFORM%d", i + 1);
        card_str(&b, k, cols[i].form);
        if (cols[i].unit[0]) {
            snprintf(k, sizeof(k), "TUNIT%d", i + 1);
            card_str(&b, k, cols[i].unit);
        }
    }
    char tdim[48];
    snprintf(tdim, sizeof(tdim), "(%ld,%ld,%d)", nc, np, h->nsblk);
    card_str(&b, "TDIM17", tdim);
    card_str(&b, "INT_TYPE", "TIME");
    card_str(&b, "INT_UNIT", "SEC");
    card_str(&b, "SCALE", "FluxDen");
    card_str(&b, "POL_TYPE", np == 4 ? "IQUV" : (np == 1 ? "AA+BB" : "AABB"));
    card_int(&b, "NPOL", np);
    card_dbl(&b, "TBIN", h->dt);
    card_int(&b, "NBIN", 1);
    card_int(&b, "NCHAN", nc);
    card_dbl(&b, "CHAN_BW", h->df);
    card_int(&b, "NCH_FILE", nc);
    card_int(&b, "NCH_STRT", 0);
    card_int(&b, "NSBLK", h->nsblk);
    card_int(&b, "NBITS", h->nbits);
    card_int(&b, "ZERO_OFF", 0);
    card_int(&b, "NSUBOFFS", pf->tot_rows);
    pad_cards(&b);
    if (fwrite(b.buf, 1, b.n, f) != b.n) {
        free(b.buf);
        fclose(f);
        return -1;
    }
    free(b.buf);
    (void)prim;
    pf->fptr = f;
    pf->naxis2_pos = (long)naxis2_card;
    pf->data_start = (long)b.n;
    pf->rows_in_file = 0;
    return 0;
}

int vmWritePsrfitsSubint(vm_psrfits *pf)
{
    if (!pf->sub.data || pf->hdr.nchan <= 0) VM_FATAL("vmWritePsrfitsSubint: header not populated");
    if (pf->fptr && pf->rows_per_file > 0 && pf->rows_in_file >= pf->rows_per_file) vmClosePsrfits(pf);
    if (!pf->fptr && open_next_file(pf) != 0) return -1;
    const long nc = pf->hdr.nchan, np = pf->hdr.npol;
    const size_t rowb = 7 * 8 + 5 * 4 + nc * 8 + nc * 4 + 2 * nc * np * 4 + (size_t)pf->sub.bytes_per_subint;
    uint8_t *row = (uint8_t *)malloc(rowb);
    if (!row) VM_FATAL("out of host memory");
    const double d7[7] = {pf->sub.tsubint, pf->sub.offs, pf->sub.lst, pf->sub.ra, pf->sub.dec,
                          pf->sub.glon, pf->sub.glat};
    const float f5[5] = {pf->sub.feed_ang, pf->sub.pos_ang, pf->sub.par_ang, pf->sub.tel_az,
                         pf->sub.tel_zen};
    size_t o = 0;
    be_put(row + o, d7, 8, 7);
    o += 56;
    be_put(row + o, f5, 4, 5);
    o += 20;
    for (long i = 0; i < nc; i++) {
        const double fr = pf->sub.dat_freqs[i];
        be_put(row + o + 8 * i, &fr, 8, 1);
    }
    o += 8 * nc;
    be_put(row + o, pf->sub.dat_weights, 4, (size_t)nc);
    o += 4 * nc;
    be_put(row + o, pf->sub.dat_offsets, 4, (size_t)(nc * np));
    o += 4 * nc * np;
    be_put(row + o, pf->sub.dat_scales, 4, (size_t)(nc * np));
    o += 4 * nc * np;
    memcpy(row + o, pf->sub.data, (size_t)pf->sub.bytes_per_subint);
    FILE *f = (FILE *)pf->fptr;
    const int ok = fwrite(row, 1, rowb, f) == rowb;
    free(row);
    if (!ok) return -1;
    pf->rows_in_file++;
    pf->tot_rows++;
    pf->sub.offs += pf->sub.tsubint;
    return 0;
}

void vmClosePsrfits(vm_psrfits *pf)
{
    FILE *f = (FILE *)pf->fptr;
    if (!f) return;
    /* the row count into NAXIS2, then pad the table to whole records */
    const long nc = pf->hdr.nchan, np = pf->hdr.npol;
    const long rowb = 7 * 8 + 5 * 4 + nc * 8 + nc * 4 + 2 * nc * np * 4 + (long)pf->sub.bytes_per_subint;
    const long end = pf->data_start + rowb * pf->rows_in_file;
    static const char zeros[FITS_BLOCK] = {0};
    fseek(f, end, SEEK_SET);
    fwrite(zeros, 1, (size_t)((FITS_BLOCK - end % FITS_BLOCK) % FITS_BLOCK), f);
    char c[40];
    snprintf(c, sizeof(c), "%-8.8s= %20d", "NAXIS2", pf->rows_in_file);
    char card[81];
    snprintf(card, sizeof(card), "%-80.80s", c);
    fseek(f, pf->naxis2_pos, SEEK_SET);
    fwrite(card, 1, 80, f);
    fclose(f);
    pf->fptr = NULL;
    pf->rows_in_file = 0;
}
