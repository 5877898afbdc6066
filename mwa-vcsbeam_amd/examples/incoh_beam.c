/*
 * incoh_beam.c -- a C caller of the library, following
 * app/make_mwa_incoh_beam.c's call sequence:
 *
 *   incoh_beam METAFITS DATADIR OUTDIR COARSE_CHAN BEGIN NSECONDS
 *
 * Reads legacy recombined seconds (<obsid>_<gps>_ch<rec>.dat in DATADIR),
 * forms the incoherent beam with cu_form_incoh_beam and writes one PSRFITS
 * (Stokes I, the tile pointing, geometry at the scheduled start) into
 * OUTDIR, printing its base name.  It uses only include/vcsbeam.h and
 * libvcsbeam_amd.so.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vcsbeam.h"

int main(int argc, char **argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s METAFITS DATADIR OUTDIR COARSE_CHAN BEGIN NSECONDS\n", argv[0]);
        return 2;
    }
    const int max_sec_per_file = 200, nstokes = 1;
    vcsbeam_context *vm = vmInit(false);
    vmLoadObsMetafits(vm, argv[1]);
    vmBindObsData(vm, argv[4], 1, 0, argv[5], atoi(argv[6]), 0, argv[2]);
    if (vm->obs_metadata->mwa_version == VCSMWAXv2) {
        fprintf(stderr, "error: observation %u is not legacy recombined data\n",
                vm->obs_metadata->obs_id);
        return 1;
    }

    /* sizes and host/device buffers (app/make_mwa_incoh_beam.c:65-91) */
    const size_t nchans = vm->obs_metadata->num_volt_fine_chans_per_coarse;
    const size_t ninputs = vm->obs_metadata->num_rf_inputs;
    const unsigned nsamples = (unsigned)vm->sample_rate;
    const size_t data_size = (size_t)vm->bytes_per_second;
    const size_t Iscaled_size = nchans * nsamples;
    uint8_t *data, *d_data;
    allocate_input_output_arrays((void **)&data, (void **)&d_data, data_size);
    float *d_incoh = (float *)vmDeviceMalloc(nchans * nsamples * sizeof(float));
    float *d_offsets = (float *)vmDeviceMalloc(nchans * sizeof(float));
    float *d_scales = (float *)vmDeviceMalloc(nchans * sizeof(float));
    uint8_t *d_Iscaled = (uint8_t *)vmDeviceMalloc(Iscaled_size);

    /* the tile pointing at the scheduled start (app/make_mwa_incoh_beam.c:93-99) */
    beam_geom bg;
    calc_beam_geom(vm->obs_metadata->ra_tile_pointing_deg / 15.0,
                   vm->obs_metadata->dec_tile_pointing_deg, vm->obs_metadata->sched_start_mjd,
                   &bg);
    vm_psrfits pf;
    memset(&pf, 0, sizeof(pf));
    populate_psrfits_header(vm, &pf, max_sec_per_file, nstokes, &bg, NULL, false);
    char base[1024];
    snprintf(base, sizeof(base), "%s/%s", argv[3], pf.basefilename);
    strcpy(pf.basefilename, base);

    int rc = 0;
    for (int t = 0; t < vm->num_gps_seconds_to_process; t++) {
        /* read_step (app/make_mwa_incoh_beam.c:113) through the registered reader */
        if (vm->read_second(vm->read_second_user, vm->gps_seconds_to_process[t],
                            vm->coarse_chan_idxs_to_process[0], data, data_size) != 0) {
            fprintf(stderr, "no data for GPS second %u\n", vm->gps_seconds_to_process[t]);
            rc = 1;
            break;
        }
        logger_start_stopwatch(vm->log, "calc", true);
        cu_form_incoh_beam(data, d_data, data_size, d_incoh, nsamples, (int)nchans, (int)ninputs,
                           pf.sub.dat_offsets, d_offsets, pf.sub.dat_scales, d_scales,
                           pf.sub.data, d_Iscaled, Iscaled_size);
        logger_stop_stopwatch(vm->log, "calc");
        if (vmWritePsrfitsSubint(&pf) != 0) {
            rc = 1;
            break;
        }
    }
    if (rc == 0) printf("%s\n", pf.basefilename);
    free_psrfits(&pf);
    vmDeviceFree(d_Iscaled);
    vmDeviceFree(d_scales);
    vmDeviceFree(d_offsets);
    vmDeviceFree(d_incoh);
    free_input_output_arrays(data, d_data);
    destroy_vcsbeam_context(vm);
    return rc;
}
