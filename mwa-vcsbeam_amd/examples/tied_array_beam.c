/*
 * tied_array_beam.c -- a C caller of the library, following
 * app/make_mwa_tied_array_beam.c's call sequence with the by-file set-up:
 *
 *   tied_array_beam METAFITS DATADIR POINTINGS CALFILE OUTDIR COARSE_CHAN BEGIN NSECONDS
 *
 * Reads MWAX subfiles, an Offringa calibration and a pointings file; writes
 * one PSRFITS (Stokes IQUV) per pointing into OUTDIR.  It uses only
 * include/vcsbeam.h and libvcsbeam_amd.so.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vcsbeam.h"

int main(int argc, char **argv)
{
    if (argc != 9) {
        fprintf(stderr, "usage: %s METAFITS DATADIR POINTINGS CALFILE OUTDIR COARSE_CHAN BEGIN "
                        "NSECONDS\n", argv[0]);
        return 2;
    }
    const int nseconds = atoi(argv[8]);
    vcsbeam_context *vm = vmInit(false);
    vmLoadObsMetafits(vm, argv[1]);
    vmBindObsData(vm, argv[6], 1, 0, argv[7], nseconds, 0, argv[2]);
    vmLoadCalMetafits(vm, argv[1]);
    vmBindCalibrationData(vm, argv[4], CAL_OFFRINGA, false, NULL);
    vmParsePointingFile(vm, argv[3]);
    vm->out_nstokes = 4;
    vmLoadFilter(vm, "FINEPFB", ANALYSIS_FILTER, 128);
    vmInitForwardPFB(vm, 128, PFB_FULL_PRECISION | PFB_EMULATE_FPGA);
    vmSetNumPointings(vm, vm->npointing);
    vmSetMaxGPUMem(vm, 1);
    vmMallocVDevice(vm);
    vmMallocJVDevice(vm);
    vmMallocEHost(vm);
    vmMallocEDevice(vm);
    vmMallocSHost(vm);
    vmMallocSDevice(vm);
    vmMallocJHost(vm);
    vmMallocJDevice(vm);
    vmMallocDHost(vm);
    vmMallocDDevice(vm);
    vmMallocPQIdxsHost(vm);
    vmMallocPQIdxsDevice(vm);
    vmSetPolIdxLists(vm);
    vmPushPolIdxLists(vm);
    vmReadCalibration(vm);
    parse_calibration_correction_file(vm->gps_seconds_to_process[0], &vm->cal);
    vmApplyCalibrationCorrections(vm);
    vmCreatePrimaryBeam(vm);
    vmCreateGeometricDelays(vm);
    vmCreateCudaStreams(vm);

    /* header geometry at the first second (app/make_mwa_tied_array_beam.c:139-147) */
    const double sec_offset =
        vm->gps_seconds_to_process[0] - vm->obs_metadata->sched_start_gps_time_ms / 1000.0;
    const double mjd = vm->obs_metadata->sched_start_mjd + sec_offset / 86400.0;
    beam_geom *bg = (beam_geom *)calloc(vm->npointing, sizeof(beam_geom));
    mpi_psrfits *mpfs = (mpi_psrfits *)calloc(vm->npointing, sizeof(mpi_psrfits));
    if (!bg || !mpfs) return 1;
    for (unsigned p = 0; p < vm->npointing; p++) {
        calc_beam_geom(vm->ras_hours[p], vm->decs_degs[p], mjd, &bg[p]);
        populate_psrfits_header(vm, &mpfs[p].coarse_chan_pf, 200, 4, &bg[p], NULL, true);
        char base[1024];
        snprintf(base, sizeof(base), "%s/%s", argv[5], mpfs[p].coarse_chan_pf.basefilename);
        strcpy(mpfs[p].coarse_chan_pf.basefilename, base);
    }
    vmCreateStatistics(vm, mpfs);

    for (int s = 0; s < nseconds; s++) {
        if (vmReadNextSecond(vm) != VM_SUCCESS) {
            fprintf(stderr, "no data for second %d\n", s);
            return 1;
        }
        /* geometry, beam and Jones inverses at the middle of this second
         * (app/make_mwa_tied_array_beam.c:257) */
        vmCalcJonesAndDelays(vm, vm->ras_hours, vm->decs_degs, bg);
        vmPushPhi(vm);
        vmPushJ(vm);
        vmBeamformSecond(vm);
        vm->v->locked = false;
        vmSendSToFits(vm, mpfs);
        for (unsigned p = 0; p < vm->npointing; p++)
            if (vmWritePsrfitsSubint(&mpfs[p].coarse_chan_pf) != 0) return 1;
    }
    for (unsigned p = 0; p < vm->npointing; p++) {
        printf("%s\n", mpfs[p].coarse_chan_pf.basefilename);
        free_psrfits(&mpfs[p].coarse_chan_pf);
    }
    vmDestroyStatistics(vm);
    vmDestroyCudaStreams(vm);
    vmFreeVDevice(vm);
    vmFreeJVDevice(vm);
    vmFreeEHost(vm);
    vmFreeEDevice(vm);
    vmFreeSHost(vm);
    vmFreeSDevice(vm);
    vmFreeJHost(vm);
    vmFreeJDevice(vm);
    vmFreeDHost(vm);
    vmFreeDDevice(vm);
    vmFreePQIdxsDevice(vm);
    vmFreePQIdxsHost(vm);
    free_primary_beam(&vm->pb);
    free(mpfs);
    free(bg);
    destroy_vcsbeam_context(vm);
    return 0;
}
