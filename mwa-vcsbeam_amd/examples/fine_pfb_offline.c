/*
 * fine_pfb_offline.c -- a C caller of the library, following
 * app/fine_pfb_offline.c's call sequence:
 *
 *   fine_pfb_offline METAFITS DATADIR COARSE_CHAN BEGIN NSECONDS [NCHUNKS [FILTER]]
 *
 * Reads MWAX subfiles from DATADIR, runs the forward fine PFB (PFB_SMART,
 * host output) and writes each second as a legacy recombined file
 * <obsid>_<gps>_ch<rec>.dat into DATADIR, as vmWritePFBOutputToFile does.
 * It uses only include/vcsbeam.h and libvcsbeam_amd.so.
 */
#include <stdio.h>
#include <stdlib.h>

#include "vcsbeam.h"

int main(int argc, char **argv)
{
    if (argc < 6 || argc > 8) {
        fprintf(stderr, "usage: %s METAFITS DATADIR COARSE_CHAN BEGIN NSECONDS [NCHUNKS [FILTER]]\n",
                argv[0]);
        return 2;
    }
    const int nchunks = argc > 6 ? atoi(argv[6]) : 1;
    char *filter = argc > 7 ? argv[7] : "FINEPFB";
    vcsbeam_context *vm = vmInit(false);
    vmLoadObsMetafits(vm, argv[1]);
    vmBindObsData(vm, argv[3], 1, 0, argv[4], atoi(argv[5]), 0, argv[2]);
    vmPrintTitle(vm, "Offline PFB");
    if (vm->obs_metadata->mwa_version != VCSMWAXv2) {
        fprintf(stderr, "error: observation %u does not appear to be coarse-channelised MWAX "
                        "data\n", vm->obs_metadata->obs_id);
        return 1;
    }
    const int K = 128;
    vmLoadFilter(vm, filter, ANALYSIS_FILTER, K);
    vm->chunks_per_second = nchunks;
    vmInitForwardPFB(vm, K, PFB_SMART | PFB_MALLOC_HOST_OUTPUT);
    while (vmReadNextSecond(vm) == VM_SUCCESS) {
        vmExecuteForwardPFB(vm); /* unlocks the read buffer after its last chunk */
        vmWritePFBOutputToFile(vm);
    }
    vmReportPerformanceStats(vm);
    destroy_vcsbeam_context(vm);
    return 0;
}
