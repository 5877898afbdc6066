// beamform_mfma.hip -- precision many-beam tied-array beamformer on MFMA.
// (Filled in by the MFMA implementation; the entry point exists so that the
// library links and reports clearly when the form is selected.)
#include <hip/hip_runtime.h>

#include "vm_internal.h"

extern "C" void vmBeamformChunkMFMA(vcsbeam_context *vm)
{
    (void)vm;
    VM_FATAL("vmBeamformChunkMFMA: MFMA form not built");
}
