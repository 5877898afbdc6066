"""vmFormIncohBeamNoAlloc, the incoherent beam on device buffers with no
allocation (its renormalisation partials in the library's fixed device
array), against cu_form_incoh_beam, the entry point it is a variant of.

Two calls in a row, the second with a smaller input (fewer samples,
channels and inputs, the inputs not a multiple of 16), must give every
output (the sums, the offsets and scales, the 8-bit values) bit for bit
equal to cu_form_incoh_beam's on the same data; sizes past the array's
limit are rejected with -1 before anything is launched.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _no_alloc(pkg, lib, data, ns, nc, ni):
    d_data = pkg.DeviceArray.from_numpy(data)
    d_incoh, d_off, d_scl, d_Is = (pkg.DeviceArray(n) for n in (ns * nc * 4, nc * 4, nc * 4,
                                                                ns * nc))
    rc = lib.vmFormIncohBeamNoAlloc(d_data.ptr, data.nbytes, d_incoh.ptr, ns, nc, ni, d_off.ptr,
                                    d_scl.ptr, d_Is.ptr, ns * nc, None)
    assert rc == 0
    lib.vmDeviceSynchronize()
    return (d_incoh.to_numpy(np.float32, (ns, nc)), d_off.to_numpy(np.float32, nc),
            d_scl.to_numpy(np.float32, nc), d_Is.to_numpy(np.uint8, (ns, nc)))


def test_no_alloc_incoherent_beam_equals_cu_form_incoh_beam(pkg, gpu_lib):
    from vcsbeam_amd.apps import incoherent_beam
    from vcsbeam_amd.workloads import legacy_int4
    lib = gpu_lib
    for call, (ns, nc, ni, seed) in enumerate([(1000, 128, 256, 3), (213, 40, 24, 4)]):
        print(f"call {call}: {ns} samples x {nc} channels x {ni} inputs", flush=True)
        data = legacy_int4(ns * nc * ni, sigma=2.5, seed=seed)
        want = incoherent_beam(data, ns, nc, ni)
        got = _no_alloc(pkg, lib, data, ns, nc, ni)
        for w, g, name in zip(want, got, ("incoh", "offsets", "scales", "Iscaled")):
            np.testing.assert_array_equal(g, w, err_msg=name)


def test_no_alloc_rejects_sizes_past_its_limit(pkg, gpu_lib):
    lib = gpu_lib
    d = pkg.DeviceArray(4096)
    limit = pkg.VM_INCOH_SCRATCH_PAIRS
    print("rejected sizes", flush=True)
    # more channels than the partials array holds
    assert lib.vmFormIncohBeamNoAlloc(d.ptr, 1 << 40, d.ptr, 1, limit + 1, 1, d.ptr, d.ptr,
                                      d.ptr, 1 << 40, None) == -1
    # data or output smaller than the sizes say
    assert lib.vmFormIncohBeamNoAlloc(d.ptr, 100, d.ptr, 10, 4, 16, d.ptr, d.ptr, d.ptr, 40,
                                      None) == -1
    assert lib.vmFormIncohBeamNoAlloc(d.ptr, 640, d.ptr, 10, 4, 16, d.ptr, d.ptr, d.ptr, 39,
                                      None) == -1
