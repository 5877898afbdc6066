"""GPU parity of the forward fine PFB against the CPU oracle.

Workload: 128 tiles x 2 pols, one 0.1 s "second" (2 MWAX blocks = 128,000
samples per input), seeded Gaussian 8+8-bit voltages (sigma 16), 12-tap x
128-channel Kaiser filter quantised to int16 (and the reference's FINEPFB).
Bars: integer WOLA and FPGA rounding bit-exact; FFT'd output <= 1e-5
norm-wise relative to the FP64 oracle; 4+4-bit output equal except where the
FP64 value lies within 1e-4 of a rounding tie.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-5


def _obs(pkg, ntiles=128, nseconds=1, finepfb=False):
    from vcsbeam_amd.workloads import Observation
    filt = np.loadtxt(os.path.join(GOLD, "FINEPFB.dat")) if finepfb else None
    return Observation(ntiles=ntiles, blocks_per_second=2, nseconds=nseconds, filter_coeffs=filt)


def _htr(obs, t):
    """Oracle input buffer for second t: [margin block | second t] in MWAX layout."""
    cur = obs.second(t)
    margin = obs.second(t - 1)[-1:] if t > 0 else np.zeros_like(cur[:1])
    return np.ascontiguousarray(np.concatenate([margin, cur], axis=0))


@pytest.mark.parametrize("fpga,finepfb", [(False, False), (True, False), (True, True)])
def test_fused_pfb_matches_oracle(pkg, orc, gpu_lib, fpga, finepfb):
    from vcsbeam_amd.apps import fine_pfb_offline
    obs = _obs(pkg, finepfb=finepfb)
    flags = pkg.PFB_FULL_PRECISION | pkg.PFB_MALLOC_HOST_OUTPUT | (pkg.PFB_EMULATE_FPGA if fpga else 0)
    print(f"fused PFB fpga={fpga} finepfb={finepfb}", flush=True)
    got = fine_pfb_offline(obs, pfb_flags=flags)[0]
    ref = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                          obs.out_idx(), fpga=fpga)
    err = orc.rel_err(got, ref)
    print("rel err", err)
    assert err <= TOL


def test_staged_wola_and_rounding_bit_exact_and_fused_identical(pkg, orc, gpu_lib):
    lib = gpu_lib
    obs = _obs(pkg)
    flags = pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA | pkg.PFB_STAGED
    vm = lib.vmInit(False)
    v = vm.contents
    lib.vmBindObsMetadata(vm, C.byref(obs.md), C.byref(obs.vmd), 0, obs.md.obs_id, 1)
    co = np.ascontiguousarray(obs.filter)
    lib.vmSetFilter(vm, co.ctypes.data_as(C.POINTER(C.c_double)), len(co), 0, 128)
    lib.vmInitForwardPFB(vm, 128, flags)
    fp = v.fpfb.contents
    sec = obs.second(0)
    assert lib.vmReadNextSecondFrom(vm, sec.ctypes.data, sec.nbytes) == 0
    lib.vmUploadForwardPFBChunk(vm)
    lib.vmWOLAChunk(vm)
    nspc, I = fp.nspectra_per_chunk, fp.I
    b = np.empty((nspc, I, 128, 2), dtype=np.int32)
    lib.vmMemcpyDtoH(b.ctypes.data, fp.d_weighted_overlap_add, b.nbytes)
    bref = orc.wola(_htr(obs, 0), I, nspc, obs.filter.astype(np.int32))
    np.testing.assert_array_equal(b, bref)
    lib.vmFPGARoundingChunk(vm)
    r = np.empty((nspc, I, 128, 2), dtype=np.float32)
    lib.vmMemcpyDtoH(r.ctypes.data, fp.d_weighted_overlap_add, r.nbytes)
    np.testing.assert_array_equal(r, orc.fpga_round(bref))
    lib.vmFFTChunk(vm)
    lib.vmPackChunk(vm)
    staged = np.empty((128, I, nspc), dtype=np.complex64)
    lib.vmMemcpyDtoH(staged.ctypes.data, fp.d_vcs_data, staged.nbytes)
    lib.destroy_vcsbeam_context(vm)
    # fused path on the same data: identical bits
    from vcsbeam_amd.apps import fine_pfb_offline
    fused = fine_pfb_offline(obs, pfb_flags=pkg.PFB_COMPLEX_FLOAT32 | pkg.PFB_EMULATE_FPGA |
                             pkg.PFB_MALLOC_HOST_OUTPUT | 0x0C)[0]
    np.testing.assert_array_equal(staged.view(np.uint32), fused.view(np.uint32))


def test_int4_legacy_output(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import fine_pfb_offline
    obs = _obs(pkg)
    got = fine_pfb_offline(obs, pfb_flags=pkg.PFB_SMART | pkg.PFB_MALLOC_HOST_OUTPUT)[0]
    ref_c = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                            obs.out_idx(), fpga=True)
    ref = orc.pack_int4(ref_c)
    diff = got != ref
    # differences are allowed only next to a DEMOTE rounding tie (x.5) of
    # either component, where an fp32 FFT and the FP64 oracle can round apart
    def near_tie(x):
        fr = np.abs(x) - np.floor(np.abs(x))
        return np.abs(fr - 0.5) < 1e-4 * np.maximum(1.0, np.abs(x))
    tie = near_tie(ref_c.real) | near_tie(ref_c.imag)
    print("int4 mismatches", int(diff.sum()), "near ties", int(tie.sum()))
    assert not np.any(diff & ~tie)


def test_chunking_and_second_continuity(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import fine_pfb_offline
    obs = _obs(pkg, nseconds=2)
    flags = pkg.PFB_COMPLEX_FLOAT32 | pkg.PFB_EMULATE_FPGA | pkg.PFB_MALLOC_HOST_OUTPUT | 0x0C
    one = fine_pfb_offline(obs, nseconds=2, nchunks=1, pfb_flags=flags)
    two = fine_pfb_offline(obs, nseconds=2, nchunks=2, pfb_flags=flags)
    for a, b in zip(one, two):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    ref = orc.pfb_forward(_htr(obs, 1), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                          obs.out_idx(), fpga=True)
    # VM_FLT is [c][i][s]; the oracle is [s][c][i]
    got = one[1].transpose(2, 0, 1)
    assert orc.rel_err(got, ref) <= TOL


@pytest.mark.parametrize("finepfb", [False, True])
def test_fused_pfb_fp16_output_is_fp32_output_rounded(pkg, gpu_lib, finepfb):
    """PFB_COMPLEX_FLOAT16 (VM_HALF, the MFMA forms' input): the fused kernel's
    FP16 output is its FP32 output (PFB_COMPLEX_FLOAT32, held to the oracle at
    1e-5 above) rounded to nearest even, bit for bit, in the same [c][i][s]
    layout; with the int16 taps and with FINEPFB's split taps, over two
    chunks."""
    from vcsbeam_amd.apps import fine_pfb_offline
    obs = _obs(pkg, ntiles=32, finepfb=finepfb)
    base = (pkg.PFB_FULL_PRECISION & ~pkg.PFB_TYPE_MASK) | pkg.PFB_MALLOC_HOST_OUTPUT | \
        pkg.PFB_EMULATE_FPGA
    print(f"fused PFB to FP32 and FP16, finepfb={finepfb}", flush=True)
    f32 = fine_pfb_offline(obs, nchunks=2, pfb_flags=base | pkg.PFB_COMPLEX_FLOAT32)[0]
    f16 = fine_pfb_offline(obs, nchunks=2, pfb_flags=base | pkg.PFB_COMPLEX_FLOAT16)[0]
    want = np.stack([f32.real, f32.imag], axis=-1).astype(np.float16)
    assert f16.shape == want.shape
    np.testing.assert_array_equal(f16.view(np.uint16), want.view(np.uint16))
    assert np.abs(f32).max() < 6e4  # within FP16's range
