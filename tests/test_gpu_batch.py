"""GPU parity of the batched many-channel beamformer: per coarse channel the
fused forward PFB and the prep pass, then one persistent MFMA launch over
every channel's work items, captured in a hipGraph.

Workload: 3 coarse channels (rec 133-135) of 32 tiles x 2 pols, 0.1 s each,
sharing one array layout, with 0, 3 and 7 flagged tiles (D = 0 and the
metafits flag) and 1, 5 and 23 beams within 15 deg of boresight; FP16 form.
Each channel's Stokes must agree with the single-channel MFMA path on the
same context to 1e-5 norm-wise (the batch leaves flagged tiles out of the
GEMM, which only reorders FP32 sums), and with the FP64 oracle chain to the
MFMA form's bar (1e-3 on Stokes I per beam, 3e-3 on Q, U, V).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = 32
FLAGS = (0, 3, 7)
BEAMS = (1, 5, 23)


def _case():
    from vcsbeam_amd.workloads import Observation, directions_in_cap
    obs = [Observation(ntiles=NT, blocks_per_second=2, nflagged=f, rec_chan=133 + g,
                       seed_offset=1000 * g, geometry_seed_offset=0)
           for g, f in enumerate(FLAGS)]
    dirs = [directions_in_cap(n, 15.0, seed=50 + g) for g, n in enumerate(BEAMS)]
    return obs, dirs


def _oracle(orc, obs, az, za):
    from vcsbeam_amd.workloads import unit_vectors
    cur = obs.second(0)
    htr = np.ascontiguousarray(np.concatenate([np.zeros_like(cur[:1]), cur], axis=0))
    v = orc.pfb_forward(htr, obs.ninput, obs.nspectra, obs.filter.astype(np.int32), obs.out_idx(),
                        fpga=True)
    B = np.zeros((len(az), obs.ntiles, 2, 2), complex)
    B[..., 0, 0] = B[..., 1, 1] = 1
    E, N, H, L = obs.positions()
    phi = orc.calc_phi(E, N, H, L, unit_vectors(az, za), obs.freqs)
    polQ, polP = obs.pol_idx()
    S, _ = orc.beamform(v, orc.calc_J(obs.calibration(), B), phi, polQ, polP,
                        1.0 / (2 * int((~obs.flag_mask).sum())), nstokes=4)
    return S


def _errors(orc, S, Sref):
    return np.array([[orc.rel_err(S[p, :, st, :], Sref[p, :, st, :]) for st in range(4)]
                     for p in range(S.shape[0])])


@pytest.fixture(scope="module")
def batch_case(pkg):
    from vcsbeam_amd.apps import BatchSession
    obs, dirs = _case()
    print(f"batch: {len(obs)} channels, beams {BEAMS}, flagged {FLAGS}", flush=True)
    bs = BatchSession(obs, dirs, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA, form=1)
    bs.read(0)
    bs.run(use_graph=False)
    S = [ses.stokes() for ses in bs.sessions]
    yield bs, obs, dirs, S
    bs.close()


def test_batch_matches_single_channel_and_oracle(pkg, orc, gpu_lib, batch_case):
    bs, obs, dirs, S = batch_case
    for g, (o, (az, za)) in enumerate(zip(obs, dirs)):
        assert bs.active_tiles(g) == NT - FLAGS[g]
        Sref = _oracle(orc, o, az, za)
        err = _errors(orc, S[g], Sref)
        print(f"channel {g}: max per-beam rel err I,Q,U,V vs FP64:", err.max(axis=0), flush=True)
        assert err[:, 0].max() <= 1e-3
        assert err[:, 1:].max() <= 3e-3
    # the single-channel MFMA path on the same contexts (their PFBs write the
    # batch's shared output buffer in turn)
    for g, ses in enumerate(bs.sessions):
        ses.pfb()
        ses.beamform()
        one = ses.stokes()
        err = _errors(orc, S[g].astype(np.float64), one.astype(np.float64))
        print(f"channel {g}: batch vs single-channel path:", err.max(axis=0), flush=True)
        assert err.max() <= 1e-5


def test_batch_graph_replays_match_eager(pkg, orc, gpu_lib, batch_case):
    bs, obs, dirs, S = batch_case
    lib = bs.lib
    for r in range(3):
        for ses in bs.sessions:
            v = ses.v
            lib.vmMemsetDevice(v.d_S, 0, v.d_S_size_bytes)
        print(f"graph replay {r}", flush=True)
        bs.run(use_graph=True)
        for g, ses in enumerate(bs.sessions):
            np.testing.assert_array_equal(ses.stokes(), S[g])
        bs.run(use_graph=False)  # eager work between the replays
    # another schedule (variant 3: 64-row tiles, one row tile per item) is
    # captured anew and gives the same sums
    lib.vmBatchSetSchedule(bs.b, 3, 1)
    bs.run(use_graph=True)
    assert bs.info()["variant"] == 3 and bs.info()["band"] == 1
    for g, ses in enumerate(bs.sessions):
        err = _errors(orc, ses.stokes().astype(np.float64), S[g].astype(np.float64))
        assert err.max() <= 1e-6


def test_batch_tune_with_device_generated_second(pkg, orc, gpu_lib, batch_case):
    """Seconds generated in place on the device (vmSynthMWAXDevice), each
    context with its own PFB output buffer; autotuned schedule."""
    from vcsbeam_amd.apps import BatchSession
    _, obs, dirs, S = batch_case
    bs = BatchSession(obs, dirs, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA,
                      form=1, share_pfb_output=False, device_data=True)
    try:
        bs.read(0)
        ms = bs.tune(reps=1)
        info = bs.info()
        print(f"tuned: {info}, {ms:.3f} ms", flush=True)
        assert ms > 0 and info["items"] > 0
        bs.run(use_graph=True)
        for g, ses in enumerate(bs.sessions):
            got = ses.stokes_slice(0, len(ses.bg), 0, ses.v.fine_sample_rate)
            err = _errors(orc, got.astype(np.float64), S[g].astype(np.float64))
            print(f"channel {g}: device-generated vs host-read second:", err.max(axis=0))
            assert err.max() <= 1e-6
        t = bs.stage_times()
        assert t["pfb_ms"] > 0 and t["prep_ms"] > 0 and t["gemm_ms"] > 0
    finally:
        bs.close()


def test_device_generator_matches_host_generator(pkg, gpu_lib):
    """vmSynthMWAXDevice reproduces workloads.mwax_voltages (same counter-based
    stream); a transcendental differing in its last bit could move a sample
    sitting exactly on a rounding tie by one, so at most a handful may differ
    and only by 1."""
    from vcsbeam_amd.workloads import SPB, mwax_voltages
    lib = gpu_lib
    nb, ni = 3, 6
    host = mwax_voltages(ni, nb, first_block=5, seed=11, zero_blocks=1)
    d = lib.vmDeviceMalloc(host.nbytes)
    try:
        lib.vmSynthMWAXDevice(11, 5, nb, ni, SPB, 16.0, 1, d)
        dev = np.empty_like(host)
        lib.vmMemcpyDtoH(dev.ctypes.data, d, dev.nbytes)
    finally:
        lib.vmDeviceFree(d)
    diff = np.abs(dev.astype(np.int16) - host.astype(np.int16))
    assert diff.max() <= 1 and int((diff != 0).sum()) <= 4
    assert not dev[0].any()


@pytest.mark.parametrize("form", [1, 3])
def test_batch_fp16_pfb_output(pkg, orc, gpu_lib, form):
    """The batch with the contexts' forward PFBs writing FP16 (VM_HALF) into
    the shared buffer (half its bytes): each channel within 1e-5 of the
    single-channel MFMA path reading the same FP16 voltages, and within the
    MFMA bars of the FP64 oracle fed the exact PFB output."""
    from vcsbeam_amd.apps import BatchSession
    obs, dirs = _case()
    flags = (pkg.PFB_FULL_PRECISION & ~pkg.PFB_TYPE_MASK) | pkg.PFB_COMPLEX_FLOAT16 | \
        pkg.PFB_EMULATE_FPGA
    print(f"batch with FP16 PFB output, MFMA form {form}", flush=True)
    bs = BatchSession(obs, dirs, pfb_flags=flags, form=form)
    try:
        bs.read(0)
        bs.run(use_graph=False)
        S = [ses.stokes() for ses in bs.sessions]
        for g, (o, (az, za)) in enumerate(zip(obs, dirs)):
            err = _errors(orc, S[g], _oracle(orc, o, az, za))
            print(f"channel {g}: max per-beam rel err I,Q,U,V vs FP64:", err.max(axis=0),
                  flush=True)
            assert err[:, 0].max() <= 1e-3
            assert err[:, 1:].max() <= 3e-3
        for g, ses in enumerate(bs.sessions):
            assert ses.v.datatype == pkg.VM_HALF
            ses.pfb()
            ses.beamform()
            err = _errors(orc, S[g].astype(np.float64), ses.stokes().astype(np.float64))
            print(f"channel {g}: batch vs single-channel path:", err.max(axis=0), flush=True)
            assert err.max() <= 1e-5
    finally:
        bs.close()
