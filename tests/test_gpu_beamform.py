"""GPU parity of the tied-array (plain FP32) and incoherent beamformers.

Workload: 128 tiles x 2 pols, 4 tiles flagged (D = 0 and flagged), one 0.1 s
second, per-tile/per-channel Jones D (lognormal gains, leakage 0.05), tiles in
a 3 km disc, pointing at za = 20 deg.  The GPU result (its own PFB, FP32
beamformer) is compared with the FP64 oracle chain (oracle PFB, oracle J^-1
and phases, oracle beamformer): <= 1e-5 norm-wise relative error per Stokes
parameter and for the beamformed voltages e.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _obs(ntiles=128, nseconds=1, legacy=False):
    from vcsbeam_amd.workloads import Observation
    return Observation(ntiles=ntiles, blocks_per_second=2, nseconds=nseconds, legacy=legacy)


def _oracle_tab(orc, obs, v, az, za, int4=False, nstokes=4):
    D = obs.calibration()
    B = np.zeros((len(az), obs.ntiles, 2, 2), complex)
    B[..., 0, 0] = B[..., 1, 1] = 1
    J = orc.calc_J(D, B)
    from vcsbeam_amd.workloads import unit_vectors
    E, N, H, L = obs.positions()
    phi = orc.calc_phi(E, N, H, L, unit_vectors(az, za), obs.freqs)
    polQ, polP = obs.pol_idx()
    nnf = int((~obs.flag_mask).sum()) * 2
    return orc.beamform(v, J, phi, polQ, polP, 1.0 / nnf, nstokes=nstokes, int4=int4)


def _htr(obs, t):
    cur = obs.second(t)
    margin = obs.second(t - 1)[-1:] if t > 0 else np.zeros_like(cur[:1])
    return np.ascontiguousarray(np.concatenate([margin, cur], axis=0))


def _check(orc, S, e, Sref, eref, tol=TOL):
    for st in range(S.shape[2]):
        err = orc.rel_err(S[:, :, st, :], Sref[:, :, st, :])
        print(f"stokes {st}: rel err {err:.2e}")
        assert err <= tol
    if e is not None:
        err = orc.rel_err(e, eref)
        print(f"e: rel err {err:.2e}")
        assert err <= tol


@pytest.mark.parametrize("device_geometry", [False, True])
def test_tied_array_single_beam(pkg, orc, gpu_lib, device_geometry):
    from vcsbeam_amd.apps import tied_array_beam
    obs = _obs()
    az, za = np.array([1.2]), np.array([np.radians(20.0)])
    print("tied-array 1 beam, device geometry", device_geometry, flush=True)
    out = tied_array_beam(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA,
                          device_geometry=device_geometry)
    v = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                        obs.out_idx(), fpga=True)
    Sref, eref = _oracle_tab(orc, obs, v, az, za)
    _check(orc, out["S"][0], out["e"][0], Sref, eref)


def test_tied_array_many_beams_chunked(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import tied_array_beam
    from vcsbeam_amd.workloads import directions_in_cap
    obs = _obs()
    az, za = directions_in_cap(6, 15.0, centre_za_deg=20.0)
    print("tied-array 6 beams, 2 chunks", flush=True)
    out = tied_array_beam(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nchunks=2)
    v = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                        obs.out_idx(), fpga=False)
    Sref, eref = _oracle_tab(orc, obs, v, az, za)
    _check(orc, out["S"][0], out["e"][0], Sref, eref)


def test_tied_array_legacy_int4_input(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import tied_array_beam
    obs = _obs(ntiles=32, legacy=True)
    az, za = np.array([0.4, 2.0]), np.radians(np.array([20.0, 5.0]))
    print("tied-array legacy 4+4-bit input", flush=True)
    out = tied_array_beam(obs, az, za, nstokes=4)
    ns = obs.vmd.num_samples_per_voltage_block
    v = obs.legacy_second(0).reshape(ns, 128, obs.ninput)
    Sref, eref = _oracle_tab(orc, obs, v, az, za, int4=True)
    _check(orc, out["S"][0], out["e"][0], Sref, eref)


def test_stokes_I_only(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import tied_array_beam
    obs = _obs(ntiles=16)
    az, za = np.array([3.0]), np.array([0.3])
    out = tied_array_beam(obs, az, za, nstokes=1, want_e=False)
    v = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                        obs.out_idx(), fpga=False)
    Sref, _ = _oracle_tab(orc, obs, v, az, za, nstokes=1)
    _check(orc, out["S"][0], None, Sref, None)


@pytest.mark.parametrize("ns,nc,ni", [(1000, 128, 256), (37, 100, 24), (1, 3, 16), (513, 130, 48),
                                      (64, 128, 512)])
def test_incoherent_beam_exact_and_renormalised(pkg, orc, gpu_lib, ns, nc, ni):
    """The sum over inputs is exact and the renormalisation (extremes per
    channel, scale, offset, 8-bit values) equals the oracle's; ragged sizes
    (inputs not a multiple of 16, channel counts not dividing the row
    stride, one sample) take the general paths."""
    from vcsbeam_amd.apps import incoherent_beam
    from vcsbeam_amd.workloads import legacy_int4
    print(f"incoherent beam {ns} x {nc} x {ni}", flush=True)
    data = legacy_int4(ns * nc * ni, sigma=2.5)
    incoh, off, scl, Is = incoherent_beam(data, ns, nc, ni)
    ref = orc.incoh(data.reshape(ns, nc, ni), ns, nc, ni)
    np.testing.assert_array_equal(incoh, ref)
    roff, rscl, rq, valid = orc.renormalise(ref.reshape(1, ns, 1, nc), 1, ns, 1, nc)
    np.testing.assert_array_equal(off, roff.ravel())
    np.testing.assert_array_equal(scl, rscl.ravel())
    np.testing.assert_array_equal(Is.reshape(1, ns, 1, nc)[valid], rq[valid])


def test_send_S_to_fits_matches_oracle_renormalisation(pkg, orc, gpu_lib):
    import ctypes as C
    from vcsbeam_amd.apps import tied_array_beam
    obs = _obs(ntiles=16)
    az, za = np.array([1.0, 2.0]), np.array([0.2, 0.4])
    out = tied_array_beam(obs, az, za, nstokes=4, want_e=False, return_vm=True)
    lib = gpu_lib
    vm = out["vm"]
    lib.vmCreateStatistics(vm, None)
    lib.vmSendSToFits(vm, None)
    v = vm.contents
    np_, ns, nst, nc = 2, obs.nspectra, 4, 128
    q = np.ctypeslib.as_array(C.cast(v.Cscaled, C.POINTER(C.c_uint8)), shape=(np_, ns, nst, nc)).copy()
    offs = np.ctypeslib.as_array(C.cast(v.offsets, C.POINTER(C.c_float)), shape=(np_, nst, nc)).copy()
    S = out["S"][0]
    roff, rscl, rq, valid = orc.renormalise(S, np_, ns, nst, nc)
    np.testing.assert_array_equal(offs, roff)
    np.testing.assert_array_equal(q[valid], rq[valid])
    lib.vmDestroyStatistics(vm)
    lib.destroy_vcsbeam_context(vm)


def test_device_jones_shared_between_pointings(pkg, orc, gpu_lib):
    """vmCalcJDevice with the same beam for every pointing writes pointing 0's
    J^-1 only: it equals the FP64 oracle's, vmExpandJDevice gives every
    pointing the same, and the FP32 form (which expands it itself) still
    matches the oracle.  A beam that differs for one pointing gives J^-1 per
    pointing."""
    import ctypes as C
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import directions_in_cap
    obs = _obs(ntiles=16)
    az, za = directions_in_cap(6, 15.0, centre_za_deg=20.0)
    ses = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=4)
    try:
        v = ses.v
        npt, nant, nch = len(az), obs.ntiles, 128
        D = obs.calibration()

        def device_J():
            J = np.empty((npt, nant, nch, 2, 2), np.complex128)
            gpu_lib.vmMemcpyDtoH(J.ctypes.data, v.d_J, J.nbytes)
            return J

        ses.read(0)
        ses.pfb()
        ses.geometry()  # B = I for every pointing
        B = np.zeros((npt, nant, 2, 2), complex)
        B[..., 0, 0] = B[..., 1, 1] = 1
        Jref = orc.calc_J(D, B)
        J = device_J()
        assert np.abs(J[0] - Jref[0]).max() <= 1e-12 * np.abs(Jref[0]).max()
        gpu_lib.vmExpandJDevice(ses.vm)
        J = device_J()
        assert np.abs(J - Jref).max() <= 1e-12 * np.abs(Jref).max()
        # the FP32 form reads J^-1 per pointing
        ses.geometry()
        ses.beamform()
        S = ses.stokes()
        vv = orc.pfb_forward(_htr(obs, 0), obs.ninput, obs.nspectra, obs.filter.astype(np.int32),
                             obs.out_idx(), fpga=False)
        Sref, _ = _oracle_tab(orc, obs, vv, az, za)
        for p in range(npt):
            assert orc.rel_err(S[p, :, 0, :], Sref[p, :, 0, :]) <= 1e-5
        # a beam that differs for pointing 2: J^-1 per pointing
        Bh = np.ctypeslib.as_array(C.cast(v.pb.B, C.POINTER(C.c_double)),
                                   shape=(npt, nant, 4, 2))
        Bh[2, :, 0, 0] = 1.5
        B[2, :, 0, 0] = 1.5
        gpu_lib.vmCalcJDevice(ses.vm)
        Jref = orc.calc_J(D, B)
        J = device_J()
        assert np.abs(J - Jref).max() <= 1e-12 * np.abs(Jref).max()
        assert np.abs(J[2] - J[0]).max() > 1e-3 * np.abs(J[0]).max()
    finally:
        ses.close()


def test_kernel_timing_counts_every_launch(pkg, gpu_lib):
    """The kernel timers record each timed launch without waiting; 40 steps
    (more than the 32 pairs a slot holds between two reads, so the ring is
    resolved once while timing) are all counted, with positive averages, and
    the Stokes are the same as untimed."""
    import ctypes as C
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import directions_in_cap
    obs = _obs(ntiles=32)
    az, za = directions_in_cap(8, 15.0, centre_za_deg=20.0)
    ses = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=4,
                           form=pkg.VM_BF_MFMA_FP16)
    try:
        ses.read(0)
        ses.pfb()
        ses.geometry()
        ses.beamform()
        S0 = ses.stokes()
        gpu_lib.vmSetKernelTiming(ses.vm, 1)
        for _ in range(40):
            ses.geometry()
            ses.beamform()
        gpu_lib.vmSetKernelTiming(ses.vm, 0)
        n = C.c_long(0)
        for name, per_step in (("beamform_mfma", 1), ("beamform_prep", 1), ("geometry", 2)):
            ms = gpu_lib.vmKernelTimeMs(ses.vm, name.encode(), C.byref(n))
            print(f"{name}: {n.value} launches, {ms:.4f} ms each", flush=True)
            assert n.value == 40 * per_step
            assert ms > 0
        assert np.array_equal(ses.stokes(), S0)
    finally:
        ses.close()
