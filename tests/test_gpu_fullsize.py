"""GPU parity at the workloads' full sizes (128 tiles x 2 pols, one whole
second), through checks that do not need the oracle to run at that size.

* Forward PFB over a full second (20 MWAX blocks, 10 000 spectra): the fused
  kernel and the staged four-pass path (WOLA, rounding, FFT, pack) give
  identical bits.  The staged path is pinned to the oracle bit for bit at
  0.1 s (test_gpu_pfb.py); its per-element arithmetic does not depend on the
  size.
* Precision form at 512 beams (the benchmark workload): every Stokes value
  of the second is finite, whole-second rows of 8 beams (the first and last
  four, all 10 000 samples x 128 channels) agree with the FP32 form to
  <= 1e-3 on Stokes I, and 4 beams x 64 samples agree with the FP64 oracle to
  <= 1e-3 on Stokes I.
* Batched workload (24 coarse channels x 1 s of 256 tiles, flagged tiles
  U{0..40} and beams log-uniform 1..2048 per channel, seconds generated on
  the device, one replayed hipGraph): per channel, its first two and last two
  beams x 32 samples x 128 channels agree with the FP64 oracle fed that
  channel's own PFB output (1e-3 on Stokes I, 3e-3 on Q, U, V).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full_obs(pkg):
    from vcsbeam_amd.workloads import Observation
    return Observation(ntiles=128, blocks_per_second=20)


def test_full_second_pfb_fused_equals_staged(pkg, gpu_lib, full_obs):
    from vcsbeam_amd.apps import fine_pfb_offline
    flags = pkg.PFB_COMPLEX_FLOAT32 | pkg.PFB_EMULATE_FPGA | pkg.PFB_MALLOC_HOST_OUTPUT | 0x0C
    print("full-second PFB, 128 tiles: fused", flush=True)
    fused = fine_pfb_offline(full_obs, pfb_flags=flags)[0]
    print("full-second PFB, 128 tiles: staged", flush=True)
    staged = fine_pfb_offline(full_obs, pfb_flags=flags | pkg.PFB_STAGED)[0]
    assert fused.shape == staged.shape and fused.size == 128 * 256 * 10000
    np.testing.assert_array_equal(fused.view(np.uint32), staged.view(np.uint32))


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.mark.parametrize("form", [1, 3])
def test_full_second_512_beams_mfma(pkg, orc, gpu_lib, full_obs, form):
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import directions_in_cap, unit_vectors
    obs = full_obs
    az, za = directions_in_cap(512, 15.0)
    sess = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA,
                            nstokes=4, form=form)
    try:
        sess.read(0)
        sess.pfb()
        print(f"512 beams x 1 s, MFMA form {form}", flush=True)
        sess.geometry()
        sess.beamform()
        beams = [(0, 4), (508, 512)]
        m = np.concatenate([sess.stokes_slice(p0, p1, 0, 10000) for p0, p1 in beams])
        # every value of the second is finite: a device reduction would need
        # a kernel of its own, so the full S comes to the host once
        S = sess.stokes()
        assert np.isfinite(S).all()
        del S
        print("512 beams x 1 s, FP32", flush=True)
        sess.v.beamform_form = pkg.VM_BF_FP32
        sess.geometry()
        sess.beamform()
        f = np.concatenate([sess.stokes_slice(p0, p1, 0, 10000) for p0, p1 in beams])
        for b in range(m.shape[0]):
            e = _rel(m[b, :, 0, :].astype(np.float64), f[b, :, 0, :].astype(np.float64))
            print(f"beam {b}: MFMA vs FP32 Stokes I {e:.2e}", flush=True)
            assert e <= 1e-3
        # FP64 oracle on 4 beams x 64 samples of the same device voltages
        vfl = sess.voltages()
        n = 64
        v = np.ascontiguousarray(vfl[:, :, :n].transpose(2, 0, 1)).astype(np.complex128)
        del vfl
        D = obs.calibration()
        B = np.zeros((4, obs.ntiles, 2, 2), complex)
        B[..., 0, 0] = B[..., 1, 1] = 1
        E, N, H, L = obs.positions()
        phi = orc.calc_phi(E, N, H, L, unit_vectors(az[:4], za[:4]), obs.freqs)
        polQ, polP = obs.pol_idx()
        ref, _ = orc.beamform(v, orc.calc_J(D, B), phi, polQ, polP,
                              1.0 / (2 * int((~obs.flag_mask).sum())), nstokes=4)
        for b in range(4):
            e = _rel(m[b, :n, 0, :].astype(np.float64), ref[b, :, 0, :])
            print(f"beam {b}: MFMA vs FP64 Stokes I {e:.2e}", flush=True)
            assert e <= 1e-3
    finally:
        sess.close()


@pytest.mark.parametrize("form,half", [(1, False), (3, True)])
def test_full_batched_workload(pkg, orc, gpu_lib, form, half):
    from vcsbeam_amd.apps import BatchSession
    from vcsbeam_amd.workloads import (SEED_BEAMS, SEED_FLAGS, Observation, directions_in_cap,
                                       uniform, unit_vectors)
    lib = gpu_lib
    nchan, ntiles, ns = 24, 256, 32
    u = uniform(SEED_FLAGS, 2 * nchan, offset=10 ** 8)
    nflag = np.minimum((u[0::2] * 41).astype(int), 40)
    nbeam = np.floor(np.exp(u[1::2] * np.log(2049.0))).astype(int)
    obs = [Observation(ntiles=ntiles, blocks_per_second=20, nflagged=int(nflag[g]),
                       rec_chan=109 + g, seed_offset=1000 * (g + 1), geometry_seed_offset=0)
           for g in range(nchan)]
    dirs = [directions_in_cap(int(nbeam[g]), 15.0, seed=SEED_BEAMS + 100 * (g + 1))
            for g in range(nchan)]
    print(f"batch: {nchan} channels x 1 s, {ntiles} tiles, {int(nbeam.sum())} beams, MFMA form "
          f"{form}, {'FP16' if half else 'FP32'} PFB output", flush=True)
    flags = pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA
    if half:
        flags = (flags & ~pkg.PFB_TYPE_MASK) | pkg.PFB_COMPLEX_FLOAT16
    bs = BatchSession(obs, dirs, pfb_flags=flags, nstokes=4, form=form, share_pfb_output=True,
                      device_data=True)
    try:
        bs.read(0)
        bs.run(use_graph=True)
        lib.vmDeviceSynchronize()
        nspc = obs[0].nspectra
        for g, ses in enumerate(bs.sessions):
            o, (az, za) = obs[g], dirs[g]
            ses.pfb()  # the shared PFB output buffer holds the last channel's
            if half:
                # VM_HALF [c][s/32][i][s%32][re, im]: block 0 of each channel is
                # one row of ninput x 32 complex halves
                assert ns == 32
                nsb = (nspc + 31) // 32
                h = np.empty((128, o.ninput, 32, 2), np.float16)
                lib.vmMemcpy2DDtoH(h.ctypes.data, o.ninput * 128, ses.v.d_v,
                                   nsb * o.ninput * 128, o.ninput * 128, 128)
                v = h.astype(np.float32).view(np.complex64)[..., 0]
            else:
                v = np.empty((128, o.ninput, ns), np.complex64)
                lib.vmMemcpy2DDtoH(v.ctypes.data, ns * 8, ses.v.d_v, nspc * 8, ns * 8,
                                   128 * o.ninput)
            npt = len(az)
            sel = sorted(set([0, 1, npt - 2, npt - 1]) & set(range(npt)))
            B = np.zeros((len(sel), o.ntiles, 2, 2), complex)
            B[..., 0, 0] = B[..., 1, 1] = 1
            E, N, H, L = o.positions()
            phi = orc.calc_phi(E, N, H, L, unit_vectors(az[sel], za[sel]), o.freqs)
            polQ, polP = o.pol_idx()
            ref, _ = orc.beamform(np.ascontiguousarray(v.transpose(2, 0, 1)).astype(np.complex128),
                                  orc.calc_J(o.calibration(), B), phi, polQ, polP,
                                  1.0 / (2 * int((~o.flag_mask).sum())), nstokes=4)
            for i, p in enumerate(sel):
                got = ses.stokes_slice(p, p + 1, 0, ns)[0].astype(np.float64)
                err = [orc.rel_err(got[:, k, :], ref[i, :, k, :]) for k in range(4)]
                print(f"channel {g} ({npt} beams, {int(nflag[g])} flagged) beam {p}: "
                      f"IQUV {np.round(err, 6)}", flush=True)
                assert err[0] <= 1e-3 and max(err[1:]) <= 3e-3
    finally:
        bs.close()


def test_full_pipeline_two_channels_stokes_ipfb_vdif(pkg, orc, gpu_lib):
    """Workload 2's device sequence at its full per-channel size, as the
    benchmark runs it: two coarse channels of 128 tiles x 1 s (4 flagged),
    seconds generated on the device, one shared forward-PFB output buffer,
    the FP32 form towards one beam at za 20 deg (Stokes I), then the inverse
    PFB and 8-bit VDIF of both channels in one launch.  Per channel:
    - Stokes I on 64 samples x 128 fine channels within 1e-5 of the FP64
      oracle fed that channel's own PFB output;
    - the inverse PFB on 40 fine samples from mid-second (5120 outputs x 2
      pols) within 1e-5 of the direct-form oracle on those rows of e;
    - the whole second's VDIF bytes equal the oracle's requantisation of the
      kernel's floats."""
    import ctypes as C
    from vcsbeam_amd.apps import TiedArraySession, device_second, invert_batch
    from vcsbeam_amd.workloads import Observation, unit_vectors
    K, P, n_s = 128, 12, 64
    obs = [Observation(ntiles=128, blocks_per_second=20, nflagged=4, rec_chan=109 + g,
                       seed_offset=1000 * (g + 1), geometry_seed_offset=0) for g in range(2)]
    ns, ni = obs[0].nspectra, obs[0].ninput
    az, za = np.array([0.7]), np.radians([20.0])
    dv = pkg.DeviceArray(K * ni * ns * 8)
    flags = (pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA) & ~pkg.PFB_MALLOC_DEVICE_OUTPUT
    sess = []
    try:
        for o in obs:
            sess.append(TiedArraySession(o, az, za, pfb_flags=flags, nstokes=1,
                                         form=pkg.VM_BF_FP32, host_buffers=False, ipfb=True,
                                         want_e=True, d_v=dv.ptr))
        v_head, S_head = [], []
        for g, x in enumerate(sess):
            print(f"channel {g}: device second, PFB, geometry, FP32 beam", flush=True)
            device_second(x, 0)
            x.pfb()
            x.geometry()
            x.beamform()
            # the first n_s samples of every [c][i] row of the shared PFB output
            h = np.empty((K * ni, n_s), np.complex64)
            gpu_lib.vmMemcpy2DDtoH(h.ctypes.data, n_s * 8, dv.ptr, ns * 8, n_s * 8, K * ni)
            v_head.append(h.reshape(K, ni, n_s))
            S_head.append(x.stokes_slice(0, 1, 0, n_s))
        print("inverse PFB + VDIF of both channels in one launch", flush=True)
        invert_batch(sess, 0)
        sf = obs[0].filter / obs[0].filter.max()
        hs = sf * (15.0 / 7.2)
        for g, (x, o) in enumerate(zip(sess, obs)):
            D = o.calibration()
            B = np.zeros((1, o.ntiles, 2, 2), complex)
            B[..., 0, 0] = B[..., 1, 1] = 1
            E, N, H, L = o.positions()
            phi = orc.calc_phi(E, N, H, L, unit_vectors(az, za), o.freqs)
            polQ, polP = o.pol_idx()
            v = np.ascontiguousarray(v_head[g].transpose(2, 0, 1)).astype(np.complex128)
            ref, _ = orc.beamform(v, orc.calc_J(D, B), phi, polQ, polP,
                                  1.0 / (2 * int((~o.flag_mask).sum())), nstokes=1)
            e = orc.rel_err(S_head[g][0, :, 0, :], ref[0, :, 0, :])
            print(f"channel {g}: Stokes I vs FP64 {e:.2e}", flush=True)
            assert e <= 1e-5
            n = ns * K * 2 * 2
            got = np.empty(n, np.float32)
            gpu_lib.vmMemcpyDtoH(got.ctypes.data, x.gi.d_out, got.nbytes)
            q = np.empty(n, np.uint8)
            gpu_lib.vmMemcpyDtoH(q.ctypes.data, x.d_vdif, q.nbytes)
            np.testing.assert_array_equal(q, orc.vdif_requant(got))
            gpu_lib.vmPullE(x.vm)
            ev = np.ctypeslib.as_array(C.cast(x.v.e, C.POINTER(C.c_double)), (ns, K, 2, 2))
            e_c = ev[..., 0] + 1j * ev[..., 1]
            # rows r0 .. r0 + W + P of [P zero history rows | e] give outputs
            # n = r0 K + n' for n' in [P K, W K) exactly as the whole buffer does
            r0, W = 5000, 40
            sub = e_c[r0 - P:r0 + W]
            ref_i, mask = orc.ipfb(sub, W, hs)
            gz = got.reshape(ns * K, 2, 2)
            gz = (gz[..., 0] + 1j * gz[..., 1])[r0 * K:(r0 + W) * K]
            ei = orc.rel_err(gz[mask], ref_i[mask])
            print(f"channel {g}: inverse PFB vs direct form {ei:.2e} on {int(mask.sum())} "
                  f"outputs x 2 pols", flush=True)
            # valid: n' >= P K, and the P multiples of K below it
            assert mask.sum() == (W - P) * K + P and ei <= 1e-5
    finally:
        for x in sess:
            x.close()
        del dv
