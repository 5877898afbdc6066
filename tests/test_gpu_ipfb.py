"""GPU parity of the inverse PFB and the VDIF requantisation.

Two consecutive seconds (file_no 0 and 1, exercising the zeroed and the
carried-over 12-sample history) of one tied-array beam are inverted through
the reference's host interface (vmPullE -> prepare_data_buffer_fine ->
cu_invert_pfb) and through the device-resident path.  The oracle inverts the
same beamformed voltages in FP64 with the reference's direct form; outputs the
reference leaves undefined (n < P*K with n % K != 0) are masked.  Bars:
<= 1e-5 norm-wise relative error; VDIF bytes within +-1 LSB of the oracle's.

VDIF framing (src/beam_vdif.c:52-88): the device writes whole frames, header
words and requantised payload, bit for bit equal to the oracle's frame loop
across second boundaries, and vdif_write_second's file holds those frames.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-5
K = 128


def test_ipfb_host_and_device_paths(pkg, orc, gpu_lib):
    from vcsbeam_amd.apps import tied_array_beam
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=32, blocks_per_second=2, nseconds=2)
    sf = obs.filter / obs.filter.max()
    az, za = np.array([0.9]), np.array([0.35])
    print("IPFB host path, 2 seconds", flush=True)
    host = tied_array_beam(obs, az, za, nseconds=2, ipfb=True, synth_filter=sf)
    print("IPFB device path, 2 seconds", flush=True)
    dev = tied_array_beam(obs, az, za, nseconds=2, ipfb=True, synth_filter=sf, ipfb_device=True)
    ns = obs.nspectra
    h = sf * (15.0 / 7.2)
    ring = np.zeros((2 * ns, K, 2), complex)
    for t in range(2):
        e = host["e"][t][0]  # [s][c][pol]
        ring[(t % 2) * ns:(t % 2 + 1) * ns] = e
        inbuf = orc.ipfb_gather(ring, t, ns, len(h) // K)
        ref, mask = orc.ipfb(inbuf, ns, h)
        got = host["vdif"][t].reshape(ns * K, 2, 2)
        got = got[..., 0] + 1j * got[..., 1]
        err = orc.rel_err(got[mask], ref[mask])
        print(f"second {t}: rel err {err:.2e}")
        assert err <= TOL
        np.testing.assert_array_equal(dev["vdif"][t], host["vdif"][t])


def test_vdif_requant_device(pkg, orc, gpu_lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 60, 100003).astype(np.float32),
                        np.arange(-140, 140, 0.5, dtype=np.float32)])
    d_in = pkg.DeviceArray.from_numpy(x)
    d_out = pkg.DeviceArray(x.size)
    gpu_lib.vmRequantVDIFDevice(d_in.ptr, d_out.ptr, x.size, None)
    gpu_lib.vmDeviceSynchronize()
    got = d_out.to_numpy(np.uint8, x.shape)
    np.testing.assert_array_equal(got, orc.vdif_requant(x))
    # +-1 LSB against requantisation of a slightly perturbed (oracle) stream
    y = x + rng.uniform(-1e-4, 1e-4, x.size).astype(np.float32)
    d = np.abs(got.astype(int) - orc.vdif_requant(y).astype(int))
    assert d.max() <= 1


def _hdr(orc, frame):
    h = orc.vdif_header(payload_bytes=512, mjd=59000, sec_of_day=3723)
    h[1] |= frame
    return h


def test_vdif_frames_device(pkg, orc, gpu_lib):
    """10 005 frames of 512 bytes at 10 000 frames/s from frame 9997: the
    frame counter wraps twice; a second call continues from the advanced
    header."""
    rng = np.random.default_rng(5)
    rate, pb, n1, n2 = 10000, 512, 10005, 3
    x = (rng.standard_normal((n1 + n2) * pb) * 50).astype(np.float32)
    x[:64] = np.arange(-140, 140, 280 / 64, dtype=np.float32)
    hdr = _hdr(orc, rate - 3)
    want1, h1 = orc.vdif_frames(x[:n1 * pb], n1, pb, hdr, rate)
    want2, h2 = orc.vdif_frames(x[n1 * pb:], n2, pb, h1, rate)
    d_in = pkg.DeviceArray.from_numpy(x)
    d_out = pkg.DeviceArray((n1 + n2) * (32 + pb))
    vh = (C.c_uint32 * 8)(*[int(w) for w in hdr])
    print(f"VDIF frames: {n1} + {n2} frames from frame {rate - 3}", flush=True)
    gpu_lib.vmVDIFFramesDevice(d_in.ptr, n1, pb, C.byref(vh), rate, d_out.ptr, None)
    np.testing.assert_array_equal(np.array(vh[:], np.uint32), h1)
    gpu_lib.vmVDIFFramesDevice(d_in.ptr + n1 * pb * 4, n2, pb, C.byref(vh), rate,
                               d_out.ptr + n1 * (32 + pb), None)
    np.testing.assert_array_equal(np.array(vh[:], np.uint32), h2)
    gpu_lib.vmDeviceSynchronize()
    got = d_out.to_numpy(np.uint8, (n1 + n2) * (32 + pb))
    np.testing.assert_array_equal(got[:n1 * (32 + pb)], want1)
    np.testing.assert_array_equal(got[n1 * (32 + pb):], want2)


def test_vdif_write_second_file(pkg, orc, gpu_lib, tmp_path):
    """Two seconds through vdif_write_second: the .vdif file holds the
    oracle's frames, the header advances by one second per call."""
    rng = np.random.default_rng(6)
    rate, pb = 10000, 512
    f = pkg.Vdifinfo()
    f.frame_length, f.dataarraylength, f.frame_rate = 32 + pb, pb, rate
    f.block_size, f.sizeof_beam, f.sizeof_buffer = (32 + pb) * rate, pb, pb * rate
    f.bits, f.iscomplex, f.npol, f.sample_rate = 8, 1, 2, 1280000
    f.basefilename = str(tmp_path / "beam").encode()
    hdr = _hdr(orc, 0)
    vh = (C.c_uint32 * 8)(*[int(w) for w in hdr])
    want = []
    for t in range(2):
        x = (rng.standard_normal(pb * rate) * 40).astype(np.float32)
        w, hdr = orc.vdif_frames(x, rate, pb, hdr, rate)
        want.append(w)
        print(f"vdif_write_second: second {t}", flush=True)
        gpu_lib.vdif_write_second(C.byref(f), C.byref(vh), x.ctypes.data)
        np.testing.assert_array_equal(np.array(vh[:], np.uint32), hdr)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "beam.vdif", np.uint8),
                                  np.concatenate(want))
    assert (tmp_path / "beam.hdr").stat().st_size == 4096


@pytest.mark.parametrize("npass", [0, 3])
def test_ipfb_device_vdif_fused(pkg, orc, gpu_lib, npass):
    """vmInvertPFBDeviceVDIF over three seconds in order: its floats are
    vmInvertPFBDevice's (same kernel, the history carried across seconds in
    two buffers used in turn) and within the tolerance of the direct-form
    oracle on the ring of the seconds' e, and its bytes are the oracle's VDIF
    requantisation of those floats, bit for bit.  npass 0: the passes per
    workgroup the library chooses (one, for one channel of 1000 samples);
    3: runs of three passes, so most rows come after a carried history."""
    import ctypes as C
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import Observation
    lib = gpu_lib
    print(f"passes per workgroup: {npass}", flush=True)
    lib.vmSetInversePFBPasses(npass)
    obs = Observation(ntiles=32, blocks_per_second=2, nseconds=3)
    az, za = np.array([0.9]), np.array([0.35])
    ses = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=1,
                           ipfb=True, want_e=True)
    ns = obs.nspectra
    sf = obs.filter / obs.filter.max()
    h = sf * (15.0 / 7.2)
    ring = np.zeros((2 * ns, K, 2), complex)
    try:
        for t in range(3):
            print(f"second {t}: PFB, beam, fused inverse PFB + VDIF", flush=True)
            ses.read(t)
            ses.pfb()
            ses.geometry()
            ses.beamform()
            ses.invert(t)
            n = ns * K * 2 * 2
            got = np.empty(n, np.float32)
            lib.vmMemcpyDtoH(got.ctypes.data, ses.gi.d_out, got.nbytes)
            q = np.empty(n, np.uint8)
            lib.vmMemcpyDtoH(q.ctypes.data, ses.d_vdif, q.nbytes)
            np.testing.assert_array_equal(q, orc.vdif_requant(got))
            lib.vmPullE(ses.vm)
            v = ses.vm.contents
            e = np.ctypeslib.as_array(C.cast(v.e, C.POINTER(C.c_double)), (1, ns, K, 2, 2))[0]
            ring[(t % 2) * ns:(t % 2 + 1) * ns] = e[..., 0] + 1j * e[..., 1]
            ref, mask = orc.ipfb(orc.ipfb_gather(ring, t, ns, len(h) // K), ns, h)
            gz = got.reshape(ns * K, 2, 2)
            err = orc.rel_err((gz[..., 0] + 1j * gz[..., 1])[mask], ref[mask])
            print(f"   rel err vs direct form {err:.2e}", flush=True)
            assert err <= TOL
    finally:
        ses.close()
        lib.vmSetInversePFBPasses(0)


def test_ipfb_device_vdif_batch_equals_per_channel(pkg, gpu_lib):
    """vmInvertPFBDeviceVDIFBatch over three coarse channels of two seconds in
    order: every channel's floats, VDIF bytes and carried history equal, bit
    for bit, those of the same channel's own vmInvertPFBDeviceVDIF calls in
    a separate session."""
    from vcsbeam_amd.apps import TiedArraySession, invert_batch
    from vcsbeam_amd.workloads import Observation
    lib = gpu_lib
    az, za = np.array([0.4]), np.array([0.3])
    obs = [Observation(ntiles=16, blocks_per_second=2, nseconds=2, rec_chan=120 + g,
                       seed_offset=100 * g) for g in range(3)]
    one = [TiedArraySession(o, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=1, ipfb=True)
           for o in obs]
    many = [TiedArraySession(o, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=1, ipfb=True)
            for o in obs]
    ns = obs[0].nspectra
    n = ns * K * 2 * 2
    try:
        for t in range(2):
            print(f"second {t}: three channels, per-channel and batched inverse PFB", flush=True)
            for x in one + many:
                x.read(t)
                x.pfb()
                x.geometry()
                x.beamform()
            for x in one:
                x.invert(t)
            invert_batch(many, t)
            for g in range(3):
                outs = []
                for x in (one[g], many[g]):
                    f = np.empty(n, np.float32)
                    lib.vmMemcpyDtoH(f.ctypes.data, x.gi.d_out, f.nbytes)
                    q = np.empty(n, np.uint8)
                    lib.vmMemcpyDtoH(q.ctypes.data, x.d_vdif, q.nbytes)
                    outs.append((f, q))
                np.testing.assert_array_equal(outs[0][0], outs[1][0])
                np.testing.assert_array_equal(outs[0][1], outs[1][1])
                assert np.abs(outs[0][0]).max() > 0
            # the channels differ, so a batch row reading another's job would show
            f0 = np.empty(n, np.float32)
            f1 = np.empty(n, np.float32)
            lib.vmMemcpyDtoH(f0.ctypes.data, many[0].gi.d_out, f0.nbytes)
            lib.vmMemcpyDtoH(f1.ctypes.data, many[1].gi.d_out, f1.nbytes)
            assert not np.array_equal(f0, f1)
    finally:
        for x in one + many:
            x.close()


def test_ipfb_run_length_does_not_change_bits(pkg, gpu_lib):
    """The batched inverse PFB of three coarse channels over two seconds with
    1, 2, 5 and 16 passes per workgroup (runs of 21, 53, 149 and 501 output
    samples against 1000 per second, the last run of each channel partial):
    floats and VDIF bytes equal bit for bit.  A row carried wrongly from one
    pass to the next would change the outputs inside runs but not at their
    starts, so the run lengths would disagree."""
    from vcsbeam_amd.apps import TiedArraySession, invert_batch
    from vcsbeam_amd.workloads import Observation
    lib = gpu_lib
    az, za = np.array([0.6]), np.array([0.25])
    obs = [Observation(ntiles=16, blocks_per_second=2, nseconds=2, rec_chan=110 + g,
                       seed_offset=50 * g) for g in range(3)]
    sess = [TiedArraySession(o, az, za, pfb_flags=pkg.PFB_FULL_PRECISION, nstokes=1, ipfb=True)
            for o in obs]
    ns = obs[0].nspectra
    n = ns * K * 2 * 2
    try:
        for t in range(2):
            for x in sess:
                x.read(t)
                x.pfb()
                x.geometry()
                x.beamform()
            first = None
            for npass in (1, 2, 5, 16):
                print(f"second {t}: {npass} passes per workgroup", flush=True)
                lib.vmSetInversePFBPasses(npass)
                # second t reads the history buffer second t - 1 wrote and
                # writes the other, so repeating it gives the same inputs
                invert_batch(sess, t)
                outs = []
                for x in sess:
                    f = np.empty(n, np.float32)
                    lib.vmMemcpyDtoH(f.ctypes.data, x.gi.d_out, f.nbytes)
                    q = np.empty(n, np.uint8)
                    lib.vmMemcpyDtoH(q.ctypes.data, x.d_vdif, q.nbytes)
                    outs.append((f, q))
                if first is None:
                    first = outs
                    assert all(np.abs(f).max() > 0 for f, _ in outs)
                    continue
                for (f0, q0), (f1, q1) in zip(first, outs):
                    np.testing.assert_array_equal(f0, f1)
                    np.testing.assert_array_equal(q0, q1)
    finally:
        lib.vmSetInversePFBPasses(0)
        for x in sess:
            x.close()
