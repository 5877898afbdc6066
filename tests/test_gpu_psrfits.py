"""GPU parity of the one-process coarse-channel splice: three coarse
channels (rec 133-135, 8 tiles x 2 pols, 0.1 s, the same 3 beams, IQUV)
renormalised on the device straight into one spliced buffer must equal,
byte for byte, each channel's own vmSendSToFits output placed at its
channel offset (the layout the reference's MPI_Igather builds,
src/beam_psrfits.c:608-642); then one second of the spliced band goes
through the PSRFITS writer."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_splice_equals_per_channel_renormalise(pkg, gpu_lib, tmp_path):
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import Observation, directions_in_cap
    lib = gpu_lib
    obs = [Observation(ntiles=8, blocks_per_second=2, rec_chan=133 + g, seed_offset=100 * g,
                       geometry_seed_offset=0) for g in range(3)]
    az, za = directions_in_cap(3, 15.0, seed=9)
    sessions = [TiedArraySession(o, az, za, nstokes=4) for o in obs]
    try:
        per = []
        for g, ses in enumerate(sessions):
            print(f"channel {g}: read, PFB, beamform", flush=True)
            ses.read(0)
            ses.pfb()
            ses.geometry()
            ses.beamform()
            lib.vmCreateStatistics(ses.vm, None)
            lib.vmSendSToFits(ses.vm, None)
            v = ses.v
            ns, nf, np_ = v.fine_sample_rate, v.nfine_chan, v.npointing
            cs = np.ctypeslib.as_array(C.cast(v.Cscaled, C.POINTER(C.c_uint8)),
                                       (np_, ns, 4, nf)).copy()
            off = np.ctypeslib.as_array(C.cast(v.offsets, C.POINTER(C.c_float)), (np_, 4, nf)).copy()
            scl = np.ctypeslib.as_array(C.cast(v.scales, C.POINTER(C.c_float)), (np_, 4, nf)).copy()
            per.append((cs, off, scl))
        ctot = 3 * nf
        pfs = (pkg.Psrfits * np_)()
        bufs = []
        for p in range(np_):
            d = np.zeros((ns, 4, ctot), np.uint8)
            o = np.zeros((4, ctot), np.float32)
            s = np.zeros((4, ctot), np.float32)
            bufs.append((d, o, s))
            pfs[p].sub.data, pfs[p].sub.dat_offsets, pfs[p].sub.dat_scales = (
                d.ctypes.data, o.ctypes.data, s.ctypes.data)
            pfs[p].sub.bytes_per_subint = d.nbytes
        vms = (C.c_void_p * 3)(*[C.cast(ses.vm, C.c_void_p).value for ses in sessions])
        print("splice 3 channels", flush=True)
        lib.vmSpliceSToFits(vms, 3, pfs)
        for p in range(np_):
            d, o, s = bufs[p]
            for g in range(3):
                cs, off, scl = per[g]
                np.testing.assert_array_equal(d[:, :, g * nf:(g + 1) * nf], cs[p])
                np.testing.assert_array_equal(o[:, g * nf:(g + 1) * nf], off[p])
                np.testing.assert_array_equal(s[:, g * nf:(g + 1) * nf], scl[p])
        # the spliced band of pointing 0 through the writer
        pf = pkg.Psrfits()
        v0 = sessions[0].v
        md = obs[0].md
        v0.ncoarse_chans = 1
        lib.populate_psrfits_header(sessions[0].vm, C.byref(pf), 1, 4, C.byref(sessions[0].bg[0]),
                                    None, True)
        pf.basefilename = str(tmp_path / "b").encode()
        C.memmove(pf.sub.data, bufs[0][0].ctypes.data, ns * 4 * nf)
        assert lib.vmWritePsrfitsSubint(C.byref(pf)) == 0
        lib.vmClosePsrfits(C.byref(pf))
        assert (tmp_path / "b_0001.fits").stat().st_size % 2880 == 0
        lib.free_psrfits(C.byref(pf))
    finally:
        for ses in sessions:
            if ses.v.d_offsets:
                lib.vmDestroyStatistics(ses.vm)
            ses.close()
