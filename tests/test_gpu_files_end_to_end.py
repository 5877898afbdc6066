"""End to end from files to files (app/make_mwa_tied_array_beam.c's
sequence): a metafits (8 tiles, one flagged), one second of MWAX voltages in
a subfile, an Offringa calibration file and a J2000 pointing go through
make_tied_array_beam_files, which writes PSRFITS (IQUV) and VDIF.  The same
second, metadata, filters and D are also run through the in-memory path
(tied_array_beam), which the other GPU tests hold to the FP64 oracle:
- Stokes agree to 1e-6 (the only difference is the unit vectors, from
  calc_beam_geom's az/el at the middle of the second here and from az/za
  there);
- the PSRFITS rows dequantise to those Stokes within one quantisation step;
- the VDIF payload is the oracle's requantisation of the inverse-PFB floats,
  its headers the oracle's for that second.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from test_obsfiles import _bintable, _card, _long_string, _pad
from test_psrfits import _col, _columns, read_fits

pytestmark = pytest.mark.gpu
NT = 8
GPS = 1300000000


def _metafits(path, east, north, height, cable, flagged, mode="MWAX_VCS", delays=None):
    """delays: optional [NT][16] dipole delay steps per tile (32 = dead),
    written as the TILEDATA column Delays."""
    n = 2 * NT
    ant = np.repeat(np.arange(NT), 2)
    pol = ["X", "Y"] * NT
    inp = np.arange(n)
    i2 = lambda a: np.asarray(a, ">i2")
    f4 = lambda a: np.asarray(a, ">f4")
    prim = [_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True),
            _card("GPSTIME", GPS), _card("MJD", 59000.5), _card("PROJECT", "G0001"),
            _card("MODE", mode), _card("NCHANS", 24), _card("BANDWDTH", 30.72),
            _card("FINECHAN", 10.0), _card("NINPUTS", n), _card("RA", 80.0), _card("DEC", -26.7),
            _long_string("CHANNELS", ",".join(str(c) for c in range(109, 133))), _card("END")]
    cols = [("Input", "I", i2(inp)), ("Antenna", "I", i2(ant)), ("Tile", "I", i2(1000 + ant)),
            ("TileName", "8A", [f"Tile{a:03d}".encode().ljust(8) for a in ant]),
            ("Pol", "1A", [p.encode() for p in pol]),
            ("Flag", "I", i2([1 if a == flagged else 0 for a in ant])),
            ("Length", "14A", [f"EL_{float(cable[a]):.3f}".encode().ljust(14) for a in ant]),
            ("North", "E", f4(north[ant])), ("East", "E", f4(east[ant])),
            ("Height", "E", f4(height[ant]))]
    if delays is not None:
        cols.append(("Delays", "16I", np.ascontiguousarray(np.asarray(delays)[ant], ">i2")))
    path.write_bytes(_pad(b"".join(prim)) + _bintable("TILEDATA", cols, n))


def _offringa(path, sol):
    import struct
    hdr = b"MWAOCAL\0" + struct.pack("<IIIIII", 0, 0, 1, sol.shape[0], sol.shape[1], 4)
    hdr += struct.pack("<dd", 0.0, 1.0)
    path.write_bytes(hdr + np.ascontiguousarray(sol).view(np.float64).tobytes())


class _Obs:
    """The in-memory path's view of the same observation."""

    def __init__(self, md, vmd, data, filt, D):
        self.md, self.vmd, self.data, self.filter, self.D = md, vmd, data, filt, D
        self.legacy = False

    def second(self, t):
        return self.data

    def calibration(self, nchan=128):
        return self.D


def test_make_tied_array_beam_files(pkg, orc, gpu_lib, tmp_path):
    from vcsbeam_amd.apps import make_tied_array_beam_files, tied_array_beam
    from vcsbeam_amd.workloads import mwax_voltages
    lib = gpu_lib
    rng = np.random.default_rng(5)
    east, north = rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT)
    height = 377.0 + rng.uniform(-2, 2, NT)
    cable = np.round(rng.uniform(50, 300, NT), 3)
    flagged = 5
    _metafits(tmp_path / "obs.metafits", east, north, height, cable, flagged)
    data = mwax_voltages(2 * NT, 20, first_block=0, seed=31)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    (tmp_path / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    sol = rng.standard_normal((NT, 128, 2, 2)) + 1j * rng.standard_normal((NT, 128, 2, 2))
    _offringa(tmp_path / "cal.bin", sol)
    print("files -> PSRFITS + VDIF", flush=True)
    (tmp_path / "pointings.txt").write_text("05:34:30.00 -26:42:00.00\n")  # 5.575 h, -26.7 deg
    out = make_tied_array_beam_files(tmp_path / "obs.metafits", tmp_path, tmp_path / "out",
                                     str(tmp_path / "pointings.txt"), None,
                                     cal=("offringa", tmp_path / "cal.bin"))
    assert abs(out["bg"][0].mean_ra - 5.575 * math.pi / 12) < 1e-12
    bg = out["bg"][0]
    assert 0 < bg.el < math.pi / 2
    # the in-memory path with the same inputs
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(tmp_path / "obs.metafits").encode(), C.byref(md), C.byref(vmd))
    D = sol.copy()
    D[..., 0, 1] = D[..., 1, 0] = 0           # cross terms zeroed (keep_cross_terms off)
    # the metafits-flagged tile keeps its solution (it changes w only)
    runtime = pkg.RUNTIME_DIR
    fine = np.loadtxt(os.path.join(runtime, "FINEPFB.dat"))
    mirror = np.loadtxt(os.path.join(runtime, "MIRROR.dat"))
    print("same second in memory", flush=True)
    az0, el0 = out["bg_per_second"][0][0]      # the middle of the second
    ref = tied_array_beam(_Obs(md, vmd, data, fine, D), np.array([az0]),
                          np.array([math.pi / 2 - el0]), pfb_flags=557, ipfb=True,
                          synth_filter=mirror, ipfb_device=True, want_e=False)
    lib.vmFreeMetafits(C.byref(md))
    S, Sref = out["S"][0][0], ref["S"][0][0]            # [s][st][c]
    assert orc.rel_err(S[:, 0, :], Sref[:, 0, :]) < 1e-6
    # PSRFITS: one row, dequantised within one step of the Stokes
    (prim, _), (sub, tab) = read_fits(pathlib_of(out["psrfits"][0] + "_0001.fits"))
    assert sub["NAXIS2"] == 1 and prim["OBSNCHAN"] == 128 and prim["STT_IMJD"] == 59000
    cols = _columns(sub)
    row = tab[:sub["NAXIS1"]]
    q = _col(row, cols["DATA"]).reshape(10000, 4, 128).astype(np.float64)
    off = _col(row, cols["DAT_OFFS"]).reshape(4, 128).astype(np.float64)
    scl = _col(row, cols["DAT_SCL"]).reshape(4, 128).astype(np.float64)
    deq = q * scl[None] + off[None]
    assert np.all(np.abs(deq - S) <= scl[None] * 1.0001 + 1e-6 * np.abs(S))
    # VDIF: one second of pointing 0, frames as the oracle builds them
    vd = np.fromfile(out["vdif_base"] + ".vdif", np.uint8)
    assert vd.size == 10000 * 544
    h0 = orc.vdif_header(mjd=59000, sec_of_day=round(bg.fracmjd * 86400))
    want, _ = orc.vdif_frames(out["vdif"][0], 10000, 512, h0, 10000)
    np.testing.assert_array_equal(vd, want)
    assert os.path.getsize(out["vdif_base"] + ".hdr") == 4096


def pathlib_of(p):
    import pathlib
    return pathlib.Path(p)


def test_make_incoh_beam_files(pkg, gpu_lib, tmp_path):
    """app/make_mwa_incoh_beam.c's sequence: a legacy recombined second in a
    .dat file through make_incoh_beam_files; the PSRFITS row holds exactly
    what cu_form_incoh_beam gives for the same bytes from memory."""
    from vcsbeam_amd.apps import incoherent_beam, make_incoh_beam_files
    from vcsbeam_amd.workloads import legacy_int4
    rng = np.random.default_rng(8)
    _metafits(tmp_path / "legacy.metafits", rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT),
              np.full(NT, 377.0), np.full(NT, 100.0), flagged=2, mode="VOLTAGE_START")
    data = legacy_int4(10000 * 128 * 2 * NT)
    data.tofile(tmp_path / f"{GPS}_{GPS}_ch109.dat")
    print("legacy file -> incoherent PSRFITS", flush=True)
    out = make_incoh_beam_files(tmp_path / "legacy.metafits", tmp_path, tmp_path / "out")
    off, scl, Is = out["seconds"][0]
    print("same second from memory", flush=True)
    _, off_m, scl_m, Is_m = incoherent_beam(data, 10000, 128, 2 * NT)
    np.testing.assert_array_equal(Is, Is_m)
    np.testing.assert_array_equal(off, off_m)
    np.testing.assert_array_equal(scl, scl_m)
    assert os.path.basename(out["psrfits"]) == f"G0001_{GPS}_incoh_ch109"
    (prim, _), (sub, tab) = read_fits(pathlib_of(out["psrfits"] + "_0001.fits"))
    assert prim["RA"] == "05:20:00.00" and prim["DEC"] == "-26:42:00.00"
    assert sub["NPOL"] == 1 and sub["POL_TYPE"] == "AA+BB" and sub["NAXIS2"] == 1
    cols = _columns(sub)
    row = tab[:sub["NAXIS1"]]
    np.testing.assert_array_equal(_col(row, cols["DATA"]), Is.ravel())
    np.testing.assert_array_equal(_col(row, cols["DAT_SCL"]), scl)


def test_offline_pfb_files_feed_the_incoherent_beam(pkg, gpu_lib, tmp_path):
    """app/fine_pfb_offline.c then app/make_mwa_incoh_beam.c on its output: an
    MWAX subfile through fine_pfb_offline_files gives the legacy recombined
    file <obsid>_<gps>_ch109.dat, byte-equal to the same second's PFB from
    memory; that file, read back as legacy data, gives the incoherent beam of
    those bytes."""
    from vcsbeam_amd.apps import (fine_pfb_offline, fine_pfb_offline_files, incoherent_beam,
                                  make_incoh_beam_files)
    from vcsbeam_amd.workloads import mwax_voltages
    lib = gpu_lib
    rng = np.random.default_rng(9)
    geo = (rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT), np.full(NT, 377.0),
           np.full(NT, 100.0))
    _metafits(tmp_path / "mwax.metafits", *geo, flagged=-1)
    _metafits(tmp_path / "legacy.metafits", *geo, flagged=-1, mode="VOLTAGE_START")
    data = mwax_voltages(2 * NT, 20, first_block=0, seed=41)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    (tmp_path / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    print("subfile -> offline PFB -> legacy file", flush=True)
    paths = fine_pfb_offline_files(tmp_path / "mwax.metafits", tmp_path, tmp_path / "pfb")
    assert [os.path.basename(p) for p in paths] == [f"{GPS}_{GPS}_ch109.dat"]
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(tmp_path / "mwax.metafits").encode(), C.byref(md), C.byref(vmd))
    fine = np.loadtxt(os.path.join(pkg.RUNTIME_DIR, "FINEPFB.dat"))
    print("same second's PFB from memory", flush=True)
    mem = fine_pfb_offline(_Obs(md, vmd, data, fine, None))[0]
    lib.vmFreeMetafits(C.byref(md))
    got = np.fromfile(paths[0], np.uint8)
    np.testing.assert_array_equal(got, mem.ravel())
    print("legacy file -> incoherent beam", flush=True)
    out = make_incoh_beam_files(tmp_path / "legacy.metafits", tmp_path / "pfb", tmp_path / "incoh")
    _, off_m, scl_m, Is_m = incoherent_beam(got, 10000, 128, 2 * NT)
    off, scl, Is = out["seconds"][0]
    np.testing.assert_array_equal(Is, Is_m)
    np.testing.assert_array_equal(scl, scl_m)


def test_bind_obs_data_and_read_calibration(pkg, gpu_lib, tmp_path):
    """The reference's own set-up calls by file name: vmLoadObsMetafits,
    vmLoadCalMetafits, vmBindObsData ("109", "+0", datadir) and
    vmReadNextSecond read the subfile's second into the read buffer;
    vmBindCalibrationData + vmReadCalibration put the Offringa solution
    into D."""
    from vcsbeam_amd.workloads import mwax_voltages
    lib = gpu_lib
    rng = np.random.default_rng(12)
    _metafits(tmp_path / "obs.metafits", rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT),
              np.full(NT, 377.0), np.full(NT, 100.0), flagged=-1)
    data = mwax_voltages(2 * NT, 20, first_block=0, seed=43)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    (tmp_path / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    sol = rng.standard_normal((NT, 128, 2, 2)) + 1j * rng.standard_normal((NT, 128, 2, 2))
    _offringa(tmp_path / "cal.bin", sol)
    mf = str(tmp_path / "obs.metafits").encode()
    vm = lib.vmInit(False)
    v = vm.contents
    try:
        lib.vmLoadObsMetafits(vm, mf)
        lib.vmLoadCalMetafits(vm, mf)
        print("vmBindObsData + vmReadNextSecond", flush=True)
        lib.vmBindObsData(vm, b"109", 1, 0, b"+0", 1, 0, str(tmp_path).encode())
        assert C.cast(v.gps_seconds_to_process, C.POINTER(C.c_uint32))[0] == GPS
        assert lib.vmReadNextSecond(vm) == pkg.VM_SUCCESS
        got = np.ctypeslib.as_array(C.cast(v.v.contents.read_ptr, C.POINTER(C.c_int8)),
                                    (data.size,))
        np.testing.assert_array_equal(got, data.ravel())
        lib.vmLoadFilter(vm, b"FINEPFB", pkg.ANALYSIS_FILTER, 128)
        lib.vmInitForwardPFB(vm, 128, pkg.PFB_FULL_PRECISION)
        lib.vmMallocDHost(vm)
        lib.vmBindCalibrationData(vm, str(tmp_path / "cal.bin").encode(), 2, False, None)
        lib.vmReadCalibration(vm)
        D = np.ctypeslib.as_array(C.cast(v.D, C.POINTER(C.c_double)), (NT, 128, 2, 2, 2))
        np.testing.assert_array_equal(D[..., 0] + 1j * D[..., 1], sol)
        lib.vmFreeDHost(vm)
    finally:
        lib.destroy_vcsbeam_context(vm)


def test_c_caller_matches_python_driver(pkg, gpu_lib, tmp_path):
    """The C caller examples/tied_array_beam.c (built with the library,
    include/vcsbeam.h only, the reference's by-file set-up calls) writes the
    same PSRFITS rows as the Python file driver for the same files."""
    import subprocess
    from vcsbeam_amd.apps import make_tied_array_beam_files
    from vcsbeam_amd.workloads import mwax_voltages
    rng = np.random.default_rng(14)
    _metafits(tmp_path / "obs.metafits", rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT),
              np.full(NT, 377.0), np.round(rng.uniform(50, 300, NT), 3), flagged=3)
    data = mwax_voltages(2 * NT, 20, first_block=0, seed=47)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    (tmp_path / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    sol = rng.standard_normal((NT, 128, 2, 2)) + 1j * rng.standard_normal((NT, 128, 2, 2))
    _offringa(tmp_path / "cal.bin", sol)
    (tmp_path / "pointings.txt").write_text("05:34:30.00 -26:42:00.00\n06:00:00.00 -20:00:00.00\n")
    print("Python file driver", flush=True)
    py = make_tied_array_beam_files(tmp_path / "obs.metafits", tmp_path, tmp_path / "py",
                                    str(tmp_path / "pointings.txt"), None,
                                    cal=("offringa", tmp_path / "cal.bin"), vdif=False)
    exe = os.path.join(os.path.dirname(pkg.LIBPATH), "tied_array_beam")
    (tmp_path / "c").mkdir()
    print("C caller", flush=True)
    r = subprocess.run([exe, str(tmp_path / "obs.metafits"), str(tmp_path),
                        str(tmp_path / "pointings.txt"), str(tmp_path / "cal.bin"),
                        str(tmp_path / "c"), "109", "+0", "1"],
                       capture_output=True, text=True, timeout=120, env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-2000:]
    cfiles = r.stdout.split()
    assert len(cfiles) == 2
    for p in range(2):
        (_, _), (sub_py, tab_py) = read_fits(pathlib_of(py["psrfits"][p] + "_0001.fits"))
        (_, _), (sub_c, tab_c) = read_fits(pathlib_of(cfiles[p] + "_0001.fits"))
        assert os.path.basename(cfiles[p]) == os.path.basename(py["psrfits"][p])
        assert sub_py["NAXIS2"] == sub_c["NAXIS2"] == 1
        assert tab_py == tab_c


def test_c_callers_offline_pfb_and_incoh_match_python_drivers(pkg, gpu_lib, tmp_path):
    """examples/fine_pfb_offline.c and examples/incoh_beam.c (include/vcsbeam.h
    only, the reference applications' call sequences) against the Python file
    drivers on the same two seconds: the offline PFB's legacy recombined
    files are byte-equal (the C caller writes them into the data directory,
    as vmWritePFBOutputToFile does), and the incoherent beam's PSRFITS from
    those files has the same headers (the file's creation DATE aside) and the
    same row bytes."""
    import subprocess
    from vcsbeam_amd.apps import fine_pfb_offline_files, make_incoh_beam_files
    from vcsbeam_amd.workloads import mwax_voltages
    rng = np.random.default_rng(15)
    nsec = 2
    geo = (rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT), np.full(NT, 377.0),
           np.full(NT, 100.0))
    _metafits(tmp_path / "mwax.metafits", *geo, flagged=-1)
    _metafits(tmp_path / "legacy.metafits", *geo, flagged=2, mode="VOLTAGE_START")
    data = mwax_voltages(2 * NT, 20 * nsec, first_block=0, seed=43)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    cdir = tmp_path / "c"
    cdir.mkdir()
    (cdir / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    bindir = os.path.dirname(pkg.LIBPATH)
    print("Python offline PFB driver, 2 s", flush=True)
    paths = fine_pfb_offline_files(tmp_path / "mwax.metafits", cdir, tmp_path / "py",
                                   nseconds=nsec)
    print("C offline PFB caller, 2 s", flush=True)
    r = subprocess.run([os.path.join(bindir, "fine_pfb_offline"), str(tmp_path / "mwax.metafits"),
                        str(cdir), "109", "+0", str(nsec)],
                       capture_output=True, text=True, timeout=120, env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-2000:]
    assert len(paths) == nsec
    for p in paths:
        name = os.path.basename(p)
        assert (cdir / name).read_bytes() == pathlib_of(p).read_bytes(), name
    print("Python incoherent-beam driver on the C caller's files", flush=True)
    py = make_incoh_beam_files(tmp_path / "legacy.metafits", cdir, tmp_path / "py_incoh",
                               nseconds=nsec)
    (tmp_path / "c_incoh").mkdir()
    print("C incoherent-beam caller", flush=True)
    r = subprocess.run([os.path.join(bindir, "incoh_beam"), str(tmp_path / "legacy.metafits"),
                        str(cdir), str(tmp_path / "c_incoh"), "109", "+0", str(nsec)],
                       capture_output=True, text=True, timeout=120, env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-2000:]
    cbase = r.stdout.split()
    assert len(cbase) == 1 and os.path.basename(cbase[0]) == os.path.basename(py["psrfits"])
    (prim_py, _), (sub_py, tab_py) = read_fits(pathlib_of(py["psrfits"] + "_0001.fits"))
    (prim_c, _), (sub_c, tab_c) = read_fits(pathlib_of(cbase[0] + "_0001.fits"))
    prim_py.pop("DATE", None)
    prim_c.pop("DATE", None)
    assert prim_py == prim_c and sub_py == sub_c
    assert sub_c["NAXIS2"] == nsec
    assert tab_py == tab_c


def test_tied_array_files_three_seconds_against_oracle(pkg, orc, gpu_lib, tmp_path):
    """Three seconds through make_tied_array_beam_files against the FP64
    oracle chain, each second's expected geometry computed here, not by the
    driver:
    - the beam geometry of second t is calc_beam_geom at
      MJD = sched_start_mjd + (t + first GPS second - obs_id + 0.5) / 86400
      (src/jones.c:170-188), so the phases follow the sky over the seconds;
    - the metafits carries dipole delays with dead dipoles on three tiles
      (two configurations besides all-live), so B per tile is the analytic
      beam of the tile's configuration times the parallactic rotation, built
      here from the FP64 oracle's beam;
    - tile 5 is flagged in the metafits and its Offringa solution is finite:
      it keeps D (metafits flags change w only, src/metadata.c:1554-1561);
    - second t's PFB reads the last block of second t - 1 as its margin.
    The FP32 Stokes must be within 1e-5 of the oracle in every second."""
    from vcsbeam_amd import BeamGeom
    from vcsbeam_amd.apps import make_tied_array_beam_files
    from vcsbeam_amd.workloads import mwax_voltages, unit_vectors
    lib = gpu_lib
    rng = np.random.default_rng(21)
    nsec = 3
    east, north = rng.uniform(-700, 700, NT), rng.uniform(-700, 700, NT)
    height = 377.0 + rng.uniform(-2, 2, NT)
    cable = np.round(rng.uniform(50, 300, NT), 3)
    flagged = 5
    delays = np.tile(np.array([0, 1, 2, 3] * 4), (NT, 1))
    delays[2, 6] = 32                     # one dead dipole
    delays[4, 6] = 32                     # the same configuration
    delays[6, [1, 14]] = 32               # two dead dipoles
    _metafits(tmp_path / "obs.metafits", east, north, height, cable, flagged, delays=delays)
    data = mwax_voltages(2 * NT, 20 * nsec, first_block=0, seed=33)
    hdr = f"HDR_SIZE 4096\nNINPUTS {2 * NT}\n".encode().ljust(4096, b"\0")
    (tmp_path / f"{GPS}_{GPS}_109.sub").write_bytes(hdr + bytes(data[0].nbytes) + data.tobytes())
    sol = rng.standard_normal((NT, 128, 2, 2)) + 1j * rng.standard_normal((NT, 128, 2, 2))
    _offringa(tmp_path / "cal.bin", sol)
    ra_h, dec_d = 11.0, -20.0             # about 70 deg elevation at MJD 59000.5
    print(f"files -> PSRFITS, {nsec} seconds", flush=True)
    out = make_tied_array_beam_files(tmp_path / "obs.metafits", tmp_path, tmp_path / "out",
                                     [ra_h], [dec_d], nseconds=nsec,
                                     cal=("offringa", tmp_path / "cal.bin"), vdif=False)
    D = sol.copy()
    D[..., 0, 1] = D[..., 1, 0] = 0       # keep_cross_terms off; tile 5 keeps D
    fine = np.loadtxt(os.path.join(pkg.RUNTIME_DIR, "FINEPFB.dat")).astype(np.int32)
    centre = 109 * 1280000
    freqs = centre - 640000 + 10000.0 * np.arange(128)
    ninput = 2 * NT
    out_idx = np.arange(ninput, dtype=np.int32)           # vcs_order = input for 16 inputs
    polQ, polP = np.arange(0, ninput, 2, dtype=np.uint32), np.arange(1, ninput, 2, dtype=np.uint32)
    invw = 1.0 / (2 * (NT - 1))
    els = []
    for t in range(nsec):
        mjd = 59000.5 + (t + 0.5) / 86400.0
        bg = BeamGeom()
        lib.calc_beam_geom(ra_h, dec_d, mjd, C.byref(bg))
        az, el = bg.az, bg.el
        els.append(el)
        assert abs(out["bg_per_second"][t][0][0] - az) < 1e-12
        assert abs(out["bg_per_second"][t][0][1] - el) < 1e-12
        za = math.pi / 2 - el
        P = np.empty(4)
        lib.parallactic_angle_correction(P.ctypes.data, C.c_double(-0.4660608448386394),
                                         C.c_double(az), C.c_double(za))
        B = np.empty((1, NT, 2, 2), complex)
        for a in range(NT):
            amps = (delays[a] != 32).astype(float)
            cfg = np.zeros(96)
            for q in range(2):
                cfg[q * 32:(q + 1) * 32:2] = amps
                cfg[64 + 16 * q:64 + 16 * (q + 1)] = np.where(amps != 0, delays[a] * 435e-12, 0)
            B[0, a] = orc.dipole_beam(np.array([az]), np.array([za]), np.array([float(centre)]),
                                      cfg)[0] @ P.reshape(2, 2)
        margin = np.zeros_like(data[:1]) if t == 0 else data[20 * t - 1:20 * t]
        htr = np.ascontiguousarray(np.concatenate([margin, data[20 * t:20 * (t + 1)]], axis=0))
        v = orc.pfb_forward(htr, ninput, 10000, fine, out_idx, fpga=True)
        # positions as the metafits stores them (float32)
        f32 = lambda x: np.asarray(x, np.float32).astype(np.float64)
        phi = orc.calc_phi(f32(east), f32(north), f32(height), cable, unit_vectors([az], [za]),
                           freqs)
        Sref, _ = orc.beamform(v, orc.calc_J(D, B), phi, polQ, polP, invw, 4)
        err = orc.rel_err(out["S"][t][0], Sref[0])
        print(f"second {t}: el {el:.9f} rad, Stokes rel err vs FP64 {err:.2e}", flush=True)
        assert err <= 1e-5
    # the sky moved: 1 s of Earth rotation is ~7.3e-5 rad
    assert abs(els[2] - els[0]) > 1e-5
    # the beams of the three configurations differ, so a wrong B would show
    assert np.abs(B[0, 0] - B[0, 2]).max() > 1e-3 and np.abs(B[0, 2] - B[0, 6]).max() > 1e-3
