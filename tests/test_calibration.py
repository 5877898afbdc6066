"""Calibration into vm->D (src/calibration.c): the Offringa binary and RTS
text loaders, the corrections (reference-tile phase, cross terms, PQ phase
slope), the PQ correction table and tile-flag lists.  Host code only: these
tests run without a GPU.

The loaders are checked against files written here to the layouts the
reference reads (src/calibration.c:426-582 for Offringa, 227-409 for RTS),
with the expected D computed by numpy indexing and matrix products.  The
corrections are checked through invariants that do not restate the code: the
reference tile's corrected PP element carries exactly the slope phase and
its QQ element is real and positive, and every magnitude is kept.  The PQ
table is the reference's own file (pq_phase_correction.txt), with the
answers its rows fix.
"""
import ctypes as C
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

NCH = 128


def _ctx(pkg, obs, cal_md=None):
    lib = pkg.load()
    vm = lib.vmInit(False)
    v = vm.contents
    v.obs_metadata = C.pointer(obs.md)
    v.cal_metadata = C.cast(C.pointer(cal_md if cal_md is not None else obs.md), C.c_void_p)
    v.nfine_chan = NCH
    v.coarse_chan_idx = 0
    keep = {"idx": (C.c_int * 1)(0)}
    v.cal_coarse_chan_idxs_to_process = C.cast(keep["idx"], C.c_void_p)
    D = np.zeros((obs.ntiles, NCH, 2, 2), np.complex128)
    keep["D"] = D
    v.D = D.ctypes.data
    v.D_size_bytes = D.nbytes
    return lib, vm, D, keep


def _release(lib, vm):
    v = vm.contents
    v.D = None
    v.obs_metadata = None
    v.cal_metadata = None
    v.cal_coarse_chan_idxs_to_process = None
    v.cal = type(v.cal)()
    lib.destroy_vcsbeam_context(vm)


def _set_str(keep, cal, field, value):
    buf = C.create_string_buffer(value)
    keep.setdefault("strs", []).append(buf)
    setattr(cal, field, C.cast(buf, C.c_void_p))


def _cj(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _offringa_file(path, sol):
    nint, nant, nchan = sol.shape[:3]
    hdr = b"MWAOCAL\0" + struct.pack("<II", 0, 0) + struct.pack("<IIII", nint, nant, nchan, 4)
    hdr += struct.pack("<dd", 0.0, 1.0)
    assert len(hdr) == 48
    body = np.ascontiguousarray(sol.reshape(nint, nant, nchan, 4)).view(np.float64)
    path.write_bytes(hdr + body.tobytes())


@pytest.mark.parametrize("nchan", [32, 16])
def test_offringa_loader(pkg, tmp_path, nchan):
    """32 solution channels (the count the metadata asks for) and 16 (a
    mismatch the loader accepts with a warning): each covers 128 / nchan
    fine channels; a NaN matrix becomes zero; only interval 0 is used."""
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=4, blocks_per_second=1)
    obs.md.num_corr_fine_chans_per_coarse = 32
    rng = np.random.default_rng(3)
    sol = _cj(rng, 2, 4, nchan, 2, 2)
    sol[0, 2, 5] = np.nan
    f = tmp_path / "cal.bin"
    _offringa_file(f, sol)
    lib, vm, D, keep = _ctx(pkg, obs)
    try:
        _set_str(keep, vm.contents.cal, "caldir", str(f).encode())
        lib.vmLoadOffringaSolution(vm)
        want = np.repeat(sol[0], NCH // nchan, axis=1)
        want[2, 5 * (NCH // nchan):6 * (NCH // nchan)] = 0
        np.testing.assert_array_equal(D, want)
    finally:
        _release(lib, vm)


def _rts_files(d, node, amp, A, Dd, Jf, chans, width_hz, skip_ant=None):
    with open(d / f"DI_JonesMatrices_node{node:03d}.dat", "w") as f:
        f.write(f"{amp!r}\n")
        f.write(", ".join(repr(float(x)) for x in A.reshape(4).view(np.float64)) + "\n")
        for a in range(Dd.shape[0]):
            f.write(", ".join(repr(float(x)) for x in Dd[a].reshape(4).view(np.float64)) + "\n")
    with open(d / f"BandpassCalibration_node{node:03d}.dat", "w") as f:
        f.write(",".join(repr(float(c * width_hz / 1e6)) for c in chans) + "\n")
        for a in range(Jf.shape[0]):
            if a == skip_ant:
                continue
            for row in range(8):
                pol = row // 2
                J = Jf[a] if row % 2 else 2 * Jf[a]  # measured rows: ignored
                vals = []
                for c in chans:
                    z = J[c].reshape(4)[pol]
                    vals += [abs(z), np.angle(z)]
                f.write(f"{a + 1}," + ",".join(repr(float(x)) for x in vals) + "\n")


@pytest.mark.parametrize("use_bandpass", [True, False])
def test_rts_loader(pkg, tmp_path, use_bandpass):
    """D[ant][ch] = reverse(Dd[c] Jf[c][ch / 4] A^-1), c = the calibration
    observation's input for the same tile (tile ids in reverse order there)
    halved; a bandpass channel or antenna missing from the file is zero."""
    from vcsbeam_amd.workloads import Observation
    nt, ncal, width = 4, 32, 40000
    obs = Observation(ntiles=nt, blocks_per_second=1)
    cal = Observation(ntiles=nt, blocks_per_second=1)
    for i in range(2 * nt):
        cal.rf[i].tile_id = 1000 + (nt - 1 - i // 2)
    for o in (obs, cal):
        o.md.num_corr_fine_chans_per_coarse = ncal
        o.md.corr_fine_chan_width_hz = width
    rng = np.random.default_rng(4)
    A = _cj(rng, 2, 2)
    Dd = _cj(rng, nt, 2, 2)
    Jf = _cj(rng, nt, ncal, 2, 2)
    chans = [c for c in range(ncal) if c not in (3, 17)]
    node = obs.cc[0].corr_chan_number + 1
    _rts_files(tmp_path, node, 1.5, A, Dd, Jf, chans, width, skip_ant=1)
    Jf[:, [3, 17]] = 0
    Jf[1] = 0
    lib, vm, D, keep = _ctx(pkg, obs, cal.md)
    try:
        v = vm.contents
        _set_str(keep, v.cal, "caldir", str(tmp_path).encode())
        v.cal.use_bandpass = use_bandpass
        lib.vmLoadRTSSolution(vm)
        Ainv = np.linalg.inv(A)
        for a in range(nt):
            c = nt - 1 - a
            for ch in range(NCH):
                M = Dd[c] @ Jf[c][ch // 4] if use_bandpass else Dd[c]
                want = (M @ Ainv).reshape(4)[::-1].reshape(2, 2)
                np.testing.assert_allclose(D[a, ch], want, rtol=1e-12, atol=1e-12)
    finally:
        _release(lib, vm)


def test_calibration_corrections(pkg):
    """Reference tile Tile001, cross terms zeroed, phase slope: the
    reference tile's PP ends with phase slope * f + offset, its QQ real and
    positive; all kept magnitudes unchanged; a zero element of the reference
    leaves that element of every tile unrotated."""
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=4, blocks_per_second=1)
    rng = np.random.default_rng(5)
    lib, vm, D, keep = _ctx(pkg, obs)
    try:
        D[:] = _cj(rng, 4, NCH, 2, 2)
        D[1, 7, 1, 1] = 0  # reference QQ zero in channel 7
        D0 = D.copy()
        v = vm.contents
        _set_str(keep, v.cal, "ref_ant", b"Tile001")
        v.cal.keep_cross_terms = False
        v.cal.phase_slope, v.cal.phase_offset = 7.7258e-09, 0.4898
        lib.vmApplyCalibrationCorrections(vm)
        f = obs.cc[0].chan_start_hz + np.arange(NCH) * (obs.md.coarse_chan_width_hz // NCH)
        phi = 7.7258e-09 * f + 0.4898
        err = np.angle(D[1, :, 0, 0] * np.exp(-1j * phi))
        assert np.abs(err).max() < 1e-12
        q = np.delete(D[1, :, 1, 1], 7)
        assert np.abs(q.imag).max() < 1e-12 * np.abs(q).max() and (q.real > 0).all()
        assert not D[:, :, 0, 1].any() and not D[:, :, 1, 0].any()
        np.testing.assert_allclose(np.abs(D[:, :, 0, 0]), np.abs(D0[:, :, 0, 0]), rtol=1e-13)
        np.testing.assert_allclose(np.abs(D[:, :, 1, 1]), np.abs(D0[:, :, 1, 1]), rtol=1e-13)
        np.testing.assert_allclose(D[:, 7, 1, 1], D0[:, 7, 1, 1], rtol=1e-15)
        # relative phases between tiles are kept
        rel = D[:, :, 0, 0] * np.conj(D[0, :, 0, 0])
        rel0 = D0[:, :, 0, 0] * np.conj(D0[0, :, 0, 0])
        np.testing.assert_allclose(rel, rel0, rtol=1e-12)
    finally:
        _release(lib, vm)


def test_corrections_off_leave_D(pkg):
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=4, blocks_per_second=1)
    lib, vm, D, keep = _ctx(pkg, obs)
    try:
        D[:] = _cj(np.random.default_rng(6), 4, NCH, 2, 2)
        D0 = D.copy()
        v = vm.contents
        _set_str(keep, v.cal, "ref_ant", b"NONE")
        v.cal.keep_cross_terms = True
        lib.vmApplyCalibrationCorrections(vm)
        np.testing.assert_array_equal(D, D0)
    finally:
        _release(lib, vm)


def test_unknown_reference_tile_exits(pkg):
    """An unknown reference tile ends the process with a message, as in the
    reference (run in a child process)."""
    code = (
        "import ctypes as C, sys; sys.path.insert(0, 'tests');"
        "from conftest import load_package; pkg = load_package();"
        "from test_calibration import _ctx, _set_str;"
        "from vcsbeam_amd.workloads import Observation;"
        "obs = Observation(ntiles=4, blocks_per_second=1);"
        "lib, vm, D, keep = _ctx(pkg, obs);"
        "_set_str(keep, vm.contents.cal, 'ref_ant', b'Tile999');"
        "lib.vmApplyCalibrationCorrections(vm)")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "reference tile Tile999 is not in the observation" in r.stderr


@pytest.mark.parametrize("gps, ref, slope, offset", [
    (1260302416, b"HexE8", 0.0, 0.0),        # also inside the next row's range: first row wins
    (1201478418, b"LBG8", 7.7258e-09, 0.4898),
    (1280275218, b"LBG8", 7.7258e-09, 0.4898),
    (1165000000, b"Tile011", 5.6e-09, -0.06283),
    (1100000000, b"NONE", 0.25, 0.5),        # no row: slope and offset untouched
])
def test_pq_correction_table(pkg, gps, ref, slope, offset):
    lib = pkg.load()
    cal = pkg.Calibration()
    lib.init_calibration(C.byref(cal))
    cal.phase_slope, cal.phase_offset = 0.25, 0.5
    lib.parse_calibration_correction_file(gps, C.byref(cal))
    assert C.cast(cal.ref_ant, C.c_char_p).value == ref
    assert (cal.phase_slope, cal.phase_offset) == (slope, offset)
    lib.free_calibration(C.byref(cal))


def test_pq_correction_keeps_caller_choices(pkg):
    lib = pkg.load()
    cal = pkg.Calibration()
    lib.init_calibration(C.byref(cal))
    mine = C.create_string_buffer(b"Tile005")
    cal.ref_ant = C.cast(mine, C.c_void_p)
    cal.custom_pq_correction = True
    cal.phase_slope = 1e-9
    lib.parse_calibration_correction_file(1201478418, C.byref(cal))
    assert C.cast(cal.ref_ant, C.c_char_p).value == b"Tile005"
    assert cal.phase_slope == 1e-9 and cal.phase_offset == 0.0


def test_custom_tile_flags(pkg, tmp_path):
    """Listed tiles' D become zero in every fine channel; an unknown name is
    skipped with a warning; the unflagged-input count is refreshed."""
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=6, blocks_per_second=1, nflagged=1)
    flags = tmp_path / "flags.txt"
    flags.write_text("Tile001\n  Tile004 NoSuchTile\n")
    lib, vm, D, keep = _ctx(pkg, obs)
    try:
        D[:] = _cj(np.random.default_rng(7), 6, NCH, 2, 2)
        D0 = D.copy()
        v = vm.contents
        _set_str(keep, v.cal, "flags_file", str(flags).encode())
        lib.vmSetCustomTileFlags(vm)
        assert v.cal.nflags == 3
        assert lib.tilename_is_flagged(b"Tile004", C.byref(v.cal))
        assert not lib.tilename_is_flagged(b"Tile002", C.byref(v.cal))
        for a in range(6):
            if a in (1, 4):
                assert not D[a].any()
            else:
                np.testing.assert_array_equal(D[a], D0[a])
        assert v.num_not_flagged == 2 * 6 - 2  # the metadata flag (one tile, two inputs)
        v.cal.flags_file = None  # a Python buffer: free_calibration frees only the names here
        lib.free_calibration(C.byref(v.cal))
    finally:
        _release(lib, vm)
