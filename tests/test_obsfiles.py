"""Observation files without mwalib (src/metadata.c:930-993, 1297-1424):
a metafits subset read from a FITS file written here to the FITS standard
(primary header with a CONTINUE'd long string, an unrelated table HDU before
TILEDATA), and one second of voltages read from MWAX subfiles and legacy
recombined files laid out as the instrument writes them.  Host code only.

Parity unpinned against mwalib itself (not available): its ordering of
rf_inputs (by antenna then polarisation here) and its legacy VCS order
formula; the file layouts follow the MWA's published formats.
"""
import ctypes as C
import struct

import numpy as np
import pytest


def _card(key, value=None, comment=""):
    if value is None:
        s = key.ljust(80)
    else:
        if isinstance(value, str):
            v = "'" + value.replace("'", "''").ljust(8) + "'"
            v = v.ljust(20)
        elif isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, int):
            v = str(value).rjust(20)
        else:
            v = repr(float(value)).upper().rjust(20)
        s = f"{key:<8}= {v}" + (f" / {comment}" if comment else "")
    assert len(s) <= 80, s
    return s.ljust(80).encode()


def _long_string(key, value):
    """The CONTINUE convention: parts of at most 60 characters ending in '&'."""
    parts = [value[i:i + 60] for i in range(0, len(value), 60)]
    out = []
    for i, p in enumerate(parts):
        last = i == len(parts) - 1
        q = "'" + p + ("" if last else "&") + "'"
        out.append((f"{key:<8}= " if i == 0 else "CONTINUE  ") + q)
    return b"".join(s.ljust(80).encode() for s in out)


def _pad(b, fill=b" "):
    return b + fill * (-len(b) % 2880)


def _bintable(name, cols, nrows):
    """cols: list of (ttype, tform, numpy big-endian array or bytes rows)."""
    rowbytes = []
    for r in range(nrows):
        row = b""
        for _, form, data in cols:
            row += data[r] if isinstance(data, list) else data[r:r + 1].tobytes()
        rowbytes.append(row)
    nax1 = len(rowbytes[0])
    hdr = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
           _card("NAXIS1", nax1), _card("NAXIS2", nrows), _card("PCOUNT", 0),
           _card("GCOUNT", 1), _card("TFIELDS", len(cols))]
    for i, (tt, form, _) in enumerate(cols, 1):
        hdr += [_card(f"TTYPE{i}", tt), _card(f"TFORM{i}", form)]
    hdr += [_card("EXTNAME", name), _card("END")]
    return _pad(b"".join(hdr)) + _pad(b"".join(rowbytes), b"\0")


CHANNELS = [133, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123,
            124, 125, 126, 127, 128, 129, 130, 131]


def write_metafits(path, ntiles=4, mode="MWAX_VCS", gps=1300000000, mjd=True):
    rng = np.random.default_rng(1)
    n = 2 * ntiles
    order = rng.permutation(n)  # rows in a shuffled order
    ant = np.array([i // 2 for i in range(n)])[order]
    pol = [("X", "Y")[i % 2] for i in range(n)]
    pol = [pol[i] for i in order]
    inp = rng.permutation(n)  # metafits input numbers: a permutation of 0 .. n-1
    north = rng.uniform(-1500, 1500, ntiles)
    east = rng.uniform(-1500, 1500, ntiles)
    height = 377.0 + rng.uniform(-3, 3, ntiles)
    cable = rng.uniform(0, 300, ntiles)
    flag = np.zeros(n, int)
    flag[[k for k in range(n) if ant[k] == 2]] = 1
    prim = [_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True),
            _card("GPSTIME", gps, "obs id"), _card("EXPOSURE", 600),
            _card("GOODTIME", float(gps + 8 + 315964800 - 18), "UNIX time of good data"),
            _card("PROJECT", "G0057"), _card("MODE", mode), _card("NCHANS", 24),
            _card("BANDWDTH", 30.72), _card("FINECHAN", 10.0), _card("NINPUTS", n),
            _card("RA", 83.6331), _card("DEC", 22.0145)]
    if mjd:
        prim.append(_card("MJD", 59000.5))
    prim.append(_long_string("CHANNELS", ",".join(str(c) for c in CHANNELS)))
    prim.append(_card("END"))
    other = _bintable("OTHER", [("X", "1J", np.arange(3, dtype=">i4"))], 3)
    i2 = lambda a: np.asarray(a, ">i2")
    f4 = lambda a: np.asarray(a, ">f4")
    cols = [("Input", "I", i2(inp)), ("Antenna", "I", i2(ant)), ("Tile", "I", i2(1000 + ant)),
            ("TileName", "8A", [f"Tile{a:03d}".encode().ljust(8) for a in ant]),
            ("Pol", "1A", [p.encode() for p in pol]), ("Rx", "I", i2(ant // 8)),
            ("Slot", "I", i2(ant % 8)), ("Flag", "I", i2(flag)),
            ("Length", "14A", [f"EL_{cable[a]:.3f}".encode().ljust(14) for a in ant]),
            ("North", "E", f4(north[ant])), ("East", "E", f4(east[ant])),
            ("Height", "E", f4(height[ant])), ("Gains", "24I", [bytes(48)] * n)]
    path.write_bytes(_pad(b"".join(prim)) + other + _bintable("TILEDATA", cols, n))
    return dict(ant=ant, pol=pol, inp=inp, north=north, east=east, height=height, cable=cable,
                flag=flag)


def test_load_metafits(pkg, tmp_path):
    lib = pkg.load()
    f = tmp_path / "1300000000.metafits"
    t = write_metafits(f)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    assert lib.vmLoadMetafits(str(f).encode(), C.byref(md), C.byref(vmd)) == 0
    try:
        assert md.obs_id == 1300000000 and md.sched_start_gps_time_ms == 1300000000000
        assert md.sched_start_mjd == 59000.5 and md.mwa_version == pkg.VCSMWAXv2
        assert md.project_id == b"G0057" and md.num_rf_inputs == 8 and md.num_ants == 4
        assert md.coarse_chan_width_hz == 1280000 and md.corr_fine_chan_width_hz == 10000
        assert md.num_corr_fine_chans_per_coarse == 128 and md.num_volt_fine_chans_per_coarse == 1
        assert (md.ra_tile_pointing_deg, md.dec_tile_pointing_deg) == (83.6331, 22.0145)
        cc = [md.metafits_coarse_chans[i] for i in range(md.num_metafits_coarse_chans)]
        assert [c.rec_chan_number for c in cc] == sorted(CHANNELS)
        assert [c.corr_chan_number for c in cc] == list(range(24))
        assert cc[0].chan_centre_hz == 109 * 1280000 and cc[0].chan_start_hz == 109 * 1280000 - 640000
        for i in range(8):
            rf = md.rf_inputs[i]
            a, p = i // 2, "XY"[i % 2]
            k = [j for j in range(8) if t["ant"][j] == a and t["pol"][j] == p][0]
            assert (rf.ant, rf.pol, rf.tile_id, rf.tile_name) == (a, p.encode(), 1000 + a,
                                                                  f"Tile{a:03d}".encode())
            assert rf.input == t["inp"][k] and rf.subfile_order == i
            assert rf.flagged == (a == 2)
            assert abs(rf.north_m - t["north"][a]) < 1e-3 and abs(rf.height_m - t["height"][a]) < 1e-4
            assert abs(rf.electrical_length_m - round(t["cable"][a], 3)) < 1e-9
            assert rf.vcs_order == t["inp"][k]  # a partial 64-input board: the input
        assert vmd.num_samples_per_voltage_block == 64000 and vmd.num_voltage_blocks_per_second == 20
        assert vmd.voltage_block_size_bytes == 64000 * 2 * 8
    finally:
        lib.vmFreeMetafits(C.byref(md))


def test_vcs_order_interleaves_whole_boards(pkg, tmp_path):
    """With whole 64-input boards the legacy order interleaves each board's
    inputs (parity unpinned against mwalib); it is a permutation."""
    lib = pkg.load()
    f = tmp_path / "big.metafits"
    t = write_metafits(f, ntiles=64)
    md = pkg.MetafitsMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), None)
    try:
        order = [md.rf_inputs[i].vcs_order for i in range(128)]
        for i in range(128):
            inp = md.rf_inputs[i].input
            assert order[i] == (inp & ~63) | ((inp & 15) << 2) | ((inp >> 4) & 3)
        assert sorted(order) == list(range(128))
    finally:
        lib.vmFreeMetafits(C.byref(md))


def test_metafits_legacy_and_gps_mjd(pkg, tmp_path):
    """Legacy VCS mode; without an MJD card the start MJD comes from GPSTIME
    (GPS epoch 1980-01-06, 18 leap seconds: GPS 1300000000 is
    2021-03-17 07:06:22 UTC)."""
    import datetime
    lib = pkg.load()
    f = tmp_path / "legacy.metafits"
    write_metafits(f, mode="VOLTAGE_START", mjd=False)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), C.byref(vmd))
    try:
        assert md.mwa_version == pkg.VCSLegacyRecombined and md.num_volt_fine_chans_per_coarse == 128
        utc = datetime.datetime(1980, 1, 6) + datetime.timedelta(seconds=1300000000 - 18)
        assert utc == datetime.datetime(2021, 3, 17, 7, 6, 22)
        want = (utc - datetime.datetime(1858, 11, 17)).total_seconds() / 86400
        assert abs(md.sched_start_mjd - want) < 1e-9
        assert vmd.voltage_block_size_bytes == 10000 * 128 * 8
        assert [md.metafits_coarse_chans[i].gpubox_number for i in range(3)] == [1, 2, 3]
    finally:
        lib.vmFreeMetafits(C.byref(md))


def _subfile(path, ninputs, blocks, nin_hdr=None):
    hdr = (f"HDR_SIZE 4096\nOBS_ID 1300000000\nNINPUTS {nin_hdr or ninputs}\n"
           f"NTIMESAMPLES 64000\n").encode()
    path.write_bytes(hdr.ljust(4096, b"\0") + blocks.tobytes())


def test_read_seconds_from_subfiles(pkg, tmp_path):
    """Two 8-second subfiles (delay block first); seconds 6..9 of the
    observation span the file boundary."""
    lib = pkg.load()
    f = tmp_path / "obs.metafits"
    write_metafits(f, ntiles=2)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), C.byref(vmd))
    try:
        I, bsz = 4, 64000 * 2 * 4
        rng = np.random.default_rng(2)
        data = {}
        for sub in (1300000000, 1300000008):
            blocks = rng.integers(-128, 128, (161, I, 64000, 2), dtype=np.int8)
            data[sub] = blocks
            _subfile(tmp_path / f"1300000000_{sub}_{md.metafits_coarse_chans[5].rec_chan_number}.sub",
                     I, blocks)
        vf = pkg.VoltageFiles(str(tmp_path).encode(), C.pointer(md), C.pointer(vmd))
        buf = np.empty(20 * bsz, np.int8)
        for gps in range(1300000006, 1300000010):
            rc = lib.vmReadSecondFromFiles(C.byref(vf), gps, 5, buf.ctypes.data, buf.nbytes)
            assert rc == 0
            sub = 1300000000 + (gps - 1300000000) // 8 * 8
            s = gps - sub
            want = data[sub][1 + 20 * s:1 + 20 * (s + 1)]
            np.testing.assert_array_equal(buf.reshape(want.shape), want)
        # a second with no file, a wrong size and a wrong channel are errors
        assert lib.vmReadSecondFromFiles(C.byref(vf), 1300000016, 5, buf.ctypes.data, buf.nbytes) != 0
        assert lib.vmReadSecondFromFiles(C.byref(vf), 1300000001, 5, buf.ctypes.data, 100) != 0
        assert lib.vmReadSecondFromFiles(C.byref(vf), 1300000001, 6, buf.ctypes.data, buf.nbytes) != 0
    finally:
        lib.vmFreeMetafits(C.byref(md))


def test_subfile_input_count_checked(pkg, tmp_path):
    lib = pkg.load()
    f = tmp_path / "obs.metafits"
    write_metafits(f, ntiles=2)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), C.byref(vmd))
    try:
        blocks = np.zeros((161, 4, 64000, 2), np.int8)
        _subfile(tmp_path / f"1300000000_1300000000_{md.metafits_coarse_chans[0].rec_chan_number}.sub",
                 4, blocks, nin_hdr=6)
        vf = pkg.VoltageFiles(str(tmp_path).encode(), C.pointer(md), C.pointer(vmd))
        buf = np.empty(20 * 64000 * 2 * 4, np.int8)
        assert lib.vmReadSecondFromFiles(C.byref(vf), 1300000000, 0, buf.ctypes.data, buf.nbytes) != 0
    finally:
        lib.vmFreeMetafits(C.byref(md))


def test_read_legacy_recombined_file(pkg, tmp_path):
    lib = pkg.load()
    f = tmp_path / "legacy.metafits"
    write_metafits(f, ntiles=2, mode="VOLTAGE_START")
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), C.byref(vmd))
    try:
        sec = np.random.default_rng(3).integers(0, 256, 10000 * 128 * 4, dtype=np.uint8)
        rec = md.metafits_coarse_chans[2].rec_chan_number
        sec.tofile(tmp_path / f"1300000000_1300000003_ch{rec:03d}.dat")
        vf = pkg.VoltageFiles(str(tmp_path).encode(), C.pointer(md), C.pointer(vmd))
        buf = np.empty_like(sec)
        assert lib.vmReadSecondFromFiles(C.byref(vf), 1300000003, 2, buf.ctypes.data, buf.nbytes) == 0
        np.testing.assert_array_equal(buf, sec)
    finally:
        lib.vmFreeMetafits(C.byref(md))


@pytest.mark.parametrize("field, value, message", [
    ("vcs_order", 999, "vcs_order 999"),
    ("ant", 77, "antenna 77"),
])
def test_bind_rejects_indices_outside_the_buffers(field, value, message):
    """Metadata whose RF-input indices would send the kernels outside their
    buffers ends the process with a message at vmBindObsMetadata, before any
    device work (run in a child process)."""
    import os
    import subprocess
    import sys
    code = ("import sys, ctypes as C; sys.path.insert(0, 'tests');"
            "from conftest import load_package; pkg = load_package();"
            "from vcsbeam_amd.workloads import Observation;"
            "o = Observation(ntiles=4, blocks_per_second=1);"
            f"o.rf[3].{field} = {value};"
            "lib = pkg.load(); vm = lib.vmInit(False);"
            "lib.vmBindObsMetadata(vm, C.byref(o.md), C.byref(o.vmd), 0, o.md.obs_id, 1)")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and message in r.stderr, r.stderr[-500:]


def test_begin_and_channel_strings(pkg, tmp_path):
    """parse_begin_string / parse_coarse_chan_string (src/metadata.c:1448-1545):
    "+n" from the first good second (GOODTIME, a UNIX time, 8 s after the
    start here) or channel index n, "-n" from the end, else absolute."""
    lib = pkg.load()
    f = tmp_path / "obs.metafits"
    write_metafits(f)
    md = pkg.MetafitsMetadata()
    lib.vmLoadMetafits(str(f).encode(), C.byref(md), None)
    try:
        assert md.good_time_gps_ms == 1300000008000 and md.num_metafits_timesteps == 600
        assert lib.parse_begin_string(C.byref(md), b"+0") == 1300000008
        assert lib.parse_begin_string(C.byref(md), b"+3") == 1300000011
        assert lib.parse_begin_string(C.byref(md), b"-1") == 1300000599
        assert lib.parse_begin_string(C.byref(md), b"1300000123") == 1300000123
        assert lib.parse_coarse_chan_string(C.byref(md), b"+2") == 2
        assert lib.parse_coarse_chan_string(C.byref(md), b"-1") == 23
        assert lib.parse_coarse_chan_string(C.byref(md), b"110") == 1
        assert lib.parse_coarse_chan_string(C.byref(md), b"133") == 23
    finally:
        lib.vmFreeMetafits(C.byref(md))


def test_context_owned_metafits_and_file_names(pkg, tmp_path):
    """vmLoadObsMetafits / vmLoadCalMetafits load into the context, which
    releases them; the MWAX and legacy file names of a second."""
    lib = pkg.load()
    f = tmp_path / "obs.metafits"
    write_metafits(f)
    vm = lib.vmInit(False)
    try:
        lib.vmLoadObsMetafits(vm, str(f).encode())
        lib.vmLoadCalMetafits(vm, str(f).encode())
        v = vm.contents
        assert v.obs_metadata.contents.obs_id == 1300000000
        cal = C.cast(v.cal_metadata, C.POINTER(pkg.MetafitsMetadata)).contents
        assert cal.num_rf_inputs == 8
        name = C.create_string_buffer(256)
        lib.vmGetVoltFilename(vm, 0, 1300000013, name)
        assert name.value == b"1300000000_1300000008_109.sub"
        lib.vmGetLegacyVoltFilename(vm, 23, 1300000013, name)
        assert name.value == b"1300000000_1300000013_ch133.dat"
        lib.vmPrintTitle(vm, b"test")
        lib.vmBindCalibrationData(vm, b"/cal/dir", 1, True, None)
        assert C.cast(v.cal.caldir, C.c_char_p).value == b"/cal/dir" and v.cal.cal_type == 1
    finally:
        lib.destroy_vcsbeam_context(vm)
