"""PSRFITS output (src/beam_psrfits.c:39-488): header population for one
coarse channel and for a spliced band, and the SUBINT writer, read back
with a FITS parser written here to the FITS standard; J2000 to galactic
coordinates against the defining directions.  Host code only.

Parity unpinned against psrfits_utils (not available): its choice of
primary-header cards beyond those the reference sets.
"""
import ctypes as C
import math
import struct

import numpy as np
import pytest

from test_obsfiles import write_metafits


def _gal(lib, ra_deg, dec_deg):
    l, b = C.c_double(), C.c_double()
    lib.vmEqToGal(math.radians(ra_deg), math.radians(dec_deg), C.byref(l), C.byref(b))
    return math.degrees(l.value), math.degrees(b.value)


def test_galactic_defining_directions(pkg):
    lib = pkg.load()
    l, b = _gal(lib, 192.85948, 27.12825)  # north galactic pole
    assert abs(b - 90.0) < 1e-4
    l, b = _gal(lib, 266.40500, -28.93617)  # galactic centre
    assert min(abs(l), abs(l - 360)) < 1e-4 and abs(b) < 1e-4
    l, b = _gal(lib, 86.40500, 28.93617)  # anticentre
    assert abs(l - 180.0) < 1e-4 and abs(b) < 1e-4


def test_galactic_rotation_keeps_angles(pkg):
    lib = pkg.load()
    rng = np.random.default_rng(0)
    def vec(lon, lat):
        return np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])
    for _ in range(20):
        ra, dec = rng.uniform(0, 360, 2), np.degrees(np.arcsin(rng.uniform(-1, 1, 2)))
        g = [_gal(lib, ra[k], dec[k]) for k in range(2)]
        e = [vec(math.radians(ra[k]), math.radians(dec[k])) for k in range(2)]
        gv = [vec(math.radians(g[k][0]), math.radians(g[k][1])) for k in range(2)]
        assert abs(e[0] @ e[1] - gv[0] @ gv[1]) < 1e-12


def read_fits(path):
    """[(cards dict, data bytes)] for every HDU of a FITS file."""
    raw = path.read_bytes()
    assert len(raw) % 2880 == 0
    hdus, pos = [], 0
    while pos < len(raw):
        cards = {}
        while True:
            card = raw[pos:pos + 80].decode()
            pos += 80
            key = card[:8].strip()
            if key == "END":
                break
            if card[8:10] == "= ":
                v = card[10:].strip()
                if v.startswith("'"):
                    cards[key] = v[1:v.rindex("'")].replace("''", "'").rstrip()
                elif v in ("T", "F"):
                    cards[key] = v == "T"
                else:
                    v = v.split("/")[0].strip()
                    cards[key] = float(v) if any(ch in v for ch in ".EeD") else int(v)
        pos = (pos + 2879) // 2880 * 2880
        n = 0
        if cards.get("NAXIS", 0):
            n = cards["NAXIS1"] * cards["NAXIS2"]
        hdus.append((cards, raw[pos:pos + n]))
        pos += (n + 2879) // 2880 * 2880
    return hdus


def _columns(cards):
    cols, off = {}, 0
    sizes = {"D": 8, "E": 4, "B": 1}
    for i in range(1, cards["TFIELDS"] + 1):
        form = cards[f"TFORM{i}"]
        n, t = int(form[:-1] or 1), form[-1]
        cols[cards[f"TTYPE{i}"]] = (off, n, t)
        off += n * sizes[t]
    assert off == cards["NAXIS1"]
    return cols


def _col(row, spec):
    off, n, t = spec
    fmt = {"D": ">f8", "E": ">f4", "B": "u1"}[t]
    return np.frombuffer(row[off:off + n * np.dtype(fmt).itemsize], fmt)


def _context(pkg, lib, md, ns, nfine, first, nco):
    vm = lib.vmInit(False)
    v = vm.contents
    v.obs_metadata = C.pointer(md)
    idx = (C.c_int * 1)(first)
    v.coarse_chan_idxs_to_process = C.cast(idx, C.c_void_p)
    v.fine_sample_rate, v.nfine_chan, v.ncoarse_chans = ns, nfine, nco
    return vm, idx


def _release(lib, vm):
    vm.contents.obs_metadata = None
    vm.contents.coarse_chan_idxs_to_process = None
    lib.destroy_vcsbeam_context(vm)


def _beam(pkg):
    bg = pkg.BeamGeom()
    bg.mean_ra, bg.mean_dec = math.radians(83.633), math.radians(22.0145)
    bg.az, bg.el, bg.lmst = math.radians(10.0), math.radians(45.0), 1.5
    bg.intmjd, bg.fracmjd = 59000.0, 3723.4 / 86400
    return bg


def test_spliced_header_and_writer(pkg, tmp_path):
    """Three coarse channels spliced, 4 Stokes, 100 samples per second;
    three seconds with two per file make two files."""
    lib = pkg.load()
    mf = tmp_path / "obs.metafits"
    write_metafits(mf, ntiles=2)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(mf).encode(), C.byref(md), C.byref(vmd))
    ns, nf, nco, first = 100, 128, 3, 4
    vm, idx = _context(pkg, lib, md, ns, nf, first, nco)
    pf = pkg.Psrfits()
    bg = _beam(pkg)
    try:
        lib.populate_spliced_psrfits_header(vm, C.byref(pf), 2, 4, C.byref(bg), None, True)
        h = pf.hdr
        cc = [md.metafits_coarse_chans[first + k] for k in range(nco)]
        assert h.nchan == nf * nco and h.npol == 4 and h.nsblk == ns and h.nbits == 8
        assert h.fctr == (cc[0].chan_centre_hz + cc[2].chan_centre_hz) / 2e6
        assert h.BW == (cc[2].chan_end_hz - cc[0].chan_start_hz) / 1e6 and h.df == 0.01
        assert (h.ra_str, h.dec_str) == (b"05:34:31.92", b"+22:00:52.20")
        assert h.start_day == 59000 and h.start_sec == 3723 and h.summed_polns == 0
        assert pf.sub.bytes_per_subint == ns * 4 * nf * nco
        want_base = (f"G0057_1300000000_05:34:31.92_+22:00:52.20_ch{cc[0].rec_chan_number:03d}-"
                     f"{cc[2].rec_chan_number:03d}")
        assert pf.basefilename.decode() == want_base
        freqs = np.ctypeslib.as_array(C.cast(pf.sub.dat_freqs, C.POINTER(C.c_float)), (nf * nco,))
        np.testing.assert_allclose(freqs, (cc[0].chan_start_hz + 10000 * np.arange(nf * nco)) / 1e6,
                                   rtol=1e-7)
        pf.basefilename = str(tmp_path / "beam").encode()
        rng = np.random.default_rng(1)
        data = np.ctypeslib.as_array(C.cast(pf.sub.data, C.POINTER(C.c_uint8)),
                                     (pf.sub.bytes_per_subint,))
        offs = np.ctypeslib.as_array(C.cast(pf.sub.dat_offsets, C.POINTER(C.c_float)), (4 * nf * nco,))
        scl = np.ctypeslib.as_array(C.cast(pf.sub.dat_scales, C.POINTER(C.c_float)), (4 * nf * nco,))
        rows = []
        for t in range(3):
            data[:] = rng.integers(0, 256, data.size)
            offs[:] = rng.standard_normal(offs.size)
            scl[:] = rng.uniform(0.1, 2, scl.size)
            rows.append((data.copy(), offs.copy(), scl.copy()))
            assert lib.vmWritePsrfitsSubint(C.byref(pf)) == 0
        lib.vmClosePsrfits(C.byref(pf))
        files = [tmp_path / "beam_0001.fits", tmp_path / "beam_0002.fits"]
        t = 0
        for k, f in enumerate(files):
            (prim, _), (sub, tab) = read_fits(f)
            assert prim["FITSTYPE"] == "PSRFITS" and prim["OBS_MODE"] == "SEARCH"
            assert prim["OBSNCHAN"] == nf * nco and prim["STT_IMJD"] == 59000
            assert prim["STT_SMJD"] == 3723 + 2 * k and prim["TELESCOP"] == "MWA"
            assert abs(prim["OBSFREQ"] - h.fctr) < 1e-9 and prim["RA"] == "05:34:31.92"
            assert sub["EXTNAME"] == "SUBINT" and sub["NAXIS2"] == (2, 1)[k]
            assert sub["NSBLK"] == ns and sub["NPOL"] == 4 and sub["NCHAN"] == nf * nco
            assert sub["TDIM17"] == f"({nf * nco},4,{ns})" and sub["NSUBOFFS"] == 2 * k
            cols = _columns(sub)
            for r in range(sub["NAXIS2"]):
                row = tab[r * sub["NAXIS1"]:(r + 1) * sub["NAXIS1"]]
                d, o, s = rows[t]
                np.testing.assert_array_equal(_col(row, cols["DATA"]), d)
                np.testing.assert_array_equal(_col(row, cols["DAT_OFFS"]), o)
                np.testing.assert_array_equal(_col(row, cols["DAT_SCL"]), s)
                np.testing.assert_array_equal(_col(row, cols["DAT_FREQ"]), freqs.astype(np.float64))
                assert _col(row, cols["TSUBINT"])[0] == 1.0
                assert _col(row, cols["OFFS_SUB"])[0] == t + 0.5
                l, b = _gal(lib, 83.633, 22.0145)
                assert abs(_col(row, cols["GLON_SUB"])[0] - l) < 1e-9
                assert abs(_col(row, cols["TEL_ZEN"])[0] - 45.0) < 1e-4
                t += 1
        assert t == 3
    finally:
        lib.free_psrfits(C.byref(pf))
        _release(lib, vm)
        lib.vmFreeMetafits(C.byref(md))


def test_single_channel_and_incoherent_headers(pkg, tmp_path):
    lib = pkg.load()
    mf = tmp_path / "obs.metafits"
    write_metafits(mf, ntiles=2)
    md, vmd = pkg.MetafitsMetadata(), pkg.VoltageMetadata()
    lib.vmLoadMetafits(str(mf).encode(), C.byref(md), C.byref(vmd))
    vm, idx = _context(pkg, lib, md, 100, 128, 2, 1)
    pf, pi = pkg.Psrfits(), pkg.Psrfits()
    bg = _beam(pkg)
    try:
        lib.populate_psrfits_header(vm, C.byref(pf), 10, 1, C.byref(bg), None, True)
        cc = md.metafits_coarse_chans[2]
        assert pf.hdr.nchan == 128 and pf.hdr.npol == 1 and pf.hdr.fctr == cc.chan_centre_hz / 1e6
        freqs = np.ctypeslib.as_array(C.cast(pf.sub.dat_freqs, C.POINTER(C.c_float)), (128,))
        np.testing.assert_allclose(freqs, (cc.chan_start_hz + 5000 + 10000 * np.arange(128)) / 1e6,
                                   rtol=1e-7)
        assert pf.basefilename.decode().endswith(f"_ch{cc.rec_chan_number:03d}")
        lib.populate_psrfits_header(vm, C.byref(pi), 10, 1, C.byref(bg), None, False)
        assert pi.hdr.summed_polns == 1 and pi.hdr.ra2000 == 83.6331 and pi.hdr.dec2000 == 22.0145
        assert pi.basefilename.decode() == f"G0057_1300000000_incoh_ch{cc.rec_chan_number:03d}"
    finally:
        lib.free_psrfits(C.byref(pf))
        lib.free_psrfits(C.byref(pi))
        _release(lib, vm)
        lib.vmFreeMetafits(C.byref(md))
