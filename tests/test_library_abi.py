"""CPU checks of the C-ABI library: it loads, exports every function declared
in include/vcsbeam.h, and the Python struct mirrors match the C layouts.
No device work is issued."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "vcsbeam.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", src)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}()]*(\([^)]*\))?[^;{}()]*)\)\s*;", src):
        name = m.group(1)
        if name in ("if", "while", "for", "return", "sizeof"):
            continue
        pre = src[max(0, m.start() - 40):m.start()]
        if "typedef" in pre.split(";")[-1]:
            continue
        names.add(name)
    return sorted(names)


def test_header_symbols_exported(pkg):
    lib = pkg.load()
    names = _declared_functions()
    assert len(names) > 120
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", ["vcsbeam_context", "forward_pfb", "MetafitsMetadata",
                                  "Rfinput", "gpu_ipfb_arrays", "vdifinfo", "vm_psrfits"])
def test_struct_layout_matches(pkg, name):
    c_side, py_side = pkg.struct_layout(name)
    assert c_side == py_side


def test_host_jones_ops_match_numpy(pkg):
    lib = pkg.load()
    rng = np.random.default_rng(0)
    A = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))).astype(np.complex128)
    B = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))).astype(np.complex128)
    out = np.zeros((2, 2), np.complex128)
    P = C.c_void_p
    for fn in ("mult2x2d", "inv2x2", "inv2x2S", "conj2x2", "calc_hermitian"):
        getattr(lib, fn).restype = None
    lib.mult2x2d(P(A.ctypes.data), P(B.ctypes.data), P(out.ctypes.data))
    np.testing.assert_allclose(out, A @ B, atol=1e-13)
    lib.inv2x2S(P(A.ctypes.data), P(out.ctypes.data))
    np.testing.assert_allclose(out, np.linalg.inv(A), atol=1e-12)
    lib.calc_hermitian(P(A.ctypes.data), P(out.ctypes.data))
    np.testing.assert_allclose(out, A.conj().T)
    lib.norm2x2.restype = C.c_double
    F = lib.norm2x2(P(A.ctypes.data), P(out.ctypes.data))
    assert abs(F - np.linalg.norm(A)) < 1e-12
    np.testing.assert_allclose(out, A / F, atol=1e-14)


def test_host_vdif_matches_oracle(pkg, orc):
    lib = pkg.load()
    x = np.linspace(-140, 140, 4001).astype(np.float32)
    q = np.zeros(x.size, np.int8)
    xx = x.copy()
    lib.float2int8_trunc(xx.ctypes.data, x.size, -126.0, 127.0, q.ctypes.data)
    lib.to_offset_binary(q.ctypes.data, x.size)
    np.testing.assert_array_equal(q.view(np.uint8), orc.vdif_requant(x))


@pytest.mark.gpu
def test_read_buffer_margin_and_lock(pkg, gpu_lib):
    """vmReadNextSecondFrom copies the previous second's last block into the
    margin, refuses a second read while locked and reports the end of the
    data (src/buffer.c:106-124, src/metadata.c:941-950).  Host logic over a
    pinned host buffer, which needs a HIP device to allocate."""
    lib = gpu_lib
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=2, blocks_per_second=2, nseconds=3, nflagged=0)
    vm = lib.vmInit(False)
    lib.vmBindObsMetadata(vm, C.byref(obs.md), C.byref(obs.vmd), 0, obs.md.obs_id, 2)
    s0, s1 = obs.second(0), obs.second(1)
    assert lib.vmReadNextSecondFrom(vm, s0.ctypes.data, s0.nbytes) == pkg.VM_SUCCESS
    assert lib.vmReadNextSecondFrom(vm, s1.ctypes.data, s1.nbytes) == pkg.VM_READ_BUFFER_LOCKED
    hb = vm.contents.v.contents
    hb.locked = False
    assert lib.vmReadNextSecondFrom(vm, s1.ctypes.data, s1.nbytes) == pkg.VM_SUCCESS
    buf = np.ctypeslib.as_array(C.cast(hb.buffer, C.POINTER(C.c_int8)), shape=(hb.buffer_size,))
    np.testing.assert_array_equal(buf[:hb.copy_size], s0.reshape(-1)[-hb.copy_size:])
    hb.locked = False
    assert lib.vmReadNextSecondFrom(vm, s1.ctypes.data, s1.nbytes) == pkg.VM_END_OF_DATA
    lib.destroy_vcsbeam_context(vm)


def _vdif_context(pkg, lib, obs, bg):
    """A context bound by hand to obs's metadata (pure host state): the
    fields vmPopulateVDIFHeader reads."""
    vm = lib.vmInit(False)
    v = vm.contents
    v.obs_metadata = C.pointer(obs.md)
    idx = (C.c_int * 1)(0)
    v.coarse_chan_idxs_to_process = C.cast(idx, C.c_void_p)
    v.npointing = len(bg)
    v.fine_sample_rate = 10000
    v.nfine_chan = 128
    return vm, idx


def test_populate_vdif_header(pkg, orc):
    """vmPopulateVDIFHeader (src/beam_vdif.c:150-246): the per-pointing
    frame description, file name and the context's frame-0 header."""
    from vcsbeam_amd.workloads import Observation
    lib = pkg.load()
    obs = Observation(ntiles=4, blocks_per_second=1, rec_chan=109)
    bg = (pkg.BeamGeom * 2)()
    # RA 05:34:31.94, Dec +22:00:52.2 and RA 0:00:00.00, Dec -0:30:00.00
    bg[0].mean_ra = np.deg2rad(15 * (5 + 34 / 60 + 31.94 / 3600))
    bg[0].mean_dec = np.deg2rad(22 + 0 / 60 + 52.2 / 3600)
    bg[1].mean_ra = 0.0
    bg[1].mean_dec = np.deg2rad(-0.5)
    bg[0].intmjd, bg[0].fracmjd = 59000.0, 3723.2 / 86400
    vm, idx = _vdif_context(pkg, lib, obs, bg)
    try:
        lib.vmPopulateVDIFHeader(vm, bg, 59000.25, 1.0)
        v = vm.contents
        vf = C.cast(v.vf, C.POINTER(pkg.Vdifinfo))
        f = vf[0]
        assert (f.bits, f.iscomplex, f.nchan, f.samples_per_frame) == (8, 1, 2, 128)
        assert (f.dataarraylength, f.frame_length, f.frame_rate) == (512, 544, 10000)
        assert (f.block_size, f.sizeof_beam, f.sizeof_buffer) == (5440000, 512, 5120000)
        assert f.sample_rate == 1280000 and f.BW == 1.28 and f.npol == 2
        assert f.fctr == 109 * 1.28 and f.MJD_start == 59000.25 and f.sec_offset == 1.0
        assert abs(f.MJD_epoch - (59000 + 3723.2 / 86400)) < 1e-12
        assert f.stationid == b"mw" and f.telescope == b"MWA" and f.obs_mode == b"PSR"
        assert f.source == b"unset" and f.scan_name == str(obs.md.obs_id).encode()
        assert f.ra_str == b"05:34:31.94" and f.dec_str == b"+22:00:52.20"
        assert vf[1].ra_str == b"00:00:00.00" and vf[1].dec_str == b"-00:30:00.00"
        assert f.date_obs == b"2020-05-31T12:00:00"  # MJD 59000.5
        proj = C.cast(obs.md.project_id, C.c_char_p).value
        assert f.basefilename == proj + b"_" + f.scan_name + b"_05:34:31.94_+22:00:52.20_ch109"
        want = orc.vdif_header(mjd=59000, sec_of_day=3723)
        np.testing.assert_array_equal(np.array(v.vhdr[:], np.uint32), want)
    finally:
        vm.contents.coarse_chan_idxs_to_process = None
        vm.contents.obs_metadata = None
        lib.destroy_vcsbeam_context(vm)


def test_vdif_write_data_files(pkg, tmp_path):
    """vdif_write_data appends one second of frames to <base>.vdif and
    rewrites <base>.hdr, a 4096-byte "KEY value" header."""
    lib = pkg.load()
    f = pkg.Vdifinfo()
    f.frame_length, f.frame_rate, f.block_size = 40, 3, 120
    f.bits, f.iscomplex, f.npol, f.sample_rate = 8, 1, 2, 1280000
    f.telescope, f.obs_mode, f.source = b"MWA", b"PSR", b"unset"
    f.ra_str, f.dec_str = b"05:34:31.94", b"+22:00:52.20"
    f.MJD_start, f.MJD_epoch, f.sec_offset, f.fctr, f.BW = 59000.25, 59000.5, 2.0, 139.52, 1.28
    f.basefilename = str(tmp_path / "beam").encode()
    a = np.arange(120, dtype=np.uint8)
    lib.vdif_write_data(C.byref(f), a.ctypes.data)
    lib.vdif_write_data(C.byref(f), (a[::-1].copy()).ctypes.data)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "beam.vdif", np.uint8),
                                  np.concatenate([a, a[::-1]]))
    raw = (tmp_path / "beam.hdr").read_bytes()
    assert len(raw) == 4096
    kv = dict(line.split(None, 1) for line in raw.rstrip(b"\0").decode().splitlines() if line)
    assert kv["INSTRUMENT"] == "VDIF" and kv["DATAFILE"] == str(tmp_path / "beam.vdif")
    assert kv["TELESCOPE"] == "MWA" and kv["MODE"] == "PSR" and kv["SOURCE"] == "unset"
    assert kv["RA"] == "05:34:31.94" and kv["DEC"] == "+22:00:52.20"
    assert float(kv["MJD_START"]) == 59000.25 and float(kv["MJD_EPOCH"]) == 59000.5
    assert float(kv["SEC_OFFSET"]) == 2.0 and float(kv["FREQ"]) == 139.52
    assert float(kv["BW"]) == 1.28 and abs(float(kv["TSAMP"]) - 1 / 1.28) < 1e-6
    assert (kv["NBIT"], kv["NDIM"], kv["NPOL"]) == ("8", "2", "2")


@pytest.mark.gpu
@pytest.mark.parametrize("nbytes", [16, 4096 * 16 + 48, 1 << 24])
def test_device_copy_kernel(pkg, gpu_lib, nbytes):
    """vmDeviceCopy (the plain copy kernel the bench's copy rate and the
    memory-counter calibration use) copies every 16-byte element, including
    sizes that do not fill its last workgroup, and nothing past the end."""
    rng = np.random.default_rng(11)
    src = rng.integers(0, 2 ** 32, size=nbytes // 4, dtype=np.uint32)
    a = pkg.DeviceArray.from_numpy(src)
    b = pkg.DeviceArray(nbytes + 64)
    gpu_lib.vmMemsetDevice(b.ptr, 0xAB, nbytes + 64)
    gpu_lib.vmDeviceCopy(b.ptr, a.ptr, nbytes)
    out = b.to_numpy(np.uint8, (nbytes + 64,))
    np.testing.assert_array_equal(out[:nbytes].view(np.uint32), src)
    assert (out[nbytes:] == 0xAB).all()
