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
                                  "Rfinput", "gpu_ipfb_arrays"])
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


def test_read_buffer_margin_and_lock(pkg):
    """vmReadNextSecondFrom copies the previous second's last block into the
    margin and refuses a second read while locked (src/buffer.c:106-124,
    src/metadata.c:941-950).  Pure host logic: only pinned host memory is
    allocated, so this runs without a GPU when the runtime allows pinned
    allocations; otherwise it is skipped."""
    lib = pkg.load()
    if lib.vmDeviceCount() < 1:
        pytest.skip("pinned host memory needs a HIP device")
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
