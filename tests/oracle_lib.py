"""numpy front end of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used by the tests, the smoke check and the CPU-baseline
leg of the benchmark only.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        _lib = C.CDLL(LIB)
        _lib.orc_array_factor.restype = C.c_double
        _lib.orc_rel_err.restype = C.c_double
        _lib.orc_demote.restype = C.c_uint8
        _lib.orc_vdif_epoch_mjd.restype = C.c_long
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def wola(htr, ninput, nspectra, h, K=128):
    """htr int8 (MWAX layout, margin first); h int32 K*P taps -> int32 [m][i][n][2]."""
    P = len(h) // K
    h = np.ascontiguousarray(h, dtype=np.int32)
    b = np.empty((nspectra, ninput, K, 2), dtype=np.int32)
    lib().orc_wola(_p(htr), ninput, nspectra, K, P, _p(h), _p(b))
    return b


def fpga_round(b):
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.empty(b.shape, dtype=np.float32)
    lib().orc_fpga_round(_p(b), C.c_size_t(b.size), _p(out))
    return out


def int2float(b):
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.empty(b.shape, dtype=np.float32)
    lib().orc_int2float(_p(b), C.c_size_t(b.size), _p(out))
    return out


def dft(x, fast=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    (lib().orc_fft if fast else lib().orc_dft)(_p(x), _p(out), len(x))
    return out


def pfb_forward(htr, ninput, nspectra, h, out_idx, fpga=True, imag_first=False, use_fft=True,
                K=128):
    """complex128 [m][k'][vcs(i)] (reference layout, Y/K, swapped unless imag_first)."""
    P = len(h) // K
    h = np.ascontiguousarray(h, dtype=np.int32)
    oi = np.ascontiguousarray(out_idx, dtype=np.int32)
    v = np.empty((nspectra, K, ninput), dtype=np.complex128)
    lib().orc_pfb_forward(_p(htr), ninput, nspectra, K, P, _p(h), _p(oi), int(fpga),
                          int(imag_first), int(use_fft), _p(v))
    return v


def pack_int4(v):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.empty(v.shape, dtype=np.uint8)
    lib().orc_pack_int4(_p(v), C.c_size_t(v.size), _p(out))
    return out


def demote(x):
    return int(lib().orc_demote(C.c_double(x)))


def incoh(v, nsample, nchan, ninput):
    out = np.empty((nsample, nchan), dtype=np.float32)
    lib().orc_incoh(_p(np.ascontiguousarray(v, dtype=np.uint8)), nsample, nchan, ninput, _p(out))
    return out


def calc_J(D, B):
    """D [a][c][2][2], B [p][a][2][2] complex128 -> J [p][a][c][2][2]."""
    D = np.ascontiguousarray(D, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    np_, na = B.shape[0], B.shape[1]
    nc = D.shape[1]
    J = np.empty((np_, na, nc, 2, 2), dtype=np.complex128)
    lib().orc_calc_J(_p(D), _p(B), np_, na, nc, _p(J))
    return J


def calc_phi(E, N, H, L, unit, freqs, Href=377.827):
    E, N, H, L = (np.ascontiguousarray(a, dtype=np.float64) for a in (E, N, H, L))
    unit = np.ascontiguousarray(unit, dtype=np.float64).reshape(-1, 3)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    phi = np.empty((len(unit), len(E), len(freqs)), dtype=np.complex128)
    lib().orc_calc_phi(_p(E), _p(N), _p(H), _p(L), len(E), _p(unit), len(unit), _p(freqs),
                       len(freqs), C.c_double(Href), _p(phi))
    return phi


def array_factor(E, N, H, L, unit1, unit2, freq, Href=377.827):
    E, N, H, L = (np.ascontiguousarray(a, dtype=np.float64) for a in (E, N, H, L))
    u1 = np.ascontiguousarray(unit1, dtype=np.float64)
    u2 = np.ascontiguousarray(unit2, dtype=np.float64)
    return lib().orc_array_factor(_p(E), _p(N), _p(H), _p(L), len(E), _p(u1), _p(u2),
                                  C.c_double(freq), C.c_double(Href))


def beamform(v, J, phi, polQ, polP, invw, nstokes=4, int4=False):
    """v: uint8 [s][c][i] (int4) or complex128 [s][c][i]; J [p][a][c][2][2];
    phi [p][a][c].  Returns (S [p][s][st][c] float64, e [p][s][c][2] complex128)."""
    ns, nc, ni = v.shape
    np_, na = J.shape[0], J.shape[1]
    v = np.ascontiguousarray(v, dtype=np.uint8 if int4 else np.complex128)
    J = np.ascontiguousarray(J, dtype=np.complex128)
    phi = np.ascontiguousarray(phi, dtype=np.complex128)
    polQ = np.ascontiguousarray(polQ, dtype=np.uint32)
    polP = np.ascontiguousarray(polP, dtype=np.uint32)
    S = np.empty((np_, ns, nstokes, nc), dtype=np.float64)
    e = np.empty((np_, ns, nc, 2), dtype=np.complex128)
    lib().orc_beamform(_p(v), 0 if int4 else 1, _p(J), _p(phi), _p(polQ), _p(polP), np_, na, nc,
                       ns, ni, C.c_double(invw), nstokes, _p(S), _p(e))
    return S, e


def renormalise(S, npointing, nstep, nstokes, nchan):
    S = np.ascontiguousarray(S, dtype=np.float32)
    off = np.empty((npointing, nstokes, nchan), np.float32)
    scl = np.empty((npointing, nstokes, nchan), np.float32)
    q = np.empty(S.shape, np.uint8)
    valid = np.empty(S.shape, np.uint8)
    lib().orc_renormalise(_p(S), npointing, nstep, nstokes, nchan, _p(off), _p(scl), _p(q),
                          _p(valid))
    return off, scl, q, valid.astype(bool)


def ipfb_gather(ring, file_no, nsamples, P, K=128, npol=2):
    ring = np.ascontiguousarray(ring, dtype=np.complex128)
    out = np.empty((nsamples + P, K, npol), dtype=np.complex128)
    lib().orc_ipfb_gather(_p(ring), file_no, nsamples, K, P, npol, _p(out))
    return out


def ipfb(inbuf, nsamples, h, K=128, npol=2):
    """inbuf complex [nsamples+P][K][npol]; h F doubles -> (out complex [n][pol], mask [n])."""
    P = len(h) // K
    inbuf = np.ascontiguousarray(inbuf, dtype=np.complex128)
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.empty((nsamples * K, npol, 2), dtype=np.float64)
    mask = np.empty(nsamples * K, dtype=np.uint8)
    lib().orc_ipfb(_p(inbuf), nsamples, K, P, npol, _p(h), _p(out), _p(mask))
    return out[..., 0] + 1j * out[..., 1], mask.astype(bool)


def vdif_requant(f):
    f = np.ascontiguousarray(f, dtype=np.float32)
    out = np.empty(f.shape, np.uint8)
    lib().orc_vdif_requant(_p(f), C.c_size_t(f.size), _p(out))
    return out


def vdif_epoch_mjd(e):
    return int(lib().orc_vdif_epoch_mjd(C.c_int(e)))


def vdif_header(payload_bytes=512, threadid=0, bits=8, nchan=2, iscomplex=1, station=b"mw",
                mjd=58849, sec_of_day=0):
    """The eight 32-bit words of a VDIF 1.0 header at frame 0 of that second."""
    w = np.zeros(8, np.uint32)
    lib().orc_vdif_header(payload_bytes, threadid, bits, nchan, iscomplex, C.c_char_p(station),
                          C.c_long(mjd), C.c_long(sec_of_day), _p(w))
    return w


def vdif_frames(f, nframes, payload_bytes, hdr, frame_rate):
    """Frames (uint8, nframes x (32 + payload_bytes)) and the header after them."""
    f = np.ascontiguousarray(f, dtype=np.float32)
    assert f.size >= nframes * payload_bytes
    h = np.array(hdr, np.uint32)
    out = np.empty(nframes * (32 + payload_bytes), np.uint8)
    lib().orc_vdif_frames(_p(f), C.c_size_t(nframes), payload_bytes, _p(h), frame_rate, _p(out))
    return out, h


# ------------------------------------------- CPU baseline (cpu_pipeline.c) --
def set_threads(n):
    lib().orc_set_threads(int(n))


def cpu_pfb(htr, ninput, nspectra, h, out_idx, fpga=True, imag_first=False, K=128):
    """Forward PFB stage by stage -> (complex64 [m][k'][vcs(i)], seconds per
    stage {pfb-wola, pfb-round, pfb-fft, pfb-pack})."""
    P = len(h) // K
    h = np.ascontiguousarray(h, dtype=np.int32)
    oi = np.ascontiguousarray(out_idx, dtype=np.int32)
    v = np.empty((nspectra, K, ninput), dtype=np.complex64)
    t = np.zeros(4)
    rc = lib().orc_cpu_pfb(_p(htr), ninput, nspectra, K, P, _p(h), _p(oi), int(fpga),
                           int(imag_first), _p(v), _p(t))
    if rc != 0:
        raise MemoryError("orc_cpu_pfb: out of memory")
    return v, dict(zip(("pfb-wola", "pfb-round", "pfb-fft", "pfb-pack"), t.tolist()))


def cpu_ipfb_table(h, K=128):
    P = len(h) // K
    h = np.ascontiguousarray(h, dtype=np.float64)
    re = np.empty(K * P * K, np.float32)
    im = np.empty(K * P * K, np.float32)
    lib().orc_cpu_ipfb_table(_p(h), K, P, _p(re), _p(im))
    return re, im


def cpu_ipfb(inbuf, n0, nout, table, P, K=128, npol=2):
    """FP32 direct-form inverse PFB of outputs [n0, n0 + nout): inbuf complex64
    [row][K][npol] (row = m + P) -> complex64 [nout][npol]."""
    inbuf = np.ascontiguousarray(inbuf, dtype=np.complex64)
    if n0 < P * K or (n0 + nout - 1) // K + P >= inbuf.shape[0]:
        raise ValueError("cpu_ipfb: outputs outside the rows given")
    out = np.empty((nout, npol), np.complex64)
    lib().orc_cpu_ipfb(_p(inbuf), C.c_long(n0), C.c_long(nout), K, P, npol, _p(table[0]),
                       _p(table[1]), _p(out))
    return out


def cpu_vdif(f):
    f = np.ascontiguousarray(f, dtype=np.float32)
    out = np.empty(f.shape, np.uint8)
    lib().orc_cpu_vdif(_p(f), C.c_size_t(f.size), _p(out))
    return out


def cpu_delay(E, N, H, L, unit, freqs, Href=377.827):
    E, N, H, L = (np.ascontiguousarray(x, dtype=np.float64) for x in (E, N, H, L))
    unit = np.ascontiguousarray(unit, dtype=np.float64).reshape(-1, 3)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    phi = np.empty((unit.shape[0], len(E), len(freqs)), np.complex128)
    lib().orc_cpu_delay(_p(E), _p(N), _p(H), _p(L), len(E), _p(unit), unit.shape[0],
                        _p(freqs), len(freqs), C.c_double(Href), _p(phi))
    return phi


def rel_err(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    if np.iscomplexobj(x) or np.iscomplexobj(ref):
        x = np.stack([np.real(x), np.imag(x)], -1)
        ref = np.stack([np.real(ref), np.imag(ref)], -1)
    d = (x.astype(np.float64) - ref.astype(np.float64)).ravel()
    r = ref.astype(np.float64).ravel()
    den = np.sqrt(np.sum(r * r))
    return float(np.sqrt(np.sum(d * d)) / den) if den > 0 else float(np.sqrt(np.sum(d * d)))


# ------------------------------------------------ primary beam + surrogate --
MLP_DIMS = (16, 256, 256, 256, 8)


def dipole_beam(az, za, freq, cfg):
    """Analytic dipole-array beam; cfg = 64 gain doubles (re, im) + 32 delays.
    Returns J [n][2][2] complex."""
    az = np.ascontiguousarray(az, np.float64)
    za = np.ascontiguousarray(za, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    cfg = np.ascontiguousarray(cfg, np.float64)
    n = az.size
    J = np.empty((n, 8), np.float64)
    lib().orc_dipole_beam(_p(az), _p(za), _p(freq), C.c_size_t(n), _p(cfg[:64]), _p(cfg[64:]),
                          _p(J))
    return (J[:, 0::2] + 1j * J[:, 1::2]).reshape(n, 2, 2)


def nominal_cfg():
    cfg = np.zeros(96)
    cfg[0:64:2] = 1.0
    return cfg


def mlp_features(az, za, freq):
    az = np.ascontiguousarray(az, np.float64)
    za = np.ascontiguousarray(za, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    X = np.empty((az.size, 16), np.float64)
    lib().orc_mlp_features(_p(az), _p(za), _p(freq), C.c_size_t(az.size), _p(X))
    return X


def _mlp_pack(W, b):
    return (np.ascontiguousarray(np.concatenate([w.ravel() for w in W]), np.float64),
            np.ascontiguousarray(np.concatenate([x.ravel() for x in b]), np.float64))


def _mlp_unpack(flatW, flatb, dims):
    W, b, wo, bo = [], [], 0, 0
    for i in range(len(dims) - 1):
        n = dims[i] * dims[i + 1]
        W.append(flatW[wo:wo + n].reshape(dims[i + 1], dims[i]))
        b.append(flatb[bo:bo + dims[i + 1]])
        wo += n
        bo += dims[i + 1]
    return W, b


def mlp_forward(X, W, b):
    dims = np.array([W[0].shape[1]] + [w.shape[0] for w in W], np.int32)
    fw, fb = _mlp_pack(W, b)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.empty((X.shape[0], dims[-1]), np.float64)
    lib().orc_mlp_forward(_p(X), X.shape[0], len(W), _p(dims), _p(fw), _p(fb), _p(Y))
    return Y


def mlp_grad(X, T, W, b):
    """Mean-squared loss and its gradients (lists like W, b), FP64."""
    lib().orc_mlp_grad.restype = C.c_double
    dims = np.array([W[0].shape[1]] + [w.shape[0] for w in W], np.int32)
    fw, fb = _mlp_pack(W, b)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    dW = np.empty_like(fw)
    db = np.empty_like(fb)
    loss = lib().orc_mlp_grad(_p(X), _p(T), X.shape[0], len(W), _p(dims), _p(fw), _p(fb),
                              _p(dW), _p(db))
    gW, gb = _mlp_unpack(dW, db, list(dims))
    return loss, gW, gb


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)
