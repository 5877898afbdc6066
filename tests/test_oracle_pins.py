"""Pins of the CPU oracle against golden data, closed forms and library routines.

Each test checks an oracle function against something other than its own
formula: the committed reference filter files, impulse responses plus
linearity, numpy's FFT, exhaustive enumeration, mathematical invariants of the
beamformer (coherent gain, array factor, single-tile cancellation, V sign) and
an independently factorised inverse PFB.
"""
import os
from decimal import ROUND_HALF_UP, Decimal

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K = 128


def test_finepfb_golden_properties():
    # doc/UsersGuide_Beamforming.md:151 and pfb_filter/README.md of the reference
    h = np.loadtxt(os.path.join(GOLD, "FINEPFB.dat"))
    assert h.shape == (1536,)
    assert h.sum() == 15106424
    assert h.max() == 117964 and int(np.argmax(h)) == 767
    assert np.all(h[:9] == 0) and np.all(h[1531:] == 0)
    m = np.loadtxt(os.path.join(GOLD, "MIRROR.dat"))
    nz = h != 0
    assert np.allclose(m[nz] / h[nz], 1 / 117964.8, rtol=1e-9)
    lsq = np.loadtxt(os.path.join(GOLD, "LSQ12.dat"))
    assert lsq.shape == (1536,) and abs(lsq.sum() - 124.861) < 1e-2


def _impulse_buffer(ninput, nblocks, i, s, value=(1, 0)):
    buf = np.zeros((nblocks, ninput, 64000, 2), dtype=np.int8)
    buf[s // 64000, i, s % 64000] = value
    return buf


@pytest.mark.parametrize("seed", [0, 1])
def test_wola_impulse_response(orc, seed):
    """A unit impulse at sample (m+500-P)K + pK + n reproduces the reversed tap
    h[KP - Kp - n - 1] in every spectrum m it reaches (src/pfb.cpp:149-161)."""
    rng = np.random.default_rng(seed)
    P = 12
    h = rng.integers(-30000, 30000, K * P).astype(np.int32)
    ninput, nspec = 3, 20
    i = int(rng.integers(0, ninput))
    n = int(rng.integers(0, K))
    q = 500 - P + int(rng.integers(P, nspec))  # block index within the buffer
    s = q * K + n
    buf = _impulse_buffer(ninput, 2, i, s, value=(1, -1))
    b = orc.wola(buf, ninput, nspec, h)
    expect = np.zeros_like(b)
    for m in range(nspec):
        p = q - (m + 500 - P)
        if 0 <= p < P:
            expect[m, i, n, 0] = h[K * P - K * p - n - 1]
            expect[m, i, n, 1] = -h[K * P - K * p - n - 1]
    np.testing.assert_array_equal(b, expect)


def test_wola_linearity_and_block_boundary(orc):
    rng = np.random.default_rng(5)
    h = np.loadtxt(os.path.join(GOLD, "FINEPFB.dat")).astype(np.int32)
    ninput, nspec = 2, 600  # crosses the 64000-sample block boundary
    a = rng.integers(-60, 60, (3, ninput, 64000, 2)).astype(np.int8)
    c = rng.integers(-60, 60, (3, ninput, 64000, 2)).astype(np.int8)
    ba, bc = orc.wola(a, ninput, nspec, h), orc.wola(c, ninput, nspec, h)
    bs = orc.wola((a + c).astype(np.int8), ninput, nspec, h)
    np.testing.assert_array_equal(bs, ba + bc)


def test_fpga_round_matches_documented_formula(orc):
    # src/pfb.cpp:185-199 documents floor(x/2^14 + 0.5) for x > 0, floor(x/2^14) otherwise
    X = np.concatenate([np.arange(-70000, 70000, 7, dtype=np.int64),
                        np.array([-2**31, 2**31 - 0x2001, 0x2000, -0x2000, 0x1fff, 0x2001,
                                  3 * 0x2000, -3 * 0x2000])])
    got = orc.fpga_round(X.astype(np.int32))
    want = [np.floor(x / 16384 + 0.5) if x > 0 else np.floor(x / 16384) for x in X.tolist()]
    np.testing.assert_array_equal(got, np.array(want, dtype=np.float32))


def test_int2float(orc):
    X = np.array([0, 1, -1, 16384, -16385, 2**24 + 1, -(2**24) - 3, 2**31 - 1, -2**31],
                 dtype=np.int32)
    got = orc.int2float(X)
    want = X.astype(np.float32) * np.float32(2.0 ** -14)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("fast", [False, True])
def test_dft_matches_numpy_fft(orc, fast):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    np.testing.assert_allclose(orc.dft(x, fast=fast), np.fft.fft(x), rtol=0, atol=1e-11)


def test_pfb_tone_lands_in_shifted_channel_with_swapped_parts(orc):
    """A complex tone on fine channel k0 lands in output channel (k0+64)%128
    (fftshift) with (im, re) order by default (src/pfb.cpp:315, 326-340)."""
    P = 12
    h = np.zeros(K * P, dtype=np.int32)
    h[K * P - 1 - np.arange(K)] = 1 << 14  # one tap: b[m][n] = 2^14 x[(m+500-P)K + n]... p=0
    ninput, nspec = 2, 4
    k0 = 5
    nblk = 2
    t = np.arange(nblk * 64000)
    tone = 40 * np.exp(2j * np.pi * k0 * t / K)
    buf = np.zeros((nblk, ninput, 64000, 2), dtype=np.int8)
    for i in range(ninput):
        buf[:, i, :, 0] = np.round(tone.real).reshape(nblk, 64000)
        buf[:, i, :, 1] = np.round(tone.imag).reshape(nblk, 64000)
    out_idx = np.array([1, 0], dtype=np.int32)
    v = orc.pfb_forward(buf, ninput, nspec, h, out_idx, fpga=True, use_fft=False)
    mag = np.abs(v[:, :, 0])
    assert np.all(np.argmax(mag, axis=1) == (k0 + 64) % K)
    # swapped: stored x = Im(Y/K), y = Re(Y/K)
    vi = orc.pfb_forward(buf, ninput, nspec, h, out_idx, fpga=True, imag_first=True, use_fft=False)
    np.testing.assert_allclose(v.real, vi.imag)
    np.testing.assert_allclose(v.imag, vi.real)
    # reorder: MWAX input 0 goes to VCS position 1
    buf2 = buf.copy()
    buf2[:, 1] = 0
    w = orc.pfb_forward(buf2, ninput, nspec, h, out_idx, use_fft=False)
    assert np.all(w[:, :, 0] == 0) and np.any(w[:, :, 1] != 0)


def test_demote_half_away_from_zero(orc):
    grid = np.concatenate([np.arange(-10, 10, 0.25), np.array([-8.5, -7.5, 6.5, 7.5, -0.5, 0.5])])
    for x in grid:
        r = int(Decimal(repr(float(x))).quantize(Decimal(1), rounding=ROUND_HALF_UP))
        r = min(max(r, -8), 7)
        assert orc.demote(float(x)) == (r & 0xF), x


def test_unpack_int4_exhaustive(orc):
    # every byte decodes to the two's-complement nibble pair that encodes it:
    # high nibble first (REAL_NIBBLE_TO_UINT8 = (X >> 4) & 0xf), low nibble
    # second, include/vcsbeam.h.in:997-1003
    lib = orc.lib()
    import ctypes as C
    re, im = C.c_double(), C.c_double()
    for byte in range(256):
        lib.orc_unpack_int4(C.c_uint8(byte), C.byref(re), C.byref(im))
        hi, lo = byte >> 4, byte & 15
        assert re.value == (hi - 16 if hi >= 8 else hi)
        assert im.value == (lo - 16 if lo >= 8 else lo)


def _nibble(n):
    """UINT8_TO_INT of a 4-bit field, stated from two's complement."""
    return n - 16 if n & 8 else n


def test_pack_int4_round_trip_pins_nibble_order(orc):
    """orc_pack_int4 packs the stored (first, second) component pair as
    PACK_NIBBLES(first, second): first in the high nibble
    (include/vcsbeam.h.in:453-456).  Unpacking with the reference's nibble
    macros must give back (clip(round(first)), clip(round(second))), with
    round half away from zero and clip to [-8, 7], over every grid pair that
    covers the whole nibble range, every +-0.5 tie and both clips."""
    grid = np.unique(np.concatenate([np.arange(-9.5, 9.51, 0.25), [-8.5, -8.49, 7.49, 7.5, 100.0, -100.0]]))

    def expect(x):
        r = int(Decimal(repr(float(x))).quantize(Decimal(1), rounding=ROUND_HALF_UP))
        return min(max(r, -8), 7)

    first, second = np.meshgrid(grid, grid, indexing="ij")
    stored = (first + 1j * second).ravel()
    packed = orc.pack_int4(stored)
    hi = np.array([_nibble(int(b) >> 4) for b in packed])
    lo = np.array([_nibble(int(b) & 0xF) for b in packed])
    np.testing.assert_array_equal(hi, [expect(x) for x in first.ravel()])
    np.testing.assert_array_equal(lo, [expect(x) for x in second.ravel()])
    # the byte is exactly (hi << 4) | lo: no carry from the low nibble
    assert all(int(b) == ((h & 0xF) << 4 | (l & 0xF)) for b, h, l in zip(packed, hi, lo))


def test_pfb_int4_output_keeps_imag_in_high_nibble(orc):
    """With PFB_IMAG_PART_FIRST unset the reference stores PACK_NIBBLES(im, re)
    (src/pfb.cpp:325-331): the high nibble holds the imaginary part of the
    fine channel.  A tone whose fine-channel value is known in closed form
    fixes which nibble is which: an input x[t] = A e^{2 pi i t k0 / K} through
    a one-tap-window filter gives, in channel k0, A h_sum (times a known
    phase) over K, so putting all the energy in the imaginary part leaves the
    low nibble (real part) zero and the high nibble non-zero."""
    K, P = 128, 12
    ninput, nspec = 2, 24
    h = np.zeros(K * P, np.int32)
    # only the p = 0 taps of the reversed filter, h[K P - n - 1] for n < K; a
    # tap of 2^14 cancels the int2float scale of 2^-14 exactly
    h[K * (P - 1): K * P] = 16384
    # MWAX layout [block][input][64000][re, im]; two blocks cover spectra
    # m + 500 - P .. for m < 24
    htr = np.zeros((2, ninput, 64000, 2), np.int8)
    htr[:, 0, :, 1] = 3                   # input 0: the constant 0 + 3i
    out_idx = np.arange(ninput, dtype=np.int32)
    v = orc.pfb_forward(htr, ninput, nspec, h, out_idx, fpga=False, imag_first=False,
                        use_fft=False, K=K)
    dc = v[:, K // 2, 0]                  # fftshift puts DC at channel 64
    # every spectrum's DC bin is K * 3i before the /K; stored as (im, re)
    np.testing.assert_allclose(dc.real, 3.0, atol=1e-12)
    np.testing.assert_allclose(dc.imag, 0.0, atol=1e-12)
    # every other channel is empty
    assert np.abs(np.delete(v[:, :, 0], K // 2, axis=1)).max() < 1e-12
    packed = orc.pack_int4(dc)
    assert all(_nibble(int(b) >> 4) == 3 and (int(b) & 0xF) == 0 for b in packed)
    # with PFB_IMAG_PART_FIRST the components come out as (re, im)
    w = orc.pfb_forward(htr, ninput, nspec, h, out_idx, fpga=False, imag_first=True,
                        use_fft=False, K=K)
    np.testing.assert_allclose(w[:, K // 2, 0], 3j, atol=1e-12)


def test_incoh_invariants(orc):
    rng = np.random.default_rng(1)
    ns, nc, ni = 3, 4, 16
    v = np.full((ns, nc, ni), 0x11, dtype=np.uint8)  # (1, 1) everywhere
    np.testing.assert_array_equal(orc.incoh(v, ns, nc, ni), 2 * ni)
    v = np.full((ns, nc, ni), 0x88, dtype=np.uint8)  # (-8, -8)
    np.testing.assert_array_equal(orc.incoh(v, ns, nc, ni), 128 * ni)
    w = rng.integers(0, 256, (ns, nc, ni), dtype=np.uint8)
    perm = rng.permutation(ni)
    np.testing.assert_array_equal(orc.incoh(w, ns, nc, ni), orc.incoh(w[:, :, perm].copy(), ns, nc, ni))


def test_calc_J_inverts_normalised_DB(orc):
    rng = np.random.default_rng(2)
    na, nc, np_ = 5, 3, 2
    D = rng.standard_normal((na, nc, 2, 2)) + 1j * rng.standard_normal((na, nc, 2, 2))
    B = rng.standard_normal((np_, na, 2, 2)) + 1j * rng.standard_normal((np_, na, 2, 2))
    D[1, 2] = 0
    J = orc.calc_J(D, B)
    for p in range(np_):
        for a in range(na):
            for c in range(nc):
                M = D[a, c] @ B[p, a]
                F = np.linalg.norm(M)
                if F == 0:
                    assert np.all(J[p, a, c] == 0)
                    continue
                np.testing.assert_allclose(J[p, a, c] @ (M / F), np.eye(2), atol=1e-12)


def test_calc_phi_against_extended_precision(orc):
    rng = np.random.default_rng(4)
    na = 16
    E, N = rng.uniform(-1500, 1500, na), rng.uniform(-1500, 1500, na)
    H = 377.827 + rng.uniform(-3, 3, na)
    L = rng.uniform(0, 300, na)
    az, za = 1.1, np.radians(20.0)
    el = np.pi / 2 - za
    unit = np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)])
    freqs = np.array([169.6e6 + 1e4 * c for c in range(0, 128, 17)])
    phi = orc.calc_phi(E, N, H, L, unit, freqs)
    ld = np.longdouble
    for a in range(na):
        w = ld(E[a]) * ld(unit[0]) + ld(N[a]) * ld(unit[1]) + (ld(H[a]) - ld(377.827)) * ld(unit[2])
        dt = (w - ld(L[a])) / ld(299792458.0)
        for c, f in enumerate(freqs):
            cyc = dt * ld(f)
            fr = cyc - np.round(cyc)
            ang = 2 * np.pi * float(fr)
            assert abs(phi[0, a, c] - np.exp(1j * ang)) < 5e-12


def _geom(rng, na):
    E, N = rng.uniform(-1500, 1500, na), rng.uniform(-1500, 1500, na)
    H = np.full(na, 377.827)
    L = rng.uniform(0, 300, na)
    return E, N, H, L


def _unit(az, za):
    el = np.pi / 2 - za
    return np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)])


def test_beamform_identity_jones_sums_voltages(orc):
    rng = np.random.default_rng(7)
    ns, nc, na = 3, 2, 6
    ni = 2 * na
    v = rng.standard_normal((ns, nc, ni)) + 1j * rng.standard_normal((ns, nc, ni))
    polQ = np.arange(0, ni, 2, dtype=np.uint32)
    polP = np.arange(1, ni, 2, dtype=np.uint32)
    J = np.zeros((1, na, nc, 2, 2), complex)
    J[..., 0, 0] = J[..., 1, 1] = 1
    phi = np.ones((1, na, nc), complex)
    S, e = orc.beamform(v, J, phi, polQ, polP, 1.0)
    np.testing.assert_allclose(e[0, :, :, 0], v[:, :, polQ].sum(-1), atol=1e-12)
    np.testing.assert_allclose(e[0, :, :, 1], v[:, :, polP].sum(-1), atol=1e-12)


def test_beamform_coherent_gain_single_tile_and_v_sign(orc):
    rng = np.random.default_rng(8)
    na, nc, ns = 8, 3, 2
    ni = 2 * na
    E, N, H, L = _geom(rng, na)
    freqs = 169.6e6 + 1e4 * np.arange(nc)
    u = _unit(0.7, np.radians(20))
    phi = orc.calc_phi(E, N, H, L, u, freqs)
    J = np.zeros((1, na, nc, 2, 2), complex)
    J[..., 0, 0] = J[..., 1, 1] = 1
    polQ = np.arange(0, ni, 2, dtype=np.uint32)
    polP = np.arange(1, ni, 2, dtype=np.uint32)
    x = rng.standard_normal((ns, nc)) + 1j * rng.standard_normal((ns, nc))
    # plane wave from the pointing direction, circular polarisation vp = i vq
    v = np.zeros((ns, nc, ni), complex)
    for a in range(na):
        v[:, :, 2 * a] = np.conj(phi[0, a])[None, :] * x
        v[:, :, 2 * a + 1] = 1j * np.conj(phi[0, a])[None, :] * x
    w = 1.0 / ni
    S, e = orc.beamform(v, J, phi, polQ, polP, w)
    g = na * (na - 1) * np.abs(x) ** 2
    np.testing.assert_allclose(S[0, :, 0, :], w * 2 * g, rtol=1e-10)   # I
    np.testing.assert_allclose(S[0, :, 1, :], 0, atol=1e-9 * g.max())  # Q
    np.testing.assert_allclose(S[0, :, 3, :], w * 2 * g, rtol=1e-10)   # V = -2w Im(bnXY) > 0
    # a signal present in one tile only is removed by the auto-correlation subtraction
    v1 = np.zeros_like(v)
    v1[:, :, 6:8] = rng.standard_normal((ns, nc, 2))
    S1, _ = orc.beamform(v1, J, phi, polQ, polP, w)
    np.testing.assert_allclose(S1, 0, atol=1e-12)


def test_beamform_pattern_equals_array_factor(orc):
    rng = np.random.default_rng(9)
    na, nc = 12, 1
    ni = 2 * na
    E, N, H, L = _geom(rng, na)
    f = 180e6
    u1, u2 = _unit(0.3, np.radians(10)), _unit(0.31, np.radians(10.2))
    phi1 = orc.calc_phi(E, N, H, L, u1, [f])
    phi2 = orc.calc_phi(E, N, H, L, u2, [f])
    v = np.zeros((1, 1, ni), complex)
    v[0, 0, 0::2] = np.conj(phi2[0, :, 0])
    J = np.zeros((1, na, nc, 2, 2), complex)
    J[..., 0, 0] = J[..., 1, 1] = 1
    polQ = np.arange(0, ni, 2, dtype=np.uint32)
    polP = np.arange(1, ni, 2, dtype=np.uint32)
    S, e = orc.beamform(v, J, phi1, polQ, polP, 1.0)
    af = orc.array_factor(E, N, H, L, u2, u1, f)
    assert abs(abs(e[0, 0, 0, 0]) ** 2 - na * na * af) < 1e-9 * na * na
    assert abs(S[0, 0, 0, 0] - (na * na * af - na)) < 1e-9 * na * na


def test_renormalise_reconstructs(orc):
    rng = np.random.default_rng(11)
    np_, ns, nst, nc = 2, 50, 4, 3
    S = rng.standard_normal((np_, ns, nst, nc)).astype(np.float32)
    off, scl, q, valid = orc.renormalise(S, np_, ns, nst, nc)
    # the maximum sample quantises to 256 (undefined in the reference): masked
    assert valid.sum() == S.size - np_ * nst * nc
    rec = off[:, None] + q.astype(np.float64) * scl[:, None]
    err = np.abs(rec.transpose(0, 1, 2, 3) - S)
    assert np.all(err[valid] <= 0.5 * np.broadcast_to(scl[:, None], S.shape)[valid] * 1.0001)


def _ipfb_factorised(inbuf, nsamples, h, K=128):
    """Independent restatement: per fine sample an inverse DFT with (-1)^r,
    then a P-tap polyphase FIR (module docstring of csrc/ipfb.hip)."""
    P = len(h) // K
    Z = np.fft.ifft(inbuf, axis=1) * K  # [s_in][r][pol]
    Z *= ((-1.0) ** np.arange(K))[None, :, None]
    out = np.zeros((nsamples * K, inbuf.shape[2]), complex)
    for j in range(nsamples):
        acc = 0
        for q in range(P):
            acc = acc + h[q * K:(q + 1) * K, None] * Z[j - q + P]
        out[j * K:(j + 1) * K] = acc / K
    return out


def test_ipfb_direct_form_equals_factorised(orc):
    rng = np.random.default_rng(12)
    P, ns = 12, 20
    h = rng.standard_normal(K * P)
    inbuf = rng.standard_normal((ns + P, K, 2)) + 1j * rng.standard_normal((ns + P, K, 2))
    out, mask = orc.ipfb(inbuf, ns, h)
    ref = _ipfb_factorised(inbuf, ns, h)
    n = np.arange(ns * K)
    np.testing.assert_array_equal(mask, ~((n < K * P) & (n % K != 0)))
    np.testing.assert_allclose(out[mask], ref[mask], atol=1e-10)


def test_ipfb_gather_ring(orc):
    rng = np.random.default_rng(13)
    ns, P = 10, 3
    ring = rng.standard_normal((2 * ns, K, 2)) + 0j
    g0 = orc.ipfb_gather(ring, 0, ns, P)
    assert np.all(g0[:P] == 0) and np.array_equal(g0[P:], ring[:ns])
    g1 = orc.ipfb_gather(ring, 1, ns, P)
    np.testing.assert_array_equal(g1, ring[ns - P:2 * ns])
    g2 = orc.ipfb_gather(ring, 2, ns, P)
    np.testing.assert_array_equal(g2[:P], ring[2 * ns - P:])
    np.testing.assert_array_equal(g2[P:], ring[:ns])


def test_vdif_requant_grid(orc):
    x = np.concatenate([np.arange(-140, 140, 0.25, dtype=np.float32),
                        np.array([-126.5, 127.5, -0.5, 0.5, 1.5, 2.5, np.inf, -np.inf],
                                 dtype=np.float32)])
    got = orc.vdif_requant(x)
    want = (np.rint(np.clip(x, -126, 127)).astype(np.int8).view(np.uint8) ^ 0x80)
    np.testing.assert_array_equal(got, want)


def test_vdif_epoch_mjd_against_calendar(orc):
    """VDIF reference epochs start on 1 January and 1 July of 2000 + e/2."""
    import datetime
    mjd0 = datetime.date(1858, 11, 17)
    for e in range(64):
        d = datetime.date(2000 + e // 2, 7 if e % 2 else 1, 1)
        assert orc.vdif_epoch_mjd(e) == (d - mjd0).days


def _vdif_fields(w):
    """Decode a VDIF 1.0 header from its 32 bytes as a reader does."""
    b = np.asarray(w, np.uint32).astype("<u4").tobytes()
    import struct
    w0, w1, w2, w3 = struct.unpack("<4I", b[:16])
    return dict(invalid=w0 >> 31, legacy=(w0 >> 30) & 1, seconds=w0 & (2 ** 30 - 1),
                epoch=(w1 >> 24) & 63, frame=w1 & (2 ** 24 - 1), version=w2 >> 29,
                log2nchan=(w2 >> 24) & 31, length8=w2 & (2 ** 24 - 1), complex=w3 >> 31,
                bitsm1=(w3 >> 26) & 31, thread=(w3 >> 16) & 1023,
                station=bytes([b[13], b[12]]), edv=b[16:32])


def test_vdif_header_known_answers(orc):
    """The MWA beam's header (8-bit complex, 2 channels, 512-byte payload,
    station "mw") at MJD 59000 (2020-05-31) + 3723 s: epoch 40 starts on MJD
    58849 (2020-01-01), 151 days earlier."""
    h = _vdif_fields(orc.vdif_header(mjd=59000, sec_of_day=3723))
    assert h == dict(invalid=0, legacy=0, seconds=151 * 86400 + 3723, epoch=40, frame=0,
                     version=0, log2nchan=1, length8=544 // 8, complex=1, bitsm1=7, thread=0,
                     station=b"mw", edv=bytes(16))
    # the first day of an epoch and the day before it
    assert _vdif_fields(orc.vdif_header(mjd=58849))["epoch"] == 40
    h = _vdif_fields(orc.vdif_header(mjd=58848, sec_of_day=86399))
    assert h["epoch"] == 39 and h["seconds"] == (58848 - 58665) * 86400 + 86399
    h = _vdif_fields(orc.vdif_header(payload_bytes=8000, threadid=5, bits=2, nchan=16,
                                     iscomplex=0, station=b"Ab"))
    assert (h["length8"], h["thread"], h["bitsm1"], h["log2nchan"], h["complex"],
            h["station"]) == (1004, 5, 1, 4, 0, b"Ab")


def test_vdif_frames_counter_wraps_and_payload(orc):
    rng = np.random.default_rng(21)
    pb, rate, n = 64, 7, 5
    x = (rng.standard_normal(n * pb) * 60).astype(np.float32)
    hdr = orc.vdif_header(payload_bytes=pb, mjd=59000, sec_of_day=10)
    hdr[1] |= rate - 2
    out, after = orc.vdif_frames(x, n, pb, hdr, rate)
    fr = out.reshape(n, 32 + pb)
    got = [_vdif_fields(fr[f, :32].view(np.uint32)) for f in range(n)]
    s0 = 151 * 86400 + 10
    assert [g["frame"] for g in got] == [5, 6, 0, 1, 2]
    assert [g["seconds"] for g in got] == [s0, s0, s0 + 1, s0 + 1, s0 + 1]
    for f in range(n):
        np.testing.assert_array_equal(fr[f, 8:32], hdr.view(np.uint8)[8:32])
        np.testing.assert_array_equal(fr[f, 32:], orc.vdif_requant(x[f * pb:(f + 1) * pb]))
    assert _vdif_fields(after)["frame"] == 3 and _vdif_fields(after)["seconds"] == s0 + 1
