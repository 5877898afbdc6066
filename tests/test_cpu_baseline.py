"""The CPU baseline's stages (oracle/cpu_pipeline.c) against the pinned
oracle functions they batch: the staged forward PFB equals orc_pfb_forward to
FP32 rounding, the FP32 direct-form inverse PFB over the reference's
filter-twiddle table equals the FP64 orc_ipfb to FP32 accuracy, and the
threaded VDIF and delay stages equal their single-call oracle functions
exactly.  These fix that the baseline times the same arithmetic the oracle
checks the GPU against."""
import numpy as np
import pytest

K = 128


@pytest.mark.parametrize("fpga,imag_first", [(True, False), (False, True)])
def test_cpu_pfb_stages_equal_oracle(orc, fpga, imag_first):
    rng = np.random.default_rng(21)
    ninput, nspec, P = 6, 120, 12
    h = rng.integers(-30000, 30000, K * P).astype(np.int32)
    htr = rng.integers(-127, 128, (3, ninput, 64000, 2)).astype(np.int8)
    out_idx = rng.permutation(ninput).astype(np.int32)
    ref = orc.pfb_forward(htr, ninput, nspec, h, out_idx, fpga=fpga, imag_first=imag_first)
    nthreads = orc.lib().orc_num_threads()
    orc.set_threads(3)
    try:
        v, t = orc.cpu_pfb(htr, ninput, nspec, h, out_idx, fpga=fpga, imag_first=imag_first)
    finally:
        orc.set_threads(nthreads)
    assert set(t) == {"pfb-wola", "pfb-round", "pfb-fft", "pfb-pack"}
    assert all(x >= 0 for x in t.values())
    np.testing.assert_array_equal(v, ref.astype(np.complex64))


def test_cpu_ipfb_equals_fp64_oracle(orc):
    rng = np.random.default_rng(22)
    P, ns, npol = 12, 24, 2
    h = rng.standard_normal(K * P)
    inbuf = rng.standard_normal((ns + P, K, npol)) + 1j * rng.standard_normal((ns + P, K, npol))
    ref, mask = orc.ipfb(inbuf, ns, h)
    n0 = P * K
    got = orc.cpu_ipfb(inbuf.astype(np.complex64), n0, ns * K - n0, orc.cpu_ipfb_table(h), P)
    assert mask[n0:].all()
    # FP32 products and sums over 1536 terms: relative error a few 1e-7
    assert orc.rel_err(got, ref[n0:]) < 2e-6
    # a dropped term (one tap row) would show far above that bar
    assert orc.rel_err(ref[n0 + K:], ref[n0:-K]) > 0.5
    with pytest.raises(ValueError):
        orc.cpu_ipfb(inbuf, n0 - 1, 10, orc.cpu_ipfb_table(h), P)


def test_cpu_vdif_and_delay_equal_single_calls(orc):
    rng = np.random.default_rng(23)
    f = (rng.standard_normal(300_001) * 90).astype(np.float32)
    np.testing.assert_array_equal(orc.cpu_vdif(f), orc.vdif_requant(f))
    E, N, H, L = (rng.uniform(-1500, 1500, 16) for _ in range(4))
    u = rng.standard_normal((5, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    freqs = 170e6 + 1e4 * np.arange(K)
    np.testing.assert_array_equal(orc.cpu_delay(E, N, H, L, u, freqs),
                                  orc.calc_phi(E, N, H, L, u, freqs))
