"""CPU pins of the oracle's primary-beam model and MLP reference.

The analytic dipole-array beam is this library's own definition (the reference
has no analytic beam), so it is pinned by what its definition fixes: the
short-dipole projections at zenith, the Dirichlet-kernel array factor of the
nominal 4x4 grid, a single dipole's phase (position indexing and delay sign)
and the ground-plane null at the horizon.  The FP64 MLP gradient is pinned by
central finite differences.
"""
import numpy as np

import oracle_lib as orc

C_LIGHT = 299792458.0


def test_nominal_beam_at_zenith_is_the_dipole_projection():
    az = np.linspace(0.0, 2 * np.pi, 13)
    J = orc.dipole_beam(az, np.zeros_like(az), np.full(az.size, 200e6), orc.nominal_cfg())
    ref = np.stack([np.stack([np.sin(az), np.cos(az)], -1),
                    np.stack([np.cos(az), -np.sin(az)], -1)], 1)
    assert np.abs(J - ref).max() < 1e-13


def test_nominal_array_factor_is_a_product_of_dirichlet_kernels():
    rng = np.random.default_rng(11)
    n = 200
    az, za = rng.uniform(0, 2 * np.pi, n), np.arccos(rng.uniform(0.2, 0.999, n))
    f = rng.uniform(170e6, 230e6, n)
    J = orc.dipole_beam(az, za, f, orc.nominal_cfg())
    k = 2 * np.pi * f / C_LIGHT
    u = k * 0.55 * np.sin(za) * np.sin(az)
    v = k * 0.55 * np.sin(za) * np.cos(az)
    af = np.sin(4 * u) / np.sin(u) * np.sin(4 * v) / np.sin(v)
    gp = 2 * np.sin(k * 0.278 * np.cos(za)) / (32 * np.sin(k * 0.278))
    Xt = gp * af * np.cos(za) * np.sin(az)
    Yp = -gp * af * np.sin(az)
    assert np.abs(J[:, 0, 0] - Xt).max() < 1e-12
    assert np.abs(J[:, 1, 1] - Yp).max() < 1e-12
    assert np.abs(J.imag).max() < 1e-12


def test_single_dipole_phase_fixes_position_and_delay_sign():
    d, tau, g = 6, 3.0e-10, 0.8 - 0.3j          # dipole 6: E = (6%4-1.5)*1.1, N = (6//4-1.5)*1.1
    cfg = np.zeros(96)
    cfg[2 * d], cfg[2 * d + 1] = g.real, g.imag  # X pol only
    cfg[64 + d] = tau
    az, za, f = np.array([0.7]), np.array([0.3]), np.array([190e6])
    J = orc.dipole_beam(az, za, f, cfg)
    k = 2 * np.pi * f[0] / C_LIGHT
    E, N = (d % 4 - 1.5) * 1.1, (d // 4 - 1.5) * 1.1
    l, m = np.sin(za[0]) * np.sin(az[0]), np.sin(za[0]) * np.cos(az[0])
    af = g * np.exp(1j * (k * (E * l + N * m) - 2 * np.pi * f[0] * tau))
    gp = 2 * np.sin(k * 0.278 * np.cos(za[0])) / (32 * np.sin(k * 0.278))
    assert abs(J[0, 0, 1] - gp * af * np.cos(az[0])) < 1e-14
    assert np.abs(J[0, 1]).max() == 0.0                     # Y pol has no dipoles


def test_ground_plane_null_at_horizon():
    J = orc.dipole_beam(np.array([0.3, 2.0]), np.full(2, np.pi / 2), np.full(2, 180e6),
                        orc.nominal_cfg())
    assert np.abs(J).max() < 1e-12


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    dims = [16, 12, 10, 8]
    W = [rng.normal(0, 0.4, (dims[i + 1], dims[i])) for i in range(3)]
    b = [rng.normal(0, 0.1, dims[i + 1]) for i in range(3)]
    X, T = rng.normal(size=(7, 16)), rng.normal(size=(7, 8))
    loss, gW, gb = orc.mlp_grad(X, T, W, b)
    Y = orc.mlp_forward(X, W, b)
    assert abs(loss - np.mean((Y - T) ** 2)) < 1e-12
    h = 1e-6
    for L in range(3):
        for (o, i) in [(0, 0), (dims[L + 1] - 1, dims[L] - 1), (3, 5)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[L][o, i] += h
            Wm[L][o, i] -= h
            fd = (np.mean((orc.mlp_forward(X, Wp, b) - T) ** 2) -
                  np.mean((orc.mlp_forward(X, Wm, b) - T) ** 2)) / (2 * h)
            assert abs(fd - gW[L][o, i]) < 1e-7 * max(1.0, abs(fd))
        bp = [x.copy() for x in b]
        bm = [x.copy() for x in b]
        bp[L][1] += h
        bm[L][1] -= h
        fd = (np.mean((orc.mlp_forward(X, W, bp) - T) ** 2) -
              np.mean((orc.mlp_forward(X, W, bm) - T) ** 2)) / (2 * h)
        assert abs(fd - gb[L][1]) < 1e-7 * max(1.0, abs(fd))
