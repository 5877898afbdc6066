"""Primary beam per dipole configuration (src/primary_beam.c): the
configuration hash, the delays/amplitudes from the metadata, vmCalcB and
calc_normalised_beam_response.  Host code only: the context's beam arrays are
set up by hand so no device is needed.

Pins that do not restate the library's formulas:
- hash_dipole_config against the ordering the reference documents
  (src/primary_beam.c:318-345): none dead, then one dead in dipole order,
  then the pairs in lexicographic order (itertools.combinations), then all;
- vmCalcB against the FP64 oracle's analytic beam (oracle/oracle.c) times
  the parallactic rotation on the right, with the reference's cache keyed by
  configuration alone (a later tile of a cached configuration receives the
  first tile's beam even when its delays differ);
- calc_normalised_beam_response against J S J^H formed in numpy for the four
  pure-Stokes skies, and P J for the rotation.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

MWA_LAT = -0.4660608448386394


def _amps(dead):
    a = np.ones(16, np.float64)
    a[list(dead)] = 0.0
    return a


def test_hash_dipole_config_matches_documented_order(pkg):
    lib = pkg.load()
    order = [()] + [(d,) for d in range(16)] + list(itertools.combinations(range(16), 2))
    for idx, dead in enumerate(order):
        assert lib.hash_dipole_config(_amps(dead).ctypes.data) == idx, dead
    assert len(order) == 137
    assert lib.hash_dipole_config(_amps(range(16)).ctypes.data) == 137
    rng = np.random.default_rng(3)
    for n in range(3, 16):
        dead = rng.choice(16, n, replace=False)
        assert lib.hash_dipole_config(_amps(dead).ctypes.data) == -1  # MANY_DEAD_DIPOLES


def test_delays_and_amps_from_metadata(pkg):
    from vcsbeam_amd.workloads import Observation
    lib = pkg.load()
    obs = Observation(ntiles=4, blocks_per_second=2, nflagged=0, dipole_delays=np.arange(16),
                      dead_dipoles={1: [0, 7], 3: range(16)})
    delays, amps = C.POINTER(C.POINTER(C.c_uint32))(), C.POINTER(C.POINTER(C.c_double))()
    lib.create_delays_amps_from_metafits(C.byref(obs.md), C.byref(delays), C.byref(amps))
    try:
        for i in range(obs.ninput):
            a = i // 2
            d = np.ctypeslib.as_array(delays[i], (16,))
            np.testing.assert_array_equal(d, obs.dipole_delays[a])
            m = np.ctypeslib.as_array(amps[i], (16,))
            np.testing.assert_array_equal(m, (obs.dipole_delays[a] != 32).astype(float))
    finally:
        lib.free_delays_amps(C.byref(obs.md), delays, amps)


def _oracle_tile_beam(orc, az, za, f, delays, amps):
    """The analytic beam of one tile from the FP64 oracle: gains = amps on
    both polarisations, delays in 435 ps steps (dead dipoles carry none)."""
    cfg = np.zeros(96)
    for P in range(2):
        for d in range(16):
            cfg[2 * (P * 16 + d)] = amps[d]
            cfg[64 + P * 16 + d] = delays[d] * 435e-12 if amps[d] != 0 else 0.0
    return orc.dipole_beam(np.array([az]), np.array([za]), np.array([f]), cfg)[0]


def _parallactic(lib, az, za):
    P = np.empty(4)
    lib.parallactic_angle_correction(P.ctypes.data, C.c_double(MWA_LAT), C.c_double(az),
                                     C.c_double(za))
    return P.reshape(2, 2)


def test_vmcalcb_per_configuration(pkg, orc):
    """Five tiles: all live (0 and 4), one dead dipole (1), the same dead
    dipole with other delays (3: the cache hands it tile 1's beam), more than
    two dead (2: its own beam), all dead (5: zero)."""
    from vcsbeam_amd import BeamGeom
    from vcsbeam_amd.workloads import Observation
    lib = pkg.load()
    base = np.array([0, 1, 2, 3] * 4)
    obs = Observation(ntiles=6, blocks_per_second=2, nflagged=0, dipole_delays=base,
                      dead_dipoles={1: [5], 2: [1, 2, 9], 3: [5], 5: range(16)})
    obs.dipole_delays[3, 0] = 7                      # other delays, same configuration
    vm = lib.vmInit(False)
    v = vm.contents
    v.obs_metadata = C.pointer(obs.md)
    np_ = 3
    B = np.zeros((np_, obs.ntiles, 2, 2), np.complex128)
    pb = v.pb
    pb.B = B.ctypes.data
    pb.npointings, pb.nant, pb.npol = np_, obs.ntiles, 4
    pb.obs_metadata = C.cast(C.pointer(obs.md), C.c_void_p)
    pb.freq_hz = 170_000_000
    delays, amps = C.POINTER(C.POINTER(C.c_uint32))(), C.POINTER(C.POINTER(C.c_double))()
    lib.create_delays_amps_from_metafits(C.byref(obs.md), C.byref(delays), C.byref(amps))
    pb.delays = C.cast(delays, C.c_void_p)
    pb.amps = C.cast(amps, C.c_void_p)
    v.pb = pb
    bg = (BeamGeom * np_)()
    dirs = [(0.3, 0.2), (2.0, 0.5), (4.0, 0.05)]
    for p, (az, za) in enumerate(dirs):
        bg[p].az, bg[p].el = az, math.pi / 2 - za
    try:
        lib.vmCalcB(vm, bg)
    finally:
        lib.free_delays_amps(C.byref(obs.md), delays, amps)
        v.pb.delays = v.pb.amps = None
        v.pb.B = None
        v.obs_metadata = None
        lib.destroy_vcsbeam_context(vm)
    f = 170e6
    for p, (az, za) in enumerate(dirs):
        P = _parallactic(lib, az, za)
        beam = {}
        for a, src in [(0, 0), (1, 1), (2, 2), (3, 1), (4, 0)]:
            dl = obs.dipole_delays[src]
            am = (dl != 32).astype(float)
            beam[a] = _oracle_tile_beam(orc, az, za, f, dl, am) @ P
        for a in range(5):
            np.testing.assert_allclose(B[p, a], beam[a], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(B[p, 5], 0)
        # the configurations really differ (the test would not see a wrong
        # cache otherwise)
        assert np.abs(beam[0] - beam[1]).max() > 1e-3
        assert np.abs(beam[1] - beam[2]).max() > 1e-3
        # tile 3's own delays would give another beam: the cache is by
        # configuration only
        own = _oracle_tile_beam(orc, az, za, f, obs.dipole_delays[3],
                                (obs.dipole_delays[3] != 32).astype(float)) @ P
        assert np.abs(own - beam[3]).max() > 1e-6


@pytest.mark.parametrize("pa", [False, True])
def test_calc_normalised_beam_response(pkg, orc, pa):
    lib = pkg.load()
    az, za, f = 1.1, 0.4, 190e6
    delays = np.array([0, 2, 4, 6] * 4, np.uint32)
    amps = _amps([3])
    J = np.zeros((2, 2), np.complex128)
    iquv = np.zeros(4)
    lib.calc_normalised_beam_response(None, C.c_double(az), C.c_double(za), C.c_double(f),
                                      delays.ctypes.data, amps.ctypes.data, iquv.ctypes.data,
                                      J.ctypes.data, pa)
    Jr = _oracle_tile_beam(orc, az, za, f, delays, amps)
    if pa:
        Jr = _parallactic(lib, az, za) @ Jr
    np.testing.assert_allclose(J, Jr, rtol=0, atol=1e-13)
    skies = [np.diag([0.5, 0.5]), np.diag([0.5, -0.5]), np.array([[0, 0.5], [0.5, 0]]),
             np.array([[0, -0.5j], [0.5j, 0]])]
    C_ = [Jr @ s @ Jr.conj().T for s in skies]
    want = [0.5 * (C_[0][0, 0] + C_[0][1, 1]).real, 0.5 * (C_[1][0, 0] - C_[1][1, 1]).real,
            C_[2][0, 1].real, -C_[3][0, 1].imag]
    np.testing.assert_allclose(iquv, want, rtol=0, atol=1e-13)
    # the reference's unpolarised sky is diag(1/2, 1/2) and I takes half the
    # trace of J S J^H: a quarter of the squared Frobenius norm of J
    assert abs(iquv[0] - 0.25 * np.sum(np.abs(Jr) ** 2)) < 1e-13
