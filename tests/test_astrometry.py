"""Pointing geometry (src/geometry.c:226-420): sidereal time, nutation,
precession, aberration and horizon coordinates against the worked examples
of J. Meeus, "Astronomical Algorithms" (2nd ed., 1998), whose IAU models PAL
implements for the reference; then calc_beam_geom's invariants and the RA/Dec
string round trips.  Host code only: no GPU."""
import ctypes as C
import math

import numpy as np
import pytest

AS = math.pi / 180 / 3600


def _hms(h, m, s):
    return (h + m / 60 + s / 3600) * math.pi / 12


def _dms(d, m, s, sign=1):
    return sign * (d + m / 60 + s / 3600) * math.pi / 180


def _angdiff(a, b):
    return abs((a - b + math.pi) % (2 * math.pi) - math.pi)


@pytest.mark.parametrize("mjd, want", [
    (46895.0, _hms(13, 10, 46.3668)),              # Example 12.a, 1987 Apr 10 0h UT
    (46895.0 + 19.35 / 24, _hms(8, 34, 57.0896)),  # Example 12.b, 19h21m00s UT
])
def test_gmst_meeus(pkg, mjd, want):
    lib = pkg.load()
    assert _angdiff(lib.vmGMST(mjd), want) < 0.0002 / 3600 * math.pi / 12  # 0.2 ms of time


def test_lmst_is_gmst_plus_longitude(pkg):
    lib = pkg.load()
    lst = C.c_double()
    lib.mjd2lst(59000.3, C.byref(lst))
    assert _angdiff(lst.value, lib.vmGMST(59000.3) + 2.0362898668561042) < 1e-14


def test_nutation_meeus(pkg):
    """Example 22.a, 1987 Apr 10 0h TD: dpsi = -3.788", deps = +9.443",
    eps0 = 23 deg 26' 27.407"."""
    lib = pkg.load()
    dp, de, e0 = C.c_double(), C.c_double(), C.c_double()
    lib.vmNutation(46895.0, C.byref(dp), C.byref(de), C.byref(e0))
    assert abs(dp.value / AS - (-3.788)) < 0.0015
    assert abs(de.value / AS - 9.443) < 0.0015
    assert abs(e0.value - _dms(23, 26, 27.407)) < 0.002 * AS


def _apparent(lib, ra, dec, mjd, stages):
    a, d = C.c_double(), C.c_double()
    lib.vmMeanToApparent(ra, dec, mjd, stages, C.byref(a), C.byref(d))
    return a.value, d.value


# theta Persei: J2000 place carried by its proper motion to 2028 Nov 13.19 TD
# (Example 21.b), JDE 2462088.69 = MJD 62088.19
TP_RA, TP_DEC, TP_MJD = _hms(2, 44, 12.975), _dms(49, 13, 39.90), 62088.19


def test_precession_meeus(pkg):
    """Example 21.b: mean place of date 2h46m11.331s, +49 deg 20' 54.54"."""
    a, d = _apparent(pkg.load(), TP_RA, TP_DEC, TP_MJD, 1)
    assert _angdiff(a, _hms(2, 46, 11.331)) * math.cos(d) < 0.03 * AS
    assert abs(d - _dms(49, 20, 54.54)) < 0.03 * AS


def test_apparent_place_meeus(pkg):
    """Example 23.a: apparent place 2h46m14.390s, +49 deg 21' 07.45"
    (precession, nutation and annual aberration)."""
    a, d = _apparent(pkg.load(), TP_RA, TP_DEC, TP_MJD, 7)
    assert _angdiff(a, _hms(2, 46, 14.390)) * math.cos(d) < 0.05 * AS
    assert abs(d - _dms(49, 21, 7.45)) < 0.05 * AS


def test_aberration_size_and_direction(pkg):
    """The Earth's velocity lies in the ecliptic, so the ecliptic pole is
    always 90 degrees from the apex and moves by kappa |v| / v_mean, between
    kappa (1 - e) ~ 20.15" and kappa (1 + e) ~ 20.84", every month of a year."""
    lib = pkg.load()
    ra, dec = _hms(18, 0, 0), _dms(66, 33, 38.55)  # J2000 north ecliptic pole
    shifts = []
    for k in range(12):
        mjd = 51544.5 + k * 365.25 / 12
        a0, d0 = _apparent(lib, ra, dec, mjd, 3)
        a1, d1 = _apparent(lib, ra, dec, mjd, 7)
        c = math.sin(d0) * math.sin(d1) + math.cos(d0) * math.cos(d1) * math.cos(a1 - a0)
        shifts.append(math.acos(min(1.0, c)) / AS)
    assert 20.1 < min(shifts) and max(shifts) < 20.9


def test_hadec_to_azel_meeus(pkg):
    """Example 13.b: Venus from Washington (phi = +38 deg 55' 17"), hour
    angle 64.352133 deg, Dec -6 deg 43' 11.61": azimuth 68.0337 deg from
    South (248.0337 deg from North through East), elevation 15.1249 deg."""
    lib = pkg.load()
    az, el = C.c_double(), C.c_double()
    lib.vmHaDecToAzEl(math.radians(64.352133), _dms(6, 43, 11.61, -1), _dms(38, 55, 17),
                      C.byref(az), C.byref(el))
    assert abs(math.degrees(az.value) - 248.0337) < 5e-4
    assert abs(math.degrees(el.value) - 15.1249) < 5e-4


@pytest.mark.parametrize("ha, dec, phi, az, el", [
    (0.0, -0.4660608448386394, -0.4660608448386394, None, 90.0),  # zenith
    (0.0, 0.3, -0.4, 0.0, 90.0 - math.degrees(0.7)),                # meridian, north of zenith
    (0.0, -1.0, -0.4, 180.0, 90.0 - math.degrees(0.6)),             # meridian, south
    (math.pi / 2, 0.0, 0.0, 270.0, 0.0),                            # equator, 6h west: setting due west
])
def test_hadec_to_azel_closed_forms(pkg, ha, dec, phi, az, el):
    lib = pkg.load()
    a, e = C.c_double(), C.c_double()
    lib.vmHaDecToAzEl(ha, dec, phi, C.byref(a), C.byref(e))
    assert abs(math.degrees(e.value) - el) < 1e-9
    if az is not None:
        assert _angdiff(a.value, math.radians(az)) < 1e-9


def test_calc_beam_geom(pkg):
    """The look direction's unit vector, the split MJD, and a source put on
    the meridian at the array's latitude (apparent place of date) is at the
    zenith."""
    lib = pkg.load()
    bg = pkg.BeamGeom()
    lib.calc_beam_geom(5.575, -26.7, 59000.3, C.byref(bg))
    assert bg.intmjd == 59000.0 and abs(bg.fracmjd - 0.3) < 1e-9
    assert bg.mean_ra == 5.575 * math.pi / 12 and bg.mean_dec == math.radians(-26.7)
    n = np.array([bg.unit_N, bg.unit_E, bg.unit_H])
    assert abs(np.linalg.norm(n) - 1) < 1e-14
    assert abs(bg.unit_H - math.sin(bg.el)) < 1e-15
    # put a source at the zenith: apparent RA = LMST, apparent Dec = latitude,
    # then find its J2000 place by fixed-point iteration on the forward model
    mjd, lat = 59000.3, -0.4660608448386394
    lst = C.c_double()
    lib.mjd2lst(mjd, C.byref(lst))
    ra, dec = lst.value, lat
    for _ in range(6):
        a, d = _apparent(lib, ra, dec, mjd, 7)
        ra += lst.value - a
        dec += lat - d
    lib.calc_beam_geom(ra * 12 / math.pi, math.degrees(dec), mjd, C.byref(bg))
    assert abs(math.degrees(bg.el) - 90.0) < 1e-6


@pytest.mark.parametrize("text, hours", [("05:34:31.94", 5 + 34 / 60 + 31.94 / 3600),
                                         ("00:00:00", 0.0), ("23:59:59.99", 24 - 0.01 / 3600)])
def test_parse_ra(pkg, text, hours):
    assert abs(pkg.load().parse_ra(text.encode()) - hours) < 1e-12


@pytest.mark.parametrize("text, degs", [("+22:00:52.20", 22 + 52.2 / 3600),
                                        ("-00:30:00", -0.5), ("-26:42:11.95", -(26 + 42 / 60 + 11.95 / 3600)),
                                        ("45:00:00", 45.0)])
def test_parse_dec_and_round_trip(pkg, text, degs):
    lib = pkg.load()
    v = lib.parse_dec(text.encode())
    assert abs(v - degs) < 1e-12
    out = C.create_string_buffer(32)
    lib.dec2hms(out, v, 1)
    want = text if text[0] in "+-" else "+" + text
    assert out.value.decode() == (want if "." in want else want + ".00")


def test_parse_pointing_file(pkg, tmp_path):
    """vmParsePointingFile (src/metadata.c:1656-1705): RA/Dec pairs set the
    pointing count and the decimal RAs (hours) and Decs (degrees)."""
    from vcsbeam_amd.workloads import Observation
    lib = pkg.load()
    f = tmp_path / "pointings.txt"
    f.write_text("05:34:31.94 +22:00:52.20\n  00:00:00 -00:30:00\n12:00:00.0  -45:00:00.0\n")
    obs = Observation(ntiles=4, blocks_per_second=1)
    vm = lib.vmInit(False)
    v = vm.contents
    v.obs_metadata = C.pointer(obs.md)
    try:
        lib.vmParsePointingFile(vm, str(f).encode())
        assert v.npointing == 3
        ras = np.ctypeslib.as_array(C.cast(v.ras_hours, C.POINTER(C.c_double)), (3,))
        decs = np.ctypeslib.as_array(C.cast(v.decs_degs, C.POINTER(C.c_double)), (3,))
        np.testing.assert_allclose(ras, [5 + 34 / 60 + 31.94 / 3600, 0.0, 12.0], rtol=0, atol=1e-12)
        np.testing.assert_allclose(decs, [22 + 52.2 / 3600, -0.5, -45.0], rtol=0, atol=1e-12)
    finally:
        v.obs_metadata = None
        lib.destroy_vcsbeam_context(vm)


@pytest.mark.parametrize("ha, dec", [(0.7, -0.9), (-1.2, 0.3), (2.5, -0.2), (0.0, -1.2), (0.0, 0.2)])
def test_parallactic_angle(pkg, ha, dec):
    """parallactic_angle_correction (src/primary_beam.c:436-456) from the
    az/za of a known hour angle and declination: the angle between the
    directions to the celestial pole and to the zenith at the source, measured
    here from the spherical triangle's vectors (not the closed form), and
    positive west of the meridian."""
    lib = pkg.load()
    lat = -0.4660608448386394
    az, el = C.c_double(), C.c_double()
    lib.vmHaDecToAzEl(ha, dec, lat, C.byref(az), C.byref(el))
    P = np.zeros(4)
    lib.parallactic_angle_correction(P.ctypes.data, lat, az.value, math.pi / 2 - el.value)
    chi = math.atan2(-P[1], -P[0])
    assert abs(P[2] - math.sin(chi)) < 1e-15 and abs(P[3] - P[0]) < 1e-15
    # source s, pole n, zenith z as unit vectors in (HA=0 meridian) equatorial axes
    s = np.array([math.cos(dec) * math.cos(-ha), math.cos(dec) * math.sin(-ha), math.sin(dec)])
    n = np.array([0.0, 0.0, 1.0])
    z = np.array([math.cos(lat), 0.0, math.sin(lat)])
    tn = n - (n @ s) * s      # towards the pole, in the sky plane at s
    tz = z - (z @ s) * s      # towards the zenith
    # the angle from tn to tz, counted positive north-through-east on the
    # observer's sky (clockwise seen from outside the sphere), so positive
    # for a source west of the meridian
    ang = math.atan2(np.cross(tz, tn) @ s, tn @ tz)
    assert abs((chi - ang + math.pi) % (2 * math.pi) - math.pi) < 1e-9
    if abs(ha) > 1e-9:
        assert (chi > 0) == (ha > 0)
