/*
 * astrometry.c -- pointing geometry of a beam: J2000 mean RA/Dec and a UTC
 * MJD to apparent place, local sidereal time, hour angle, azimuth and
 * elevation, and the look direction's unit vector (North, East, Up).
 *
 * Reference: MWA-vcsbeam src/geometry.c:226-286 (calc_beam_geom), 301-342
 * (dec2hms), 354-361 (mjd2lst), 370-420 (parse_ra, parse_dec).  The
 * reference calls PAL (palMap, palGmst, palDe2h, palRanorm, palDtf2r,
 * palDaf2r), which is not part of its tree.  The same quantities are
 * computed here from the published IAU models PAL implements, following
 * J. Meeus, "Astronomical Algorithms" (2nd ed., 1998):
 *   - Greenwich mean sidereal time, IAU 1982 (Meeus eq. 12.4), UT1 = UTC;
 *   - precession from J2000, IAU 1976 (Meeus eq. 21.3-21.4);
 *   - nutation, IAU 1980 series terms of 0.0003" and above (Meeus
 *     table 22.A), mean obliquity Meeus eq. 22.2;
 *   - annual aberration with the FK5 e-terms (Meeus eq. 23.2), Sun's true
 *     longitude Meeus ch. 25;
 *   - HA/Dec to azimuth (from North through East) and elevation.
 * Gravitational light deflection (PAL includes it; below 0.01" more than
 * 5 degrees from the Sun) is not applied.  The MJD is used for both the
 * sidereal time and the apparent place, as the reference does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vm_internal.h"

#define AS2R (M_PI / (180.0 * 3600.0))
#define DEG2R (M_PI / 180.0)

static double ranorm(double a)
{
    a = fmod(a, 2.0 * M_PI);
    return a < 0.0 ? a + 2.0 * M_PI : a;
}

double vmGMST(double mjd)
{
    /* IAU 1982 GMST; d = days from J2000.0 (JD 2451545.0 = MJD 51544.5) */
    const double d = mjd - 51544.5, T = d / 36525.0;
    const double fd = d - floor(d);
    double deg = 280.46061837 + 360.0 * fd + 0.98564736629 * d + 0.000387933 * T * T -
                 T * T * T / 38710000.0;
    return ranorm(deg * DEG2R);
}

void mjd2lst(double mjd, double *lst)
{
    *lst = ranorm(vmGMST(mjd) + MWA_LONGITUDE_RADIANS);
}

/* IAU 1980 nutation terms >= 0.0003" (Meeus table 22.A): multiples of
 * D, M, M', F, Omega; dpsi = (s0 + s1 T) sin(arg), deps = (c0 + c1 T) cos(arg),
 * in units of 0.0001". */
static const struct { signed char d, m, mp, f, om; float s0, s1, c0, c1; } NUT[] = {
    {0, 0, 0, 0, 1, -171996, -174.2f, 92025, 8.9f},
    {-2, 0, 0, 2, 2, -13187, -1.6f, 5736, -3.1f},
    {0, 0, 0, 2, 2, -2274, -0.2f, 977, -0.5f},
    {0, 0, 0, 0, 2, 2062, 0.2f, -895, 0.5f},
    {0, 1, 0, 0, 0, 1426, -3.4f, 54, -0.1f},
    {0, 0, 1, 0, 0, 712, 0.1f, -7, 0},
    {-2, 1, 0, 2, 2, -517, 1.2f, 224, -0.6f},
    {0, 0, 0, 2, 1, -386, -0.4f, 200, 0},
    {0, 0, 1, 2, 2, -301, 0, 129, -0.1f},
    {-2, -1, 0, 2, 2, 217, -0.5f, -95, 0.3f},
    {-2, 0, 1, 0, 0, -158, 0, 0, 0},
    {-2, 0, 0, 2, 1, 129, 0.1f, -70, 0},
    {0, 0, -1, 2, 2, 123, 0, -53, 0},
    {2, 0, 0, 0, 0, 63, 0, 0, 0},
    {0, 0, 1, 0, 1, 63, 0.1f, -33, 0},
    {2, 0, -1, 2, 2, -59, 0, 26, 0},
    {0, 0, -1, 0, 1, -58, -0.1f, 32, 0},
    {0, 0, 1, 2, 1, -51, 0, 27, 0},
    {-2, 0, 2, 0, 0, 48, 0, 0, 0},
    {0, 0, -2, 2, 1, 46, 0, -24, 0},
    {2, 0, 0, 2, 2, -38, 0, 16, 0},
    {0, 0, 2, 2, 2, -31, 0, 13, 0},
    {0, 0, 2, 0, 0, 29, 0, 0, 0},
    {-2, 0, 1, 2, 2, 29, 0, -12, 0},
    {0, 0, 0, 2, 0, 26, 0, 0, 0},
    {-2, 0, 0, 2, 0, -22, 0, 0, 0},
    {0, 0, -1, 2, 1, 21, 0, -10, 0},
    {0, 2, 0, 0, 0, 17, -0.1f, 0, 0},
    {2, 0, -1, 0, 1, 16, 0, -8, 0},
    {-2, 2, 0, 2, 2, -16, 0.1f, 7, 0},
    {0, 1, 0, 0, 1, -15, 0, 9, 0},
    {-2, 0, 1, 0, 1, -13, 0, 7, 0},
    {0, -1, 0, 0, 1, -12, 0, 6, 0},
    {0, 0, 2, -2, 0, 11, 0, 0, 0},
    {2, 0, -1, 2, 1, -10, 0, 5, 0},
    {2, 0, 1, 2, 2, -8, 0, 3, 0},
    {0, 1, 0, 2, 2, 7, 0, -3, 0},
    {-2, 1, 1, 0, 0, -7, 0, 0, 0},
    {0, -1, 0, 2, 2, -7, 0, 3, 0},
    {2, 0, 0, 2, 1, -7, 0, 3, 0},
    {2, 0, 1, 0, 0, 6, 0, 0, 0},
    {-2, 0, 2, 2, 2, 6, 0, -3, 0},
    {-2, 0, 1, 2, 1, 6, 0, -3, 0},
    {2, 0, -2, 0, 1, -6, 0, 3, 0},
    {2, 0, 0, 0, 1, -6, 0, 3, 0},
    {0, -1, 1, 0, 0, 5, 0, 0, 0},
    {-2, -1, 0, 2, 1, -5, 0, 3, 0},
    {-2, 0, 0, 0, 1, -5, 0, 3, 0},
    {0, 0, 2, 2, 1, -5, 0, 3, 0},
    {-2, 0, 2, 0, 1, 4, 0, 0, 0},
    {-2, 1, 0, 2, 1, 4, 0, 0, 0},
    {0, 0, 1, -2, 0, 4, 0, 0, 0},
    {-1, 0, 1, 0, 0, -4, 0, 0, 0},
    {-2, 1, 0, 0, 0, -4, 0, 0, 0},
    {1, 0, 0, 0, 0, -4, 0, 0, 0},
    {0, 0, 1, 2, 0, 3, 0, 0, 0},
    {0, 0, -2, 2, 2, -3, 0, 0, 0},
    {-1, -1, 1, 0, 0, -3, 0, 0, 0},
    {0, 1, 1, 0, 0, -3, 0, 0, 0},
    {0, -1, 1, 2, 2, -3, 0, 0, 0},
    {2, -1, -1, 2, 2, -3, 0, 0, 0},
    {0, 0, 3, 2, 2, -3, 0, 0, 0},
    {2, -1, 0, 2, 2, -3, 0, 0, 0},
};

void vmNutation(double mjd, double *dpsi, double *deps, double *eps0)
{
    /* T in Julian centuries from J2000.0; angles in radians */
    const double T = (mjd - 51544.5) / 36525.0, T2 = T * T, T3 = T2 * T;
    const double D = (297.85036 + 445267.111480 * T - 0.0019142 * T2 + T3 / 189474.0) * DEG2R;
    const double M = (357.52772 + 35999.050340 * T - 0.0001603 * T2 - T3 / 300000.0) * DEG2R;
    const double Mp = (134.96298 + 477198.867398 * T + 0.0086972 * T2 + T3 / 56250.0) * DEG2R;
    const double F = (93.27191 + 483202.017538 * T - 0.0036825 * T2 + T3 / 327270.0) * DEG2R;
    const double Om = (125.04452 - 1934.136261 * T + 0.0020708 * T2 + T3 / 450000.0) * DEG2R;
    double sp = 0.0, se = 0.0;
    for (size_t k = 0; k < sizeof(NUT) / sizeof(NUT[0]); k++) {
        const double a = NUT[k].d * D + NUT[k].m * M + NUT[k].mp * Mp + NUT[k].f * F + NUT[k].om * Om;
        sp += (NUT[k].s0 + NUT[k].s1 * T) * sin(a);
        se += (NUT[k].c0 + NUT[k].c1 * T) * cos(a);
    }
    *dpsi = sp * 1e-4 * AS2R;
    *deps = se * 1e-4 * AS2R;
    *eps0 = (84381.448 - 46.8150 * T - 0.00059 * T2 + 0.001813 * T3) * AS2R;
}

/* x <- R v for a 3x3 row-major R */
static void rot(const double R[9], double v[3])
{
    double o[3];
    for (int i = 0; i < 3; i++) o[i] = R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
    memcpy(v, o, sizeof(o));
}

/* frame rotations: R1 about x, R2 about y, R3 about z (positive angle
 * rotates the frame anticlockwise seen from the axis' tip) */
static void R1(double a, double R[9])
{
    const double c = cos(a), s = sin(a);
    const double M[9] = {1, 0, 0, 0, c, s, 0, -s, c};
    memcpy(R, M, sizeof(M));
}
static void R2(double a, double R[9])
{
    const double c = cos(a), s = sin(a);
    const double M[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
    memcpy(R, M, sizeof(M));
}
static void R3(double a, double R[9])
{
    const double c = cos(a), s = sin(a);
    const double M[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
    memcpy(R, M, sizeof(M));
}

void vmMeanToApparent(double ra, double dec, double mjd, int stages, double *ra_ap, double *dec_ap)
{
    /* stages: bit 0 precession, bit 1 nutation, bit 2 aberration (7 = all;
     * the partial forms exist for checking each model on its own) */
    const double T = (mjd - 51544.5) / 36525.0, T2 = T * T, T3 = T2 * T;
    double v[3] = {cos(dec) * cos(ra), cos(dec) * sin(ra), sin(dec)}, R[9];
    if (stages & 1) {
        const double zeta = (2306.2181 * T + 0.30188 * T2 + 0.017998 * T3) * AS2R;
        const double z = (2306.2181 * T + 1.09468 * T2 + 0.018203 * T3) * AS2R;
        const double th = (2004.3109 * T - 0.42665 * T2 - 0.041833 * T3) * AS2R;
        R3(-zeta, R); rot(R, v);
        R2(th, R); rot(R, v);
        R3(-z, R); rot(R, v);
    }
    double dpsi, deps, eps0;
    vmNutation(mjd, &dpsi, &deps, &eps0);
    const double eps = eps0 + deps;
    if (stages & 2) {
        R1(eps0, R); rot(R, v);   /* to the mean ecliptic of date */
        R3(-dpsi, R); rot(R, v);  /* longitude + dpsi */
        R1(-eps, R); rot(R, v);   /* back with the true obliquity */
    }
    if (stages & 4) {
        /* Earth's velocity / c in the ecliptic of date (kappa = 20.49552"),
         * from the Sun's true longitude, the orbit's eccentricity e and the
         * longitude of perihelion pi */
        const double L0 = 280.46646 + 36000.76983 * T + 0.0003032 * T2;
        const double Ms = (357.52911 + 35999.05029 * T - 0.0001537 * T2) * DEG2R;
        const double C = (1.914602 - 0.004817 * T - 0.000014 * T2) * sin(Ms) +
                         (0.019993 - 0.000101 * T) * sin(2 * Ms) + 0.000289 * sin(3 * Ms);
        const double sun = (L0 + C) * DEG2R;
        const double e = 0.016708634 - 0.000042037 * T - 0.0000001267 * T2;
        const double pi = (102.93735 + 1.71946 * T + 0.00046 * T2) * DEG2R;
        const double k = 20.49552 * AS2R;
        const double ex = k * (sin(sun) - e * sin(pi)), ey = -k * (cos(sun) - e * cos(pi));
        const double ce = cos((stages & 2) ? eps : eps0), se = sin((stages & 2) ? eps : eps0);
        const double w[3] = {ex, ey * ce, ey * se};
        double n = 0.0;
        for (int i = 0; i < 3; i++) {
            v[i] += w[i];
            n += v[i] * v[i];
        }
        n = sqrt(n);
        for (int i = 0; i < 3; i++) v[i] /= n;
    }
    *ra_ap = ranorm(atan2(v[1], v[0]));
    *dec_ap = asin(fmax(-1.0, fmin(1.0, v[2])));
}

void vmHaDecToAzEl(double ha, double dec, double phi, double *az, double *el)
{
    /* azimuth from North through East, elevation above the horizon */
    const double ch = cos(ha), sh = sin(ha), cd = cos(dec), sd = sin(dec);
    const double cp = cos(phi), sp = sin(phi);
    const double x = -ch * cd * sp + sd * cp; /* North */
    const double y = -sh * cd;                /* East */
    const double z = ch * cd * cp + sd * sp;  /* Up */
    const double r = hypot(x, y);
    *az = (r != 0.0) ? ranorm(atan2(y, x)) : 0.0;
    *el = atan2(z, r);
}

void parallactic_angle_correction(double *P, double lat, double az, double za)
{
    /* src/primary_beam.c:436-456: the rotation by the parallactic angle chi
     * of the direction (az from North through East, za) at latitude lat:
     * az/el back to hour angle and declination, then
     * chi = atan2(cos lat sin H, sin lat cos dec - cos lat sin dec cos H) */
    const double el = M_PI / 2 - za;
    const double sd = sin(el) * sin(lat) + cos(el) * cos(lat) * cos(az);
    const double dec = asin(fmax(-1.0, fmin(1.0, sd)));
    const double ha = atan2(-sin(az) * cos(el), sin(el) * cos(lat) - cos(el) * sin(lat) * cos(az));
    const double chi = atan2(cos(lat) * sin(ha), sin(lat) * cos(dec) - cos(lat) * sin(dec) * cos(ha));
    const double s = sin(chi), c = cos(chi);
    P[0] = -c;
    P[1] = -s;
    P[2] = s;
    P[3] = -c;
}

void calc_beam_geom(double ras_hours, double decs_degs, double mjd, beam_geom *bg)
{
    /* src/geometry.c:226-286 */
    const double intmjd = floor(mjd);
    double lmst, ra_ap, dec_ap, az, el;
    mjd2lst(mjd, &lmst);
    const double mean_ra = ras_hours * M_PI / 12.0, mean_dec = decs_degs * DEG2R;
    vmMeanToApparent(mean_ra, mean_dec, mjd, 7, &ra_ap, &dec_ap);
    const double ha = ranorm(lmst - ra_ap);
    vmHaDecToAzEl(ha, dec_ap, MWA_LATITUDE_RADIANS, &az, &el);
    bg->mean_ra = mean_ra;
    bg->mean_dec = mean_dec;
    bg->az = az;
    bg->el = el;
    bg->lmst = lmst;
    bg->fracmjd = mjd - intmjd;
    bg->intmjd = intmjd;
    bg->unit_N = cos(el) * cos(az);
    bg->unit_E = cos(el) * sin(az);
    bg->unit_H = sin(el);
}

void dec2hms(char *out, double in, int sflag)
{
    /* [sign]HH:MM:SS.SS of decimal hours or degrees; '+' for non-negative
     * values when sflag (src/geometry.c:301-342) */
    const int neg = in < 0.0;
    if (neg) in = -in;
    int h = (int)in;
    in = (in - h) * 60.0;
    int m = (int)in;
    double s = (in - m) * 60.0;
    if (m >= 60) {
        h += 1;
        m -= 60;
    }
    if (s >= 59.995) {
        m += 1;
        s = 0.0;
    }
    sprintf(out, "%s%2.2d:%2.2d:%05.2f", neg ? "-" : (sflag ? "+" : ""), h, m, s);
}

void vmParsePointingFile(vcsbeam_context *vm, const char *filename)
{
    /* src/metadata.c:1656-1705: whitespace-separated "RA Dec" pairs
     * (HH:MM:SS.S, DD:MM:SS.S) into vm->ras_hours / vm->decs_degs; the
     * pointing count is set from the file */
    FILE *f = fopen(filename, "r");
    if (!f) VM_FATAL("cannot open pointings file %s", filename);
    char word[64], ra[64], dec[64];
    int nwords = 0;
    while (fscanf(f, "%63s", word) == 1) nwords++;
    if (nwords % 2 != 0 || nwords == 0) VM_FATAL("cannot parse pointings file %s", filename);
    vmSetNumPointings(vm, (unsigned)(nwords / 2));
    free(vm->ras_hours);
    free(vm->decs_degs);
    vm->ras_hours = (double *)malloc(sizeof(double) * vm->npointing);
    vm->decs_degs = (double *)malloc(sizeof(double) * vm->npointing);
    if (!vm->ras_hours || !vm->decs_degs) VM_FATAL("out of host memory");
    rewind(f);
    for (unsigned p = 0; p < vm->npointing; p++) {
        if (fscanf(f, "%63s %63s", ra, dec) != 2) VM_FATAL("cannot parse pointings file %s", filename);
        vm->ras_hours[p] = parse_ra(ra);
        vm->decs_degs[p] = parse_dec(dec);
    }
    fclose(f);
}

double parse_ra(char *ra_hhmmss)
{
    /* "HH:MM:SS.S" -> decimal hours; fields out of range end the process */
    if (!ra_hhmmss) return 0.0;
    int h = 0, m = 0;
    double s = 0.0;
    sscanf(ra_hhmmss, "%d:%d:%lf", &h, &m, &s);
    if (h < 0 || h > 23 || m < 0 || m > 59 || s < 0.0 || s >= 60.0)
        VM_FATAL("Error parsing %s as hhmmss (h = %d, m = %d, s = %f)", ra_hhmmss, h, m, s);
    return h + m / 60.0 + s / 3600.0;
}

double parse_dec(char *dec_ddmmss)
{
    /* "[+-]DD:MM:SS.S" -> decimal degrees; the sign is taken from the text,
     * so "-00:30:00" is -0.5 */
    if (!dec_ddmmss) return 0.0;
    const char *p = dec_ddmmss;
    while (*p == ' ' || *p == '\t') p++;
    const double sign = (*p == '-') ? -1.0 : 1.0;
    int d = 0, m = 0;
    double s = 0.0;
    sscanf(p, "%d:%d:%lf", &d, &m, &s);
    d = abs(d);
    if (d > 359 || m < 0 || m > 59 || s < 0.0 || s >= 60.0)
        VM_FATAL("Error parsing %s as dd:mm:ss - got %d:%d:%f", dec_ddmmss, d, m, s);
    return sign * (d + m / 60.0 + s / 3600.0);
}
