"""Seeded synthetic observations for the voltage-beamforming path.

Builds everything the C-ABI consumes in place of an MWA observation:
metadata (the mwalib-subset structs of ``include/vcsbeam.h``), MWAX 8+8-bit
voltages, an integer analysis filter, calibration Jones matrices D, tile
geometry and beam directions.  All randomness comes from the counter-based
generator in ``csrc/vmsynth.c`` with fixed seeds (data 1, filter 2, Jones 3,
geometry 4, beams 5, flags 6), so the same observation is byte-identical on
every host.
"""
import ctypes as C
import math

import numpy as np

from . import (CoarseChannel, MetafitsMetadata, Rfinput, VoltageMetadata, VCSMWAXv2,
               VCSLegacyRecombined, MWA_ALTITUDE_METRES, synth)

SEED_DATA, SEED_FILTER, SEED_JONES, SEED_GEOM, SEED_BEAMS, SEED_FLAGS = 1, 2, 3, 4, 5, 6
SPB = 64000  # samples per MWAX voltage block
K = 128


def uniform(seed, n, offset=0):
    out = np.empty(n, dtype=np.float64)
    synth().vmsynth_uniform(seed, offset, n, out.ctypes.data)
    return out


def normal(seed, n, offset=0):
    out = np.empty(n, dtype=np.float64)
    synth().vmsynth_normal(seed, offset, n, out.ctypes.data)
    return out


def kaiser_filter(ntaps_total=1536, k=K, beta=5.0, peak=32767.0, seed=SEED_FILTER):
    """Integer Kaiser-windowed sinc prototype (float64 array of integers)."""
    out = np.empty(ntaps_total, dtype=np.float64)
    synth().vmsynth_kaiser_filter(seed, ntaps_total, k, beta, peak, out.ctypes.data)
    return out


def mwax_voltages(ninput, nblocks, first_block=0, sigma=16.0, zero_blocks=0, seed=SEED_DATA):
    """MWAX voltages [nblocks][ninput][64000][2] int8 (re, im)."""
    out = np.empty((nblocks, ninput, SPB, 2), dtype=np.int8)
    synth().vmsynth_mwax_voltages(seed, first_block, nblocks, ninput, SPB, sigma, zero_blocks,
                                  out.ctypes.data)
    return out


def legacy_int4(n, sigma=2.0, seed=SEED_DATA):
    out = np.empty(n, dtype=np.uint8)
    synth().vmsynth_legacy_int4(seed, n, sigma, out.ctypes.data)
    return out


def directions_in_cap(n, radius_deg, centre_za_deg=0.0, seed=SEED_BEAMS):
    """n (az, za) pairs uniform on the spherical cap of the given radius around
    (az=0, za=centre_za); returns radians."""
    u = uniform(seed, 2 * n)
    cosr = math.cos(math.radians(radius_deg))
    ct = 1.0 - u[0::2] * (1.0 - cosr)  # cos of offset angle
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2.0 * np.pi * u[1::2]
    # offset vector around the +z axis, then rotate to the cap centre
    x, y, z = st * np.cos(ph), st * np.sin(ph), ct
    c0 = math.radians(centre_za_deg)
    x2 = x * math.cos(c0) + z * math.sin(c0)
    z2 = -x * math.sin(c0) + z * math.cos(c0)
    za = np.arccos(np.clip(z2, -1, 1))
    az = np.mod(np.arctan2(y, x2), 2 * np.pi)
    return az, za


def unit_vectors(az, za):
    """(uE, uN, uH) per direction (src/geometry.c:271-273 with el = 90 - za)."""
    el = np.pi / 2 - np.asarray(za)
    az = np.asarray(az)
    return np.stack([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)], axis=-1)


class Observation:
    """A synthetic MWA observation of one coarse channel.

    ntiles tiles (2 RF inputs each), `blocks_per_second` MWAX voltage blocks
    of 64000 samples per "second" (20 for a real second; 2 gives the 0.1 s
    workload), one coarse channel.
    """

    def __init__(self, ntiles=128, blocks_per_second=20, nseconds=1, nflagged=4,
                 rec_chan=133, legacy=False, filter_coeffs=None, obs_id=1300000000,
                 seed_offset=0, geometry_seed_offset=None, dipole_delays=None,
                 dead_dipoles=None):
        """dipole_delays: 16 analog beamformer delay steps shared by every
        tile (None: the metadata carries no dipole delays, so the primary
        beam stays the identity); dead_dipoles: {tile: [dipole, ...]}, each
        listed dipole's delay set to 32 (dead)."""
        self.ntiles = ntiles
        self.ninput = 2 * ntiles
        self.blocks_per_second = blocks_per_second
        self.nseconds = nseconds
        self.legacy = legacy
        self.seed_offset = seed_offset
        I = self.ninput
        # geometry: uniform in a disc of radius 1.5 km, small height scatter,
        # electrical lengths 0..300 m
        # geometry_seed_offset: the array layout (positions, cables, VCS order)
        # shared by coarse channels that differ in data, flags and Jones
        gso = seed_offset if geometry_seed_offset is None else geometry_seed_offset
        g = uniform(SEED_GEOM + gso, 4 * ntiles)
        r = 1500.0 * np.sqrt(g[0::4])
        th = 2 * np.pi * g[1::4]
        self.east = r * np.cos(th)
        self.north = r * np.sin(th)
        self.height = MWA_ALTITUDE_METRES + 3.0 * (g[2::4] - 0.5)
        self.cable = 300.0 * g[3::4]
        # VCS order: seeded permutation of the RF inputs
        perm_keys = uniform(SEED_GEOM + gso, I, offset=10 ** 6)
        self.vcs_order = np.argsort(perm_keys, kind="stable").astype(np.uint32)
        # flags: nflagged tiles (both pols)
        fk = uniform(SEED_FLAGS + seed_offset, ntiles)
        self.flagged_tiles = np.sort(np.argsort(fk, kind="stable")[:nflagged])
        self.flag_mask = np.zeros(ntiles, dtype=bool)
        self.flag_mask[self.flagged_tiles] = True

        self._names = [f"Tile{a:03d}".encode() for a in range(ntiles)]
        self._pols = [b"X", b"Y"]
        rf = (Rfinput * I)()
        for i in range(I):
            a, p = i // 2, i % 2
            rf[i].input = i
            rf[i].ant = a
            rf[i].tile_id = 1000 + a
            rf[i].tile_name = self._names[a]
            rf[i].pol = self._pols[p]
            rf[i].electrical_length_m = self.cable[a]
            rf[i].north_m = self.north[a]
            rf[i].east_m = self.east[a]
            rf[i].height_m = self.height[a]
            rf[i].vcs_order = int(self.vcs_order[i])
            rf[i].subfile_order = i
            rf[i].flagged = bool(self.flag_mask[a])
        self.dipole_delays = None
        if dipole_delays is not None:
            dd = np.tile(np.asarray(dipole_delays, np.uint32).reshape(1, 16), (ntiles, 1))
            for a, dead in (dead_dipoles or {}).items():
                dd[a, list(dead)] = 32
            self.dipole_delays = np.ascontiguousarray(dd)
            for i in range(I):
                rf[i].dipole_delays = self.dipole_delays[i // 2].ctypes.data_as(
                    C.POINTER(C.c_uint32))
        self.rf = rf
        cc = (CoarseChannel * 1)()
        cc[0].corr_chan_number = 0
        cc[0].rec_chan_number = rec_chan
        cc[0].gpubox_number = 1
        cc[0].chan_width_hz = 1280000
        cc[0].chan_centre_hz = rec_chan * 1280000
        cc[0].chan_start_hz = rec_chan * 1280000 - 640000
        cc[0].chan_end_hz = rec_chan * 1280000 + 640000
        self.cc = cc
        self._proj = b"G0000"
        md = MetafitsMetadata()
        md.obs_id = obs_id
        md.mwa_version = VCSLegacyRecombined if legacy else VCSMWAXv2
        md.sched_start_mjd = 59000.5
        md.sched_start_gps_time_ms = obs_id * 1000
        md.num_ants = ntiles
        md.num_rf_inputs = I
        md.rf_inputs = C.cast(rf, C.POINTER(Rfinput))
        md.num_ant_pols = 2
        md.num_visibility_pols = 4
        md.num_metafits_coarse_chans = 1
        md.metafits_coarse_chans = C.cast(cc, C.POINTER(CoarseChannel))
        md.coarse_chan_width_hz = 1280000
        md.num_volt_fine_chans_per_coarse = K if legacy else 1
        md.ra_tile_pointing_deg = 0.0
        md.dec_tile_pointing_deg = -26.7
        md.project_id = self._proj
        md.num_delays = 16 if dipole_delays is not None else 0
        self.md = md
        vmd = VoltageMetadata()
        vmd.mwa_version = md.mwa_version
        if legacy:
            # legacy recombined: 10000 samples/s of 128 fine channels, 4+4 bit
            vmd.num_samples_per_voltage_block = 10000 * blocks_per_second // 20
            vmd.num_voltage_blocks_per_second = 1
            vmd.voltage_block_size_bytes = vmd.num_samples_per_voltage_block * K * I
        else:
            vmd.num_samples_per_voltage_block = SPB
            vmd.num_voltage_blocks_per_second = blocks_per_second
            vmd.voltage_block_size_bytes = SPB * I * 2
        self.vmd = vmd
        self.freqs = np.array([cc[0].chan_start_hz + (1280000 // K) * c for c in range(K)],
                              dtype=np.float64)
        self.filter = kaiser_filter()
        if filter_coeffs is not None:
            self.filter = np.asarray(filter_coeffs, dtype=np.float64)

    # ----------------------------------------------------------------- data
    @property
    def nspectra(self):
        return self.blocks_per_second * SPB // K

    def second(self, t):
        """Voltages of second t as read by mwalib: [blocks][I][64000][2] int8."""
        return mwax_voltages(self.ninput, self.blocks_per_second,
                             first_block=t * self.blocks_per_second,
                             seed=SEED_DATA + self.seed_offset)

    def legacy_second(self, t):
        n = self.vmd.num_samples_per_voltage_block * K * self.ninput
        return legacy_int4(n, seed=SEED_DATA + 100 * (t + 1) + self.seed_offset)

    def calibration(self, nchan=K, seed=SEED_JONES):
        """D [a][c][2][2] complex128: diagonal gains lognormal(0, 0.2) with
        uniform phases, off-diagonal leakage 0.05 CN(0,1); zero for flagged
        tiles."""
        T = self.ntiles
        n = T * nchan
        z = normal(seed + self.seed_offset, 6 * n).reshape(T, nchan, 6)
        u = uniform(seed + self.seed_offset, 2 * n, offset=10 ** 7).reshape(T, nchan, 2)
        D = np.zeros((T, nchan, 2, 2), dtype=np.complex128)
        D[:, :, 0, 0] = np.exp(0.2 * z[..., 0]) * np.exp(2j * np.pi * u[..., 0])
        D[:, :, 1, 1] = np.exp(0.2 * z[..., 1]) * np.exp(2j * np.pi * u[..., 1])
        D[:, :, 0, 1] = 0.05 * (z[..., 2] + 1j * z[..., 3]) / np.sqrt(2)
        D[:, :, 1, 0] = 0.05 * (z[..., 4] + 1j * z[..., 5]) / np.sqrt(2)
        D[self.flag_mask] = 0.0
        return D

    def pol_idx(self):
        """(polQ, polP) per antenna = VCS order of its X / Y input."""
        polQ = self.vcs_order[0::2].astype(np.uint32)
        polP = self.vcs_order[1::2].astype(np.uint32)
        return polQ, polP

    def out_idx(self):
        """i_output_idx[mwax input] = VCS order."""
        return self.vcs_order.astype(np.int32)

    def positions(self):
        return self.east, self.north, self.height, self.cable
