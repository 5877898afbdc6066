"""Host-side drivers mirroring the reference applications' call sequences.

Each function drives the C-ABI exactly in the order the corresponding
MWA-vcsbeam application does (minus command-line parsing and file I/O):

* :func:`tied_array_beam` -- app/make_mwa_tied_array_beam.c:75-385
* :func:`incoherent_beam` -- app/make_mwa_incoh_beam.c:53-189
* :func:`fine_pfb_offline` -- app/fine_pfb_offline.c:38-94

Inputs come from a :class:`workloads.Observation`; outputs are returned as
numpy arrays in the reference host layouts.
"""
import ctypes as C

import numpy as np

from . import (ANALYSIS_FILTER, SYNTHESIS_FILTER, BeamGeom, GpuIpfbArrays, MpiPsrfits,
               PFB_EMULATE_FPGA, PFB_FULL_PRECISION, PFB_MALLOC_HOST_OUTPUT, PFB_SMART, VCSMWAXv2, VM_BF_FP32,
               VM_HALF, VM_SUCCESS, VcsbeamContext, Vdifinfo, load)
from .workloads import unit_vectors

# samples per block of the device Stokes layout d_S [c][s / 64][p][st][s % 64]
S_BLOCK = 64


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _beam_geoms(az, za):
    u = unit_vectors(az, za)
    bg = (BeamGeom * len(u))()
    for p in range(len(u)):
        bg[p].az = float(az[p])
        bg[p].el = float(np.pi / 2 - za[p])
        bg[p].unit_E, bg[p].unit_N, bg[p].unit_H = (float(x) for x in u[p])
    return bg


def tied_array_beam(obs, az, za, *, pfb_flags=PFB_FULL_PRECISION, nstokes=4, nchunks=1,
                    form=VM_BF_FP32, nseconds=1, device_geometry=False, want_e=True,
                    ipfb=False, synth_filter=None, ipfb_device=False, return_vm=False,
                    variant=None):
    """Tied-array beamforming of `nseconds` seconds of `obs` towards the
    directions (az, za) [radians].  Returns dict with per-second lists:
    'S' [p][s][st][c] float32, 'e' [p][s][c][pol] complex128 (if want_e),
    'vdif' float32 [n][pol][2] (if ipfb).  variant: the MFMA form's GEMM
    variant (vmSetBeamformerVariant, tile order 0) instead of the default."""
    lib = load()
    az = np.atleast_1d(np.asarray(az, dtype=np.float64))
    za = np.atleast_1d(np.asarray(za, dtype=np.float64))
    npointing = len(az)
    vm = lib.vmInit(False)
    v = vm.contents
    lib.vmBindObsMetadata(vm, C.byref(obs.md), C.byref(obs.vmd), 0, obs.md.obs_id, nseconds)
    v.out_nstokes = nstokes
    v.chunks_per_second = nchunks
    v.beamform_form = form
    coeffs = np.ascontiguousarray(obs.filter, dtype=np.float64)
    if v.do_forward_pfb:
        lib.vmSetFilter(vm, coeffs.ctypes.data_as(C.POINTER(C.c_double)), len(coeffs),
                        ANALYSIS_FILTER, 128)
        lib.vmInitForwardPFB(vm, 128, pfb_flags)
    lib.vmSetNumPointings(vm, npointing)
    bg = _beam_geoms(az, za)
    nchans, npols, nsamples = v.nfine_chan, 2, v.fine_sample_rate
    if ipfb:
        v.do_inverse_pfb = True
        sf = np.ascontiguousarray(synth_filter if synth_filter is not None else coeffs / coeffs.max(),
                                  dtype=np.float64)
        lib.vmSetFilter(vm, sf.ctypes.data_as(C.POINTER(C.c_double)), len(sf), SYNTHESIS_FILTER,
                        nchans)
        lib.vmScaleFilterCoeffs(vm, SYNTHESIS_FILTER, 15.0 / 7.2)
        data_buffer_fine = lib.create_data_buffer_fine(npointing, 2 * nsamples, nchans, npols)
    lib.vmSetMaxGPUMem(vm, nchunks)
    for f in ("vmMallocVDevice", "vmMallocJVDevice", "vmMallocEHost", "vmMallocEDevice",
              "vmMallocSHost", "vmMallocSDevice", "vmMallocJHost", "vmMallocJDevice",
              "vmMallocDHost", "vmMallocDDevice", "vmMallocPQIdxsHost", "vmMallocPQIdxsDevice"):
        getattr(lib, f)(vm)
    lib.vmSetPolIdxLists(vm)
    lib.vmPushPolIdxLists(vm)
    gi = GpuIpfbArrays()
    vdif = None
    if ipfb:
        vdif = np.zeros(nsamples * nchans * npols * npointing * 2, dtype=np.float32)
        lib.malloc_ipfb(C.byref(gi), v.synth_filter, nsamples, npols, npointing)
        lib.cu_load_ipfb_filter(v.synth_filter, C.byref(gi))
    # calibration D (flagged tiles zeroed), primary beam B = identity
    D = np.ascontiguousarray(obs.calibration(nchans))
    C.memmove(v.D, D.ctypes.data, D.nbytes)
    lib.vmCreatePrimaryBeam(vm)
    lib.vmCreateGeometricDelays(vm)
    lib.vmCreateCudaStreams(vm)
    if device_geometry:
        lib.vmPushD(vm)
    if variant is not None:
        lib.vmSetBeamformerVariant(vm, variant, 0)
    out = {"S": [], "e": [], "vdif": []}
    for t in range(nseconds):
        sec = obs.second(t) if not obs.legacy else obs.legacy_second(t)
        err = lib.vmReadNextSecondFrom(vm, _ptr(sec), sec.nbytes)
        if err != VM_SUCCESS:
            raise RuntimeError(f"vmReadNextSecondFrom: {err}")
        if device_geometry:
            lib.vmCalcPhiDevice(vm, bg)
            lib.vmCalcJDevice(vm)
        else:
            lib.vmCalcPhi(vm, bg)
            lib.vmCalcJ(vm)
            lib.vmPushPhi(vm)
            lib.vmPushJ(vm)
        lib.vmBeamformSecond(vm)
        if ipfb:
            if ipfb_device:
                lib.vmInvertPFBDevice(vm, C.byref(gi), t, _ptr(vdif))
            else:
                lib.vmPullE(vm)
                lib.prepare_data_buffer_fine(data_buffer_fine, vm, t)
                lib.cu_invert_pfb(data_buffer_fine, t, npointing, nsamples, nchans, npols, 0,
                                  C.byref(gi), _ptr(vdif))
            out["vdif"].append(vdif.copy())
        lib.vmPullS(vm)
        S = np.ctypeslib.as_array(C.cast(v.S, C.POINTER(C.c_float)),
                                  shape=(npointing, nsamples, nstokes, nchans)).copy()
        out["S"].append(S)
        if want_e:
            if not ipfb or ipfb_device:
                lib.vmPullE(vm)
            e = np.ctypeslib.as_array(C.cast(v.e, C.POINTER(C.c_double)),
                                      shape=(npointing, nsamples, nchans, 2, 2)).copy()
            out["e"].append(e[..., 0] + 1j * e[..., 1])
    if return_vm:
        out["vm"] = vm
        return out
    if ipfb:
        lib.free_ipfb(C.byref(gi))
    lib.vmDestroyCudaStreams(vm)
    for f in ("vmFreeVDevice", "vmFreeJVDevice", "vmFreeEHost", "vmFreeEDevice", "vmFreeSHost",
              "vmFreeSDevice", "vmFreeJDevice", "vmFreeJHost", "vmFreeDHost", "vmFreeDDevice",
              "vmFreePQIdxsDevice", "vmFreePQIdxsHost"):
        getattr(lib, f)(vm)
    lib.free_primary_beam(C.addressof(v.pb))
    lib.destroy_vcsbeam_context(vm)
    return out


def half_blocks_to_cis(h, nchan, ninput, ns):
    """VM_HALF device voltages (float16 (re, im) pairs in 32-sample blocks
    [c][s/32][i][s%32], the last block padded) -> float16 [c][i][s][2]."""
    nsb = (ns + 31) // 32
    a = np.asarray(h).reshape(nchan, nsb, ninput, 32, 2).transpose(0, 2, 1, 3, 4)
    return np.ascontiguousarray(a.reshape(nchan, ninput, nsb * 32, 2)[:, :, :ns])


def fine_pfb_offline(obs, *, nseconds=1, nchunks=1, pfb_flags=PFB_SMART | PFB_MALLOC_HOST_OUTPUT,
                     voltage_dir=None, first_gps=None):
    """Forward PFB of `nseconds` seconds to the legacy recombined format.
    Returns a list of per-second host outputs (uint8 [s][c][i] for SMART,
    complex128 [s][c][i] for FULL_PRECISION + host output, complex64
    [c][i][s] for PFB_COMPLEX_FLOAT32, float16 [c][i][s][re, im] for
    PFB_COMPLEX_FLOAT16).  With
    `voltage_dir`, the seconds from `first_gps` on are read from MWAX
    subfiles there (vmReadSecondFromFiles) instead of generated."""
    from . import VoltageFiles
    lib = load()
    vm = lib.vmInit(False)
    v = vm.contents
    gps0 = obs.md.obs_id if first_gps is None else first_gps
    lib.vmBindObsMetadata(vm, C.byref(obs.md), C.byref(obs.vmd), 0, gps0, nseconds)
    if voltage_dir is not None:
        files = VoltageFiles(str(voltage_dir).encode(), C.pointer(obs.md), C.pointer(obs.vmd))
        v.read_second = C.cast(lib.vmReadSecondFromFiles, C.c_void_p).value
        v.read_second_user = C.addressof(files)
    coeffs = np.ascontiguousarray(obs.filter, dtype=np.float64)
    lib.vmSetFilter(vm, coeffs.ctypes.data_as(C.POINTER(C.c_double)), len(coeffs),
                    ANALYSIS_FILTER, 128)
    v.chunks_per_second = nchunks
    lib.vmInitForwardPFB(vm, 128, pfb_flags)
    fp = v.fpfb.contents
    outs = []
    for t in range(nseconds):
        if voltage_dir is not None:
            err = lib.vmReadNextSecond(vm)
        else:
            sec = obs.second(t)
            err = lib.vmReadNextSecondFrom(vm, _ptr(sec), sec.nbytes)
        if err != VM_SUCCESS:
            raise RuntimeError("read failed")
        lib.vmExecuteForwardPFB(vm)
        v.v.contents.locked = False
        raw = np.ctypeslib.as_array(C.cast(fp.vcs_data, C.POINTER(C.c_uint8)),
                                    shape=(fp.vcs_size,)).copy()
        if pfb_flags & 0x10:
            outs.append(raw.reshape(fp.nspectra, fp.K, fp.I))
        elif pfb_flags & 0xC0:
            # device layout per chunk, chunks concatenated: complex64 [c][i][s],
            # or (0x80) float16 (re, im) pairs [c][s/32][i][s%32]; returned as
            # [c][i][s] (float16 with a trailing (re, im) axis)
            nch = fp.nspectra // fp.nspectra_per_chunk
            if pfb_flags & 0x80:
                a = np.stack([half_blocks_to_cis(x, fp.K, fp.I, fp.nspectra_per_chunk)
                              for x in raw.view(np.float16).reshape(nch, -1)])
            else:
                a = raw.view(np.complex64).reshape(nch, fp.K, fp.I, fp.nspectra_per_chunk)
            a = np.ascontiguousarray(np.moveaxis(a, 0, 2))
            outs.append(a.reshape((fp.K, fp.I, fp.nspectra) + a.shape[4:]))
        else:
            outs.append(raw.view(np.complex128).reshape(fp.nspectra, fp.K, fp.I))
    lib.destroy_vcsbeam_context(vm)
    return outs


def make_tied_array_beam_files(metafits, datadir, outdir, ras_hours, decs_degs, *,
                               coarse_chan_idx=0, first_gps=None, nseconds=1, cal=None,
                               ref_ant=None, keep_cross_terms=False, flags_file=None,
                               psrfits=True, vdif=True, max_sec_per_file=200,
                               pfb_flags=PFB_FULL_PRECISION | PFB_EMULATE_FPGA):
    """app/make_mwa_tied_array_beam.c's sequence over files, with this
    library's readers and writers in place of mwalib, PAL, psrfits_utils and
    vdifio: metafits + MWAX subfiles in, calibration from cal = ("offringa",
    file) or ("rts", directory) (None: identity D for unflagged tiles),
    pointings as J2000 RA (hours) / Dec (degrees) (or ras_hours = a pointings
    file of "RA Dec" strings, decs_degs ignored), header geometry at the
    first second (lines 139-147) and the beamformer's geometry, primary
    beam and Jones inverses at the middle of every second
    (vmCalcJonesAndDelays, line 257), FINEPFB analysis and MIRROR synthesis
    filters from the runtime directory, PSRFITS (IQUV, 8-bit, one file per
    pointing and max_sec_per_file seconds) and VDIF (pointing 0, as the
    reference's inverse PFB) written under outdir.  Returns the output base
    names, the per-second Stokes [p][s][st][c] and VDIF floats."""
    import os
    from . import MetafitsMetadata, MpiPsrfits, VoltageFiles, VoltageMetadata
    lib = load()
    md, vmd = MetafitsMetadata(), VoltageMetadata()
    lib.vmLoadMetafits(str(metafits).encode(), C.byref(md), C.byref(vmd))
    if isinstance(ras_hours, (str, os.PathLike)):
        # a pointings file of "RA Dec" string pairs (vmParsePointingFile)
        tmp = lib.vmInit(False)
        tmp.contents.obs_metadata = C.pointer(md)
        lib.vmParsePointingFile(tmp, str(ras_hours).encode())
        n = tmp.contents.npointing
        ras_hours = np.ctypeslib.as_array(C.cast(tmp.contents.ras_hours, C.POINTER(C.c_double)),
                                          (n,)).copy()
        decs_degs = np.ctypeslib.as_array(C.cast(tmp.contents.decs_degs, C.POINTER(C.c_double)),
                                          (n,)).copy()
        tmp.contents.obs_metadata = None
        lib.destroy_vcsbeam_context(tmp)
    ras = np.atleast_1d(np.asarray(ras_hours, np.float64))
    decs = np.atleast_1d(np.asarray(decs_degs, np.float64))
    npointing = len(ras)
    if md.mwa_version != VCSMWAXv2:
        raise ValueError("make_tied_array_beam_files reads MWAX subfiles")
    vm = lib.vmInit(False)
    v = vm.contents
    gps0 = md.obs_id if first_gps is None else first_gps
    lib.vmBindObsMetadata(vm, C.byref(md), C.byref(vmd), coarse_chan_idx, gps0, nseconds)
    files = VoltageFiles(str(datadir).encode(), C.pointer(md), C.pointer(vmd))
    v.read_second = C.cast(lib.vmReadSecondFromFiles, C.c_void_p).value
    v.read_second_user = C.addressof(files)
    v.out_nstokes = 4
    v.chunks_per_second = 1
    lib.vmLoadFilter(vm, b"FINEPFB", ANALYSIS_FILTER, 128)
    lib.vmInitForwardPFB(vm, 128, pfb_flags)
    lib.vmSetNumPointings(vm, npointing)
    nchans, npols, nsamples = v.nfine_chan, 2, v.fine_sample_rate
    if vdif:
        v.do_inverse_pfb = True
        lib.vmLoadFilter(vm, b"MIRROR", SYNTHESIS_FILTER, nchans)
        lib.vmScaleFilterCoeffs(vm, SYNTHESIS_FILTER, 15.0 / 7.2)
    lib.vmSetMaxGPUMem(vm, 1)
    allocs = ("vmMallocVDevice", "vmMallocJVDevice", "vmMallocEHost", "vmMallocEDevice",
              "vmMallocSHost", "vmMallocSDevice", "vmMallocJHost", "vmMallocJDevice",
              "vmMallocDHost", "vmMallocDDevice", "vmMallocPQIdxsHost", "vmMallocPQIdxsDevice")
    for f in allocs:
        getattr(lib, f)(vm)
    lib.vmSetPolIdxLists(vm)
    lib.vmPushPolIdxLists(vm)
    # calibration into D
    keep = []
    v.cal_metadata = C.cast(C.pointer(md), C.c_void_p)
    D = np.ctypeslib.as_array(C.cast(v.D, C.POINTER(C.c_double)),
                              shape=(md.num_ants, nchans, 2, 2, 2))
    if cal is None:
        D[:] = 0
        D[..., 0, 0, 0] = D[..., 1, 1, 0] = 1.0
    else:
        kind, path = cal
        buf = C.create_string_buffer(str(path).encode())
        keep.append(buf)
        v.cal.caldir = C.cast(buf, C.c_void_p)
        if kind == "offringa":
            lib.vmLoadOffringaSolution(vm)
        elif kind == "rts":
            lib.vmLoadRTSSolution(vm)
        else:
            raise ValueError(f"calibration kind {kind!r}")
    if ref_ant is not None:
        rb = C.create_string_buffer(ref_ant.encode())
        keep.append(rb)
        v.cal.ref_ant = C.cast(rb, C.c_void_p)
    v.cal.keep_cross_terms = keep_cross_terms
    lib.parse_calibration_correction_file(gps0, C.byref(v.cal))
    lib.vmApplyCalibrationCorrections(vm)
    if flags_file is not None:
        fb = C.create_string_buffer(str(flags_file).encode())
        keep.append(fb)
        v.cal.flags_file = C.cast(fb, C.c_void_p)
    # custom-flagged tiles get D = 0; tiles the metadata flags keep their
    # solutions and only change w (src/calibration.c:952-1013,
    # src/metadata.c:1554-1561)
    lib.vmSetCustomTileFlags(vm)
    lib.vmCreatePrimaryBeam(vm)
    lib.vmCreateGeometricDelays(vm)
    lib.vmCreateCudaStreams(vm)
    # pointing geometry at the first second, for the output headers (lines
    # 139-147); the beamformer's geometry is recomputed every second
    sec_offset = gps0 - md.sched_start_gps_time_ms / 1000.0
    mjd = md.sched_start_mjd + sec_offset / 86400.0
    bg = (BeamGeom * npointing)()
    for p in range(npointing):
        lib.calc_beam_geom(float(ras[p]), float(decs[p]), mjd, C.byref(bg[p]))
    os.makedirs(outdir, exist_ok=True)
    mpfs = (MpiPsrfits * npointing)()
    if psrfits:
        for p in range(npointing):
            lib.populate_psrfits_header(vm, C.byref(mpfs[p].coarse_chan_pf), max_sec_per_file, 4,
                                        C.byref(bg[p]), None, True)
            base = os.path.join(str(outdir), mpfs[p].coarse_chan_pf.basefilename.decode())
            mpfs[p].coarse_chan_pf.basefilename = base.encode()
        lib.vmCreateStatistics(vm, mpfs)
    gi = GpuIpfbArrays()
    vdif_buf = None
    if vdif:
        lib.malloc_ipfb(C.byref(gi), v.synth_filter, nsamples, npols, 1)
        lib.cu_load_ipfb_filter(v.synth_filter, C.byref(gi))
        lib.vmPopulateVDIFHeader(vm, bg, md.sched_start_mjd, sec_offset)
        vf = C.cast(v.vf, C.POINTER(Vdifinfo))
        base = os.path.join(str(outdir), vf[0].basefilename.decode())
        vf[0].basefilename = base.encode()
        vdif_buf = np.zeros(nsamples * nchans * npols * 2, np.float32)
    out = {"S": [], "vdif": [], "psrfits": [], "vdif_base": None, "bg_per_second": []}
    ras_c = np.ascontiguousarray(ras, np.float64)
    decs_c = np.ascontiguousarray(decs, np.float64)
    bg_t = (BeamGeom * npointing)()
    for t in range(nseconds):
        if lib.vmReadNextSecond(vm) != VM_SUCCESS:
            raise RuntimeError("vmReadNextSecond failed")
        # geometry, primary beam and Jones inverses at the middle of this
        # second (app/make_mwa_tied_array_beam.c:257)
        lib.vmCalcJonesAndDelays(vm, _ptr(ras_c), _ptr(decs_c), bg_t)
        out["bg_per_second"].append([(bg_t[p].az, bg_t[p].el) for p in range(npointing)])
        lib.vmPushPhi(vm)
        lib.vmPushJ(vm)
        lib.vmBeamformSecond(vm)
        v.v.contents.locked = False
        lib.vmPullS(vm)
        out["S"].append(np.ctypeslib.as_array(C.cast(v.S, C.POINTER(C.c_float)),
                                              shape=(npointing, nsamples, 4, nchans)).copy())
        if psrfits:
            lib.vmSendSToFits(vm, mpfs)
            for p in range(npointing):
                if lib.vmWritePsrfitsSubint(C.byref(mpfs[p].coarse_chan_pf)) != 0:
                    raise RuntimeError("PSRFITS write failed")
        if vdif:
            lib.vmInvertPFBDevice(vm, C.byref(gi), t, _ptr(vdif_buf))
            out["vdif"].append(vdif_buf.copy())
            lib.vdif_write_second(vf, C.byref(v, VcsbeamContext.vhdr.offset), _ptr(vdif_buf))
    if psrfits:
        for p in range(npointing):
            out["psrfits"].append(mpfs[p].coarse_chan_pf.basefilename.decode())
            lib.free_psrfits(C.byref(mpfs[p].coarse_chan_pf))
        lib.vmDestroyStatistics(vm)
    if vdif:
        out["vdif_base"] = vf[0].basefilename.decode()
        lib.free_ipfb(C.byref(gi))
    lib.vmDestroyCudaStreams(vm)
    for f in allocs:
        getattr(lib, f.replace("Malloc", "Free"))(vm)
    lib.free_primary_beam(C.addressof(v.pb))
    # detach the buffers Python owns; free_calibration then releases what the
    # library allocated (the PQ table's reference tile, the flagged-tile list)
    v.cal.caldir = None
    v.cal.flags_file = None
    if ref_ant is not None:
        v.cal.ref_ant = None
    lib.free_calibration(C.byref(v.cal))
    v.cal_metadata = None
    lib.destroy_vcsbeam_context(vm)
    lib.vmFreeMetafits(C.byref(md))
    out["bg"] = bg
    return out


def fine_pfb_offline_files(metafits, datadir, outdir, *, coarse_chan_idx=0, first_gps=None,
                           nseconds=1, nchunks=1, filter_name="FINEPFB"):
    """app/fine_pfb_offline.c's sequence over files: MWAX subfiles (metafits +
    <obsid>_<subobs>_<rec>.sub) through the forward PFB (PFB_SMART, host
    output) into legacy recombined files <obsid>_<gps>_ch<rec>.dat in
    outdir.  Returns the file paths."""
    import os
    from . import MetafitsMetadata, VoltageFiles, VoltageMetadata
    lib = load()
    md, vmd = MetafitsMetadata(), VoltageMetadata()
    lib.vmLoadMetafits(str(metafits).encode(), C.byref(md), C.byref(vmd))
    if md.mwa_version != VCSMWAXv2:
        raise ValueError(f"observation {md.obs_id} is not coarse-channelised MWAX data")
    vm = lib.vmInit(False)
    v = vm.contents
    gps0 = md.obs_id if first_gps is None else first_gps
    lib.vmBindObsMetadata(vm, C.byref(md), C.byref(vmd), coarse_chan_idx, gps0, nseconds)
    files = VoltageFiles(str(datadir).encode(), C.pointer(md), C.pointer(vmd))
    v.read_second = C.cast(lib.vmReadSecondFromFiles, C.c_void_p).value
    v.read_second_user = C.addressof(files)
    os.makedirs(outdir, exist_ok=True)
    outbuf = C.create_string_buffer(str(outdir).encode())
    v.datadir = C.cast(outbuf, C.c_void_p)
    lib.vmLoadFilter(vm, filter_name.encode(), ANALYSIS_FILTER, 128)
    v.chunks_per_second = nchunks
    lib.vmInitForwardPFB(vm, 128, PFB_SMART | PFB_MALLOC_HOST_OUTPUT)
    rec = md.metafits_coarse_chans[coarse_chan_idx].rec_chan_number
    paths = []
    try:
        while lib.vmReadNextSecond(vm) == VM_SUCCESS:
            lib.vmExecuteForwardPFB(vm)
            v.v.contents.locked = False
            lib.vmWritePFBOutputToFile(vm)
            gps = gps0 + len(paths)
            paths.append(os.path.join(str(outdir), f"{md.obs_id}_{gps}_ch{rec:03d}.dat"))
    finally:
        v.datadir = None
        lib.destroy_vcsbeam_context(vm)
        lib.vmFreeMetafits(C.byref(md))
    return paths


def make_incoh_beam_files(metafits, datadir, outdir, *, coarse_chan_idx=0, first_gps=None,
                          nseconds=1, max_sec_per_file=200):
    """app/make_mwa_incoh_beam.c's sequence over files: legacy recombined
    seconds (metafits + <obsid>_<gps>_ch<rec>.dat) through
    cu_form_incoh_beam into an incoherent-beam PSRFITS (Stokes I, the tile
    pointing, geometry at the scheduled start).  Returns the PSRFITS base
    name and, per second, (offsets, scales, Iscaled [s][c])."""
    import os
    from . import DeviceArray, MetafitsMetadata, Psrfits, VoltageFiles, VoltageMetadata
    lib = load()
    md, vmd = MetafitsMetadata(), VoltageMetadata()
    lib.vmLoadMetafits(str(metafits).encode(), C.byref(md), C.byref(vmd))
    if md.mwa_version == VCSMWAXv2:
        raise ValueError("make_incoh_beam_files reads legacy recombined files")
    vm = lib.vmInit(False)
    v = vm.contents
    gps0 = md.obs_id if first_gps is None else first_gps
    lib.vmBindObsMetadata(vm, C.byref(md), C.byref(vmd), coarse_chan_idx, gps0, nseconds)
    files = VoltageFiles(str(datadir).encode(), C.pointer(md), C.pointer(vmd))
    nchans, ninputs, nsamples = md.num_volt_fine_chans_per_coarse, md.num_rf_inputs, v.sample_rate
    data_size = v.bytes_per_second
    h, d = C.c_void_p(), C.c_void_p()
    lib.allocate_input_output_arrays(C.byref(h), C.byref(d), data_size)
    d_incoh = DeviceArray(nchans * nsamples * 4)
    d_off, d_scl = DeviceArray(nchans * 4), DeviceArray(nchans * 4)
    d_Is = DeviceArray(nchans * nsamples)
    bg = BeamGeom()
    lib.calc_beam_geom(md.ra_tile_pointing_deg / 15.0, md.dec_tile_pointing_deg,
                       md.sched_start_mjd, C.byref(bg))
    pf = Psrfits()
    lib.populate_psrfits_header(vm, C.byref(pf), max_sec_per_file, 1, C.byref(bg), None, False)
    os.makedirs(outdir, exist_ok=True)
    pf.basefilename = os.path.join(str(outdir), pf.basefilename.decode()).encode()
    out = {"psrfits": pf.basefilename.decode(), "seconds": []}
    try:
        for t in range(nseconds):
            if lib.vmReadSecondFromFiles(C.byref(files), gps0 + t, coarse_chan_idx, h,
                                         data_size) != 0:
                raise RuntimeError(f"no data for GPS second {gps0 + t}")
            lib.cu_form_incoh_beam(h, d, data_size, d_incoh.ptr, nsamples, nchans, ninputs,
                                   pf.sub.dat_offsets, d_off.ptr, pf.sub.dat_scales, d_scl.ptr,
                                   pf.sub.data, d_Is.ptr, nchans * nsamples)
            off = np.ctypeslib.as_array(C.cast(pf.sub.dat_offsets, C.POINTER(C.c_float)),
                                        (nchans,)).copy()
            scl = np.ctypeslib.as_array(C.cast(pf.sub.dat_scales, C.POINTER(C.c_float)),
                                        (nchans,)).copy()
            Is = np.ctypeslib.as_array(C.cast(pf.sub.data, C.POINTER(C.c_uint8)),
                                       (nsamples, nchans)).copy()
            out["seconds"].append((off, scl, Is))
            if lib.vmWritePsrfitsSubint(C.byref(pf)) != 0:
                raise RuntimeError("PSRFITS write failed")
    finally:
        lib.free_psrfits(C.byref(pf))
        lib.free_input_output_arrays(h, d)
        lib.destroy_vcsbeam_context(vm)
        lib.vmFreeMetafits(C.byref(md))
    return out


def incoherent_beam(data, nsample, nchan, ninput):
    """cu_form_incoh_beam on one second of legacy 4+4-bit data [s][c][i].
    Returns (incoh float [s][c] before scaling, offsets, scales, Iscaled)."""
    from . import DeviceArray
    lib = load()
    size = nsample * nchan * ninput
    h, d = C.c_void_p(), C.c_void_p()
    lib.allocate_input_output_arrays(C.byref(h), C.byref(d), size)
    C.memmove(h, data.ctypes.data, size)
    d_incoh = DeviceArray(nsample * nchan * 4)
    d_off = DeviceArray(nchan * 4)
    d_scl = DeviceArray(nchan * 4)
    d_Is = DeviceArray(nsample * nchan)
    off = np.zeros(nchan, np.float32)
    scl = np.zeros(nchan, np.float32)
    Is = np.zeros(nsample * nchan, np.uint8)
    lib.cu_form_incoh_beam(h, d, size, d_incoh.ptr, nsample, nchan, ninput, _ptr(off), d_off.ptr,
                           _ptr(scl), d_scl.ptr, _ptr(Is), d_Is.ptr, Is.nbytes)
    incoh = d_incoh.to_numpy(np.float32, (nsample, nchan))
    lib.free_input_output_arrays(h, d)
    return incoh, off, scl, Is.reshape(nsample, nchan)


class TiedArraySession:
    """A tied-array beamforming context held open across steps.

    Set-up follows app/make_mwa_tied_array_beam.c:85-240; the per-second
    device work is split into :meth:`pfb` (forward PFB of the resident
    second), :meth:`geometry` (FP64 phases and Jones inverses on the GPU) and
    :meth:`beamform` (weights + beamformer), so that each can be timed.
    d_e is only allocated when `want_e` (the reference always writes e; the
    many-beam Stokes workloads do not need it).
    """

    def __init__(self, obs, az, za, *, pfb_flags=PFB_FULL_PRECISION, nstokes=4, form=VM_BF_FP32,
                 want_e=False, host_buffers=True, ipfb=False, d_v=None):
        """host_buffers=False leaves out the host S and J buffers and frees the
        pinned read buffer (for seconds generated on the device and outputs
        read back in slices).  ipfb=True adds the inverse PFB of pointing 0
        (mirror synthesis filter x 15/7.2, app/make_mwa_tied_array_beam.c:
        155-165) and device VDIF requantisation (:meth:`invert`).  d_v: a
        caller-owned device buffer for the forward PFB output when pfb_flags
        lacks PFB_MALLOC_DEVICE_OUTPUT."""
        lib = self.lib = load()
        self.obs = obs
        az = np.atleast_1d(np.asarray(az, dtype=np.float64))
        za = np.atleast_1d(np.asarray(za, dtype=np.float64))
        self.npointing = len(az)
        vm = self.vm = lib.vmInit(False)
        v = self.v = vm.contents
        lib.vmBindObsMetadata(vm, C.byref(obs.md), C.byref(obs.vmd), 0, obs.md.obs_id, 1 << 20)
        v.out_nstokes = nstokes
        v.chunks_per_second = 1
        v.beamform_form = form
        self.coeffs = np.ascontiguousarray(obs.filter, dtype=np.float64)
        lib.vmSetFilter(vm, self.coeffs.ctypes.data_as(C.POINTER(C.c_double)), len(self.coeffs),
                        ANALYSIS_FILTER, 128)
        lib.vmInitForwardPFB(vm, 128, pfb_flags)
        if d_v is not None:
            v.fpfb.contents.d_vcs_data = d_v
        lib.vmSetNumPointings(vm, self.npointing)
        self.bg = _beam_geoms(az, za)
        self.ipfb = ipfb
        if ipfb:
            sf = np.ascontiguousarray(self.coeffs / self.coeffs.max())
            lib.vmSetFilter(vm, sf.ctypes.data_as(C.POINTER(C.c_double)), len(sf),
                            SYNTHESIS_FILTER, 128)
            lib.vmScaleFilterCoeffs(vm, SYNTHESIS_FILTER, 15.0 / 7.2)
        lib.vmSetMaxGPUMem(vm, 1)
        names = ["vmMallocVDevice", "vmMallocSDevice", "vmMallocJDevice", "vmMallocDHost",
                 "vmMallocDDevice", "vmMallocPQIdxsHost", "vmMallocPQIdxsDevice"]
        if host_buffers:
            names += ["vmMallocSHost", "vmMallocJHost"]
        else:
            lib.vmFreeVHost(vm)
        if want_e:
            names += ["vmMallocEHost", "vmMallocEDevice"]
        elif ipfb:
            names += ["vmMallocEDevice"]
        for f in names:
            getattr(lib, f)(vm)
        if ipfb:
            self.gi = GpuIpfbArrays()
            lib.malloc_ipfb(C.byref(self.gi), v.synth_filter, v.fine_sample_rate, 2, 1)
            lib.cu_load_ipfb_filter(v.synth_filter, C.byref(self.gi))
            self.nvdif = v.fine_sample_rate * v.nfine_chan * 2 * 2
            self.d_vdif = lib.vmDeviceMalloc(self.nvdif)
        lib.vmSetPolIdxLists(vm)
        lib.vmPushPolIdxLists(vm)
        D = np.ascontiguousarray(obs.calibration(v.nfine_chan))
        C.memmove(v.D, D.ctypes.data, D.nbytes)
        lib.vmCreatePrimaryBeam(vm)
        lib.vmCreateGeometricDelays(vm)
        lib.vmPushD(vm)
        self.stream = lib.vmGetStream(vm)

    def read(self, t):
        sec = self.obs.second(t)
        self.v.v.contents.locked = False
        err = self.lib.vmReadNextSecondFrom(self.vm, _ptr(sec), sec.nbytes)
        if err != VM_SUCCESS:
            raise RuntimeError(f"vmReadNextSecondFrom: {err}")
        self.lib.vmUploadForwardPFBChunk(self.vm)

    def pfb(self):
        for f in ("vmWOLAChunk", "vmFPGARoundingChunk", "vmFFTChunk", "vmPackChunk"):
            getattr(self.lib, f)(self.vm)

    def invert(self, t):
        """Inverse PFB of pointing 0 from the device e (second t) and its 8-bit
        VDIF requantisation, both on the device in one kernel."""
        self.lib.vmInvertPFBDeviceVDIF(self.vm, C.byref(self.gi), t, self.d_vdif)
        self.lib.vmSynchronize(self.vm)

    def geometry(self):
        self.lib.vmCalcPhiDevice(self.vm, self.bg)
        self.lib.vmCalcJDevice(self.vm)

    def beamform(self):
        self.lib.vmApplyJChunk(self.vm)
        self.lib.vmBeamformChunk(self.vm)

    def stokes(self):
        """Pull S to the host layout [p][s][st][c]."""
        self.lib.vmPullS(self.vm)
        v = self.v
        return np.ctypeslib.as_array(C.cast(v.S, C.POINTER(C.c_float)),
                                     shape=(self.npointing, v.fine_sample_rate, v.out_nstokes,
                                            v.nfine_chan)).copy()

    def voltages(self):
        """The device voltages d_v (VM_FLT [c][i][s], or VM_HALF widened) as
        complex64."""
        v = self.v
        n = v.nfine_chan * self.obs.ninput * v.fine_sample_rate
        if v.datatype == VM_HALF:
            nsb = (v.fine_sample_rate + 31) // 32
            h = np.empty(v.nfine_chan * nsb * 32 * self.obs.ninput * 2, np.float16)
            self.lib.vmMemcpyDtoH(h.ctypes.data, v.d_v, h.nbytes)
            h = half_blocks_to_cis(h, v.nfine_chan, self.obs.ninput, v.fine_sample_rate)
            out = h.astype(np.float32).view(np.complex64)
        else:
            out = np.empty(n, np.complex64)
            self.lib.vmMemcpyDtoH(out.ctypes.data, v.d_v, out.nbytes)
        return out.reshape(v.nfine_chan, self.obs.ninput, v.fine_sample_rate)

    def stokes_slice(self, p0, p1, s0, s1):
        """Device S for pointings p0:p1 and samples s0:s1, read without a full
        host copy; returned as [p][s][st][c] like stokes().  The device layout
        is [c][s / 64][p][st][s % 64]: per channel, the sample blocks covering
        s0:s1 are rows of a strided copy."""
        v = self.v
        nst, nch, ns, npt = v.out_nstokes, v.nfine_chan, v.fine_sample_rate, self.npointing
        nblk = (ns + S_BLOCK - 1) // S_BLOCK
        b0, b1 = s0 // S_BLOCK, (s1 + S_BLOCK - 1) // S_BLOCK
        width = (p1 - p0) * nst * S_BLOCK * 4
        pitch = npt * nst * S_BLOCK * 4
        buf = np.empty((nch, b1 - b0, p1 - p0, nst, S_BLOCK), np.float32)
        base = C.cast(v.d_S, C.c_void_p).value
        for c in range(nch):
            src = base + 4 * ((c * nblk + b0) * npt + p0) * nst * S_BLOCK
            self.lib.vmMemcpy2DDtoH(buf[c].ctypes.data, width, src, pitch, width, b1 - b0)
        out = buf.transpose(2, 1, 4, 3, 0).reshape(p1 - p0, (b1 - b0) * S_BLOCK, nst, nch)
        return np.ascontiguousarray(out[:, s0 - b0 * S_BLOCK:s1 - b0 * S_BLOCK])

    def close(self):
        lib, vm = self.lib, self.vm
        if self.ipfb:
            lib.free_ipfb(C.byref(self.gi))
            lib.vmDeviceFree(self.d_vdif)
            if not self.v.e and self.v.d_e:
                lib.vmFreeEDevice(vm)
        for f in ("vmFreeVDevice", "vmFreeSDevice", "vmFreeJDevice",
                  "vmFreeDHost", "vmFreeDDevice", "vmFreePQIdxsDevice", "vmFreePQIdxsHost"):
            getattr(lib, f)(vm)
        if self.v.S:
            lib.vmFreeSHost(vm)
        if self.v.J:
            lib.vmFreeJHost(vm)
        if self.v.e:
            lib.vmFreeEHost(vm)
            lib.vmFreeEDevice(vm)
        lib.free_primary_beam(C.addressof(self.v.pb))
        lib.destroy_vcsbeam_context(vm)
        self.vm = None


def invert_batch(sessions, t):
    """Inverse PFB and 8-bit VDIF of second t for several sessions (coarse
    channels of the same second) in one launch: vmInvertPFBDeviceVDIFBatch
    leaves each session's outputs and history as its own invert(t) would."""
    n = len(sessions)
    lib = sessions[0].lib
    ctx = type(sessions[0].vm)
    vms = (ctx * n)(*[x.vm for x in sessions])
    gis = (C.POINTER(GpuIpfbArrays) * n)(*[C.pointer(x.gi) for x in sessions])
    outs = (C.c_void_p * n)(*[x.d_vdif for x in sessions])
    lib.vmInvertPFBDeviceVDIFBatch(vms, gis, n, t, outs)


def device_second(ses, t):
    """Generate second t of ses.obs in place in the session's forward-PFB
    input buffer (margin block + the second's blocks) with the device copy of
    the workload generator, as vmUploadForwardPFBChunk would have left it."""
    from .workloads import SEED_DATA, SPB
    o, fp = ses.obs, ses.v.fpfb.contents
    base = C.cast(fp.d_htr_data, C.c_void_p).value
    nb = o.blocks_per_second
    block = o.ninput * SPB * 2
    # margin block = last block of the previous second (zeros at t = 0)
    ses.lib.vmSynthMWAXDevice(SEED_DATA + o.seed_offset, max(t * nb - 1, 0), 1, o.ninput, SPB,
                              16.0, 1 if t == 0 else 0, base)
    ses.lib.vmSynthMWAXDevice(SEED_DATA + o.seed_offset, t * nb, nb, o.ninput, SPB, 16.0, 0,
                              base + block)


class BatchSession:
    """Several coarse channels, each its own context, beamformed together by
    the batched persistent MFMA launch (vmBatch*).

    obs_list[g] and dirs[g] = (az, za) describe channel g; channels may
    differ in beam count and flags.  With share_pfb_output the contexts'
    forward PFBs are initialised without PFB_MALLOC_DEVICE_OUTPUT and share
    one output buffer owned by the batch.  With device_data each second is
    generated in place on the device (vmSynthMWAXDevice, the same generator
    as Observation.second) instead of read from the host.
    """

    def __init__(self, obs_list, dirs, *, pfb_flags=PFB_FULL_PRECISION, nstokes=4, form=1,
                 share_pfb_output=True, device_data=False):
        self.lib = lib = load()
        flags = pfb_flags & ~0x08 if share_pfb_output else pfb_flags
        self.device_data = device_data
        self.sessions = [TiedArraySession(o, az, za, pfb_flags=flags, nstokes=nstokes, form=form,
                                          host_buffers=not device_data)
                         for o, (az, za) in zip(obs_list, dirs)]
        for ses in self.sessions:
            ses.geometry()
        arr = (type(self.sessions[0].vm) * len(self.sessions))(*[x.vm for x in self.sessions])
        self.b = lib.vmBatchCreate(arr, len(self.sessions))

    def read(self, t):
        for ses in self.sessions:
            if self.device_data:
                device_second(ses, t)
            else:
                ses.read(t)

    def run(self, use_graph=True):
        self.lib.vmBatchBeamformSecond(self.b, 1 if use_graph else 0)

    def tune(self, reps=3):
        return self.lib.vmBatchTune(self.b, reps)

    def info(self):
        v, band, n = C.c_int(), C.c_int(), C.c_int()
        self.lib.vmBatchInfo(self.b, C.byref(v), C.byref(band), C.byref(n))
        return {"variant": v.value, "band": band.value, "items": n.value}

    def stage_times(self):
        t = [C.c_float() for _ in range(3)]
        self.lib.vmBatchStageTimes(self.b, *[C.byref(x) for x in t])
        return {"pfb_ms": t[0].value, "prep_ms": t[1].value, "gemm_ms": t[2].value}

    def active_tiles(self, g):
        return self.lib.vmBatchActiveTiles(self.b, g)

    def close(self):
        if self.b:
            self.lib.vmBatchDestroy(self.b)
            self.b = None
        for ses in self.sessions:
            ses.close()
