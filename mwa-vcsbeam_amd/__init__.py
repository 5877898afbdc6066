"""ctypes binding of the MI355X voltage-beamforming C-ABI (``include/vcsbeam.h``).

This module is plumbing for the test suite and the benchmark: it loads the
in-tree ``_lib/libvcsbeam_amd.so`` (HIP kernels + host C), mirrors the C
structs the callers touch, and declares the exported entry points.  It never
falls back to another implementation: if the library is missing it raises.

The directory name contains a hyphen, so load it with :func:`importlib`::

    spec = importlib.util.spec_from_file_location("vcsbeam_amd", ".../mwa-vcsbeam_amd/__init__.py")
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "_lib")
LIBPATH = os.path.join(LIBDIR, "libvcsbeam_amd.so")
SYNTHPATH = os.path.join(LIBDIR, "libvmsynth.so")
RUNTIME_DIR = os.path.join(HERE, "pfb_filter")

# ------------------------------------------------------------------ enums --
VCSLegacyRecombined = 4
VCSMWAXv2 = 5
VM_INT4, VM_DBL, VM_FLT, VM_HALF = 0, 1, 2, 3
VM_SUCCESS, VM_END_OF_DATA, VM_READ_BUFFER_NOT_SET, VM_READ_BUFFER_LOCKED = 0, 1, 2, 3
VM_INCOH_SCRATCH_PAIRS = 65536  # partial extremes the no-allocation incoherent beam holds
ANALYSIS_FILTER, SYNTHESIS_FILTER = 0, 1
PFB_MALLOC_HOST_INPUT = 0x01
PFB_MALLOC_HOST_OUTPUT = 0x02
PFB_MALLOC_DEVICE_INPUT = 0x04
PFB_MALLOC_DEVICE_OUTPUT = 0x08
PFB_TYPE_MASK = 0xF0
PFB_COMPLEX_INT4 = 0x10
PFB_COMPLEX_FLOAT64 = 0x20
PFB_COMPLEX_FLOAT32 = 0x40
PFB_COMPLEX_FLOAT16 = 0x80
PFB_IMAG_PART_FIRST = 0x100
PFB_EMULATE_FPGA = 0x200
PFB_STAGED = 0x400
PFB_SMART = 0x21D
PFB_FULL_PRECISION = 0x2D
VM_BF_FP32, VM_BF_MFMA_FP16, VM_BF_MFMA_BF16, VM_BF_MFMA_FP16_SINGLE = 0, 1, 2, 3
MWA_ALTITUDE_METRES = 377.827
SPEED_OF_LIGHT = 299792458.0


# ---------------------------------------------------------------- structs --
class Rfinput(C.Structure):
    _fields_ = [("input", C.c_uint32), ("ant", C.c_uint32), ("tile_id", C.c_uint32),
                ("tile_name", C.c_char_p), ("pol", C.c_char_p),
                ("electrical_length_m", C.c_double), ("north_m", C.c_double),
                ("east_m", C.c_double), ("height_m", C.c_double),
                ("vcs_order", C.c_uint32), ("subfile_order", C.c_uint32), ("flagged", C.c_bool),
                ("dipole_delays", C.POINTER(C.c_uint32))]


class CoarseChannel(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("corr_chan_number", "rec_chan_number", "gpubox_number",
                                          "chan_width_hz", "chan_start_hz", "chan_centre_hz",
                                          "chan_end_hz")]


class MetafitsMetadata(C.Structure):
    _fields_ = [("obs_id", C.c_uint32), ("mwa_version", C.c_int), ("sched_start_mjd", C.c_double),
                ("sched_start_gps_time_ms", C.c_uint64), ("num_ants", C.c_size_t),
                ("num_rf_inputs", C.c_size_t), ("rf_inputs", C.POINTER(Rfinput)),
                ("num_ant_pols", C.c_size_t), ("num_visibility_pols", C.c_size_t),
                ("num_metafits_coarse_chans", C.c_size_t),
                ("metafits_coarse_chans", C.POINTER(CoarseChannel)),
                ("coarse_chan_width_hz", C.c_uint32), ("num_volt_fine_chans_per_coarse", C.c_size_t),
                ("ra_tile_pointing_deg", C.c_double), ("dec_tile_pointing_deg", C.c_double),
                ("project_id", C.c_char_p), ("num_corr_fine_chans_per_coarse", C.c_size_t),
                ("corr_fine_chan_width_hz", C.c_uint32), ("good_time_gps_ms", C.c_uint64),
                ("num_metafits_timesteps", C.c_size_t), ("num_delays", C.c_size_t)]


class VoltageMetadata(C.Structure):
    _fields_ = [("mwa_version", C.c_int), ("num_samples_per_voltage_block", C.c_uint64),
                ("num_voltage_blocks_per_second", C.c_uint64),
                ("voltage_block_size_bytes", C.c_uint64)]


class Calibration(C.Structure):
    _fields_ = [("metafits", C.c_void_p), ("caldir", C.c_void_p), ("cal_type", C.c_int),
                ("ref_ant", C.c_void_p), ("keep_cross_terms", C.c_bool),
                ("phase_offset", C.c_double), ("phase_slope", C.c_double),
                ("use_bandpass", C.c_bool), ("custom_pq_correction", C.c_bool),
                ("flags_file", C.c_void_p), ("nflags", C.c_int), ("flagged_tilenames", C.c_void_p)]


class PfbFilter(C.Structure):
    _fields_ = [("coeffs", C.POINTER(C.c_double)), ("ncoeffs", C.c_int), ("ntaps", C.c_int),
                ("nchans", C.c_int), ("twiddles", C.c_void_p), ("type", C.c_int)]


class GeometricDelays(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("d_phi", C.c_void_p), ("npointings", C.c_size_t),
                ("nant", C.c_size_t), ("nchan", C.c_size_t),
                ("chan_freqs_hz", C.POINTER(C.c_double)), ("obs_metadata", C.c_void_p)]


class PrimaryBeam(C.Structure):
    _fields_ = [("B", C.c_void_p), ("beam", C.c_void_p), ("delays", C.c_void_p),
                ("amps", C.c_void_p), ("npointings", C.c_size_t), ("nant", C.c_size_t),
                ("npol", C.c_size_t), ("freq_hz", C.c_uint32), ("obs_metadata", C.c_void_p)]


class BeamGeom(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mean_ra", "mean_dec", "az", "el", "lmst", "fracmjd",
                                          "intmjd", "unit_N", "unit_E", "unit_H")]


class Vdifinfo(C.Structure):
    """struct vdifinfo (one pointing's VDIF output description)."""
    _fields_ = [("frame_length", C.c_int), ("frame_rate", C.c_int),
                ("samples_per_frame", C.c_size_t), ("dataarraylength", C.c_int),
                ("bits", C.c_int), ("nchan", C.c_int), ("chan_width", C.c_int),
                ("sample_rate", C.c_int), ("npol", C.c_int), ("iscomplex", C.c_int),
                ("threadid", C.c_int), ("stationid", C.c_char * 3), ("exp_name", C.c_char * 17),
                ("scan_name", C.c_char * 17), ("block_size", C.c_size_t),
                ("sizeof_buffer", C.c_size_t), ("sizeof_beam", C.c_size_t),
                ("b_scales", C.c_void_p), ("b_offsets", C.c_void_p),
                ("telescope", C.c_char * 24), ("source", C.c_char * 24),
                ("obs_mode", C.c_char * 8), ("fctr", C.c_double), ("ra_str", C.c_char * 16),
                ("dec_str", C.c_char * 16), ("BW", C.c_double), ("MJD_start", C.c_double),
                ("sec_offset", C.c_double), ("MJD_epoch", C.c_double),
                ("date_obs", C.c_char * 24), ("basefilename", C.c_char * 1024),
                ("got_scales", C.c_int)]


class VoltageFiles(C.Structure):
    """vm_voltage_files: the reader state of vmReadSecondFromFiles."""
    _fields_ = [("dir", C.c_char_p), ("md", C.POINTER(MetafitsMetadata)),
                ("vmd", C.POINTER(VoltageMetadata))]


class HostBuffer(C.Structure):
    _fields_ = [("buffer", C.c_void_p), ("buffer_size", C.c_size_t), ("read_ptr", C.c_void_p),
                ("read_size", C.c_size_t), ("copy_from_ptr", C.c_void_p),
                ("copy_to_ptr", C.c_void_p), ("copy_size", C.c_size_t), ("locked", C.c_bool)]


class ForwardPfb(C.Structure):
    _fields_ = [("d_htr_data", C.c_void_p), ("vcs_data", C.c_void_p), ("d_vcs_data", C.c_void_p),
                ("d_htr_size", C.c_size_t), ("htr_stride", C.c_size_t), ("vcs_size", C.c_size_t),
                ("d_vcs_size", C.c_size_t), ("vcs_stride", C.c_size_t),
                ("char2s_per_second", C.c_size_t), ("bytes_per_block", C.c_size_t),
                ("chunks_per_second", C.c_size_t), ("chunk", C.c_size_t),
                ("ninputs_per_cufft_batch", C.c_int), ("cufft_batch_size", C.c_int),
                ("d_weighted_overlap_add", C.c_void_p), ("weighted_overlap_add_size", C.c_size_t),
                ("filter_coeffs", C.c_void_p), ("d_filter_coeffs", C.c_void_p),
                ("nspectra", C.c_int), ("nspectra_per_chunk", C.c_int), ("M", C.c_int),
                ("K", C.c_int), ("I", C.c_int), ("P", C.c_int), ("i_output_idx", C.c_void_p),
                ("d_i_output_idx", C.c_void_p), ("flags", C.c_int), ("plan", C.c_void_p),
                ("d_flt_out", C.c_void_p), ("filter_fits_int16", C.c_int)]


class PsrfitsSub(C.Structure):
    _fields_ = [("dat_offsets", C.c_void_p), ("dat_scales", C.c_void_p), ("data", C.c_void_p),
                ("bytes_per_subint", C.c_int), ("dat_freqs", C.c_void_p),
                ("dat_weights", C.c_void_p)] + \
        [(n, C.c_double) for n in ("tsubint", "offs", "lst", "ra", "dec", "glon", "glat")] + \
        [(n, C.c_float) for n in ("feed_ang", "pos_ang", "par_ang", "tel_az", "tel_zen")]


class PsrfitsHdr(C.Structure):
    _fields_ = [(n, C.c_char * k) for n, k in (
        ("project_id", 24), ("obs_mode", 8), ("observer", 24), ("telescope", 24),
        ("frontend", 24), ("backend", 24), ("source", 24), ("date_obs", 64), ("poln_type", 8),
        ("track_mode", 16), ("cal_mode", 8), ("feed_mode", 8), ("ra_str", 16), ("dec_str", 16))] + \
        [(n, C.c_double) for n in ("scanlen", "dt", "fctr", "BW", "df", "ra2000", "dec2000",
                                   "azimuth", "zenith_ang", "beam_FWHM", "start_lst", "MJD_epoch",
                                   "fd_sang", "fd_xyph", "chan_dm")] + \
        [(n, C.c_int) for n in ("npol", "nchan", "nbits", "nsblk", "summed_polns", "rcvr_polns",
                                "fd_hand", "be_phase", "start_day", "start_sec", "scan_number")]


class Psrfits(C.Structure):
    """vm_psrfits: PSRFITS header, one sub-integration and the writer state."""
    _fields_ = [("sub", PsrfitsSub), ("hdr", PsrfitsHdr), ("basefilename", C.c_char * 1024),
                ("rows_per_file", C.c_int), ("filenum", C.c_int), ("tot_rows", C.c_int),
                ("rows_in_file", C.c_int), ("fptr", C.c_void_p), ("naxis2_pos", C.c_long),
                ("data_start", C.c_long)]


class MpiPsrfits(C.Structure):
    # mpi_psrfits { vm_psrfits coarse_chan_pf; int writer_id; vm_psrfits spliced_pf }
    _fields_ = [("coarse_chan_pf", Psrfits), ("writer_id", C.c_int), ("spliced_pf", Psrfits)]


class GpuIpfbArrays(C.Structure):
    _fields_ = [("ntaps", C.c_int), ("in_size", C.c_int), ("ft_size", C.c_int),
                ("out_size", C.c_int)] + [(n, C.c_void_p) for n in (
                    "in_real", "in_imag", "ft_real", "ft_imag", "d_in_real", "d_in_imag",
                    "d_ft_real", "d_ft_imag", "d_out")]


_P = C.c_void_p
_SZ = C.c_size_t


class VcsbeamContext(C.Structure):
    _fields_ = [
        ("use_mpi", C.c_bool), ("mpi_size", C.c_int), ("mpi_rank", C.c_int),
        ("ncoarse_chans", C.c_int), ("coarse_chan_idx", C.c_int), ("writer", C.c_int),
        ("obs_context", _P), ("obs_metadata", C.POINTER(MetafitsMetadata)), ("vcs_context", _P),
        ("vcs_metadata", C.POINTER(VoltageMetadata)), ("cal_context", _P), ("cal_metadata", _P),
        ("obs_context_legacy", _P), ("obs_metadata_legacy", _P),
        ("datadir", _P), ("filenames", _P), ("nfiles", C.c_int), ("nfiletimes", C.c_uint),
        ("seconds_per_file", C.c_uint),
        ("cal", Calibration),
        ("ras_hours", _P), ("decs_degs", _P),
        ("analysis_filter", C.POINTER(PfbFilter)), ("synth_filter", C.POINTER(PfbFilter)),
        ("fpfb", C.POINTER(ForwardPfb)),
        ("v", C.POINTER(HostBuffer)), ("d_v", _P), ("v_size_bytes", _SZ), ("d_v_size_bytes", _SZ),
        ("S", _P), ("d_S", _P), ("S_size_bytes", _SZ), ("d_S_size_bytes", _SZ),
        ("e", _P), ("d_e", _P), ("e_size_bytes", _SZ), ("d_e_size_bytes", _SZ),
        ("J", _P), ("d_J", _P), ("J_size_bytes", _SZ), ("d_J_size_bytes", _SZ),
        ("Jv_P", _P), ("d_Jv_P", _P), ("Jv_Q", _P), ("d_Jv_Q", _P), ("Jv_size_bytes", _SZ),
        ("d_Jv_size_bytes", _SZ),
        ("D", _P), ("d_D", _P), ("D_size_bytes", _SZ), ("d_D_size_bytes", _SZ),
        ("polP_idxs", _P), ("d_polP_idxs", _P), ("polQ_idxs", _P), ("d_polQ_idxs", _P),
        ("pol_idxs_size_bytes", _SZ), ("d_pol_idxs_size_bytes", _SZ),
        ("gdelays", GeometricDelays), ("pb", PrimaryBeam),
        ("npointing", C.c_uint),
        ("max_gpu_mem_bytes", _SZ), ("chunks_per_second", C.c_uint32),
        ("num_coarse_chans_to_process", C.c_int), ("coarse_chan_idxs_to_process", _P),
        ("cal_coarse_chan_idxs_to_process", _P),
        ("num_gps_seconds_to_process", C.c_int), ("gps_seconds_to_process", _P),
        ("output_fine_channels", C.c_bool), ("output_coarse_channels", C.c_bool),
        ("current_gps_idx", _SZ),
        ("do_forward_pfb", C.c_bool), ("do_inverse_pfb", C.c_bool),
        ("sample_rate", C.c_int), ("fine_sample_rate", C.c_int), ("bytes_per_second", C.c_int),
        ("num_not_flagged", C.c_int), ("nchan", C.c_int), ("nfine_chan", C.c_int),
        ("out_nstokes", C.c_int),
        ("vhdr", C.c_uint32 * 8), ("vf", _P),
        ("d_offsets", _P), ("offsets", _P), ("d_scales", _P), ("scales", _P),
        ("d_Cscaled", _P), ("Cscaled", _P),
        ("offsets_size", _SZ), ("scales_size", _SZ), ("Cscaled_size", _SZ),
        ("chunk_to_load", C.c_int), ("datatype", C.c_int),
        ("streams", _P), ("log", _P),
        ("log_message", C.c_char * 1024), ("error_message", C.c_char * 1024),
        ("dev", _P), ("beamform_form", C.c_int), ("read_second", _P), ("read_second_user", _P),
        ("bf_err", _P),
    ]


# Field lists matching vmStructLayout() in csrc/vm_host.c
LAYOUT_FIELDS = {
    "vcsbeam_context": (VcsbeamContext, ["obs_metadata", "cal", "fpfb", "v", "d_v", "S", "e", "J",
                                         "D", "polP_idxs", "gdelays", "pb", "npointing",
                                         "chunks_per_second", "gps_seconds_to_process",
                                         "do_inverse_pfb", "sample_rate", "out_nstokes", "vhdr",
                                         "d_offsets", "offsets_size", "chunk_to_load", "streams",
                                         "log_message", "dev", "beamform_form", "bf_err"]),
    "forward_pfb": (ForwardPfb, ["vcs_data", "d_vcs_data", "vcs_size", "d_weighted_overlap_add",
                                 "nspectra", "P", "flags", "d_flt_out", "filter_fits_int16"]),
    "MetafitsMetadata": (MetafitsMetadata, ["sched_start_mjd", "rf_inputs",
                                            "metafits_coarse_chans", "coarse_chan_width_hz",
                                            "project_id", "num_corr_fine_chans_per_coarse",
                                            "corr_fine_chan_width_hz", "good_time_gps_ms",
                                            "num_metafits_timesteps", "num_delays"]),
    "Rfinput": (Rfinput, ["pol", "electrical_length_m", "vcs_order", "flagged", "dipole_delays"]),
    "gpu_ipfb_arrays": (GpuIpfbArrays, ["in_real", "d_out"]),
    "vm_psrfits": (Psrfits, ["hdr", "basefilename", "fptr", "data_start"]),
    "vdifinfo": (Vdifinfo, ["samples_per_frame", "stationid", "block_size", "b_scales", "fctr",
                            "ra_str", "BW", "date_obs", "basefilename", "got_scales"]),
}

_CTX = C.POINTER(VcsbeamContext)

# name: (restype, argtypes)
_SIGS = {
    "vmInit": (_CTX, [C.c_bool]),
    "destroy_vcsbeam_context": (None, [_CTX]),
    "vmBindObsMetadata": (None, [_CTX, C.POINTER(MetafitsMetadata), C.POINTER(VoltageMetadata),
                                 C.c_int, C.c_uint32, C.c_int]),
    "vmReadNextSecondFrom": (C.c_int, [_CTX, C.c_void_p, C.c_size_t]),
    "vmReadNextSecond": (C.c_int, [_CTX]),
    "vmReadBufferCopyMargin": (C.c_int, [C.POINTER(HostBuffer)]),
    "vmSetFilter": (None, [_CTX, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]),
    "vmLoadFilter": (None, [_CTX, C.c_char_p, C.c_int, C.c_int]),
    "vmScaleFilterCoeffs": (None, [_CTX, C.c_int, C.c_double]),
    "vmInitForwardPFB": (None, [_CTX, C.c_int, C.c_int]),
    "vmSetMaxGPUMem": (None, [_CTX, C.c_int]),
    "vmSetNumPointings": (None, [_CTX, C.c_uint]),
    "vmSetPolIdxLists": (None, [_CTX]),
    "vmPushPolIdxLists": (None, [_CTX]),
    "vmCalcJ": (None, [_CTX]),
    "vmFormIncohBeamNoAlloc": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint, C.c_int,
                                         C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                         C.c_void_p]),
    "vmCalcB": (None, [_CTX, C.POINTER(BeamGeom)]),
    "vmCalcJonesAndDelays": (None, [_CTX, C.c_void_p, C.c_void_p, C.POINTER(BeamGeom)]),
    "hash_dipole_config": (C.c_int, [C.c_void_p]),
    "calc_normalised_beam_response": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_bool]),
    "create_delays_amps_from_metafits": (None, [C.POINTER(MetafitsMetadata), C.c_void_p,
                                                C.c_void_p]),
    "free_delays_amps": (None, [C.POINTER(MetafitsMetadata), C.c_void_p, C.c_void_p]),
    "vmCalcPhi": (None, [_CTX, C.POINTER(BeamGeom)]),
    "vmCalcPhiDevice": (None, [_CTX, C.POINTER(BeamGeom)]),
    "vmCalcJDevice": (None, [_CTX]),
    "vmExpandJDevice": (None, [_CTX]),
    "vmCalcJDeviceChanB": (None, [_CTX, C.c_void_p]),
    "vmSurrogateCreate": (C.c_void_p, [C.c_uint64, C.c_double, C.c_double]),
    "vmSurrogateDestroy": (None, [C.c_void_p]),
    "vmSurrogateTrain": (C.c_double, [C.c_void_p, C.c_long, C.c_int, C.c_int, C.c_float,
                                      C.c_double, C.c_uint64, C.c_void_p]),
    "vmSurrogateGrad": (C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_void_p]),
    "vmSurrogateParams": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "vmSurrogateConfig": (None, [C.c_void_p, C.c_void_p]),
    "vmSurrogateJones": (C.c_long, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "vmPushD": (None, [_CTX]),
    "vmPushJ": (None, [_CTX]),
    "vmPushPhi": (None, [_CTX]),
    "vmCreateGeometricDelays": (None, [_CTX]),
    "vmCreatePrimaryBeam": (None, [_CTX]),
    "free_primary_beam": (None, [C.c_void_p]),
    "vmCreateCudaStreams": (None, [_CTX]),
    "vmDestroyCudaStreams": (None, [_CTX]),
    "vmCreateStatistics": (None, [_CTX, C.c_void_p]),
    "vmDestroyStatistics": (None, [_CTX]),
    "vmBeamformSecond": (None, [_CTX]),
    "vmBeamformChunk": (None, [_CTX]),
    "vmApplyJChunk": (None, [_CTX]),
    "vmBeamformChunkMFMA": (None, [_CTX]),
    "vmPullE": (None, [_CTX]),
    "vmPullS": (None, [_CTX]),
    "vmSendSToFits": (None, [_CTX, C.c_void_p]),
    "vmExecuteForwardPFB": (None, [_CTX]),
    "vmWritePFBOutputToFile": (None, [_CTX]),
    "vmUploadForwardPFBChunk": (None, [_CTX]),
    "vmWOLAChunk": (None, [_CTX]),
    "vmFPGARoundingChunk": (None, [_CTX]),
    "vmFFTChunk": (None, [_CTX]),
    "vmPackChunk": (None, [_CTX]),
    "vmDownloadForwardPFBChunk": (None, [_CTX]),
    "vmSynchronize": (None, [_CTX]),
    "vmGetStream": (C.c_void_p, [_CTX]),
    "cu_form_incoh_beam": (None, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint,
                                  C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "allocate_input_output_arrays": (None, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                            C.c_size_t]),
    "free_input_output_arrays": (None, [C.c_void_p, C.c_void_p]),
    "malloc_ipfb": (None, [C.POINTER(GpuIpfbArrays), C.POINTER(PfbFilter), C.c_int, C.c_int,
                           C.c_int]),
    "free_ipfb": (None, [C.POINTER(GpuIpfbArrays)]),
    "cu_load_ipfb_filter": (None, [C.POINTER(PfbFilter), C.POINTER(GpuIpfbArrays)]),
    "cu_invert_pfb": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                             C.POINTER(GpuIpfbArrays), C.c_void_p]),
    "vmInvertPFBDevice": (None, [_CTX, C.POINTER(GpuIpfbArrays), C.c_int, C.c_void_p]),
    "vmInvertPFBDeviceVDIF": (None, [_CTX, C.POINTER(GpuIpfbArrays), C.c_int, C.c_void_p]),
    "vmInvertPFBDeviceVDIFBatch": (None, [C.POINTER(_CTX), C.POINTER(C.POINTER(GpuIpfbArrays)),
                                          C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "vmRequantVDIFDevice": (None, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "create_data_buffer_fine": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "prepare_data_buffer_fine": (None, [C.c_void_p, _CTX, C.c_size_t]),
    "vmSpliceSToFits": (None, [C.c_void_p, C.c_int, C.POINTER(Psrfits)]),
    "populate_spliced_psrfits_header": (None, [_CTX, C.POINTER(Psrfits), C.c_int, C.c_int,
                                               C.POINTER(BeamGeom), C.c_char_p, C.c_bool]),
    "populate_psrfits_header": (None, [_CTX, C.POINTER(Psrfits), C.c_int, C.c_int,
                                       C.POINTER(BeamGeom), C.c_char_p, C.c_bool]),
    "free_psrfits": (None, [C.POINTER(Psrfits)]),
    "vmWritePsrfitsSubint": (C.c_int, [C.POINTER(Psrfits)]),
    "vmClosePsrfits": (None, [C.POINTER(Psrfits)]),
    "vmEqToGal": (None, [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "vmLoadMetafits": (C.c_int, [C.c_char_p, C.POINTER(MetafitsMetadata),
                                 C.POINTER(VoltageMetadata)]),
    "vmFreeMetafits": (None, [C.POINTER(MetafitsMetadata)]),
    "vmReadSecondFromFiles": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_size_t]),
    "vmParsePointingFile": (None, [_CTX, C.c_char_p]),
    "parallactic_angle_correction": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "vmLoadObsMetafits": (None, [_CTX, C.c_char_p]),
    "vmLoadCalMetafits": (None, [_CTX, C.c_char_p]),
    "vmGetVoltageMetadata": (None, [_CTX]),
    "vmBindObsData": (None, [_CTX, C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_int,
                             C.c_char_p]),
    "vmBindCalibrationData": (None, [_CTX, C.c_char_p, C.c_int, C.c_bool, C.c_char_p]),
    "vmReadCalibration": (None, [_CTX]),
    "vmPrintTitle": (None, [_CTX, C.c_char_p]),
    "vmGetVoltFilename": (None, [_CTX, C.c_uint, C.c_uint64, C.c_char_p]),
    "vmGetLegacyVoltFilename": (None, [_CTX, C.c_uint, C.c_uint64, C.c_char_p]),
    "get_relative_gps": (C.c_ulong, [C.POINTER(MetafitsMetadata), C.c_long]),
    "parse_begin_string": (C.c_ulong, [C.POINTER(MetafitsMetadata), C.c_char_p]),
    "parse_coarse_chan_string": (C.c_size_t, [C.POINTER(MetafitsMetadata), C.c_char_p]),
    "calc_beam_geom": (None, [C.c_double, C.c_double, C.c_double, C.POINTER(BeamGeom)]),
    "dec2hms": (None, [C.c_char_p, C.c_double, C.c_int]),
    "mjd2lst": (None, [C.c_double, C.POINTER(C.c_double)]),
    "parse_ra": (C.c_double, [C.c_char_p]),
    "parse_dec": (C.c_double, [C.c_char_p]),
    "vmGMST": (C.c_double, [C.c_double]),
    "vmNutation": (None, [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_double)]),
    "vmMeanToApparent": (None, [C.c_double, C.c_double, C.c_double, C.c_int,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "vmHaDecToAzEl": (None, [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                             C.POINTER(C.c_double)]),
    "vmLoadOffringaSolution": (None, [_CTX]),
    "vmLoadRTSSolution": (None, [_CTX]),
    "vmApplyCalibrationCorrections": (None, [_CTX]),
    "vmSetCustomTileFlags": (None, [_CTX]),
    "parse_calibration_correction_file": (None, [C.c_uint32, C.POINTER(Calibration)]),
    "tilename_is_flagged": (C.c_bool, [C.c_char_p, C.POINTER(Calibration)]),
    "vmParseFlaggedTilenamesFile": (None, [C.c_char_p, C.POINTER(Calibration)]),
    "init_calibration": (None, [C.POINTER(Calibration)]),
    "free_calibration": (None, [C.POINTER(Calibration)]),
    "remove_reference_phase": (None, [C.c_void_p, C.c_void_p]),
    "zero_PQ_and_QP": (None, [C.c_void_p]),
    "vmVDIFFramesDevice": (None, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_void_p]),
    "vmPopulateVDIFHeader": (None, [_CTX, C.POINTER(BeamGeom), C.c_double, C.c_double]),
    "vdif_write_second": (None, [C.POINTER(Vdifinfo), C.c_void_p, C.c_void_p]),
    "vdif_write_data": (None, [C.POINTER(Vdifinfo), C.c_void_p]),
    "float2int8_trunc": (None, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_void_p]),
    "to_offset_binary": (None, [C.c_void_p, C.c_int]),
    "calc_array_factor": (C.c_double, [C.POINTER(MetafitsMetadata), C.c_uint32,
                                       C.POINTER(BeamGeom), C.POINTER(BeamGeom)]),
    "vmStructLayout": (C.c_int, [C.c_char_p, C.POINTER(C.c_size_t), C.c_int]),
    "vmSetKernelTiming": (None, [_CTX, C.c_int]),
    "vmTuneBeamformer": (C.c_int, [_CTX, C.c_int, C.POINTER(C.c_float)]),
    "vmSetBeamformerVariant": (None, [_CTX, C.c_int, C.c_int]),
    "vmSetInversePFBPasses": (None, [C.c_int]),
    "vmBeamformerSelection": (C.c_int, [_CTX, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vmBeamformerVariantInfo": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]),
    "vmBeamformerVariantKernel": (C.c_char_p, [C.c_int]),
    "vmKernelTimeMs": (C.c_double, [_CTX, C.c_char_p, C.POINTER(C.c_long)]),
    "vmDeviceMalloc": (C.c_void_p, [C.c_size_t]),
    "vmDeviceFree": (None, [C.c_void_p]),
    "vmMemcpyHtoD": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "vmMemcpyDtoH": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "vmMemsetDevice": (None, [C.c_void_p, C.c_int, C.c_size_t]),
    "vmDeviceSynchronize": (None, []),
    "vmDeviceCopy": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "vmMFMAPeakTFLOPS": (C.c_double, []),
    "vmMemcpy2DDtoH": (None, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                              C.c_size_t]),
    "vmSynthMWAXDevice": (None, [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_int, C.c_void_p]),
    "vmBatchCreate": (C.c_void_p, [C.POINTER(_CTX), C.c_int]),
    "vmBatchBeamformSecond": (None, [C.c_void_p, C.c_int]),
    "vmBatchTune": (C.c_float, [C.c_void_p, C.c_int]),
    "vmBatchSetSchedule": (None, [C.c_void_p, C.c_int, C.c_int]),
    "vmBatchStageTimes": (None, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
    "vmBatchInfo": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int)]),
    "vmBatchActiveTiles": (C.c_int, [C.c_void_p, C.c_int]),
    "vmBatchDestroy": (None, [C.c_void_p]),
    "vmDeviceCount": (C.c_int, []),
    "vmEventCreate": (C.c_void_p, []),
    "vmEventRecord": (None, [C.c_void_p, C.c_void_p]),
    "vmEventElapsedMs": (C.c_float, [C.c_void_p, C.c_void_p]),
    "vmEventDestroy": (None, [C.c_void_p]),
    "vmLibraryVersion": (C.c_char_p, []),
}

for _n in ("vmMallocVDevice", "vmMallocJVDevice", "vmMallocEHost", "vmMallocEDevice",
           "vmMallocSHost", "vmMallocSDevice", "vmMallocJHost", "vmMallocJDevice", "vmMallocDHost",
           "vmMallocDDevice", "vmMallocPQIdxsHost", "vmMallocPQIdxsDevice", "vmFreeVDevice",
           "vmFreeJVDevice", "vmFreeEHost", "vmFreeEDevice", "vmFreeSHost", "vmFreeSDevice",
           "vmFreeJHost", "vmFreeJDevice", "vmFreeDHost", "vmFreeDDevice", "vmFreePQIdxsHost",
           "vmFreePQIdxsDevice", "vmMallocJVHost", "vmFreeJVHost", "vmMallocVHost",
           "vmFreeVHost"):
    _SIGS[_n] = (None, [_CTX])

_lib = None
_synth = None


def load():
    """Load the C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} not built: run python mwa-vcsbeam_amd/build.py")
    os.environ.setdefault("VCSBEAM_RUNTIME_DIR", RUNTIME_DIR)
    lib = C.CDLL(LIBPATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def synth():
    """Load the synthetic-input generator library."""
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(SYNTHPATH):
        raise RuntimeError(f"{SYNTHPATH} not built: run python mwa-vcsbeam_amd/build.py")
    s = C.CDLL(SYNTHPATH)
    s.vmsynth_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_void_p]
    s.vmsynth_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_void_p]
    s.vmsynth_mwax_voltages.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_int, C.c_void_p]
    s.vmsynth_legacy_int4.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_void_p]
    s.vmsynth_kaiser_filter.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_void_p]
    for f in ("vmsynth_uniform", "vmsynth_normal", "vmsynth_mwax_voltages",
              "vmsynth_legacy_int4", "vmsynth_kaiser_filter"):
        getattr(s, f).restype = None
    _synth = s
    return s


class DeviceArray:
    """A device allocation owned by Python (freed on garbage collection)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = load().vmDeviceMalloc(self.nbytes)

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        load().vmMemcpyHtoD(d.ptr, a.ctypes.data, a.nbytes)
        return d

    def to_numpy(self, dtype, shape):
        import numpy as np
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        load().vmMemcpyDtoH(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def zero(self):
        load().vmMemsetDevice(self.ptr, 0, self.nbytes)

    def __del__(self):
        if getattr(self, "ptr", None):
            load().vmDeviceFree(self.ptr)
            self.ptr = None


def struct_layout(name):
    """(C sizes/offsets from the library, ctypes sizes/offsets) for a struct."""
    lib = load()
    buf = (C.c_size_t * 64)()
    n = lib.vmStructLayout(name.encode(), buf, 64)
    cls, fields = LAYOUT_FIELDS[name]
    mine = [C.sizeof(cls)] + [getattr(cls, f).offset for f in fields]
    return list(buf[:n]), mine
