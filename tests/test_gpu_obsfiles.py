"""End to end from files: three seconds of an MWAX observation (8 tiles x 2
pols, 0.1 s per second of data) written to two 8-second subfiles as the
instrument lays them out, read through vmReadSecondFromFiles across the
subfile boundary (GPS seconds 6, 7 and 8 of the observation), then through
the forward PFB: bit for bit the output of the same seconds fed from
memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_pfb_from_subfiles_equals_from_memory(pkg, gpu_lib, tmp_path):
    from vcsbeam_amd.apps import fine_pfb_offline
    from vcsbeam_amd.workloads import Observation
    obs = Observation(ntiles=8, blocks_per_second=2, nseconds=3)
    gps0 = obs.md.obs_id
    bps, rec = obs.vmd.num_voltage_blocks_per_second, obs.cc[0].rec_chan_number
    secs = [obs.second(t) for t in range(3)]
    files = {gps0: np.full((1 + 8 * bps,) + secs[0].shape[1:], 7, np.int8),
             gps0 + 8: np.full((1 + 8 * bps,) + secs[0].shape[1:], -7, np.int8)}
    for t, gps in enumerate(range(gps0 + 6, gps0 + 9)):
        sub = gps0 + (gps - gps0) // 8 * 8
        s = gps - sub
        files[sub][1 + bps * s:1 + bps * (s + 1)] = secs[t]
    for sub, blocks in files.items():
        hdr = f"HDR_SIZE 4096\nNINPUTS {obs.ninput}\n".encode().ljust(4096, b"\0")
        (tmp_path / f"{gps0}_{sub}_{rec}.sub").write_bytes(hdr + blocks.tobytes())
    flags = pkg.PFB_COMPLEX_FLOAT32 | pkg.PFB_EMULATE_FPGA | pkg.PFB_MALLOC_HOST_OUTPUT | 0x0C
    print("PFB of 3 seconds from subfiles", flush=True)
    got = fine_pfb_offline(obs, nseconds=3, pfb_flags=flags, voltage_dir=tmp_path,
                           first_gps=gps0 + 6)
    print("PFB of 3 seconds from memory", flush=True)
    want = fine_pfb_offline(obs, nseconds=3, pfb_flags=flags)
    for t in range(3):
        np.testing.assert_array_equal(got[t].view(np.uint32), want[t].view(np.uint32))
