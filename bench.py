#!/usr/bin/env python3
"""Benchmark of the MI355X voltage-beamforming path.

Workload (one "step"): one coarse channel x 1 s of 128-tile x 2-pol MWAX
voltages, already channelised by the fused forward PFB and resident in HBM
(VM_FLT, 128 fine channels x 10,000 samples x 256 inputs), beamformed towards
512 directions uniform within 15 deg of boresight: per-second FP64 phases and
Jones inverses on the GPU, weights, the beamformer and full-Stokes detection.
Metric: beam-samples/s = 512 beams x 128 channels x 10,000 samples / step time.

Also reported: the fused forward-PFB kernel on the same second (HBM roofline),
the dominant kernel's roofline, and the FP64 CPU oracle timed on a bounded
sample of the same beamforming work on this host's cores.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "mwa-vcsbeam_amd")
METRIC = "beam-samples/s at 128 tiles x 512 beams, MFMA vs FP32 at <=1e-3 rel Stokes-I err"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP16_PEAK_TFLOPS = 2500.0


def load_pkg():
    if "vcsbeam_amd" in sys.modules:
        return sys.modules["vcsbeam_amd"]
    spec = importlib.util.spec_from_file_location("vcsbeam_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vcsbeam_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def event_time(lib, stream, fn, reps):
    """Average ms per call of fn() measured with HIP events on `stream`."""
    a, b = lib.vmEventCreate(), lib.vmEventCreate()
    lib.vmEventRecord(a, stream)
    for _ in range(reps):
        fn()
    lib.vmEventRecord(b, stream)
    ms = lib.vmEventElapsedMs(a, b) / reps
    lib.vmEventDestroy(a)
    lib.vmEventDestroy(b)
    return ms


def cpu_baseline(sess, obs, az, za, target_s=10.0):
    """FP64 oracle beamformer on the first n samples of the same second."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the checker, used here only as the CPU baseline
    from vcsbeam_amd.workloads import unit_vectors
    vfl = sess.voltages()  # [c][i][s]
    D = obs.calibration()
    B = np.zeros((len(az), obs.ntiles, 2, 2), complex)
    B[..., 0, 0] = B[..., 1, 1] = 1
    J = orc.calc_J(D, B)
    E, N, H, L = obs.positions()
    phi = orc.calc_phi(E, N, H, L, unit_vectors(az, za), obs.freqs)
    polQ, polP = obs.pol_idx()
    invw = 1.0 / (2 * int((~obs.flag_mask).sum()))

    def run(n):
        v = np.ascontiguousarray(vfl[:, :, :n].transpose(2, 0, 1)).astype(np.complex128)
        t0 = time.perf_counter()
        S, _ = orc.beamform(v, J, phi, polQ, polP, invw, nstokes=4)
        return time.perf_counter() - t0, S

    t, _ = run(2)
    n = int(max(2, min(400, round(target_s / max(t / 2, 1e-6)))))
    t, S = run(n)
    nbeam = len(az)
    return {"seconds": t, "nsamp": n, "S": S, "rate": nbeam * 128 * n / t,
            "cores": int(orc.lib().orc_num_threads())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--beams", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the beamforming path shards over coarse channels "
              "(one GPU per channel) and this build runs on a single MI355X; use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    lib = pkg.load()
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import Observation, directions_in_cap

    print("[bench] generating 1 s of 128-tile MWAX voltages", file=sys.stderr, flush=True)
    obs = Observation(ntiles=128, blocks_per_second=20)
    az, za = directions_in_cap(args.beams, 15.0)
    sess = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA,
                            nstokes=4, form=pkg.VM_BF_FP32)
    sess.read(0)
    stream = sess.stream
    print("[bench] forward PFB", file=sys.stderr, flush=True)
    sess.pfb()
    pfb_ms = event_time(lib, stream, lambda: lib.vmWOLAChunk(sess.vm), 5)
    nsamp_in = obs.ninput * obs.nspectra * 128
    pfb_bytes = nsamp_in * 10  # 2 B in + 8 B out per input complex sample
    sess.geometry()
    sess.beamform()

    def step():
        sess.geometry()
        sess.beamform()

    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    lib.vmDeviceSynchronize()
    print("[bench] timed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    lib.vmDeviceSynchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    beam_samples = args.beams * 128 * obs.nspectra
    value = beam_samples / ((t1 - t0) / args.steps)

    bf_ms = event_time(lib, stream, lambda: lib.vmBeamformChunk(sess.vm), 3)
    flops = 4096.0 * beam_samples  # complex GEMM: 8 x 2 pols x 256 tile-pols per beam-sample
    achieved = flops / (bf_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "k_beamform_fp32", "kernel_ms": round(bf_ms, 4)}
    S_gpu = sess.stokes()

    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "beam-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian 8+8-bit voltages, sigma 16)",
        "config": {"workload": "128 tiles x 2 pols, 1 coarse channel x 1 s, 512 beams within "
                               "15 deg of boresight, plain FP32 beamformer, Stokes IQUV",
                   "tiles": 128, "beams": args.beams, "fine_channels": 128,
                   "samples": obs.nspectra, "form": "fp32"},
        "roofline": roofline,
        "stages": {"pfb_fused_ms_per_ch_s": round(pfb_ms, 4),
                   "pfb_GBps": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9, 1),
                   "pfb_hbm_frac": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu:
        print("[bench] CPU baseline (FP64 oracle)", file=sys.stderr, flush=True)
        cb = cpu_baseline(sess, obs, az, za)
        n = cb["nsamp"]
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        errs = [orc.rel_err(S_gpu[:, :n, st, :], cb["S"][:, :n, st, :]) for st in range(4)]
        result["cpu_baseline"] = {"value": round(cb["rate"], 1), "unit": "beam-samples/s",
                                  "cores": cb["cores"], "kind": "port",
                                  "sample": f"{args.beams} beams x 128 ch x {n} samples of the same "
                                            f"second, FP64 oracle beamformer ({cb['seconds']:.1f} s)"}
        result["parity"] = {"stokes_rel_err_vs_fp64": [float(f"{e:.3e}") for e in errs]}
    sess.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
