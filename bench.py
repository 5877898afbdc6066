#!/usr/bin/env python3
"""Benchmark of the MI355X voltage-beamforming path.

Workload (one "step"): one coarse channel x 1 s of 128-tile x 2-pol MWAX
voltages, already channelised by the fused forward PFB and resident in HBM
(VM_FLT, 128 fine channels x 10,000 samples x 256 inputs), beamformed towards
512 directions uniform within 15 deg of boresight: per-second FP64 phases and
Jones inverses on the GPU, weights, the beamformer and full-Stokes detection.
Metric: beam-samples/s = 512 beams x 128 channels x 10,000 samples / step time,
for the precision MFMA form (FP16 voltages, split hi+lo FP16 weights, FP32
accumulation); the plain FP32 form is timed beside it on the same data.

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


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def _sized(fn, n_probe, target_s, n_max, quantum=1):
    """Run fn(n_probe), then fn(n) with n sized so the run takes about
    target_s; returns (n, seconds, result of the last run)."""
    t, _ = fn(n_probe)
    n = int(max(n_probe, min(n_max, target_s / max(t / n_probe, 1e-9))))
    n = max(quantum, n // quantum * quantum)
    t, r = fn(n)
    return n, t, r


def cpu_baseline(sess, obs, az, za, budget_s=12.0):
    """The reference arithmetic on the host cores, stage by stage, under the
    reference's stopwatch names (src/metadata.c:127-139), each stage run on a
    bounded sample of one coarse channel-second of the 128-tile workload and
    its time scaled to the whole channel-second: the forward PFB (pfb-wola,
    pfb-round, pfb-fft, pfb-pack), the FP64 geometric phases of every beam
    (delay), the FP64 beamformer (calc), and for one beam the FP32
    direct-form inverse PFB (ipfb) and its VDIF requantisation (vdif).  Run
    with all OpenMP threads and again with one."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the oracle, used here only as the CPU baseline
    from vcsbeam_amd.workloads import SEED_DATA, mwax_voltages, unit_vectors
    K, P = 128, 12
    nspc, nbeam = obs.nspectra, len(az)
    vfl = sess.voltages()  # [c][i][s] complex64
    D = obs.calibration()
    B = np.zeros((nbeam, obs.ntiles, 2, 2), complex)
    B[..., 0, 0] = B[..., 1, 1] = 1
    J = orc.calc_J(D, B)
    E, N, H, L = obs.positions()
    unit = unit_vectors(az, za)
    polQ, polP = obs.pol_idx()
    invw = 1.0 / (2 * int((~obs.flag_mask).sum()))
    h16 = obs.filter.astype(np.int32)
    out_idx = obs.out_idx()
    # PFB input: a zero margin block and the first blocks of second 0
    # (orc_wola reads (nspec + 500) K samples of the margin-first buffer)
    nb = 3
    htr = np.ascontiguousarray(np.concatenate(
        [np.zeros((1, obs.ninput, 64000, 2), np.int8),
         mwax_voltages(obs.ninput, nb - 1, first_block=0, seed=SEED_DATA + obs.seed_offset)]))
    nspec_max = nb * 64000 // K - 500
    rng = np.random.default_rng(7)
    synth = np.ascontiguousarray(rng.standard_normal(K * P) / K)
    table = orc.cpu_ipfb_table(synth)
    vdif_in = (rng.standard_normal(nspc * K * 2 * 2) * 30).astype(np.float32)
    phi_all = None
    S = None
    nthreads = int(orc.lib().orc_num_threads())
    legs = {}
    for threads in (nthreads, 1):
        orc.set_threads(threads)
        share = budget_s * (1.0 if threads == nthreads else 0.5)
        stages, sample = {}, {}

        def pfb(n):
            t0 = time.perf_counter()
            _, st = orc.cpu_pfb(htr, obs.ninput, n, h16, out_idx)
            return time.perf_counter() - t0, st
        n, _, st = _sized(pfb, 2, 0.15 * share, nspec_max)
        for k, v in st.items():
            stages[k] = v * nspc / n
        sample["pfb"] = f"{n} of {nspc} spectra x {obs.ninput} inputs"

        def delay(nb_):
            t0 = time.perf_counter()
            ph = orc.cpu_delay(E, N, H, L, unit[:nb_], obs.freqs)
            return time.perf_counter() - t0, ph
        n, t, ph = _sized(delay, min(4, nbeam), 0.05 * share, nbeam)
        stages["delay"] = t * nbeam / n
        sample["delay"] = f"{n} of {nbeam} beams x {obs.ntiles} tiles x {K} channels"
        if phi_all is None:
            phi_all = ph if n == nbeam else orc.calc_phi(E, N, H, L, unit, obs.freqs)

        def calc(ns):
            v = np.ascontiguousarray(vfl[:, :, :ns].transpose(2, 0, 1)).astype(np.complex128)
            t0 = time.perf_counter()
            out, _ = orc.beamform(v, J, phi_all, polQ, polP, invw, nstokes=4)
            return time.perf_counter() - t0, out
        n, t, Sn = _sized(calc, 2, 0.6 * share, 400)
        stages["calc"] = t * nspc / n
        sample["calc"] = f"{nbeam} beams x {K} channels x {n} of {nspc} samples"
        if threads == nthreads:
            S, n_calc, t_calc = Sn, n, t

        def ipfb(nrow):
            nout = nrow * K
            buf = (rng.standard_normal((nrow + 2 * P + 2, K, 2))
                   + 1j * rng.standard_normal((nrow + 2 * P + 2, K, 2))).astype(np.complex64)
            t0 = time.perf_counter()
            out = orc.cpu_ipfb(buf, P * K, nout, table, P)
            return time.perf_counter() - t0, out
        n, t, _ = _sized(ipfb, 1, 0.15 * share, nspc)
        stages["ipfb"] = t * nspc / n
        sample["ipfb"] = f"{n * K} of {nspc * K} output samples x 2 pols, one beam"
        t0 = time.perf_counter()
        orc.cpu_vdif(vdif_in)
        stages["vdif"] = time.perf_counter() - t0
        sample["vdif"] = f"{vdif_in.size} values, one beam (whole channel-second)"
        pfb_s = sum(stages[k] for k in ("pfb-wola", "pfb-round", "pfb-fft", "pfb-pack"))
        legs[threads] = {
            "threads": threads,
            "seconds_per_channel_second": {k: round(v, 4) for k, v in stages.items()},
            "beam_samples_per_s": round(nbeam * K * nspc / (stages["delay"] + stages["calc"]), 1),
            "pfb_input_samples_per_s": round(obs.ninput * nspc * K / pfb_s, 1),
            "end_to_end_s_per_channel_second_one_beam": round(
                pfb_s + (stages["delay"] + stages["calc"]) / nbeam + stages["ipfb"]
                + stages["vdif"], 3),
            "sample": sample}
    orc.set_threads(nthreads)
    return {"rate": legs[nthreads]["beam_samples_per_s"], "cores": nthreads, "S": S,
            "nsamp": n_calc, "seconds": t_calc, "legs": legs, "info": cpu_info()}


def kernel_ms(lib, sess, name):
    import ctypes as C
    n = C.c_long(0)
    ms = lib.vmKernelTimeMs(sess.vm, name.encode(), C.byref(n))
    return ms, n.value


def time_form(lib, sess, form, steps, warmup):
    """Wall-clock ms per step and live per-kernel event timing for one form."""
    sess.v.beamform_form = form
    sess.geometry()          # marks the weights dirty for the new form
    sess.beamform()
    for _ in range(warmup):
        sess.geometry()
        sess.beamform()
    lib.vmDeviceSynchronize()
    lib.vmSetKernelTiming(sess.vm, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        sess.geometry()
        sess.beamform()
    lib.vmDeviceSynchronize()
    t1 = time.perf_counter()
    lib.vmSetKernelTiming(sess.vm, 0)
    return (t1 - t0) * 1e3 / steps


COPY_BYTES = 1 << 30  # calibration copy for the memory counters


def rocprof_summary(dirs, title, command, out_md, out_json):
    """Summarise rocprofv3 CSV output (a --kernel-trace --stats run and
    --pmc FETCH_SIZE / WRITE_SIZE runs of the same command) into a short
    markdown table and a traffic record.  The counters are calibrated with the
    plain copy kernel that ran in the same command (vmDeviceCopy of
    COPY_BYTES): factor = COPY_BYTES / (counter KiB * 1024)."""
    import csv
    import glob
    stats, counters = [], {}
    for d in dirs:
        for f in glob.glob(os.path.join(d, "*_kernel_stats.csv")):
            stats += list(csv.DictReader(open(f)))
        for f in glob.glob(os.path.join(d, "*_counter_collection.csv")):
            for r in csv.DictReader(open(f)):
                name = r["Kernel_Name"].split("(")[0]
                counters.setdefault(r["Counter_Name"], {}).setdefault(name, []).append(
                    float(r["Counter_Value"]))
    lines = [f"# {title}", "", "Command (from the repository root, on one MI355X):", "",
             "```", command.strip(), "```", ""]
    if stats:
        lines += ["## Kernel trace statistics", "",
                  "| kernel | calls | avg (us) | total (ms) | % |", "|---|---:|---:|---:|---:|"]
        for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"])):
            nm = r["Name"].split("(")[0]
            lines.append(f"| `{nm}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                         f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")
        lines.append("")
    traffic = {}
    fet, wri = counters.get("FETCH_SIZE", {}), counters.get("WRITE_SIZE", {})
    cp = "k_stream_copy"
    if fet and wri and cp in fet and cp in wri:
        ff = COPY_BYTES / (sum(fet[cp]) / len(fet[cp]) * 1024)
        wf = COPY_BYTES / (sum(wri[cp]) / len(wri[cp]) * 1024)
        lines += ["## HBM traffic per dispatch", "",
                  f"Calibration: the 1 GiB copy kernel reads FETCH_SIZE x {ff:.3f} and "
                  f"WRITE_SIZE x {wf:.3f} of its true bytes.", "",
                  "| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | read (GB) | write (GB) |",
                  "|---|---:|---:|---:|---:|---:|"]
        for nm in sorted(set(fet) & set(wri)):
            fk = sum(fet[nm]) / len(fet[nm])
            wk = sum(wri[nm]) / len(wri[nm])
            rb, wb = fk * 1024 * ff, wk * 1024 * wf
            traffic[nm] = {"read_bytes": rb, "write_bytes": wb, "dispatches": len(fet[nm])}
            lines.append(f"| `{nm}` | {len(fet[nm])} | {fk:.4g} | {wk:.4g} | {rb / 1e9:.3f} | "
                         f"{wb / 1e9:.3f} |")
        lines.append("")
    other = {k: v for k, v in counters.items() if k not in ("FETCH_SIZE", "WRITE_SIZE")}
    if other:
        lines += ["## Other counters (mean per dispatch)", "", "| counter | kernel | value |",
                  "|---|---|---:|"]
        for c, per in sorted(other.items()):
            for nm, v in sorted(per.items()):
                lines.append(f"| {c} | `{nm}` | {sum(v) / len(v):.4g} |")
        lines.append("")
    open(out_md, "w").write("\n".join(lines))
    if out_json and traffic:
        json.dump({"title": title, "command": command.strip(), "kernels": traffic},
                  open(out_json, "w"), indent=1)


def committed_traffic(kernel, variant, order, form=1):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    summaries.  The beamformer GEMM's traffic depends on its tile shape and
    order (how often voltage fragments are re-read) and on the weight parts
    its form reads, so its records are kept per (form, variant, order):
    profiles/bench_traffic_v<variant>_o<order>.json for the split FP16 form,
    profiles/bench_traffic_f<form>_v<variant>_o<order>.json for the others."""
    tag = "" if form == 1 else f"f{form}_"
    f = os.path.join(ROOT, "profiles", f"bench_traffic_{tag}v{variant}_o{order}.json")
    if not os.path.exists(f):
        return None
    try:
        rec = json.load(open(f))["kernels"]
    except (ValueError, KeyError):
        return None
    for nm, t in rec.items():
        if nm.replace("void ", "").startswith(kernel):
            return {"bytes": round(t["read_bytes"] + t["write_bytes"]),
                    "read_bytes": round(t["read_bytes"]), "write_bytes": round(t["write_bytes"]),
                    "source": os.path.relpath(f, ROOT)}
    return None


def surrogate_stage(lib, nbeams=4096, nchan=128, steps=4000, batch=2048, nsamples=2_000_000,
                    feed=None):
    """Train the BF16 MLP primary-beam surrogate on the GPU, then produce Jones
    for nbeams directions (within 15 deg) x nchan fine channels with fallback
    to the analytic beam; errors are against the analytic (FP64) beam.  With
    feed = (pkg, session, obs, az, za) the surrogate's Jones for the session's
    pointings drive its MFMA beamformer (surrogate_feed)."""
    import ctypes as C
    s = lib.vmSurrogateCreate(7, 0.01, 50e-12)
    try:
        curve = np.zeros(steps // 100 + 1, np.float32)
        t0 = time.perf_counter()
        loss = lib.vmSurrogateTrain(s, nsamples, steps, batch, 2e-3, 20.0, 11, curve.ctypes.data)
        t_train = time.perf_counter() - t0
        rng = np.random.default_rng(5)
        za = np.arccos(1 - rng.uniform(0, 1, nbeams) * (1 - np.cos(np.radians(15.0))))
        az = rng.uniform(0, 2 * np.pi, nbeams)
        freq = 191.36e6 + 10e3 * np.arange(nchan)
        B = np.empty((nbeams, nchan, 4, 2))
        err = np.empty((nbeams, nchan), np.float32)
        args = (az.ctypes.data, za.ctypes.data, nbeams, freq.ctypes.data, nchan, B.ctypes.data)
        lib.vmSurrogateJones(s, 3, *args, err.ctypes.data)  # warm-up
        t0 = time.perf_counter()
        nfb = lib.vmSurrogateJones(s, 3, *args, err.ctypes.data)
        t_sur = time.perf_counter() - t0
        Bt = np.empty_like(B)
        lib.vmSurrogateJones(s, 0, az.ctypes.data, za.ctypes.data, nbeams, freq.ctypes.data,
                             nchan, Bt.ctypes.data, None)
        Bn = np.empty_like(B)
        lib.vmSurrogateJones(s, 1, az.ctypes.data, za.ctypes.data, nbeams, freq.ctypes.data,
                             nchan, Bn.ctypes.data, None)
        e_nom = np.sqrt(((Bn - Bt) ** 2).sum((2, 3)) / (Bt ** 2).sum((2, 3)))
        Bc = np.repeat(Bt[:, nchan // 2:nchan // 2 + 1], nchan, axis=1)
        e_cen = np.sqrt(((Bc - Bt) ** 2).sum((2, 3)) / (Bt ** 2).sum((2, 3)))
        fed = surrogate_feed(lib, s, *feed) if feed else None
        return {"feeding_mfma_beamformer": fed,
                "train_samples": nsamples, "train_steps": steps, "batch": batch,
                "train_s": round(t_train, 2), "final_loss": float(f"{loss:.3e}"),
                "jones_evaluations": nbeams * nchan,
                "infer_with_fallback_ms": round(t_sur * 1e3, 2),
                "fallbacks": int(nfb),
                "err_median": float(f"{np.median(err):.3e}"),
                "err_p99": float(f"{np.quantile(err, 0.99):.3e}"),
                "err_max": float(f"{err.max():.3e}"),
                "nominal_beam_err_median": float(f"{np.median(e_nom):.3e}"),
                "centre_frequency_beam_err_median": float(f"{np.median(e_cen):.3e}")}
    finally:
        lib.vmSurrogateDestroy(s)


def batch_stage(pkg, lib, nchan=24, ntiles=256, steps=3, nsamp_check=64, beams_check=4, form=1,
                half=False):
    """Batched uneven workload: nchan coarse channels x 1 s of ntiles tiles,
    per channel a seeded flagged-tile count ~ U{0..40} and beam count ~
    floor(exp U[0, ln 2049)) (log-uniform 1..2048) within 15 deg; forward
    PFB to FP32 (or to FP16, PFB_COMPLEX_FLOAT16, with `half`) into the
    shared buffer + the MFMA beamformer in `form` (FP16 voltages), Stokes IQUV, every channel in one persistent launch inside
    one replayed hipGraph; seconds generated on the device.  Per channel,
    beams_check beams x nsamp_check samples x 128 channels are checked
    against the FP64 oracle chain (oracle PFB of the same seeded voltages,
    then the FP64 beamformer), so the check covers the FP16 voltages too."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the checker (per-channel error only)
    from vcsbeam_amd.apps import BatchSession
    from vcsbeam_amd.workloads import (SEED_BEAMS, SEED_DATA, SEED_FLAGS, Observation,
                                       directions_in_cap, mwax_voltages, uniform, unit_vectors)
    u = uniform(SEED_FLAGS, 2 * nchan, offset=10 ** 8)
    nflag = np.minimum((u[0::2] * 41).astype(int), 40)
    nbeam = np.floor(np.exp(u[1::2] * np.log(2049.0))).astype(int)
    obs = [Observation(ntiles=ntiles, blocks_per_second=20, nflagged=int(nflag[g]),
                       rec_chan=109 + g, seed_offset=1000 * (g + 1), geometry_seed_offset=0)
           for g in range(nchan)]
    dirs = [directions_in_cap(int(nbeam[g]), 15.0, seed=SEED_BEAMS + 100 * (g + 1))
            for g in range(nchan)]
    t0 = time.perf_counter()
    flags = pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA
    if half:
        flags = (flags & ~pkg.PFB_TYPE_MASK) | pkg.PFB_COMPLEX_FLOAT16
    bs = BatchSession(obs, dirs, pfb_flags=flags, nstokes=4, form=form,
                      share_pfb_output=True, device_data=True)
    try:
        bs.read(0)
        t_setup = time.perf_counter() - t0
        print(f"[bench] batch: {nchan} channels, {int(nbeam.sum())} beams, set up in "
              f"{t_setup:.1f} s; autotune", file=sys.stderr, flush=True)
        tune_ms = bs.tune(reps=2)
        bs.run(use_graph=True)  # capture + first replay
        lib.vmDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bs.run(use_graph=True)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        st = bs.stage_times()
        info = bs.info()
        nact = [bs.active_tiles(g) for g in range(nchan)]
        nspc = obs[0].nspectra
        bsamp = int(nbeam.sum()) * 128 * nspc
        flops = sum(16.0 * nact[g] * nbeam[g] * 128 * nspc for g in range(nchan))
        wbytes = 2 if form == 3 else 4  # per weight: one FP16 part, or hi + lo
        gemm_bytes = sum(nbeam[g] * 128 * nspc * 16 + 128 * nact[g] * 2 * nspc * 4
                         + 128 * (2 * nbeam[g]) * (2 * nact[g]) * wbytes + 128 * nspc * 16
                         for g in range(nchan))
        print("[bench] batch: per-channel check against the FP64 oracle", file=sys.stderr,
              flush=True)
        per = []
        for g, ses in enumerate(bs.sessions):
            o, (az, za) = obs[g], dirs[g]
            # oracle PFB of the first nsamp_check spectra: margin block (zeros
            # in the first second) and the second's first block
            htr = np.ascontiguousarray(np.concatenate(
                [np.zeros((1, o.ninput, 64000, 2), np.int8),
                 mwax_voltages(o.ninput, 1, first_block=0, seed=SEED_DATA + o.seed_offset)]))
            v = orc.pfb_forward(htr, o.ninput, nsamp_check, o.filter.astype(np.int32),
                                o.out_idx(), fpga=True)  # [s][c][i]
            pk = min(beams_check, len(az))
            B = np.zeros((pk, o.ntiles, 2, 2), complex)
            B[..., 0, 0] = B[..., 1, 1] = 1
            E, N, H, L = o.positions()
            phi = orc.calc_phi(E, N, H, L, unit_vectors(az[:pk], za[:pk]), o.freqs)
            polQ, polP = o.pol_idx()
            Sref, _ = orc.beamform(v, orc.calc_J(o.calibration(), B), phi, polQ, polP,
                                   1.0 / (2 * int((~o.flag_mask).sum())), nstokes=4)
            got = ses.stokes_slice(0, pk, 0, nsamp_check)
            err = np.array([[orc.rel_err(got[p, :, k, :], Sref[p, :, k, :]) for k in range(4)]
                            for p in range(pk)]).max(axis=0)
            per.append({"beams": int(nbeam[g]), "flagged_tiles": int(nflag[g]),
                        "gemm_tiles": nact[g],
                        "max_beam_rel_err_IQUV": [float(f"{x:.3e}") for x in err]})
        wname = {1: "split hi+lo FP16 weights", 2: "BF16 voltages, split BF16 weights",
                 3: "one FP16 part per weight"}[form]
        return {"workload": f"{nchan} coarse channels x 1 s, {ntiles} tiles x 2 pols, flagged "
                            f"tiles U{{0..40}} and beams log-uniform 1..2048 per channel, PFB to "
                            f"{'FP16' if half else 'FP32'} + MFMA beamformer form {form} "
                            f"({wname}), Stokes IQUV, one persistent launch in one hipGraph",
                "value": round(bsamp / (ms * 1e-3), 1), "unit": "beam-samples/s",
                "beams_total": int(nbeam.sum()), "ms_per_second_of_data": round(ms, 2),
                "stages_eager_ms": {k: round(x, 2) for k, x in st.items()},
                "gemm_tflops_algorithmic": round(flops / (st["gemm_ms"] * 1e-3) / 1e12, 1),
                "gemm_GBps_algorithmic": round(gemm_bytes / (st["gemm_ms"] * 1e-3) / 1e9, 1),
                # 2 B in, 8 B (complex FP32) or 4 B (complex FP16) out per input-sample
                "pfb_hbm_frac": round(nchan * 2 * ntiles * nspc * 128 * (6 if half else 10)
                                      / (st["pfb_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "schedule": {**info, "tuned_gemm_ms": round(tune_ms, 2)},
                "setup_s": round(t_setup, 1),
                "max_channel_rel_err_I": max(p["max_beam_rel_err_IQUV"][0] for p in per),
                "per_channel": per}
    finally:
        bs.close()


def pipeline_stage(pkg, lib, nchan=24, steps=2):
    """Full plain pipeline: nchan coarse channels x 1 s of 128 tiles (4
    flagged), fused FP32 forward PFB, FP64 device geometry, the plain FP32
    tied-array beam towards one pointing (za 20 deg), Stokes I (the PSRFITS
    product, 100 us samples), inverse PFB back to 1.28 MHz complex float32 with
    its 8-bit VDIF requantisation in the same kernel (one launch over all
    channels), all device-resident; the channels' PFB and beamformer run in
    turn and share one forward-PFB output buffer owned by the caller."""
    from vcsbeam_amd.apps import TiedArraySession, device_second, invert_batch
    from vcsbeam_amd.workloads import Observation
    obs = [Observation(ntiles=128, blocks_per_second=20, nflagged=4, rec_chan=109 + g,
                       seed_offset=1000 * (g + 1), geometry_seed_offset=0) for g in range(nchan)]
    nspc = obs[0].nspectra
    az, za = np.array([0.7]), np.radians([20.0])
    dv = pkg.DeviceArray(128 * obs[0].ninput * nspc * 8)
    flags = (pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA) & ~pkg.PFB_MALLOC_DEVICE_OUTPUT
    sess = []
    try:
        for o in obs:
            sess.append(TiedArraySession(o, az, za, pfb_flags=flags, nstokes=1,
                                         form=pkg.VM_BF_FP32, host_buffers=False, ipfb=True,
                                         d_v=dv.ptr))
        for x in sess:
            device_second(x, 0)

        def second(t):
            for x in sess:
                x.pfb()
                x.geometry()
                x.beamform()
            invert_batch(sess, t)  # all channels' inverse PFB + VDIF in one launch

        second(0)
        for x in sess:
            lib.vmSetKernelTiming(x.vm, 1)
        lib.vmSetKernelTiming(None, 1)
        lib.vmDeviceSynchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            second(1 + k)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        stages = {}
        for name in ("pfb", "geometry", "beamform_fp32", "ipfb"):
            tot = 0.0
            for x in sess:
                avg, n = kernel_ms(lib, x, name)
                tot += avg * n / steps
            stages[name + "_ms"] = round(tot, 3)
        lib.vmSetKernelTiming(None, 0)
        # the inverse PFB and its VDIF requantisation run as one kernel: 8 B
        # fine in, 8 B coarse out (float) and 2 B VDIF out per sample-pol
        ipfb_bytes = nchan * nspc * 128 * 2 * 18
        return {"workload": f"{nchan} coarse channels x 1 s, 128 tiles x 2 pols, one beam: FP32 "
                            f"PFB, FP32 beamformer (Stokes I), inverse PFB, 8-bit VDIF, "
                            f"device-resident",
                "value": round(nchan * 128 * nspc / (ms * 1e-3), 1), "unit": "beam-samples/s",
                "ms_per_second_of_data": round(ms, 2),
                "kernel_ms_per_second_of_data": stages,
                "pfb_hbm_frac": round(nchan * 2 * 128 * nspc * 128 * 10 /
                                      (stages["pfb_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "ipfb_vdif_hbm_frac": round(ipfb_bytes / (stages["ipfb_ms"] * 1e-3) / 1e9
                                            / HBM_PEAK_GBS, 4)}
    finally:
        for x in sess:
            x.close()
        del dv


def aux_stage(pkg, lib, reps=5):
    """Row kernels outside the beamformer at one coarse channel-second of 128
    tiles: the incoherent beam (cu_form_incoh_beam on 4+4-bit legacy input
    [s][c][i]: the 4+4 unpack, the power sum and its 8-bit renormalisation,
    timed together) and the 8-bit renormalise of
    one beam's Stokes IQUV (vmSendSToFits), each with its HBM roofline."""
    import ctypes as C
    from vcsbeam_amd.apps import TiedArraySession, device_second
    from vcsbeam_amd.workloads import Observation, legacy_int4
    ns, nch, ni = 10000, 128, 256
    data = legacy_int4(ns * nch * ni)
    h, d = C.c_void_p(), C.c_void_p()
    lib.allocate_input_output_arrays(C.byref(h), C.byref(d), data.nbytes)
    C.memmove(h, data.ctypes.data, data.nbytes)
    d_incoh, d_off, d_scl, d_Is = (pkg.DeviceArray(n) for n in (ns * nch * 4, nch * 4, nch * 4,
                                                                ns * nch))
    off, scl, Is = np.zeros(nch, np.float32), np.zeros(nch, np.float32), np.zeros(ns * nch, np.uint8)
    call = lambda: lib.cu_form_incoh_beam(h, d, data.nbytes, d_incoh.ptr, ns, nch, ni,
                                          off.ctypes.data, d_off.ptr, scl.ctypes.data, d_scl.ptr,
                                          Is.ctypes.data, d_Is.ptr, Is.nbytes)
    call()
    lib.vmSetKernelTiming(None, 1)
    for _ in range(reps):
        call()
    incoh_ms, _ = kernel_ms_global(lib, "incoh")
    lib.vmSetKernelTiming(None, 0)
    lib.free_input_output_arrays(h, d)
    # the whole device part of the call: 4+4 data read, I written, I read
    # twice by the renormalisation (extremes, then quantisation), 8-bit out
    incoh_bytes = data.nbytes + ns * nch * (4 + 4 + 4 + 1)
    obs = Observation(ntiles=128, blocks_per_second=20)
    ses = TiedArraySession(obs, np.array([0.7]), np.radians([20.0]),
                           pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA, nstokes=4,
                           form=pkg.VM_BF_FP32, host_buffers=False)
    try:
        device_second(ses, 0)
        ses.pfb()
        ses.geometry()
        ses.beamform()
        lib.vmCreateStatistics(ses.vm, None)
        lib.vmSendSToFits(ses.vm, None)
        lib.vmSetKernelTiming(ses.vm, 1)
        for _ in range(reps):
            lib.vmSendSToFits(ses.vm, None)
        renorm_ms, _ = kernel_ms(lib, ses, "renormalise")
        lib.vmSetKernelTiming(ses.vm, 0)
        lib.vmDestroyStatistics(ses.vm)
    finally:
        ses.close()
    renorm_bytes = 2 * 4 * nch * ns * 4 + 4 * nch * ns  # S read twice, uint8 out
    return {"workload": "one coarse channel-second, 128 tiles x 2 pols",
            "incoh_ms": round(incoh_ms, 4),
            "incoh_hbm_frac": round(incoh_bytes / (incoh_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "incoh_algorithmic_bytes": incoh_bytes,
            "renormalise_ms": round(renorm_ms, 4),
            "renormalise_hbm_frac": round(renorm_bytes / (renorm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                          4),
            "renormalise_algorithmic_bytes": renorm_bytes}


def kernel_ms_global(lib, name):
    import ctypes as C
    n = C.c_long(0)
    ms = lib.vmKernelTimeMs(None, name.encode(), C.byref(n))
    return ms, n.value


def surrogate_feed(lib, s, pkg, sess, obs, az, za, steps=5, nbeam_check=16, nsamp_check=64):
    """The session's pointings with per-pointing, per-channel primary beams
    from the surrogate (with fallback) on its MFMA FP16 beamformer: a step is
    FP64 phases, J^-1 from D and those beams (vmCalcJDeviceChanB), weights,
    prep and GEMM; error against the FP64 oracle chain with the analytic beam
    on nbeam_check beams x nsamp_check samples x 128 channels."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc  # the checker (error only)
    from vcsbeam_amd.workloads import unit_vectors
    npnt, nch = len(az), 128
    B = np.empty((npnt, nch, 4, 2))
    err = np.empty((npnt, nch), np.float32)
    nfb = lib.vmSurrogateJones(s, 3, az.ctypes.data, za.ctypes.data, npnt, obs.freqs.ctypes.data,
                               nch, B.ctypes.data, err.ctypes.data)
    sess.v.beamform_form = pkg.VM_BF_MFMA_FP16

    host = {"phi": 0.0, "jones": 0.0, "beamform": 0.0}

    def step():
        t = time.perf_counter()
        lib.vmCalcPhiDevice(sess.vm, sess.bg)
        t1 = time.perf_counter()
        lib.vmCalcJDeviceChanB(sess.vm, B.ctypes.data)
        t2 = time.perf_counter()
        sess.beamform()
        t3 = time.perf_counter()
        host["phi"] += t1 - t
        host["jones"] += t2 - t1
        host["beamform"] += t3 - t2

    step()
    lib.vmDeviceSynchronize()
    lib.vmSetKernelTiming(sess.vm, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    lib.vmDeviceSynchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    kern = {k: round(kernel_ms(lib, sess, k)[0], 3)
            for k in ("geometry", "beamform_weights", "beamform_prep", "beamform_mfma")}
    lib.vmSetKernelTiming(sess.vm, 0)
    S = sess.stokes()
    pk, n = min(nbeam_check, npnt), nsamp_check
    Bt = np.empty((pk, nch, 4, 2))
    lib.vmSurrogateJones(s, 0, az.ctypes.data, za.ctypes.data, pk, obs.freqs.ctypes.data, nch,
                         Bt.ctypes.data, None)
    Bt = (Bt[..., 0] + 1j * Bt[..., 1]).reshape(pk, nch, 2, 2)
    D = obs.calibration()
    J = np.empty((pk, obs.ntiles, nch, 2, 2), complex)
    for c in range(nch):
        J[:, :, c] = orc.calc_J(D[:, c:c + 1], np.repeat(Bt[:, c][:, None], obs.ntiles, axis=1))[:, :, 0]
    v = np.ascontiguousarray(sess.voltages()[:, :, :n].transpose(2, 0, 1)).astype(np.complex128)
    E, N, H, L = obs.positions()
    phi = orc.calc_phi(E, N, H, L, unit_vectors(az[:pk], za[:pk]), obs.freqs)
    polQ, polP = obs.pol_idx()
    Sref, _ = orc.beamform(v, J, phi, polQ, polP, 1.0 / (2 * int((~obs.flag_mask).sum())), 4)
    e = np.array([[orc.rel_err(S[p, :n, k, :], Sref[p, :, k, :]) for k in range(4)]
                  for p in range(pk)]).max(axis=0)
    return {"beams": npnt, "value": round(npnt * nch * obs.nspectra / (ms * 1e-3), 1),
            "unit": "beam-samples/s", "ms_per_step": round(ms, 3), "kernel_ms": kern,
            "call_ms": {k: round(v * 1e3 / (steps + 1), 3) for k, v in host.items()},
            "fallbacks": int(nfb),
            "max_beam_rel_err_IQUV_vs_analytic_fp64": [float(f"{x:.3e}") for x in e],
            "check": f"{pk} beams x {n} samples x 128 ch"}


def calibration_copy(pkg, lib):
    """One plain copy of COPY_BYTES: the memory counters' calibration."""
    a, b = pkg.DeviceArray(COPY_BYTES), pkg.DeviceArray(COPY_BYTES)
    lib.vmDeviceCopy(b.ptr, a.ptr, COPY_BYTES)
    del a, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--beams", type=int, default=512)
    ap.add_argument("--fp32-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 24-channel uneven workload")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched 24-channel workload")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the incoherent-beam and renormalise timings")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="skip the 24-channel plain pipeline with inverse PFB")
    ap.add_argument("--pipeline-only", action="store_true",
                    help="run only the 24-channel plain pipeline with inverse PFB")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the primary-beam surrogate stage (training + 4096-beam inference)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="fix the MFMA beamformer variant instead of autotuning")
    ap.add_argument("--order", type=int, default=1, help="tile order with --variant")
    ap.add_argument("--form", type=int, default=1, choices=(1, 2, 3),
                    help="MFMA form of the headline: 1 FP16 voltages with split hi+lo FP16 "
                         "weights, 2 BF16 split, 3 FP16 with single FP16 weights")
    ap.add_argument("--no-side-form", action="store_true",
                    help="skip timing the other FP16 form beside the headline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 1 GiB plain copy (memory-counter calibration)")
    ap.add_argument("--rocprof-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 CSV output directories (no GPU needed)")
    ap.add_argument("--title", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--summary-out", default="")
    ap.add_argument("--traffic-out", default="")
    args = ap.parse_args()
    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary, args.title, args.command, args.summary_out,
                        args.traffic_out)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the beamforming path shards over coarse channels "
              "(one GPU per channel) and this benchmark runs on a single MI355X; use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    lib = pkg.load()
    if args.batch_only:
        print(json.dumps({"batch": batch_stage(pkg, lib),
                          "batch_fp16_pfb_single_weights": batch_stage(pkg, lib, form=3,
                                                                       half=True)}))
        return
    if args.pipeline_only:
        out = {"pipeline": pipeline_stage(pkg, lib), "aux": aux_stage(pkg, lib)}
        if args.calibrate_copy:
            calibration_copy(pkg, lib)
        print(json.dumps(out))
        return
    from vcsbeam_amd.apps import TiedArraySession
    from vcsbeam_amd.workloads import Observation, directions_in_cap

    print("[bench] generating 1 s of 128-tile MWAX voltages", file=sys.stderr, flush=True)
    obs = Observation(ntiles=128, blocks_per_second=20)
    az, za = directions_in_cap(args.beams, 15.0)
    sess = TiedArraySession(obs, az, za, pfb_flags=pkg.PFB_FULL_PRECISION | pkg.PFB_EMULATE_FPGA,
                            nstokes=4, form=pkg.VM_BF_MFMA_FP16)
    sess.read(0)
    print("[bench] forward PFB", file=sys.stderr, flush=True)
    for _ in range(10):  # warm-up: clocks settle under load
        sess.pfb()
    lib.vmSetKernelTiming(sess.vm, 1)
    for _ in range(10):
        lib.vmWOLAChunk(sess.vm)
    pfb_ms, _ = kernel_ms(lib, sess, "pfb")
    lib.vmSetKernelTiming(sess.vm, 0)
    # measured HBM ceilings on this device: a 2.62 GB write (hipMemset) and
    # a 1 GiB copy kernel (read + write)
    nsamp_in_bytes = obs.ninput * obs.nspectra * 128 * 8
    ceil_buf = pkg.DeviceArray(nsamp_in_bytes)
    lib.vmMemsetDevice(ceil_buf.ptr, 0, nsamp_in_bytes)
    t0 = time.perf_counter()
    for _ in range(5):
        lib.vmMemsetDevice(ceil_buf.ptr, 0, nsamp_in_bytes)
    write_GBps = 5 * nsamp_in_bytes / (time.perf_counter() - t0) / 1e9
    del ceil_buf
    ca, cb = pkg.DeviceArray(COPY_BYTES), pkg.DeviceArray(COPY_BYTES)
    lib.vmDeviceCopy(cb.ptr, ca.ptr, COPY_BYTES)
    t0 = time.perf_counter()
    for _ in range(5):
        lib.vmDeviceCopy(cb.ptr, ca.ptr, COPY_BYTES)
    copy_GBps = 5 * 2 * COPY_BYTES / (time.perf_counter() - t0) / 1e9
    del ca, cb
    # measured dense FP16 MFMA rate (register-only matrix instruction chains);
    # the better of two calls, the first of which may still see the clock rise
    mfma_peak_tf = max(lib.vmMFMAPeakTFLOPS(), lib.vmMFMAPeakTFLOPS())
    nsamp_in = obs.ninput * obs.nspectra * 128
    pfb_bytes = nsamp_in * 10  # 2 B in + 8 B out per input complex sample
    beam_samples = args.beams * 128 * obs.nspectra

    import ctypes as C
    parts = {1: 2, 2: 2, 3: 1}  # weight parts (split hi + lo, or one)
    form_name = {1: "mfma-fp16-split", 2: "mfma-bf16-split", 3: "mfma-fp16-single"}

    def run_form(form):
        """Autotune (or fix) the GEMM variant for one MFMA form, then time
        the step; per-kernel times are live HIP events on the context stream."""
        print(f"[bench] MFMA form {form} ({form_name[form]}): autotune", file=sys.stderr,
              flush=True)
        sess.v.beamform_form = form
        sess.geometry()
        best = C.c_float(0)
        if args.variant >= 0:
            variant = args.variant
            lib.vmSetBeamformerVariant(sess.vm, args.variant, args.order)
        else:
            variant = lib.vmTuneBeamformer(sess.vm, 3, C.byref(best))
        print(f"[bench] MFMA form {form}: warmup + timed steps", file=sys.stderr, flush=True)
        ms = time_form(lib, sess, form, args.steps, args.warmup)
        g_ms, g_n = kernel_ms(lib, sess, "beamform_mfma")
        sv, so = C.c_int(0), C.c_int(0)
        lib.vmBeamformerSelection(sess.vm, C.byref(sv), C.byref(so))
        return {"form": form, "ms_step": ms, "gemm_ms": g_ms, "gemm_n": g_n, "variant": variant,
                "order": so.value, "sel_variant": sv.value, "best": best.value,
                "prep_ms": kernel_ms(lib, sess, "beamform_prep")[0],
                "weights_ms": kernel_ms(lib, sess, "beamform_weights")[0],
                "S": sess.stokes()}

    head = run_form(args.form)
    side_form = 1 if args.form == 3 else 3
    side = None if args.no_side_form else run_form(side_form)
    ms_step, gemm_ms, gemm_n, variant = head["ms_step"], head["gemm_ms"], head["gemm_n"], head["variant"]
    prep_ms, weights_ms = head["prep_ms"], head["weights_ms"]
    vr, vs, vw = C.c_int(0), C.c_int(0), C.c_int(0)
    lib.vmBeamformerVariantInfo(variant, C.byref(vr), C.byref(vs), C.byref(vw))
    value = beam_samples / (ms_step * 1e-3)
    S_mfma = head["S"]

    def gemm_kernel(v):
        return lib.vmBeamformerVariantKernel(v).decode()
    # Dominant kernel (the GEMM variant's kernel, k_beamform_mfma or
    # k_beamform_res), per launch (one channel-second, 512 beams):
    # algorithmic bytes = Stokes IQUV out (4 B x 4 per beam-sample) + J^-1-applied
    # voltages in (FP16 complex, 4 B per tile-pol-sample) + phase weights
    # (2 B per beam-row x k' per weight part: hi + lo, or one) + N terms in
    # (16 B per channel-sample); algorithmic flops = 8 x 2 pols x tiles per
    # beam-sample (phase GEMM after J^-1 is applied per antenna), issued = that
    # times the weight parts.
    ntile, nch, nsamp = obs.ntiles, 128, obs.nspectra

    def gemm_algorithmic_bytes(form):
        return (beam_samples * 4 * 4 + nch * ntile * 2 * nsamp * 4
                + nch * (2 * args.beams) * (2 * ntile) * 2 * parts[form] + nch * nsamp * 16)
    gemm_bytes = gemm_algorithmic_bytes(args.form)
    gemm_flops = 8.0 * 2 * ntile * beam_samples
    achieved = gemm_bytes / (gemm_ms * 1e-3) / 1e9
    mfma_tf = gemm_flops / (gemm_ms * 1e-3) / 1e12
    issued = parts[args.form]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": committed_traffic(gemm_kernel(variant), head["sel_variant"],
                                             head["order"], args.form),
                "algorithmic_bytes": int(gemm_bytes), "algorithmic_flops": gemm_flops,
                "mfma_tflops": round(mfma_tf, 1),
                "mfma_frac_issued": round(issued * mfma_tf / FP16_PEAK_TFLOPS, 4),
                "mfma_measured_peak_tflops": round(mfma_peak_tf, 1),
                "mfma_frac_of_measured_issued": round(issued * mfma_tf / mfma_peak_tf, 4),
                # the kernel's bytes are 88 % writes: beside the spec peak, the
                # write and copy rates measured on this device in this run
                "measured_write_GBps": round(write_GBps, 1),
                "frac_of_measured_write": round(achieved / write_GBps, 4),
                "measured_copy_GBps": round(copy_GBps, 1),
                "kernel": f"{gemm_kernel(variant)}<{form_name[args.form]}>",
                "kernel_ms": round(gemm_ms, 4),
                "launches": gemm_n, "autotuned_variant": variant, "tile_order": head["order"],
                "workgroup_tile": f"{vr.value} rows x {vs.value} samples, {vw.value} waves",
                "autotune_best_ms": round(head["best"], 4)}
    dtype = {1: "f16 (split hi+lo f16 weights, f32 accumulate)",
             2: "bf16 (split hi+lo bf16 weights, f32 accumulate)",
             3: "f16 (f16 weights, f32 accumulate)"}[args.form]
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "beam-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded Gaussian 8+8-bit voltages, sigma 16)",
        "config": {"workload": "128 tiles x 2 pols, 1 coarse channel x 1 s, 512 beams within "
                               "15 deg of boresight, MFMA beamformer, Stokes IQUV",
                   "tiles": 128, "beams": args.beams, "fine_channels": 128,
                   "samples": obs.nspectra, "form": form_name[args.form]},
        "roofline": roofline,
        "stages": {"pfb_fused_ms_per_ch_s": round(pfb_ms, 4),
                   "pfb_GBps": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9, 1),
                   "pfb_hbm_frac": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   "measured_write_GBps": round(write_GBps, 1),
                   "measured_copy_GBps": round(copy_GBps, 1),
                   "measured_mfma_fp16_tflops": round(mfma_peak_tf, 1),
                   "pfb_frac_of_measured_write": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9
                                                       / write_GBps, 4),
                   "beamform_prep_ms": round(prep_ms, 4),
                   "beamform_weights_ms": round(weights_ms, 4)},
    }
    S_side = None
    if side is not None:
        sb = gemm_algorithmic_bytes(side_form)
        result["other_fp16_form"] = {
            "form": form_name[side_form], "value": round(beam_samples / (side["ms_step"] * 1e-3), 1),
            "ms_per_step": round(side["ms_step"], 4), "kernel_ms": round(side["gemm_ms"], 4),
            "hbm_frac": round(sb / (side["gemm_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "autotuned_variant": side["variant"], "tile_order": side["order"],
            "kernel": gemm_kernel(side["variant"]),
            "traffic": committed_traffic(gemm_kernel(side["variant"]), side["sel_variant"],
                                         side["order"], side_form)}
        S_side = side["S"]
    S_fp32 = None
    if not args.no_fp32:
        print("[bench] plain FP32 form", file=sys.stderr, flush=True)
        ms32 = time_form(lib, sess, pkg.VM_BF_FP32, args.fp32_steps, 1)
        k32, _ = kernel_ms(lib, sess, "beamform_fp32")
        S_fp32 = sess.stokes()
        result["fp32_form"] = {"value": round(beam_samples / (ms32 * 1e-3), 1),
                               "ms_per_step": round(ms32, 3), "kernel_ms": round(k32, 3),
                               "mfma_speedup": round(ms32 / ms_step, 2)}
    if not args.no_surrogate:
        print("[bench] primary-beam surrogate: train + 4096 beams x 128 channels",
              file=sys.stderr, flush=True)
        result["surrogate"] = surrogate_stage(lib, feed=(pkg, sess, obs, az, za))
    if not args.no_cpu:
        print("[bench] CPU baseline (FP64 oracle)", file=sys.stderr, flush=True)
        cb = cpu_baseline(sess, obs, az, za)
        n = cb["nsamp"]
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        full, one = cb["legs"][cb["cores"]], cb["legs"][1]
        result["cpu_baseline"] = {
            "value": round(cb["rate"], 1), "unit": "beam-samples/s", "cores": cb["cores"],
            "kind": "port",
            "sample": f"{args.beams} beams x 128 ch x {n} samples of the same second (FP64 "
                      f"delay + calc, the GPU step's work; {cb['seconds']:.1f} s), stages "
                      f"scaled to the channel-second",
            **cb["info"],
            "single_thread_value": one["beam_samples_per_s"],
            "stages_s_per_channel_second": full["seconds_per_channel_second"],
            "stages_s_per_channel_second_1_thread": one["seconds_per_channel_second"],
            "pfb_input_samples_per_s": full["pfb_input_samples_per_s"],
            "end_to_end_s_per_channel_second_one_beam":
                full["end_to_end_s_per_channel_second_one_beam"],
            "samples": full["sample"], "samples_1_thread": one["sample"]}
        par = {}
        for name, S in ((form_name[args.form], S_mfma),
                        (form_name[side_form], S_side), ("fp32", S_fp32)):
            if S is None:
                continue
            e = np.array([[orc.rel_err(S[p, :n, st, :], cb["S"][p, :n, st, :]) for st in range(4)]
                          for p in range(S.shape[0])])
            par[name] = {"max_beam_rel_err_IQUV": [float(f"{x:.3e}") for x in e.max(axis=0)]}
        result["parity_vs_fp64"] = par
    sess.close()
    if not args.no_batch:
        print("[bench] batched uneven workload (24 channels x 256 tiles)", file=sys.stderr,
              flush=True)
        result["batch"] = batch_stage(pkg, lib)
        print("[bench] batched workload with FP16 PFB output and single weights",
              file=sys.stderr, flush=True)
        result["batch_fp16_pfb_single_weights"] = batch_stage(pkg, lib, form=3, half=True)
    if not args.no_pipeline:
        print("[bench] plain pipeline with inverse PFB (24 channels x 128 tiles)",
              file=sys.stderr, flush=True)
        result["pipeline"] = pipeline_stage(pkg, lib)
    if not args.no_aux:
        print("[bench] incoherent beam and renormalise", file=sys.stderr, flush=True)
        result["aux"] = aux_stage(pkg, lib)
    if args.calibrate_copy:
        calibration_copy(pkg, lib)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
