"""Build the MI355X voltage-beamforming library (gfx950) in-tree.

Produces, under ``mwa-vcsbeam_amd/_lib/``:

* ``libvcsbeam_amd.so`` -- the C-ABI library (HIP kernels + host C module)
  declared in ``include/vcsbeam.h``;
* ``libvmsynth.so`` -- the seeded synthetic-input generator;
* ``tied_array_beam``, ``incoh_beam``, ``fine_pfb_offline`` -- C callers of the
  library (examples/*.c) following app/make_mwa_tied_array_beam.c,
  app/make_mwa_incoh_beam.c and app/fine_pfb_offline.c.

Run as ``python mwa-vcsbeam_amd/build.py`` or call :func:`build`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")
INC = os.path.join(ROOT, "include")

HIP_SOURCES = ["vm_device.hip", "pfb.hip", "beamform.hip", "beamform_mfma.hip",
               "ipfb.hip", "geometry.hip", "beam_model.hip"]
C_SOURCES = ["vm_host.c", "calibration.c", "astrometry.c", "obsfiles.c", "psrfits.c"]
EXAMPLES = ["tied_array_beam", "incoh_beam", "fine_pfb_offline"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code (they have given
# silently wrong results on this hardware generation).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + cmd[0])
    return r.stdout


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    obj = os.path.join(LIB, "obj")
    os.makedirs(obj, exist_ok=True)
    runtime_dir = os.path.join(HERE, "pfb_filter")
    objs = []
    for src in C_SOURCES:
        o = os.path.join(obj, src + ".o")
        _run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function",
              f'-DVCSBEAM_RUNTIME_DIR="{runtime_dir}"', "-I", INC, "-I", CSRC, "-c",
              os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    procs = []
    for src in HIP_SOURCES:
        o = os.path.join(obj, src + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-I", INC, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", o]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        objs.append(o)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + out)
            failed = True
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("HIP compilation failed")
    lib = os.path.join(LIB, "libvcsbeam_amd.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib + ".tmp"] + objs)
    os.replace(lib + ".tmp", lib)
    # C callers of the library (examples/*.c, one per reference application),
    # linked against the installed header and the .so next to them
    for name in EXAMPLES:
        example = os.path.join(LIB, name)
        _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INC,
              os.path.join(HERE, "examples", name + ".c"), "-L", LIB, "-lvcsbeam_amd",
              "-Wl,-rpath,$ORIGIN", "-lm", "-o", example + ".tmp"])
        os.replace(example + ".tmp", example)
    synth = os.path.join(LIB, "libvmsynth.so")
    _run(["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-o", synth + ".tmp",
          os.path.join(CSRC, "vmsynth.c"), "-lm"])
    os.replace(synth + ".tmp", synth)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
