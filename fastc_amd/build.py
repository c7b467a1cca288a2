"""Build the HIP library (libfasthip.so) in-tree with hipcc for gfx950.

Device code is compiled with -ffp-contract=off so that every float multiply
and add is rounded separately, as in the reference's x86-64 build; division
and sqrt stay correctly rounded (hipcc's default) and denormals are kept.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfasthip.so")
SOURCES = ["capi.cpp", "tables.cpp", "timing.cpp", "ranker.cpp", "batch.cpp", "bc7_encode.hip",
           "dxt_encode.hip", "etc1_encode.hip", "mlp.hip", "decode.hip", "mips.hip", "ktx.cpp",
           "static_mem.hip"]
ARCH = os.environ.get("FASTHIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False, jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    _run(cmd, verbose)
    os.replace(LIB + ".tmp", LIB)
    synth = os.path.join(HERE, "libsynth.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", synth + ".tmp",
          os.path.join(CSRC, "synth.c")], verbose)
    os.replace(synth + ".tmp", synth)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode != 0:
        sys.stderr.write(text)
        raise RuntimeError(f"hipcc failed on {src}")
    if text.strip():
        sys.stderr.write(text)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("link failed")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
