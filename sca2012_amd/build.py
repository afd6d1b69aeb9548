"""Build recipe for libsph_amd.so (gfx950) -- run by __graft_entry__.build().

hipcc flags that matter for parity: -ffp-contract=off (no FMA contraction, so every fp32
operation rounds like the CPU restatement) and the default correctly rounded fp32 divide
and square root. -Xclang -target-feature -Xclang -packed-fp32-ops keeps the compiler from
emitting packed fp32 (v_pk_*) arithmetic.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsph_amd.so")
SOURCES = ["sph_prims.hip", "sph_kernels.hip", "sph_capi.cpp", "sph_batch.cpp", "sph_adjoint.hip",
           "sph_system_dropin.cpp"]
BIN = os.path.join(HERE, "_bin")
DEMO = os.path.join(BIN, "sph_dropin_run")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-x", "hip", "-O3", "--offload-arch=" + ARCH, "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    # the drop-in demo application, linked against the library in place
    os.makedirs(BIN, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-o", DEMO, os.path.join(HERE, "examples", "dropin_run.cpp"),
          "-L" + HERE, "-lsph_amd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
