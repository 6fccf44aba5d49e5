"""Builds libcapow.so (gfx950 HIP kernels + C ABI) and the capow host program.

Run as `python -m capow_amd.build_lib` or through __graft_entry__.build().
Outputs stay in-tree: capow_amd/libcapow.so and capow_amd/bin/capow.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libcapow.so")
EXE = os.path.join(HERE, "bin", "capow")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INC, "-I", CSRC,
    # no packed-FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]
HIP_SOURCES = ["capow_api.hip", "deposit.hip", "fft.hip", "binning.hip", "adjoint.hip",
               "survey.hip", "noalloc.hip"]
C_SOURCES = ["capow_host.c", "capow_ascii.c", "capow_survey_ingest.c"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False, jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in C_SOURCES:
        o = os.path.join(OBJ, s.replace(".c", ".o"))
        cmds.append(["gcc", "-O2", "-fPIC", "-pthread", "-std=gnu11", "-Wall", "-I", INC, "-c",
                     os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    host_srcs = sorted(f for f in os.listdir(HOST) if f.endswith(".c")) if os.path.isdir(HOST) else []
    if host_srcs:
        _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INC,
              *[os.path.join(HOST, f) for f in host_srcs], "-o", EXE,
              "-L", HERE, "-lcapow", "-Wl,-rpath,$ORIGIN/..", "-lm"])
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
