"""Builds the native libraries in-tree.

  libheimdall_mi355x.so  the search (hd_* C ABI + gfx950 kernels), hipcc
  libhd_synth.so         synthetic filterbank generator, g++/OpenMP
  bin/heimdall           gulp driver over sigproc files (app/heimdall_main.cpp)
  bin/generate_dmlist    DM trial list tool (app/generate_dmlist.cpp)
  bin/coincidencer       multibeam coincidence of candidate files / TCP messages
  bin/candidate_profiler DM-time / frequency-time plots around one candidate

Every HIP source is compiled for gfx950 only, with -ffp-contract=off (the
reference performs no fused multiply-adds, and results are compared bit for
bit) and with packed FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libheimdall_mi355x.so")
SYNTH = os.path.join(HERE, "libhd_synth.so")

HIP_SOURCES = ["hd_api.cpp", "dedisp_plan.cpp", "formats.cpp", "k_dedisperse.hip", "k_normalise.hip",
               "k_filter.hip", "k_cluster.hip", "k_stages.hip", "k_clean.hip",
               "k_proposal.hip", "k_cnn.hip", "k_synth.hip"]
# command-line tools linked against the library: bin/<name> from app/<name>.cpp
APPS = {"heimdall": "heimdall_main.cpp", "generate_dmlist": "generate_dmlist.cpp",
        "coincidencer": "coincidencer.cpp", "candidate_profiler": "candidate_profiler.cpp"}
BIN = os.path.join(HERE, "bin")
# tools that also call the HIP runtime (device buffers and a copy stream)
APP_HIP = {"heimdall"}
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-result",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build(verbose=False):
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = SYNTH + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c++17",
          "-I" + os.path.join(ROOT, "include"), "-o", tmp, os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, SYNTH)
    os.makedirs(BIN, exist_ok=True)
    for name, src in APPS.items():
        out = os.path.join(BIN, name)
        hip = (["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]
               if name in APP_HIP else [])
        hip_lib = (["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-pthread"]
                   if name in APP_HIP else [])
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include")] + hip +
             [os.path.join(HERE, "app", src), "-o", out + ".tmp", "-L" + HERE,
              "-lheimdall_mi355x"] + hip_lib + ["-Wl,-rpath,$ORIGIN/.."])
        os.replace(out + ".tmp", out)
    if verbose:
        print("built", LIB, SYNTH, BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
