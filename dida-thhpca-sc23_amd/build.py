"""Build the LETKF engine for gfx950 in-tree.

  lib/libdida_letkf.so   the analysis library (HIP kernels + C host layer)
  lib/libdida_synth.so   the synthetic workload generator (inputs for tests
                         and the benchmark; not part of the analysis path)

Run ``python build.py`` (or ``__graft_entry__.build()``).  Every call
recompiles.  HIP sources are compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled; the C host layer with gcc against the HIP runtime
headers.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
C_FLAGS = ["-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", f"-I{ROCM}/include"]

LIB_HIP = ["letkf_select.hip", "letkf_transform.hip", "letkf_update.hip", "letkf_wave.hip",
           "letkf_obsop.hip", "letkf_noalloc.hip"]
LIB_C = ["letkf_host.c"]
SYNTH_HIP = ["synth.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass reports the disabled GPU feature as unknown: harmless
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", _obj(src)])
    else:
        _run(["gcc", *C_FLAGS, "-c", path, "-o", _obj(src)])
    return src


def build(verbose: bool = True) -> dict:
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = LIB_HIP + LIB_C + SYNTH_HIP
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for s in ex.map(_compile, srcs):
            if verbose:
                print(f"  compiled {s}", flush=True)
    lib = os.path.join(OUT, "libdida_letkf.so")
    synth = os.path.join(OUT, "libdida_synth.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib,
          *[_obj(s) for s in LIB_HIP + LIB_C]])
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", synth,
          *[_obj(s) for s in SYNTH_HIP]])
    if verbose:
        print(f"  linked {lib}\n  linked {synth}", flush=True)
    return {"lib": lib, "synth": synth}


if __name__ == "__main__":
    build()
    sys.exit(0)
