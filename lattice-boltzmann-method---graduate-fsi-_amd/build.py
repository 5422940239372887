"""Build the gfx950 library liblbm_hip.so in-tree with hipcc.

Usage: python build.py [--jobs N]

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled, into _build/, then linked into liblbm_hip.so next to this
file.  The C++ drop-in host layer (host/) is compiled into the same library.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{os.path.join(HERE, 'csrc')}",
    f"-I{os.path.join(HERE, 'host')}",
]
# gfx950 code objects; packed FP32 ops (v_pk_*_f32) are not generated
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    os.path.join(HERE, "csrc", "lbm_kernels.hip"),
    os.path.join(HERE, "csrc", "lbm_stream_fold.hip"),
    os.path.join(HERE, "csrc", "lbm_capi.hip"),
    os.path.join(HERE, "csrc", "lbm_probe.hip"),
    os.path.join(HERE, "csrc", "lbm_ibm.hip"),
    os.path.join(HERE, "csrc", "lbm_thermal.hip"),
    os.path.join(HERE, "csrc", "lbm_output.hip"),
    os.path.join(HERE, "csrc", "lbm_batch.hip"),
    os.path.join(HERE, "csrc", "lbm_q7.hip"),
    os.path.join(HERE, "csrc", "lbm_filament.hip"),
    os.path.join(HERE, "csrc", "lbm_mea.hip"),
    os.path.join(HERE, "csrc", "lbm_faces.hip"),
    os.path.join(HERE, "csrc", "lbm_stats.hip"),
    os.path.join(HERE, "csrc", "lbm_noalloc.hip"),
    os.path.join(HERE, "csrc", "lbm_checkpoint.hip"),
]
HOST_SOURCES = [
    os.path.join(HERE, "host", "LBM.cpp"),
    os.path.join(HERE, "host", "IBM.cpp"),
]
# reference-style drivers on the drop-in classes (host code only)
EXAMPLES = ["example_cavity", "example_cylinder_ibm", "example_couette_thermal", "example_cylinder_bb",
            "example_helpers"]
EXAMPLE_BIN = {"example_cavity": "lbm_example_cavity",
               "example_cylinder_ibm": "lbm_example_cylinder_ibm",
               "example_couette_thermal": "lbm_example_couette_thermal",
               "example_cylinder_bb": "lbm_example_cylinder_bb",
               "example_helpers": "lbm_example_helpers"}


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()
    )
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out and os.environ.get("LBM_BUILD_VERBOSE"):
        print(out)


def _obj(src: str) -> str:
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    for s in SOURCES:
        cmds.append([HIPCC, *COMMON, *DEVICE, "-c", s, "-o", _obj(s)])
    for s in HOST_SOURCES:
        if os.path.exists(s):
            cmds.append([HIPCC, *COMMON, "-x", "c++", "-c", s, "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    objs = [_obj(s) for s in SOURCES] + [_obj(s) for s in HOST_SOURCES if os.path.exists(s)]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for name in EXAMPLES:
        ex = os.path.join(HERE, "host", name + ".cpp")
        exe = os.path.join(HERE, "host", EXAMPLE_BIN[name])
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
              f"-I{os.path.join(HERE, 'host')}", ex, "-o", exe, f"-L{HERE}", "-llbm_hip",
              "-Wl,-rpath,$ORIGIN/.."])
    return LIB


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) under oracle/."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs))
    build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
