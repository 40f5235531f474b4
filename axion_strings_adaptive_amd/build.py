"""Build recipe for libaxs_hip.so (gfx950) and the host driver.

    python -m axion_strings_adaptive_amd.build          # library + driver
    python axion_strings_adaptive_amd/build.py --lib    # library only

Outputs stay in-tree under axion_strings_adaptive_amd/_build/ so they travel
with the source tree to the GPU machine.  Every translation unit is compiled
for gfx950 only, with packed FP32 instructions disabled.
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(BUILD, "libaxs_hip.so")
DRIVER = os.path.join(BUILD, "axion_strings")

HIP_SOURCES = [
    "axs_capi.hip",
    "axs_rk.hip",
    "axs_pair.hip",
    "axs_step.hip",
    "axs_diag.hip",
    "axs_fft.hip",
    "axs_boxes.hip",
    "axs_adjoint.hip",
    "axs_amr.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++20",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + INCLUDE,
    "-I" + CSRC,
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str | None:
    """Host C++ driver (params file -> run -> .dat files) over the C ABI."""
    src = os.path.join(CSRC, "host", "main_axion_strings.cpp")
    if not os.path.exists(src):
        return None
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++20", "-I" + INCLUDE, "-I" + CSRC, src,
          os.path.join(CSRC, "host", "AxionStringsLevel.cpp"),
          "-o", DRIVER, "-L" + BUILD, "-laxs_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    return DRIVER


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args(argv)
    print(build_library(args.j))
    if not args.lib:
        d = build_driver()
        if d:
            print(d)
    return 0


if __name__ == "__main__":
    sys.exit(main())
