"""Compile libgpe_hip.so for gfx950 from the HIP sources in csrc/.

Usage: python -m gpe_simulator_amd.build [--jobs N]

Every pass instantiation (precision x grid side x tile variant, read from the
variant list in csrc/gpe_shapes.h) is its own translation unit so the build
runs in parallel.  Objects go to gpe_simulator_amd/_build/ and
the shared library to gpe_simulator_amd/libgpe_hip.so.
"""

from __future__ import annotations

import argparse
import concurrent.futures as futures
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libgpe_hip.so"
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    # Packed FP32 instructions stay off; plain FP32 code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build libgpe_hip.so")
    return path


def variants():
    """(precision, side, variant) of every compiled pass variant: the generic
    transform of each precision plus the X(...) lines of gpe_shapes.h."""
    text = (CSRC / "gpe_shapes.h").read_text()
    block = text[text.index("#define GPE_VARIANTS(X)"):text.index("namespace gpe")]
    found = [tuple(int(v) for v in m) for m in
             re.findall(r"X\((\d+),\s*(\d+),\s*(\d+),\s*\d+,\s*\d+,\s*\d+,\s*\d+\)", block)]
    return [(0, 0, 0), (1, 0, 0)] + [f[:3] for f in found]


def _units():
    units = [(CSRC / "gpe_api.hip", BUILD / "gpe_api.o", []),
             (CSRC / "gpe_misc.hip", BUILD / "gpe_misc.o", []),
             (CSRC / "gpe_jit.hip", BUILD / "gpe_jit.o", [])]
    for prec, side, var in variants():
        units.append((CSRC / "gpe_pass_inst.hip", BUILD / f"gpe_pass_p{prec}_n{side}_v{var}.o",
                      [f"-DGPE_PREC={prec}", f"-DGPE_N={side}", f"-DGPE_V={var}"]))
    return units


def _compile(unit):
    src, obj, defs = unit
    cmd = [hipcc(), *COMMON, *defs, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile every translation unit and link libgpe_hip.so (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    units = _units()
    with futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, units))
    tmp = LIBRARY.with_suffix(".so.tmp")
    # hiprtc compiles the time-dependent potential programs at run time
    rocm_lib = Path(hipcc()).resolve().parents[1] / "lib"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects),
           f"-L{rocm_lib}", "-lhiprtc", f"-Wl,-rpath,{rocm_lib}"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}", flush=True)
    return LIBRARY


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int, default=None)
    args = parser.parse_args(argv)
    build(args.jobs, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
