"""Build libkm_hydro.so (gfx950) in-tree with hipcc.

Every translation unit is compiled for gfx950 only, with packed-FP32 instructions
disabled. The reference-order FP64 units (tasks, exact sweeps, runtime) use
-ffp-contract=off so each operation rounds on its own; the FMA and FP32-mixed sweep
units use -ffp-contract=fast.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libkm_hydro.so"
OBJ = HERE / "_obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-x",
    "hip",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unused-result",
]

# (source, contraction)
UNITS = [
    ("km_runtime.hip", "off"),
    ("km_tasks.hip", "off"),
    ("km_fused_exact.hip", "off"),
    ("km_fused_fma.hip", "fast"),
    ("km_fused_fp32.hip", "fast"),
    ("km_fused_mhd_exact.hip", "off"),
    ("km_fused_mhd_fma.hip", "fast"),
    ("km_fused_mhd_fp32.hip", "fast"),
    ("km_config.cpp", "off"),
    ("km_problems.cpp", "off"),
    ("km_amr.hip", "off"),
    ("km_amr_tree.cpp", "off"),
    ("km_history.hip", "off"),
]


def _compile(src: str, contract: str, verbose: bool, extra=(), objdir=OBJ) -> pathlib.Path:
    obj = objdir / (pathlib.Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, f"-ffp-contract={contract}", *extra, "-c", str(CSRC / src), "-o",
           str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-8000:]}")
    return obj


def build(verbose: bool = False, jobs: int | None = None, extra=(), out: pathlib.Path = OUT,
          objdir: pathlib.Path = OBJ) -> pathlib.Path:
    """Compile all units and link libkm_hydro.so (or `out`); returns its path.
    `extra` adds compiler flags (e.g. -D tuning macros) to every unit."""
    objdir.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(u[0], u[1], verbose, extra, objdir), UNITS))
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-8000:]}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
