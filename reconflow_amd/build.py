"""In-tree build of the native libraries.

* ``reconflow_amd/_build/librfhip.so``   — the HIP solver (gfx950) behind ``include/rfhip.h``;
* ``reconflow_amd/_build/librfsynth.so`` — the host-side seeded pair generator.

Both are compiled every time :func:`build_all` runs. HIP code is built with contraction off and
IEEE division / square root so that the fused and the per-step kernels agree bit for bit, and
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"

HIP_SOURCES = ["rf_kernels.hip", "rf_batch.hip", "rf_solver.hip", "rf_capi_kernels.hip", "rf_noalloc.hip",
               "rf_warmnet.hip"]
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP solver cannot be built")


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build_hip(jobs: int = 8, csrc: Path | None = None, out: Path | None = None) -> Path:
    """Compile every HIP source for gfx950 and link librfhip.so (from ``csrc`` into ``out``,
    by default the package's sources and ``_build``)."""
    csrc = CSRC if csrc is None else Path(csrc)
    out = OUT if out is None else Path(out)
    out.mkdir(parents=True, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = out / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIP_FLAGS, "-c", str(csrc / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    io_obj = out / "rf_io.o"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-c", str(csrc / "rf_io.cpp"), "-o", str(io_obj)])
    objs.append(io_obj)
    lib = out / "librfhip.so"
    tmp = out / "librfhip.so.tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          "-Wl,-rpath,/opt/rocm/lib"])
    tmp.replace(lib)
    return lib


def build_demo() -> Path:
    """Link the C++ command-line program over include/rfhip_cudaflow.hpp (needs both libraries)."""
    exe = OUT / "rf_demo"
    tmp = OUT / "rf_demo.tmp"
    _run(["g++", "-O2", "-std=c++17", str(CSRC / "rf_demo.cpp"), "-I", str(ROOT / "include"), "-L", str(OUT),
          "-lrfhip", "-lrfsynth", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)])
    tmp.replace(exe)
    return exe


def _drain(procs: list) -> None:
    while procs:
        cmd, p = procs.pop(0)
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed: {' '.join(cmd)}\n{out}\n{err}")


def build_synth() -> Path:
    """Compile the host-only pair generator."""
    OUT.mkdir(exist_ok=True)
    lib = OUT / "librfsynth.so"
    tmp = OUT / "librfsynth.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", str(CSRC / "synth.cpp"), "-o",
          str(tmp)])
    tmp.replace(lib)
    return lib


def build_all() -> None:
    build_synth()
    build_hip()
    build_demo()


if __name__ == "__main__":
    build_all()
    print("built", OUT)
