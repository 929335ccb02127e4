"""Build the native pieces in-tree.

* ``glisser_amd/libglisser_amd.so`` -- the HIP kernels and the C ABI, for
  gfx950, with ``hipcc`` (cross-compiles without a GPU).
* ``tests/cxx/executor_drive`` -- a C++ caller of the C ABI and of
  ``glx::Executor`` (g++, linked to the library) that the GPU tests run.
* ``oracle/liboracle.so`` and, where the reference tree exists,
  ``oracle/_ref/libglisser_ref.so`` -- the CPU checker (``make -C oracle``).

Run ``python -m glisser_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "glisser_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libglisser_amd.so"
ARCH = os.environ.get("GLX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["wh_kernels.hip", "compact.hip", "interp.hip", "batch.hip", "track.hip", "sincos.hip", "static_part.hip", "glx_api.cpp", "batch_api.cpp"]

# Host-only C++ (planet producer, conversions): g++ with the reference CPU
# build's arithmetic -- no FP contraction, no fast-math.
HOST_SOURCES = ["host/convert.cpp", "host/planets.cpp", "host/host_api.cpp", "host/executor.cpp", "host/formats.cpp",
                "host/formats_api.cpp", "host/history.cpp", "host/history_api.cpp"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-fno-fast-math"]

# Packed-FP32 instructions stay off: plain float code must not be fused into
# v_pk_* forms on this target.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP extension cannot be built")
    return path


def build_hip(out: Path = LIB) -> Path:
    """Compile every HIP source into one shared library (object files are
    built next to it in a temporary folder and linked)."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        _run([hipcc(), *HIP_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    cxx = shutil.which("g++") or "g++"
    for src in HOST_SOURCES:
        obj = objdir / ("host_" + Path(src).stem + ".o")
        _run([cxx, *HOST_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, out)
    return out


CXX_TEST = ROOT / "tests" / "cxx" / "executor_drive"


def build_cxx_test(out: Path = CXX_TEST) -> Path:
    """The C++ caller of the test suite: g++ against include/*.h, linked to
    the library (found next to it at run time through a relative rpath)."""
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(out.with_suffix(".cpp")),
          f"-L{PKG}", "-lglisser_amd", "-Wl,-rpath,$ORIGIN/../../glisser_amd", "-o", str(out)])
    return out


def build_oracle() -> None:
    targets = ["oracle"]
    if Path("/root/reference/src/wh.cpp").exists():
        targets.append("ref")
    _run(["make", "-C", str(ROOT / "oracle"), *targets])


def build_all() -> None:
    build_hip()
    build_cxx_test()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "hip" in what:
            build_hip()
        if "oracle" in what:
            build_oracle()
        if "cxx" in what:
            build_cxx_test()
