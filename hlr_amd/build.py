"""Build the native libraries in-tree.

    libhlrgen.so  host C++ generator (g++, OpenMP)
    libhmvm.so    HIP device library for gfx950 (hipcc, OpenMP host code)

The oracle library (oracle/liboracle.so) is test infrastructure and is built
by oracle/build.py, which __graft_entry__.build() calls as well.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "hlr_amd"
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# host code runs on the GPU box's CPU too: no -march=native
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-fPIC", "-fopenmp", "-std=c++17"]
# packed FP32 instructions are disabled for all device code
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} -> {cmd[-1]}")


def build_hlrgen() -> Path:
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libhlrgen.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{INCLUDE}",
          str(CSRC / "construct" / "hlrgen.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_hmvm() -> Path:
    LIBDIR.mkdir(exist_ok=True)
    dev = CSRC / "device"
    srcs = ("plan_build.cpp", "hmfile.cpp", "hmvm.hip", "mrhs.hip", "batch.hip", "vecops.hip", "uniform.hip",
            "construct.hip")
    objs = [str(LIBDIR / (Path(src).stem + ".o")) for src in srcs]

    def compile_one(src: str, obj: str) -> None:
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        flags = DEVICE_FLAGS + [f"--offload-arch={ARCH}"] if src.endswith(".hip") else []
        _run([HIPCC, "-O3", "-fPIC", "-fopenmp", "-std=c++17", "-march=x86-64-v3",
              f"-I{INCLUDE}", f"-I{dev}", *flags, *lang, "-c", str(dev / src), "-o", obj])

    # the translation units are independent: compile them side by side
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1)) as pool:
        for f in [pool.submit(compile_one, s, o) for s, o in zip(srcs, objs)]:
            f.result()
    out = LIBDIR / "libhmvm.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fopenmp", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_driver() -> Path:
    """mvm-hip: benchmark driver over the C++ mirror hlr::hip::mul_vec"""
    out = LIBDIR / "mvm-hip"
    tmp = LIBDIR / "mvm-hip.tmp"
    _run([HIPCC, "-O2", "-std=c++20", f"-I{INCLUDE}", f"-I{CSRC / 'host'}", str(CSRC / "host" / "mvm_hip.cpp"),
          f"-L{LIBDIR}", "-lhmvm", "-lhlrgen", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> list[Path]:
    return [build_hlrgen(), build_hmvm(), build_driver()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
