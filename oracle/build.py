"""Build the oracle (CPU FP64 restatement of HLR's mul_vec) into oracle/liboracle.so.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load the result.  The reference itself cannot be compiled
here (HLIBpro is an empty submodule), so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


# compiler flags of the restatement (quoted by bench.py's cpu_baseline)
CFLAGS = ["-O3", "-fPIC", "-fopenmp", "-march=x86-64-v3", "-std=c11"]


def build() -> Path:
    out = HERE / "liboracle.so"
    tmp = out.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, "-shared", f"-I{ROOT / 'include'}", str(HERE / "hmvm_oracle.c"), "-o", str(tmp), "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
