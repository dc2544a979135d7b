"""The C++ mirror of the reference surface (hlr::hip::mul_vec, tlr::mul_vec,
hip_arithmetic) through the mvm-hip driver program (GPU): the reference's
own self-consistency check (y_ref from one call with alpha = 100 against
50 calls with alpha = 2) must agree to FP64 rounding."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent.parent / "hlr_amd" / "lib" / "mvm-hip"


@pytest.mark.parametrize("args,tol", [
    (["--app", "logkernel", "-n", "4096"], 1e-13),
    (["--app", "invr", "-n", "8192", "--op", "transposed"], 1e-13),
    (["--app", "materncov", "-n", "4096", "--ntile", "512", "--tlr", "--xrand"], 1e-13),
    (["--app", "invr", "-n", "8192", "--prec", "valr"], 1e-6),
])
def test_driver(args, tol):
    print(f"[gpu] mvm-hip {' '.join(args)}", flush=True)
    out = subprocess.run([str(DRIVER), *args, "--nbench", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"error\s+=\s+([0-9.eE+-]+)", out.stdout)
    assert m, out.stdout
    assert float(m.group(1)) <= tol, out.stdout


def test_arithmetic_plugin_generic_algorithm():
    """hip_arithmetic as HLR's arithmetic plugin: a generic power iteration
    constrained by the arithmetic concept (norm::spectral's shape) runs with
    it on host vectors and agrees with the device spectral norm."""
    print("[gpu] mvm-hip --norm", flush=True)
    out = subprocess.run([str(DRIVER), "--app", "logkernel", "-n", "4096", "--norm"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"generic\s+([0-9.eE+-]+)\s+device\s+([0-9.eE+-]+)\s+rel\.diff\s+([0-9.eE+-]+)", out.stdout)
    assert m, out.stdout
    assert float(m.group(1)) > 0 and float(m.group(3)) <= 1e-10, out.stdout
