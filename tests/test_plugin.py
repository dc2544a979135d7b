"""hlr::hip::hip_arithmetic registers with HLR's arithmetic trait
(arith/defaults.hh:20-24) from hlr_hip.hh, whichever of the two headers
comes first: compiled here (syntax only, C++20) against the trait and
concept as that file defines them.  The other order, and a generic algorithm
running with the plugin, are exercised by the mvm-hip driver (--norm)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent

TRAIT = """
namespace hlr {
template <typename T> struct is_arithmetic { static constexpr bool value = false; };
template <typename T> inline constexpr bool is_arithmetic_v = is_arithmetic<T>::value;
template <typename T> concept arithmetic_type = is_arithmetic_v<T>;
}
"""

CHECKS = """
static_assert(hlr::arithmetic_type<hlr::hip::hip_arithmetic>);
static_assert(hlr::arithmetic_type<const hlr::hip::hip_arithmetic&>);
static_assert(hlr::arithmetic_type<hlr::hip::hip_arithmetic&>);
static_assert(!hlr::arithmetic_type<int>);
template <hlr::arithmetic_type A> int takes(A&&) { return 1; }
int use() { hlr::hip::hip_arithmetic a; const auto& c = a; return takes(a) + takes(c); }
"""


def _compile(src: str, tmp_path):
    f = tmp_path / "plugin.cpp"
    f.write_text(src)
    return subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++20", "-fsyntax-only", "-x", "hip", "--offload-arch=gfx950",
                           f"-I{ROOT / 'include'}", f"-I{ROOT / 'hlr_amd' / 'csrc' / 'host'}", str(f)],
                          capture_output=True, text=True)


def test_trait_defined_before_header(tmp_path):
    out = _compile(TRAIT + '#include "hlr_hip.hh"\n' + CHECKS, tmp_path)
    assert out.returncode == 0, out.stderr


def test_trait_defined_after_header(tmp_path):
    out = _compile('#include "hlr_hip.hh"\n' + TRAIT + CHECKS, tmp_path)
    assert out.returncode == 0, out.stderr
