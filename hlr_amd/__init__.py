"""hlr_amd: MI355X-native H-/TLR-matrix-vector product (HLR's mul_vec path).

Layers:
  include/hmvm.h, include/hlrgen.h   C-ABI (the drop-in boundary)
  hlr_amd/csrc/device                HIP kernels + plan layout (libhmvm.so)
  hlr_amd/csrc/construct             H-matrix generator (libhlrgen.so)
  hlr_amd/arith.py                   mirror of hlr::{seq,tbb}::mul_vec and the
                                     arithmetic object over the C-ABI
"""
__all__ = ["arith", "hmatrix", "plan"]
