"""hlr_amd: MI355X-native H-/TLR-matrix-vector product (HLR's mul_vec path).

Layers:
  include/hmvm.h, include/hlrgen.h   C-ABI (the drop-in boundary)
  hlr_amd/csrc/device                HIP kernels + plan layout (libhmvm.so)
  hlr_amd/csrc/construct             H-matrix generator (libhlrgen.so)
  hlr_amd/csrc/host/hlr_hip.hh       C++ mirror of hlr::{seq,tbb}::mul_vec and the
                                     arithmetic object over the C-ABI
  hlr_amd/plan.py                    Python binding of a plan (ctypes)
  hlr_amd/batch.py                   batches of products of many plans
  hlr_amd/autograd.py                torch.autograd binding of the adjoint
  hlr_amd/hmatrix.py                 generator binding and H-matrix files
"""
__all__ = ["autograd", "batch", "hmatrix", "plan"]
