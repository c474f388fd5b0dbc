"""fftwpp_amd: MI355X-native hybrid implicitly dealiased convolutions.

Modules
  fftwpp      -- the reference Python wrapper's interface (numpy, host arrays)
  torch_conv  -- device tensors, FP64/FP32, autograd with the adjoint
  _lib        -- ctypes binding of libfftwpp_amd.so
  build       -- compiles libfftwpp_amd.so in-tree for gfx950
"""
from ._lib import LIBPATH, FFTWPPError, lib  # noqa: F401

__all__ = ["lib", "LIBPATH", "FFTWPPError"]
