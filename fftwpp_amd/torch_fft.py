"""PyTorch face of the standalone FFTs (the fft1d / mfft1d / fft2d / fft3d /
rcfft1d / crfft1d / mrcfft1d / mcrfft1d transforms of fftw++.h:556-1444):
unnormalised transforms out_k = scale * sum_j in_j exp(sign 2 pi i j k / n)
of device tensors on the current HIP stream.  All arithmetic runs in
libfftwpp_amd.so."""
import ctypes

import torch

from ._lib import F32, F64, check, lib, sizes

FFT_KINDS = {"c2c": 0, "r2c": 1, "c2r": 2}


class DeviceFFT:
    """Plan of a transform of shape ``n`` (an int, or a tuple of up to three
    lengths for complex multidimensional transforms).  ``dtype`` is the
    precision (torch.float64 or torch.float32); complex data are the matching
    complex tensors.  For one dimension, ``howmany`` transforms with element
    j of transform t at j*stride + t*dist (in elements of each side's type;
    0 selects the contiguous defaults) as mfft1d / mrcfft1d / mcrfft1d."""

    def __init__(self, n, sign=-1, kind="c2c", dtype=torch.float64, howmany=1,
                 istride=0, ostride=0, idist=0, odist=0):
        n = (n,) if isinstance(n, int) else tuple(int(v) for v in n)
        self.n, self.sign, self.kind = n, sign, kind
        code = F64 if dtype in (torch.float64, torch.complex128) else F32
        if code == F32 and dtype not in (torch.float32, torch.complex64):
            raise TypeError("dtype must be float64 or float32")
        self.real = torch.float64 if code == F64 else torch.float32
        self.cplx = torch.complex128 if code == F64 else torch.complex64
        h = ctypes.c_void_p()
        check(lib().fftwpp_amd_fft_create(ctypes.byref(h), FFT_KINDS[kind], len(n), sizes(n),
                                          int(sign), code, howmany, istride, ostride, idist,
                                          odist))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().fftwpp_amd_fft_destroy(h)
            except Exception:  # interpreter shutdown: module globals already gone
                pass
            self._h = None

    def normalization(self):
        return lib().fftwpp_amd_fft_normalization(self._h)

    def __call__(self, x, out, scale=1.0):
        """out <- scale * FFT(x) on device tensors (out may be x)."""
        want_in = self.real if self.kind == "r2c" else self.cplx
        want_out = self.real if self.kind == "c2r" else self.cplx
        if not (x.is_cuda and out.is_cuda) or x.dtype != want_in or out.dtype != want_out:
            raise ValueError(f"expected GPU tensors {want_in} -> {want_out}")
        if not (x.is_contiguous() and out.is_contiguous()):
            raise ValueError("tensors must be contiguous (strides are given to the plan)")
        check(lib().fftwpp_amd_fft_execute(self._h, x.data_ptr(), out.data_ptr(), float(scale),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out
