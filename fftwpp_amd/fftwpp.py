"""Python wrapper with the interface of the reference's wrappers/fftwpp.py:
``complex_align``, ``Convolution(shape)`` and ``HConvolution(shape)`` acting on
numpy complex128 arrays through the cfftw++ C ABI of libfftwpp_amd.so.

The convolution is performed in place (``f`` is overwritten), as in the
reference (wrappers/fftwpp.py:240-250, 329-351).  The arithmetic runs on the
GPU; host arrays are staged through device memory by the library.
"""
import numpy as np

from ._lib import lib

__all__ = ["Convolution", "HConvolution", "complex_align",
           "fftwpp_set_maxthreads", "fftwpp_get_maxthreads"]


def complex_align(shape):
    """16-byte aligned complex128 array (wrappers/fftwpp.py:15-20)."""
    dtype = np.dtype(np.complex128)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    buf = np.empty(nbytes + 16, dtype=np.uint8)
    start = -buf.ctypes.data % 16
    return buf[start:start + nbytes].view(dtype).reshape(shape)


def fftwpp_set_maxthreads(nthreads):
    lib().set_fftwpp_maxthreads(nthreads)


def fftwpp_get_maxthreads():
    return lib().get_fftwpp_maxthreads()


def _ptr(a):
    if a.dtype != np.complex128 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous complex128")
    return a.ctypes.data


class Convolution:
    """Implicitly dealiased complex convolution of 1, 2 or 3 dimensions
    (wrappers/fftwpp.py:84-251)."""

    def __init__(self, shape):
        if isinstance(shape, int):
            shape = (shape,)
        self.dim = len(shape)
        self.shape = tuple(int(s) for s in shape)
        L = lib()
        if self.dim == 1:
            self.cptr = L.fftwpp_create_conv1d(*self.shape)
            self._convolve, self._delete = L.fftwpp_conv1d_convolve, L.fftwpp_conv1d_delete
        elif self.dim == 2:
            self.cptr = L.fftwpp_create_conv2d(*self.shape)
            self._convolve, self._delete = L.fftwpp_conv2d_convolve, L.fftwpp_conv2d_delete
        elif self.dim == 3:
            self.cptr = L.fftwpp_create_conv3d(*self.shape)
            self._convolve, self._delete = L.fftwpp_conv3d_convolve, L.fftwpp_conv3d_delete
        else:
            raise ValueError("invalid shape (length/dimension should be 1, 2, or 3)")
        if not self.cptr:
            raise RuntimeError(lib().fftwpp_amd_last_error().decode())

    def __del__(self):
        if getattr(self, "cptr", None):
            self._delete(self.cptr)
            self.cptr = None

    def convolve(self, f, g):
        """Compute the convolution of f and g; the result overwrites f."""
        assert f.shape == self.shape
        assert g.shape == self.shape
        self._convolve(self.cptr, _ptr(f), _ptr(g))


class HConvolution:
    """Implicitly dealiased centred Hermitian convolution
    (wrappers/fftwpp.py:253-351).  shape is the stored shape: (H,),
    (Lx, Hy) or (Lx, Ly, Hz) with H = number of non-negative modes."""

    def __init__(self, shape):
        if isinstance(shape, int):
            shape = (shape,)
        self.dim = len(shape)
        self.shape = tuple(int(s) for s in shape)
        L = lib()
        if self.dim == 1:
            self.cptr = L.fftwpp_create_hconv1d(2 * self.shape[0] - 1)
            self._convolve, self._delete = L.fftwpp_hconv1d_convolve, L.fftwpp_hconv1d_delete
        elif self.dim == 2:
            self.cptr = L.fftwpp_create_hconv2d(self.shape[0], 2 * self.shape[1] - 1)
            self._convolve, self._delete = L.fftwpp_hconv2d_convolve, L.fftwpp_hconv2d_delete
        elif self.dim == 3:
            self.cptr = L.fftwpp_create_hconv3d(self.shape[0], self.shape[1],
                                                2 * self.shape[2] - 1)
            self._convolve, self._delete = L.fftwpp_hconv3d_convolve, L.fftwpp_hconv3d_delete
        else:
            raise ValueError("invalid shape (dimension should be 1, 2, or 3)")
        if not self.cptr:
            raise RuntimeError(lib().fftwpp_amd_last_error().decode())

    def __del__(self):
        if getattr(self, "cptr", None):
            self._delete(self.cptr)
            self.cptr = None

    def convolve(self, f, g):
        """Symmetrise f and g, then overwrite f with their convolution."""
        L = lib()
        if self.dim == 1:
            L.fftwpp_HermitianSymmetrize(_ptr(f))
            L.fftwpp_HermitianSymmetrize(_ptr(g))
        elif self.dim == 2:
            Lx = f.shape[0]
            Hx = (Lx + 1) // 2
            Hy = f.shape[1]
            x0 = Lx // 2
            L.fftwpp_HermitianSymmetrizeX(Hx, Hy, x0, _ptr(f))
            L.fftwpp_HermitianSymmetrizeX(Hx, Hy, x0, _ptr(g))
        elif self.dim == 3:
            Lx, Ly, Hz = f.shape
            Hx, Hy = (Lx + 1) // 2, (Ly + 1) // 2
            x0, y0 = Lx // 2, Ly // 2
            L.fftwpp_HermitianSymmetrizeXY(Hx, Hy, Hz, x0, y0, _ptr(f))
            L.fftwpp_HermitianSymmetrizeXY(Hx, Hy, Hz, x0, y0, _ptr(g))
        self._convolve(self.cptr, _ptr(f), _ptr(g))
