"""Batches of independent 1D complex dealiased convolutions of uneven lengths
on device tensors (complex128 or complex64), on the current HIP stream.

FFTW++ has no batch interface: a user of the reference builds one
``Convolution`` per length and calls ``convolve`` once per pair
(convolve.h:982-1166).  ``BatchedConvolution`` replaces those n objects and
n calls with one plan whose items are packed back to back: item i occupies
``[off[i], off[i+1])`` of the packed tensors f, g and out.  With
``graph=True`` the first call on a set of tensors records a hipGraph that
later calls replay in one launch; with ``tune=True`` the plan times candidate
sub-FFT sizes and tile shapes per length bucket at construction.
"""
import ctypes

import numpy as np
import torch

from ._lib import F32, F64, MULT_BINARY, MULT_CORRELATION, MULT_NONE, check, lib

_MULTS = {"binary": MULT_BINARY, "correlation": MULT_CORRELATION, "none": MULT_NONE}
TUNE, GRAPH = 1, 2


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class BatchedConvolution:
    """Plan for n convolutions of lengths ``lengths`` (minimum padded lengths
    ``M``, default 2L-1) in precision ``dtype``; ``mult`` is "binary"
    (f*g), "correlation" (f*conj g) or "none"."""

    def __init__(self, lengths, M=None, dtype=torch.complex128, mult="binary",
                 tune=False, graph=True):
        L = np.asarray(lengths, dtype=np.uint64)
        if L.ndim != 1 or L.size == 0:
            raise ValueError("lengths must be a non-empty 1D sequence")
        if dtype == torch.complex128:
            code = F64
        elif dtype == torch.complex64:
            code = F32
        else:
            raise TypeError("dtype must be torch.complex128 or torch.complex64")
        n = int(L.size)
        Larr = (ctypes.c_size_t * n)(*[int(v) for v in L])
        Marr = None
        if M is not None:
            Mv = np.broadcast_to(np.asarray(M, dtype=np.uint64), L.shape)
            Marr = (ctypes.c_size_t * n)(*[int(v) for v in Mv])
        flags = (TUNE if tune else 0) | (GRAPH if graph else 0)
        h = ctypes.c_void_p()
        check(lib().fftwpp_amd_batch_create(ctypes.byref(h), n, Larr, Marr, code,
                                            _MULTS[mult], flags))
        self._h = h
        self.n = n
        self.dtype = dtype
        self.lengths = L.astype(np.int64)
        off = (ctypes.c_size_t * (n + 1))()
        check(lib().fftwpp_amd_batch_offsets(h, off))
        self.offsets = np.array(off[:], dtype=np.int64)
        self.size = int(self.offsets[-1])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().fftwpp_amd_batch_destroy(h)
            except Exception:  # interpreter shutdown: module globals already gone
                pass
            self._h = None

    def geometry(self, i):
        """(m, q, N) chosen for item i."""
        m, q, N = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().fftwpp_amd_batch_geometry(self._h, int(i), ctypes.byref(m),
                                              ctypes.byref(q), ctypes.byref(N)))
        return m.value, q.value, N.value

    def describe(self):
        """One line per launch class: m, q, item count, tile shape."""
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().fftwpp_amd_batch_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def _check(self, t, name):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a device tensor")
        if t.dtype != self.dtype:
            raise TypeError(f"{name} must be {self.dtype}")
        if not t.is_contiguous() or t.numel() != self.size:
            raise ValueError(f"{name} must be contiguous with {self.size} elements")

    def __call__(self, f, g, out=None):
        """out[off_i .. off_i+L_i) <- f_i (*) g_i for every item; returns out."""
        self._check(f, "f")
        self._check(g, "g")
        if out is None:
            out = torch.empty_like(f)
        self._check(out, "out")
        check(lib().fftwpp_amd_batch_execute(self._h, out.data_ptr(), f.data_ptr(),
                                             g.data_ptr(), _stream()))
        return out

    def pack(self, arrays, device="cuda"):
        """Pack a list of n 1D arrays into one device tensor."""
        if len(arrays) != self.n:
            raise ValueError("need one array per item")
        host = np.concatenate([np.asarray(a).reshape(-1) for a in arrays])
        if host.size != self.size:
            raise ValueError("array lengths do not match the plan")
        return torch.as_tensor(host, dtype=self.dtype, device=device)

    def unpack(self, t):
        """Split a packed tensor into the n items (numpy, on the host)."""
        h = t.detach().cpu().numpy()
        return [h[self.offsets[i]:self.offsets[i + 1]] for i in range(self.n)]
