"""PyTorch face of the convolution layer: dealiased convolutions of device
tensors (complex128 or complex64) on the current HIP stream, and a
torch.autograd.Function whose backward is the hand-written adjoint
(dealiased correlations on the same padded residues).

PyTorch is plumbing here: tensors provide device memory and the stream; all
arithmetic runs in libfftwpp_amd.so.
"""
import ctypes

import torch

from ._lib import F32, F64, check, lib, sizes

KINDS = {"complex": 0, "hermitian": 1, "centered": 2, "real": 3}


def _dtype_code(dtype, kind="complex"):
    if kind == "real":
        if dtype == torch.float64:
            return F64
        if dtype == torch.float32:
            return F32
        raise TypeError("real convolutions take torch.float64 or torch.float32")
    if dtype == torch.complex128:
        return F64
    if dtype == torch.complex64:
        return F32
    raise TypeError("dtype must be torch.complex128 or torch.complex64")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceConvolution:
    """Plan of a dealiased convolution of device tensors.

    shape: stored shape of each input.  kind="complex": (L,), (Lx,Ly) or
    (Lx,Ly,Lz) (HybridConvolution{,2,3}); kind="hermitian": (H,), (Lx,Hy) or
    (Lx,Ly,Hz) with odd full length 2H-1 on the last axis unless
    ``lengths`` gives the full lengths (HybridConvolutionHermitian{,2,3});
    kind="centered": (L,) centred 1D; kind="real": (L,), (Lx,Ly) or
    (Lx,Ly,Lz) real data (fftPadReal outer dimension, tests/hybridconvr*.cc;
    dtype torch.float64 or torch.float32).  M and m optionally give the
    minimum padded sizes and forced sub-FFT sizes per dimension.

    tune=True times candidate sub-FFT sizes on the current device when the
    plan is made (fftwpp_amd_conv_create_tuned: the reference's optimizer,
    fftBase::OptBase, convolve.h:164-236, on the GPU) and keeps the fastest;
    the choice is cached, so a later plan of the same problem takes no
    timing (retune=True times again).  It excludes m.
    """

    def __init__(self, shape, kind="complex", dtype=torch.complex128,
                 M=None, m=None, lengths=None, tune=False, retune=False):
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.kind = kind
        self.dtype = dtype
        self.dim = len(self.shape)
        if lengths is None:
            lengths = list(self.shape)
            if kind == "hermitian":
                lengths[-1] = 2 * self.shape[-1] - 1
        self.lengths = tuple(int(v) for v in lengths)
        h = ctypes.c_void_p()
        if tune:
            if m is not None:
                raise ValueError("tune=True chooses the sub-FFT sizes: give no m")
            check(lib().fftwpp_amd_conv_create_tuned(
                ctypes.byref(h), self.dim, KINDS[kind], sizes(self.lengths),
                sizes(M) if M is not None else None, _dtype_code(dtype, kind),
                1 if retune else 0, None))
        else:
            check(lib().fftwpp_amd_conv_create(
                ctypes.byref(h), self.dim, KINDS[kind], sizes(self.lengths),
                sizes(M) if M is not None else None,
                sizes(m) if m is not None else None, _dtype_code(dtype, kind)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().fftwpp_amd_conv_destroy(h)
            except Exception:  # interpreter shutdown: module globals already gone
                pass
            self._h = None

    def geometry(self):
        """(N, m, q) per dimension: padded length, sub-FFT size, residues."""
        N, m, q = sizes([0] * 3), sizes([0] * 3), sizes([0] * 3)
        check(lib().fftwpp_amd_conv_geometry(self._h, N, m, q))
        d = self.dim
        return tuple(N[:d]), tuple(m[:d]), tuple(q[:d])

    def _check(self, t):
        if not t.is_cuda:
            raise ValueError("tensors must live on the GPU")
        if t.dtype != self.dtype or tuple(t.shape) != self.shape:
            raise ValueError(f"expected {self.dtype} tensor of shape {self.shape}")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")

    def convolve_(self, a, b):
        """a <- a (*) b in place (normalised)."""
        self._check(a)
        self._check(b)
        check(lib().fftwpp_amd_conv_execute(self._h, a.data_ptr(), b.data_ptr(), _stream()))
        return a

    def convolve_out(self, out, a, b):
        """out <- a (*) b (normalised); a and b are only read."""
        self._check(out)
        self._check(a)
        self._check(b)
        check(lib().fftwpp_amd_conv_execute_out(self._h, out.data_ptr(), a.data_ptr(),
                                                b.data_ptr(), _stream()))
        return out

    def __call__(self, a, b):
        out = a.clone()
        return self.convolve_(out, b)

    def set_memory_limit(self, nbytes):
        """Keep the plan's device work buffers within nbytes (2D and 3D
        plans; 0 = no limit): a 3D plan the plane kernel takes runs its y-z
        stage there, other plans run their outer residues in blocks."""
        check(lib().fftwpp_amd_conv_set_memory_limit(self._h, int(nbytes)))

    def work_bytes(self):
        """(bytes, plane): device work bytes of the next execution and
        whether it takes the plane kernel (3D complex plans)."""
        n = ctypes.c_size_t()
        p = ctypes.c_int()
        check(lib().fftwpp_amd_conv_work_bytes(self._h, ctypes.byref(n), ctypes.byref(p)))
        return n.value, bool(p.value)

    def _plane(self):
        """Self-conjugate plane (Hx, Hy, x0, y0, Sx, Sy) of a Hermitian plan's
        stored array (the arguments of fftwpp_amd_hermitian_plane)."""
        if self.dim == 1:
            return (1, 1, 0, 0, 1, 1)
        Lx, Ly = self.lengths[0], self.lengths[1]
        if Lx % 2 == 0 or (self.dim == 3 and Ly % 2 == 0):
            raise ValueError("centred dimensions of a differentiable Hermitian "
                             "convolution must have odd length")
        if self.dim == 2:
            return ((Lx + 1) // 2, 1, Lx // 2, 0, (Ly + 1) // 2, 1)
        Hz = (self.lengths[2] + 1) // 2
        return ((Lx + 1) // 2, (Ly + 1) // 2, Lx // 2, Ly // 2, Ly * Hz, Hz)

    def symmetrized(self, t):
        """Copy of t with the self-conjugate plane made Hermitian
        (HermitianSymmetrize{,X,XY}, convolve.h:1168-1256)."""
        self._check(t)
        out = t.clone()
        Hx, Hy, x0, y0, Sx, Sy = self._plane()
        check(lib().fftwpp_amd_hermitian_plane(_dtype_code(self.dtype), 0, Hx, Hy, x0, y0,
                                               out.data_ptr(), Sx, Sy, _stream()))
        return out

    def adjoint(self, w, a, b, need_a=True, need_b=True):
        """Cotangents of (a, b) for output cotangent w (PyTorch convention
        grad = dL/d conj(z)).  Complex and real kinds: grad_a = w (x) b,
        grad_b = w (x) a (dealiased correlations; real plans whose outer
        dimension runs on the real residue transforms).  Hermitian kind: the cotangents
        of h = conv(S a, S b) with S the symmetrizer of the self-conjugate
        plane, with respect to the stored modes."""
        self._check(w)
        ga = torch.empty_like(w) if need_a else None
        gb = torch.empty_like(w) if need_b else None
        check(lib().fftwpp_amd_conv_adjoint(
            self._h, w.data_ptr(), a.data_ptr(), b.data_ptr(),
            ga.data_ptr() if ga is not None else None,
            gb.data_ptr() if gb is not None else None, _stream()))
        return ga, gb


class _ConvolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, plan):
        a = a.contiguous()
        b = b.contiguous()
        ctx.plan = plan
        ctx.save_for_backward(a, b)
        if plan.kind == "hermitian":
            return plan(plan.symmetrized(a), plan.symmetrized(b))
        return plan(a, b)

    @staticmethod
    def backward(ctx, w):
        a, b = ctx.saved_tensors
        ga, gb = ctx.plan.adjoint(w.contiguous(), a, b,
                                  ctx.needs_input_grad[0], ctx.needs_input_grad[1])
        return ga, gb, None


def convolve(a, b, plan):
    """Differentiable dealiased convolution h = a (*) b with ``plan``.  For a
    Hermitian plan the inputs are symmetrized first (h = S a (*) S b), so the
    function and its gradient are defined for any stored arrays."""
    return _ConvolveFn.apply(a, b, plan)


class UserConvolution:
    """Plan of a dealiased convolution of A device tensors into B with a user
    device multiplier (fftwpp_amd_conv_create_general; Application(A,B,mult)
    of convolve.h:84-121 with a multiplier that runs on the device).

    ``mult`` is the address of an ``fftwpp_amd_multiplier`` callback compiled
    into a HIP library (e.g. ``ctypes.cast(cdll.my_mult, ctypes.c_void_p)``);
    it is called once per convolution with the whole padded transforms.
    ``shape`` is (L,), (Lx,Ly) or (Lx,Ly,Lz) of contiguous complex data;
    ``centered`` makes every dimension centred.  ``user`` is passed through.
    """

    def __init__(self, shape, A, B, mult, dtype=torch.complex128, centered=False,
                 M=None, m=None, user=None):
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.dim = len(self.shape)
        self.A, self.B = int(A), int(B)
        self.dtype = dtype
        self._mult = ctypes.c_void_p(mult.value if isinstance(mult, ctypes.c_void_p) else int(mult))
        h = ctypes.c_void_p()
        check(lib().fftwpp_amd_conv_create_general(
            ctypes.byref(h), self.dim, int(bool(centered)), sizes(self.shape),
            sizes(M) if M is not None else None, sizes(m) if m is not None else None,
            _dtype_code(dtype), self.A, self.B, self._mult, user))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().fftwpp_amd_conv_destroy(h)
            except Exception:  # interpreter shutdown: module globals already gone
                pass
            self._h = None

    def geometry(self):
        """(N, m, q) per dimension: padded length, sub-FFT size, residues."""
        N, m, q = sizes([0] * 3), sizes([0] * 3), sizes([0] * 3)
        check(lib().fftwpp_amd_conv_geometry(self._h, N, m, q))
        d = self.dim
        return tuple(N[:d]), tuple(m[:d]), tuple(q[:d])

    def __call__(self, *arrays):
        """In place on max(A,B) device tensors: inputs arrays[:A], the
        normalised products written to arrays[:B], which are returned."""
        n = max(self.A, self.B)
        if len(arrays) != n:
            raise ValueError(f"expected {n} tensors")
        for t in arrays:
            if not t.is_cuda or t.dtype != self.dtype or tuple(t.shape) != self.shape \
                    or not t.is_contiguous():
                raise ValueError(f"expected contiguous {self.dtype} GPU tensors of shape {self.shape}")
        ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in arrays])
        check(lib().fftwpp_amd_conv_execute_general(self._h, ptrs, _stream()))
        return arrays[:self.B]
