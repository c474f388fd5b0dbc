"""Fused user multipliers (include/fftwpp_amd_fused.h): a device functor
compiled into the library's fused row kernels (tests/support/fusedmult.hip),
the multiplier running inside the residue loop as Convolution::operate calls
it (convolve.h:1120-1135).

The functor weights the product by 1/(1 + c |k|^2), k the signed frequency
of the padded grid point in every dimension, so each case checks that the
row kernels hand the functor the right padded point (l = q s + r within a
row, the outer point through RowLayout).  Oracle: explicit zero padding with
the same weight on the padded grid (oracle.explicit_general, the reference's
multiplier contract), for complex, centred, centred Hermitian (expanded to
the full array) and real plans; FP64 <= 1e-12, FP32 <= 1e-5.  With c = 0 the
fused product equals the library multiplier bit for bit, up to the full 3D
256^3 -> 512^3 workload."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TESTLIB = os.path.join(HERE, "support", "libtestmult.so")
TOL = {"f64": 1e-12, "f32": 1e-5}


def _fn():
    lib = ctypes.CDLL(TESTLIB)
    f = lib.fusedmult_weighted
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                          ctypes.c_double, ctypes.c_void_p]
    return f


def run_fused(plan, out, a, b, c):
    import torch
    from fftwpp_amd import _lib
    N = plan.geometry()[0]
    arr = (ctypes.c_int64 * 3)(*[int(n) for n in N] + [1] * (3 - len(N)))
    rc = _fn()(plan._h, out.data_ptr(), a.data_ptr(), b.data_ptr(), arr, len(N), c, None)
    torch.cuda.synchronize()
    _lib.check(rc)


def weight(N, c):
    def op(Fs, x):
        k2 = 0
        for n, xd in zip(N, x):
            k = np.where(2 * xd <= n, xd, xd - n)
            k2 = k2 + k.astype(np.float64) ** 2
        return [Fs[0] * Fs[1] / (1.0 + c * k2)]
    return op


def _rnd(rng, shape, real=False):
    if real:
        return rng.uniform(-1, 1, shape)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _check(gpu, oracle, shape, kind, want_fn, c, real=False, M=None, m=None, rng_seed=3):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(rng_seed)
    f, g = _rnd(rng, shape, real), _rnd(rng, shape, real)
    dtypes = ((torch.float64, "f64"), (torch.float32, "f32")) if real else \
        ((torch.complex128, "f64"), (torch.complex64, "f32"))
    want = None
    for dtype, key in dtypes:
        plan = DeviceConvolution(shape, kind=kind, dtype=dtype, M=M, m=m)
        if want is None:
            want = want_fn(f, g, plan.geometry()[0])
        a = torch.tensor(f, dtype=dtype, device=gpu)
        b = torch.tensor(g, dtype=dtype, device=gpu)
        out = torch.empty_like(a)
        run_fused(plan, out, a, b, c)
        got = out.cpu().numpy()
        err = oracle.rel_l2(got.astype(np.float64 if real else np.complex128), want)
        assert err < TOL[key], (shape, kind, key, err)
        # in place (out = a) gives the same values
        run_fused(plan, a, a, b, c)
        assert np.array_equal(a.cpu().numpy(), got)


# (shape, M, m): inner rows on the register row kernels (m = 256 with q = 2:
# the paired-residue kernel; other powers of two: rowreg_kernel)
COMPLEX = [((1000,), [2000], [1024]), ((64, 48), [128, 96], [0, 64]),
           ((32, 20, 24), [64, 40, 48], [0, 0, 32]), ((256, 256), [512, 512], [0, 256]),
           ((12, 10, 200), [24, 20, 1000], [0, 0, 256])]


@pytest.mark.parametrize("shape,M,m", COMPLEX)
def test_fused_complex(gpu, oracle, shape, M, m):
    c = 1e-3

    def want(f, g, N):
        return oracle.explicit_general([f, g], tuple(N), weight(N, c))[0]
    _check(gpu, oracle, shape, "complex", want, c, M=M, m=m)


def test_fused_centred(gpu, oracle):
    """Centred rows (k0 = -floor(L/2)) on the register row kernel."""
    c = 2e-3

    def want(f, g, N):
        return oracle.explicit_general([f, g], tuple(N), weight(N, c), centered=True)[0]
    _check(gpu, oracle, (201,), "centered", want, c, M=[302], m=[256])


def _hermitian_want(oracle, lengths, c):
    def want(f, g, N):
        F = oracle.expand_hermitian(f, None)
        G = oracle.expand_hermitian(g, None)
        full = oracle.explicit_general([F, G], tuple(N), weight(N, c), centered=True)[0]
        H = f.shape[-1]
        return full[..., H - 1:]
    return want


# Hermitian rows on the register row kernel: H <= m < L (p = 2, q = 3)
@pytest.mark.parametrize("lengths", [(31,), (15, 31), (255, 255), (7, 9, 31)])
def test_fused_hermitian(gpu, oracle, lengths):
    """Centred Hermitian plans: the functor gets the real residue values of
    the Hermitian rows.  Inputs are symmetrized first (the plan's own
    symmetrizer), as the reference's tests do."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    c = 1e-3
    shape = tuple(lengths[:-1]) + ((lengths[-1] + 1) // 2,)
    mm = [0] * (len(lengths) - 1) + [(lengths[-1] + 1) // 2]
    rng = np.random.default_rng(11)
    plan = DeviceConvolution(shape, kind="hermitian", lengths=lengths, m=mm)
    f = plan.symmetrized(torch.tensor(_rnd(rng, shape), device=gpu)).cpu().numpy()
    g = plan.symmetrized(torch.tensor(_rnd(rng, shape), device=gpu)).cpu().numpy()
    want = _hermitian_want(oracle, lengths, c)(f, g, plan.geometry()[0])
    for dtype, key in ((torch.complex128, "f64"), (torch.complex64, "f32")):
        p = DeviceConvolution(shape, kind="hermitian", dtype=dtype, lengths=lengths, m=mm)
        a = torch.tensor(f, dtype=dtype, device=gpu)
        b = torch.tensor(g, dtype=dtype, device=gpu)
        out = torch.empty_like(a)
        run_fused(p, out, a, b, c)
        err = oracle.rel_l2(out.cpu().numpy().astype(np.complex128), want)
        assert err < TOL[key], (lengths, key, err)


def test_fused_hermitian_config2(gpu, oracle):
    """The 2D centred Hermitian workload (2047 x 1024 stored, M = 3070) with the
    weighted functor against explicit padding at 3070 x 3070."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    c = 1e-6
    lengths = (2047, 2047)
    f, g = oracle.inputs_h2(2047, 1024)
    plan = DeviceConvolution(f.shape, kind="hermitian", lengths=lengths, M=[3070, 3070])
    N = plan.geometry()[0]
    F = oracle.expand_hermitian(f, None)
    G = oracle.expand_hermitian(g, None)
    want = oracle.explicit_general([F, G], tuple(N), weight(N, c), centered=True)[0][..., 1023:]
    del F, G
    a = torch.tensor(f, device=gpu)
    b = torch.tensor(g, device=gpu)
    out = torch.empty_like(a)
    run_fused(plan, out, a, b, c)
    err = oracle.rel_l2(out.cpu().numpy(), want)
    print(f"fused Hermitian 2047x1024 rel L2 {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("shape,M,m", [((64, 48), [128, 96], [64, 64]),
                                       ((40, 32, 36), [96, 64, 72], [32, 0, 64])])
def test_fused_real(gpu, oracle, shape, M, m):
    """Real plans on the real residue transforms (x: m = 64 with q = 2, and
    m = 32 with p = 2, q = 3): the functor sees the kept x planes, which
    RowLayout maps to their padded x points."""
    c = 1e-3

    def want(f, g, N):
        return oracle.explicit_general([f, g], tuple(N), weight(N, c))[0].real
    _check(gpu, oracle, shape, "real", want, c, real=True, M=M, m=m)


def test_fused_unit_weight_is_library_product(gpu, oracle):
    """c = 0: the fused functor computes 1.0 * (F G), which is the library's
    multBinary bit for bit, on the full 3D 256^3 -> 512^3 workload."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    L = 256
    rng = np.random.default_rng(20261202)
    f = torch.tensor(_rnd(rng, (L, L, L)), device=gpu)
    g = torch.tensor(_rnd(rng, (L, L, L)), device=gpu)
    plan = DeviceConvolution((L, L, L), M=[2 * L] * 3)
    out = torch.empty_like(f)
    run_fused(plan, out, f, g, 0.0)
    ref = torch.empty_like(f)
    plan.convolve_out(ref, f, g)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    # the plain product as a caller's functor
    fb = ctypes.CDLL(TESTLIB).fusedmult_binary
    fb.argtypes = [ctypes.c_void_p] * 5
    out.zero_()
    assert fb(plan._h, out.data_ptr(), f.data_ptr(), g.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_fused_refusals(gpu, oracle):
    """Row geometries outside the register row kernels and 1D real plans are
    refused with an error, leaving nothing half-done on the stream."""
    import torch
    from fftwpp_amd import FFTWPPError
    from fftwpp_amd.torch_conv import DeviceConvolution
    plan = DeviceConvolution((5000,))          # a row longer than 4096
    a = torch.ones(5000, dtype=torch.complex128, device=gpu)
    with pytest.raises(FFTWPPError):
        run_fused(plan, a.clone(), a, a, 1e-3)
    plan = DeviceConvolution((1000,), kind="real", dtype=torch.float64)
    a = torch.ones(1000, dtype=torch.float64, device=gpu)
    with pytest.raises(FFTWPPError):
        run_fused(plan, a.clone(), a, a, 1e-3)
