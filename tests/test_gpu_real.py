"""Real convolutions (fftPadReal as the outer or only dimension,
tests/hybridconvr{,2,3}.cc) on the GPU against the CPU oracle.

On the device a binary convolution of real f and g is the halved imaginary
part of the complex self-convolution of z = f + i g.  Inputs are the
reference's ramps f_a[j] = j + a + 1 checked against the direct sums
(tests/direct.cc), over the reference's real test grid (tests/tests.py:615-653),
and random data against explicit zero padding.  Bar: relative L2 error
<= 1e-12 in FP64, <= 1e-5 in FP32 against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def real_grid():
    # tests/tests.py:615-653: (L, M, m)
    yield 8, 16, 16
    for M in [16, 24, 64]:
        for L in [8, 3]:
            yield L, M, 8
    for M in [16, 24, 32]:
        for L in [8, 5]:
            yield L, M, 4
    for L in [8, 7]:
        yield L, 16, 2
    for L in [9, 7]:
        yield L, 63, 3
    for L in [24, 21]:
        yield L, 96, 4


def run_real(gpu, f, g, dtype=None, M=None, m=None):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    dtype = dtype or torch.float64
    plan = DeviceConvolution(f.shape, kind="real", dtype=dtype, M=M, m=m)
    a = torch.tensor(np.ascontiguousarray(f), dtype=dtype, device=gpu)
    b = torch.tensor(np.ascontiguousarray(g), dtype=dtype, device=gpu)
    out = torch.empty_like(a)
    plan.convolve_out(out, a, b)
    plan.convolve_(a, b)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64), a.cpu().numpy().astype(np.float64), plan


@pytest.mark.parametrize("L,M,m", list(real_grid()))
def test_real1d_reference_grid(gpu, oracle, L, M, m):
    f = np.arange(L, dtype=np.float64) + 1
    g = np.arange(L, dtype=np.float64) + 2
    want = oracle.direct_conv(f.astype(complex), g.astype(complex)).real
    h, h_inplace, plan = run_real(gpu, f, g, M=[M], m=[m])
    assert plan.geometry()[1][0] == m
    assert oracle.rel_l2(h, want) < TOL64
    assert np.array_equal(h, h_inplace)


@pytest.mark.parametrize("L", [1, 2, 100, 1000, 3000, 5000])
def test_real1d_sizes(gpu, oracle, L):
    import torch
    rng = np.random.default_rng(L)
    f, g = rng.uniform(-1, 1, L), rng.uniform(-1, 1, L)
    want = oracle.explicit_conv(f, g, 2 * L).real
    for dtype, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        h, _, _ = run_real(gpu, f, g, dtype=dtype)
        assert oracle.rel_l2(h, want) < tol, dtype


def test_real2d_3d(gpu, oracle):
    import torch
    # tests/hybridconvr2.cc / hybridconvr3.cc ramp inputs against direct sums
    f2 = oracle.ramp2(8, 8, 0).real.astype(np.float64)
    g2 = oracle.ramp2(8, 8, 1).real.astype(np.float64)
    h, _, _ = run_real(gpu, f2, g2)
    assert oracle.rel_l2(h, oracle.direct_conv2(f2.astype(complex), g2.astype(complex)).real) < TOL64
    f3 = oracle.ramp3(4, 5, 6, 0).real.astype(np.float64)
    g3 = oracle.ramp3(4, 5, 6, 1).real.astype(np.float64)
    h, _, _ = run_real(gpu, f3, g3)
    assert oracle.rel_l2(h, oracle.direct_conv3(f3.astype(complex), g3.astype(complex)).real) < TOL64
    rng = np.random.default_rng(29)
    for shape in [(37, 29), (16, 12, 20), (300, 256)]:
        f, g = rng.uniform(-1, 1, shape), rng.uniform(-1, 1, shape)
        want = oracle.explicit_conv(f, g, tuple(2 * L for L in shape)).real
        for dtype, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
            h, _, _ = run_real(gpu, f, g, dtype=dtype)
            assert oracle.rel_l2(h, want) < tol, (shape, dtype)


def test_real_rejects_adjoint_and_complex_dtype(gpu, oracle):
    """Complex dtypes are refused; so is the adjoint of a geometry that runs
    as the packed complex self-convolution (m = 8 here)."""
    import torch
    from fftwpp_amd import FFTWPPError
    from fftwpp_amd.torch_conv import DeviceConvolution
    with pytest.raises(TypeError):
        DeviceConvolution((16,), kind="real", dtype=torch.complex128)
    plan = DeviceConvolution((8,), kind="real", dtype=torch.float64, M=[16], m=[8])
    a = torch.ones(8, dtype=torch.float64, device=gpu)
    with pytest.raises(FFTWPPError):
        plan.adjoint(a, a, a)


def _correlate(w, b):
    import scipy.signal as ss
    # c[i] = sum_j w[i + j] b[j] over the stored box
    full = ss.correlate(w, b, mode="full")
    return full[tuple(slice(s - 1, 2 * s - 1) for s in w.shape)]


@pytest.mark.parametrize("shape", [(1000,), (64, 48), (33, 64, 40)])
def test_real_adjoint(gpu, oracle, shape):
    """The adjoint of a real convolution on the real residue transforms:
    grad_a = w (x) b and grad_b = w (x) a (correlations, scipy.signal), and
    the adjoint identity <grad_a, d> = <w, d (*) b>."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(len(shape) + 40)
    a, b, w, dd = (rng.uniform(-1, 1, shape) for _ in range(4))
    want_a, want_b = _correlate(w, b), _correlate(w, a)
    for dtype, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        plan = DeviceConvolution(shape, kind="real", dtype=dtype)
        N, m, q = plan.geometry()
        assert 16 <= m[0] <= 4096 and m[0] & (m[0] - 1) == 0 and shape[0] <= 2 * m[0]
        T = lambda x: torch.tensor(x, dtype=dtype, device=gpu)
        ga, gb = plan.adjoint(T(w), T(a), T(b))
        h = plan(T(dd), T(b))
        torch.cuda.synchronize()
        ga = ga.cpu().numpy().astype(np.float64)
        assert oracle.rel_l2(ga, want_a) < tol, dtype
        assert oracle.rel_l2(gb.cpu().numpy().astype(np.float64), want_b) < tol, dtype
        lhs = float(np.sum(ga * dd))
        rhs = float(np.sum(w * h.cpu().numpy().astype(np.float64)))
        assert abs(lhs - rhs) <= tol * np.linalg.norm(w) * np.linalg.norm(dd) * np.linalg.norm(b), dtype


def _ratio_inputs(shape, ratio, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, shape)
    g = rng.uniform(-1, 1, shape) * ratio
    return f, g


@pytest.mark.parametrize("ratio", [1.0, 1e-3, 1e-6, 1e-9, 0.0, 1e6])
def test_real_unbalanced_norms(gpu, oracle, ratio):
    """Inputs whose norms differ by up to nine orders of magnitude (and a zero
    input): the error stays relative to ||f|| ||g|| (the norms are balanced by
    an exact power of two before the complex self-convolution).  1D L = 4096
    and 2D 64 x 48 against the direct sums (tests/direct.cc), 3D 12 x 10 x 14
    too; FP64 <= 1e-12, FP32 <= 1e-5."""
    import torch
    cases = [((4096,), oracle.direct_conv), ((64, 48), oracle.direct_conv2),
             ((12, 10, 14), oracle.direct_conv3)]
    for shape, direct in cases:
        f, g = _ratio_inputs(shape, ratio, len(shape) + 7)
        want = direct(f.astype(complex), g.astype(complex)).real
        for dtype, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
            h, h_inplace, _ = run_real(gpu, f, g, dtype=dtype)
            err = oracle.rel_l2(h, want)
            assert err < tol, (shape, ratio, dtype, err)
            assert np.array_equal(h, h_inplace)
            if ratio == 0.0:
                assert not np.any(h)
