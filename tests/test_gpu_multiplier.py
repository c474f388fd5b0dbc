"""User device multipliers (the multiplier/Indices contract, convolve.h:48-82,
Convolution::operate convolve.h:1120-1135) on the GPU against the oracle.

The multipliers are compiled from tests/support/multipliers.hip, as a user of
the library writes them: a ternary product (A = 3), autoconvolution (A = 1),
two outputs (B = 2: the product and the difference, whose result must be
f - g exactly) and a product weighted by a function of the padded grid point
x (reads the layout the library hands over).  The oracle is explicit
zero padding with the same pointwise operation (oracle.explicit_general,
pinned against direct sums and the shift identity); the ternary product is
also checked against the direct triple sums.  Bar: relative L2 error
<= 1e-12 in FP64, <= 1e-5 in FP32 against the FP64 oracle."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))
TESTLIB = os.path.join(HERE, "support", "libtestmult.so")


@pytest.fixture(scope="module")
def mults(gpu):
    if not os.path.exists(TESTLIB):
        raise RuntimeError("tests/support/libtestmult.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(TESTLIB)


def addr(lib, name):
    return ctypes.cast(getattr(lib, name), ctypes.c_void_p)


def weight(x, N):
    # the weight of testmult_weighted at grid point x (tuple of index arrays)
    t = sum((d + 1) * x[d] / N[d] for d in range(len(N)))
    return 1.0 + 0.5 * np.cos(2 * np.pi * x[0] / N[0]) + 0.25 * np.sin(2 * np.pi * t)


OPS = {
    "testmult_ternary": (3, 1, lambda F, x, N: [F[0] * F[1] * F[2]]),
    "testmult_square": (1, 1, lambda F, x, N: [F[0] * F[0]]),
    "testmult_pair": (2, 2, lambda F, x, N: [F[0] * F[1], F[0] - F[1]]),
    "testmult_weighted": (2, 1, lambda F, x, N: [F[0] * F[1] * weight(x, N)]),
}


def rnd(rng, shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def run_user(gpu, mults, name, arrays, dtype=None, centered=False, M=None, m=None):
    import torch
    from fftwpp_amd.torch_conv import UserConvolution
    dtype = dtype or torch.complex128
    A, B, _ = OPS[name]
    shape = arrays[0].shape
    plan = UserConvolution(shape, A, B, addr(mults, name), dtype=dtype,
                           centered=centered, M=M, m=m)
    n = max(A, B)
    ts = [torch.tensor(np.ascontiguousarray(arrays[a] if a < A else np.zeros(shape)),
                       dtype=dtype, device=gpu) for a in range(n)]
    out = plan(*ts)
    torch.cuda.synchronize()
    return [o.cpu().numpy().astype(np.complex128) for o in out], plan


def want_user(oracle, name, arrays, N, centered=False):
    _, _, op = OPS[name]
    return oracle.explicit_general(arrays, N, lambda F, x: op(F, x, N), centered)


@pytest.mark.parametrize("L", [1, 2, 7, 100, 1000, 3000])
def test_ternary_1d_direct(gpu, oracle, mults, L):
    import torch
    rng = np.random.default_rng(L)
    f, g, h = rnd(rng, L), rnd(rng, L), rnd(rng, L)
    M = [max(3 * L - 2, 1)]
    want = oracle.direct_conv(oracle.direct_conv(f, g), h) if L <= 1000 else None
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, TOL32)):
        got, plan = run_user(gpu, mults, "testmult_ternary", [f, g, h], dtype=dtype, M=M)
        N = plan.geometry()[0]
        ref = want if want is not None else want_user(oracle, "testmult_ternary", [f, g, h], N)[0]
        assert oracle.rel_l2(got[0], ref) < tol, (L, dtype)


@pytest.mark.parametrize("name", list(OPS))
@pytest.mark.parametrize("shape,M,m", [
    ((37,), None, None), ((100,), None, [16]), ((64,), [200], [32]),
    ((16, 12), None, None), ((9, 20), [30, 41], [4, 8]),
    ((6, 5, 7), None, None), ((8, 6, 10), None, [4, 2, 8]),
])
def test_user_multipliers(gpu, oracle, mults, name, shape, M, m):
    import torch
    A, B, _ = OPS[name]
    rng = np.random.default_rng(sum(shape))
    arrays = [rnd(rng, shape) for _ in range(A)]
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, TOL32)):
        got, plan = run_user(gpu, mults, name, arrays, dtype=dtype, M=M, m=m)
        if m is not None:
            assert plan.geometry()[1] == tuple(m)
        want = want_user(oracle, name, arrays, plan.geometry()[0])
        for b in range(B):
            assert oracle.rel_l2(got[b], want[b]) < tol, (name, shape, dtype, b)
        if name == "testmult_pair":
            assert oracle.rel_l2(got[1], arrays[0] - arrays[1]) < tol


@pytest.mark.parametrize("shape,m", [((41,), None), ((40,), [4]), ((11, 14), None),
                                     ((5, 6, 9), None), ((6, 8, 200), [4, 8, 256]),
                                     ((12, 300), [16, 512])])
def test_user_multipliers_centered(gpu, oracle, mults, shape, m):
    rng = np.random.default_rng(7 + sum(shape))
    f, g = rnd(rng, shape), rnd(rng, shape)
    for name in ["testmult_weighted", "testmult_pair"]:
        got, plan = run_user(gpu, mults, name, [f, g], centered=True, m=m)
        want = want_user(oracle, name, [f, g], plan.geometry()[0], centered=True)
        for b in range(OPS[name][1]):
            assert oracle.rel_l2(got[b], want[b]) < TOL64, (name, shape, b)
    got, plan = run_user(gpu, mults, "testmult_square", [f], centered=True, m=m)
    if len(shape) == 1:
        assert oracle.rel_l2(got[0], oracle.direct_convC(f, f)) < TOL64


def test_user_multiplier_3d_large(gpu, oracle, mults):
    # 128^3 -> 255^3 minimum: the z stage runs over more than 65535 rows (as
    # the columns of one launch); the difference output is f - g at any size
    import torch
    shape = (128, 128, 128)
    rng = np.random.default_rng(128)
    f, g = rnd(rng, shape), rnd(rng, shape)
    got, plan = run_user(gpu, mults, "testmult_pair", [f, g])
    N = plan.geometry()[0]
    assert N[0] * N[1] > 65535
    assert oracle.rel_l2(got[1], f - g) < TOL64
    assert oracle.rel_l2(got[0], oracle.explicit_conv(f, g, N)) < TOL64
    got32, _ = run_user(gpu, mults, "testmult_pair", [f, g], dtype=torch.complex64)
    assert oracle.rel_l2(got32[1], f - g) < TOL32


@pytest.mark.parametrize("which,shape", [(0, (30,)), (0, (9, 11)), (0, (5, 6, 7)),
                                         (3, (33,)), (3, (12, 10)), (3, (4, 6, 5))])
def test_cxx_face_host_arrays(gpu, oracle, mults, which, shape):
    # Application(A,1,mult) + fftPad + Convolution{,2,3} on host arrays, the
    # multiplier with the reference signature reading indices->layout
    rng = np.random.default_rng(40 + sum(shape))
    A = 3 if which == 0 else 2
    arrays = [np.ascontiguousarray(rnd(rng, shape)) for _ in range(A)]
    M = [3 * L - 2 if which == 0 else 2 * L - 1 for L in shape]
    work = [a.copy() for a in arrays]
    ptrs = (ctypes.POINTER(ctypes.c_double) * A)(
        *[w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for w in work])
    Ls = (ctypes.c_size_t * 3)(*shape)
    Ms = (ctypes.c_size_t * 3)(*M)
    Ns = (ctypes.c_size_t * 3)()
    assert mults.testmult_cxx(which, len(shape), Ls, Ms, ptrs, Ns) == 0
    name = "testmult_ternary" if which == 0 else "testmult_weighted"
    N = tuple(Ns[:len(shape)])
    want = want_user(oracle, name, arrays, N)[0]
    assert oracle.rel_l2(work[0], want) < TOL64
    if which == 0 and len(shape) == 1:
        f, g, h = arrays
        assert oracle.rel_l2(work[0], oracle.direct_conv(oracle.direct_conv(f, g), h)) < TOL64


def test_user_plan_errors(gpu, mults):
    import torch
    from fftwpp_amd import FFTWPPError
    from fftwpp_amd._lib import lib
    from fftwpp_amd.torch_conv import UserConvolution
    with pytest.raises(FFTWPPError):
        UserConvolution((16,), 0, 1, addr(mults, "testmult_square"))
    with pytest.raises(FFTWPPError):
        UserConvolution((16,), 2, 9, addr(mults, "testmult_pair"))
    plan = UserConvolution((16,), 2, 1, addr(mults, "testmult_weighted"))
    a = torch.ones(16, dtype=torch.complex128, device=gpu)
    # the binary entry points do not take a user-multiplier plan
    rc = lib().fftwpp_amd_conv_execute(plan._h, a.data_ptr(), a.data_ptr(), None)
    assert rc != 0
    with pytest.raises(ValueError):
        plan(a)
