"""Parity of the register-FFT row and column engines (power-of-two sub-FFT
sizes 16..4096) against the CPU oracle.

The reference's own grids (tests/tests.py:561-613) use sub-FFT sizes of at
most 20, so they reach these engines only at m = 16; here the sub-FFT size is
forced (the reference's -m option, tests/hybridconv.cc) across the whole
range, for complex, centred and Hermitian rows and for complex and centred
columns with p = 1 and p = 2.  Every case first asserts that the dispatcher
routes its dimensions to the register engines, so a change of routing cannot
silently turn these into tests of the generic kernels.

Bar: relative L2 error <= 1e-12 in FP64, <= 1e-5 in FP32 against the FP64
oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
REGISTER = 2  # fftwpp_amd_internal_engine: register-FFT kernels

KIND = {"complex": 0, "centered": 1, "hermitian": 2}


def rnd(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def engine(L, m, q, kind, stage):
    from fftwpp_amd._lib import check, lib
    d = ctypes.c_void_p()
    check(lib().fftwpp_amd_dim_create(ctypes.byref(d), L, m, q, kind))
    try:
        return lib().fftwpp_amd_internal_engine(d, stage)
    finally:
        lib().fftwpp_amd_dim_destroy(d)


def assert_register_routing(plan, lengths, kind):
    """Leading dimensions are column transforms (stage 0), the last one the
    fused rows (stage 1)."""
    N, m, q = plan.geometry()
    d = len(lengths)
    for i in range(d):
        if kind == "hermitian":
            k = KIND["hermitian"] if i == d - 1 else KIND["centered"]
        else:
            k = KIND[kind]
        stage = 1 if i == d - 1 else 0
        assert engine(lengths[i], m[i], q[i], k, stage) == REGISTER, (i, lengths[i], m[i], q[i])


def run(gpu, kind, f, g, dtype, M=None, m=None, lengths=None):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    plan = DeviceConvolution(f.shape, kind=kind, dtype=dtype, M=M, m=m, lengths=lengths)
    a = torch.tensor(np.ascontiguousarray(f), dtype=dtype, device=gpu)
    b = torch.tensor(np.ascontiguousarray(g), dtype=dtype, device=gpu)
    plan.convolve_(a, b)
    torch.cuda.synchronize()
    return a.cpu().numpy().astype(np.complex128), plan


def dtypes():
    import torch
    return [(torch.complex128, TOL64), (torch.complex64, TOL32)]


# ---------------------------------------------------------------- 1D rows

# (L, M, m): p = 1; q = 1, 2, 5 and 8
COMPLEX_ROWS = [(16, 32, 16), (29, 58, 32), (17, 34, 64), (64, 320, 64),
                (120, 1000, 128), (300, 600, 512), (1000, 2000, 1024),
                (2048, 4096, 2048), (4000, 8000, 4096)]


@pytest.mark.parametrize("L,M,m", COMPLEX_ROWS)
def test_complex_rows(gpu, oracle, L, M, m):
    rng = np.random.default_rng(L * 7 + m)
    f, g = rnd(rng, L), rnd(rng, L)
    want = oracle.explicit_conv(f, g, 2 * L)
    for dtype, tol in dtypes():
        h, plan = run(gpu, "complex", f, g, dtype, M=[M], m=[m])
        assert plan.geometry()[1][0] == m
        assert_register_routing(plan, [L], "complex")
        assert oracle.rel_l2(h, want) < tol, dtype


# centred rows need q = 1 when p is odd (convolve.h:605-607): M <= m
@pytest.mark.parametrize("L,m", [(41, 64), (100, 256), (333, 512), (2001, 4096)])
def test_centered_rows(gpu, oracle, L, m):
    rng = np.random.default_rng(L)
    f, g = rnd(rng, L), rnd(rng, L)
    want = oracle.explicit_conv(f, g, 2 * L, centered=True)
    for dtype, tol in dtypes():
        h, plan = run(gpu, "centered", f, g, dtype, m=[m])
        assert plan.geometry()[2][0] == 1
        assert_register_routing(plan, [L], "centered")
        assert oracle.rel_l2(h, want) < tol, dtype


# H = ceil(L/2) <= m; L > m gives p = 2 (D = 2, q = 3), L <= M <= m gives q = 1
HERMITIAN_ROWS = [(15, 32), (31, 16), (127, 64), (200, 128), (1023, 512),
                  (2047, 1024), (4095, 2048), (8191, 4096)]


@pytest.mark.parametrize("L,m", HERMITIAN_ROWS)
def test_hermitian_rows(gpu, oracle, L, m):
    rng = np.random.default_rng(L + 3)
    H = (L + 1) // 2
    f, g = rnd(rng, H), rnd(rng, H)
    f[0], g[0] = f[0].real, g[0].real
    want = oracle.explicit_hconv(f, g, oracle.default_M(L, True))
    for dtype, tol in dtypes():
        h, plan = run(gpu, "hermitian", f, g, dtype, m=[m], lengths=[L])
        assert plan.geometry()[1][0] == m
        assert_register_routing(plan, [L], "hermitian")
        assert oracle.rel_l2(h, want) < tol, dtype


# ---------------------------------------------------------------- 2D / 3D columns

# (shape, m): columns with p = 2 (Lx > mx) and p = 1, rows up to m = 4096
COMPLEX_ND = [((100, 40), (64, 64)), ((30, 50), (32, 64)), ((1000, 24), (512, 32)),
              ((3000, 8), (2048, 16)), ((4096, 6), (4096, 16)), ((20, 4000), (16, 4096)),
              ((40, 60, 20), (32, 64, 32)), ((17, 9, 300), (16, 16, 512))]


@pytest.mark.parametrize("shape,m", COMPLEX_ND)
def test_complex_columns(gpu, oracle, shape, m):
    rng = np.random.default_rng(sum(shape))
    f, g = rnd(rng, *shape), rnd(rng, *shape)
    want = oracle.explicit_conv(f, g, tuple(2 * L for L in shape))
    for dtype, tol in dtypes():
        h, plan = run(gpu, "complex", f, g, dtype, m=list(m))
        assert list(plan.geometry()[1]) == list(m)
        assert_register_routing(plan, list(shape), "complex")
        assert oracle.rel_l2(h, want) < tol, dtype


# centred x columns with p = 2 (D = 2), Hermitian y rows with H = my
HERMITIAN_2D = [((63, 63), (32, 32)), ((255, 255), (128, 128)), ((1023, 63), (512, 32)),
                ((31, 4095), (16, 2048))]


@pytest.mark.parametrize("lengths,m", HERMITIAN_2D)
def test_hermitian_2d_columns(gpu, oracle, lengths, m):
    Lx, Ly = lengths
    Hy = (Ly + 1) // 2
    f, g = oracle.inputs_h2(Lx, Hy, seed=Lx + Ly)
    Ms = (oracle.default_M(Lx, True), oracle.default_M(Ly, True))
    want = oracle.explicit_hconv(f, g, Ms)
    for dtype, tol in dtypes():
        h, plan = run(gpu, "hermitian", f, g, dtype, m=list(m), lengths=list(lengths))
        assert list(plan.geometry()[1]) == list(m)
        assert_register_routing(plan, list(lengths), "hermitian")
        assert oracle.rel_l2(h, want) < tol, dtype


def test_hermitian_3d_columns(gpu, oracle):
    # x and y centred columns with p = 2, z Hermitian rows with Hz = mz
    Lx, Ly, Lz = 31, 33, 31
    Hz = (Lz + 1) // 2
    f, g = oracle.ramph3(Lx, Ly, Hz, 0), oracle.ramph3(Lx, Ly, Hz, 1)
    want = oracle.direct_convh3(f, g)
    for dtype, tol in dtypes():
        h, plan = run(gpu, "hermitian", f, g, dtype, m=[16, 32, 16], lengths=[Lx, Ly, Lz])
        assert list(plan.geometry()[1]) == [16, 32, 16]
        assert_register_routing(plan, [Lx, Ly, Lz], "hermitian")
        assert oracle.rel_l2(h, want) < tol, dtype
