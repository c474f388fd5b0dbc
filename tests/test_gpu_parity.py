"""GPU parity of the hybrid dealiased convolutions against the CPU oracle.

Bar: relative L2 error <= 1e-12 in FP64 (tests/tests.py:223-224 uses 1e-12);
<= 1e-5 in FP32 against the FP64 oracle.  Every call goes through the C ABI of
libfftwpp_amd.so (ctypes), with torch providing device memory."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rnd(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def run(gpu, kind, f, g, M=None, m=None, dtype=None, lengths=None):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    dtype = dtype or torch.complex128
    plan = DeviceConvolution(f.shape, kind=kind, dtype=dtype, M=M, m=m, lengths=lengths)
    a = torch.tensor(np.ascontiguousarray(f), dtype=dtype, device=gpu)
    b = torch.tensor(np.ascontiguousarray(g), dtype=dtype, device=gpu)
    plan.convolve_(a, b)
    torch.cuda.synchronize()
    return a.cpu().numpy().astype(np.complex128), plan


# ---------------------------------------------------------------- 1D complex

def complex_grid():
    # tests/tests.py:561-573
    L = 8
    for M in [16, 20]:
        for m in [M, L + 1, L, 4, 2]:
            yield L, M, m


@pytest.mark.parametrize("L,M,m", list(complex_grid()))
def test_conv1d_reference_grid(gpu, oracle, L, M, m):
    f, g = oracle.ramp1(L, 0), oracle.ramp1(L, 1)
    h, plan = run(gpu, "complex", f, g, M=[M], m=[m])
    N, mm, q = plan.geometry()
    assert mm[0] == m and N[0] == m * oracle.parameters(L, M, m, False)[2]
    assert oracle.rel_l2(h, oracle.direct_conv(f, g)) < TOL64


def test_conv1d_closed_form(gpu, oracle):
    # tests/hybridconv.cc:50-69 accuracy case, L=2^6..2^11 (tests/accuracy.py)
    iF, iG = complex(math.sqrt(3), math.sqrt(7)), complex(math.sqrt(5), math.sqrt(11))
    for e in range(6, 12):
        L = 2 ** e
        j = np.arange(L)
        z = np.exp(1j * j)
        h, _ = run(gpu, "complex", iF * z, iG * z)
        assert oracle.rel_l2(h, iF * iG * (j + 1) * z) < TOL64, L


def test_conv1d_workload(gpu, oracle):
    # L=1024 -> M=2048, uniform[-1,1] from the workload seed
    f, g = oracle.inputs_1d(1024)
    h, plan = run(gpu, "complex", f, g, M=[2048])
    assert plan.geometry()[0][0] >= 2048
    assert oracle.rel_l2(h, oracle.direct_conv(f, g)) < TOL64


@pytest.mark.parametrize("L", [1, 2, 3, 100, 1000, 3000, 4099])
def test_conv1d_ragged_lengths(gpu, oracle, L):
    rng = np.random.default_rng(L)
    f, g = rnd(rng, L), rnd(rng, L)
    h, _ = run(gpu, "complex", f, g)
    assert oracle.rel_l2(h, oracle.explicit_conv(f, g, 2 * L)) < TOL64


def test_conv1d_hybrid_port_agrees(gpu, oracle):
    rng = np.random.default_rng(11)
    f, g = rnd(rng, 700), rnd(rng, 700)
    for m in [700, 350, 175, 1400]:
        _, _, q = oracle.parameters(700, 1399, m, False)
        h, _ = run(gpu, "complex", f, g, M=[1399], m=[m])
        assert oracle.rel_l2(h, oracle.hybrid_conv1(f, g, m, q)) < TOL64, m


# ---------------------------------------------------------------- 1D centred

def centered_grid():
    # tests/tests.py:575-593
    for L in [8, 7]:
        L2 = (L + 1) // 2
        for M in [3 * L2 - 2 * (L % 2), 3 * L2 - 2 * (L % 2) + 1, 2 * L, 5 * L2]:
            for m in [-(-L // 4), L2, -(-(L2 + L) // 2), M]:
                yield L, M, m


@pytest.mark.parametrize("L,M,m", list(centered_grid()))
def test_centered_reference_grid(gpu, oracle, L, M, m):
    from fftwpp_amd import FFTWPPError
    f, g = oracle.ramp1(L, 0), oracle.ramp1(L, 1)
    p, n, q = oracle.parameters(L, M, m, True)
    valid = q == 1 or p % 2 == 0  # fftPadCentered::Opt::valid, convolve.h:605-607
    if not valid:
        with pytest.raises(FFTWPPError):
            run(gpu, "centered", f, g, M=[M], m=[m])
        return
    h, _ = run(gpu, "centered", f, g, M=[M], m=[m])
    assert oracle.rel_l2(h, oracle.direct_convC(f, g)) < TOL64


# ---------------------------------------------------------------- 1D Hermitian

def hermitian_grid():
    # tests/tests.py:595-613
    for L in [8, 7]:
        L2 = (L + 1) // 2
        for M in [3 * L2 - 2 * (L % 2), 2 * L, 5 * L2]:
            for m in [M, L2, -(-L // 4)]:
                yield L, M, m


@pytest.mark.parametrize("L,M,m", list(hermitian_grid()))
def test_hconv1d_reference_grid(gpu, oracle, L, M, m):
    from fftwpp_amd import FFTWPPError
    H = (L + 1) // 2
    f, g = oracle.rampH1(H, 0), oracle.rampH1(H, 1)
    p, n, q = oracle.parameters(L, M, m, True)
    D = 1 if q == 1 else 2
    valid = (D == 1 and q == 1) or (D == 2 and p % 2 == 0)  # convolve.h:721-723
    if not valid:
        with pytest.raises(FFTWPPError):
            run(gpu, "hermitian", f, g, M=[M], m=[m], lengths=[L])
        return
    h, _ = run(gpu, "hermitian", f, g, M=[M], m=[m], lengths=[L])
    assert oracle.rel_l2(h, oracle.direct_convh(f, g)) < TOL64


@pytest.mark.parametrize("L", [15, 64, 1023, 2047, 4095])
def test_hconv1d_sizes(gpu, oracle, L):
    rng = np.random.default_rng(L)
    H = (L + 1) // 2
    f, g = rnd(rng, H), rnd(rng, H)
    f[0] = f[0].real
    g[0] = g[0].real
    h, plan = run(gpu, "hermitian", f, g, lengths=[L])
    M = 3 * H - 2 * (L % 2)
    assert oracle.rel_l2(h, oracle.explicit_hconv(f, g, M)) < TOL64


# ---------------------------------------------------------------- 2D / 3D

def test_conv2d_small_and_medium(gpu, oracle):
    f, g = oracle.ramp2(8, 8, 0).astype(complex), oracle.ramp2(8, 8, 1).astype(complex)
    h, _ = run(gpu, "complex", f, g)
    assert oracle.rel_l2(h, oracle.direct_conv2(f, g)) < TOL64
    rng = np.random.default_rng(21)
    f, g = rnd(rng, 37, 29), rnd(rng, 37, 29)
    h, _ = run(gpu, "complex", f, g)
    assert oracle.rel_l2(h, oracle.explicit_conv(f, g, (73, 57))) < TOL64
    f, g = rnd(rng, 200, 120), rnd(rng, 200, 120)
    for mx, my in [(None, None), (100, 60), (50, 240)]:
        m = None if mx is None else [mx, my]
        h, _ = run(gpu, "complex", f, g, m=m)
        assert oracle.rel_l2(h, oracle.explicit_conv(f, g, (399, 239))) < TOL64


def test_conv3d_small_and_medium(gpu, oracle):
    f = oracle.ramp3(4, 5, 6, 0).astype(complex)
    g = oracle.ramp3(4, 5, 6, 1).astype(complex)
    h, _ = run(gpu, "complex", f, g)
    assert oracle.rel_l2(h, oracle.direct_conv3(f, g)) < TOL64
    rng = np.random.default_rng(31)
    f, g = rnd(rng, 16, 12, 20), rnd(rng, 16, 12, 20)
    h, _ = run(gpu, "complex", f, g)
    assert oracle.rel_l2(h, oracle.explicit_conv(f, g, (31, 23, 39))) < TOL64


@pytest.mark.parametrize("Lx,Ly", [(7, 7), (8, 8), (7, 8), (8, 7), (9, 5)])
def test_hconv2d_reference_cases(gpu, oracle, Lx, Ly):
    Hy = (Ly + 1) // 2
    f, g = oracle.ramph2(Lx, Hy, 0), oracle.ramph2(Lx, Hy, 1)
    h, _ = run(gpu, "hermitian", f, g, lengths=[Lx, Ly])
    assert oracle.rel_l2(h, oracle.direct_convh2(f, g)) < TOL64


def test_hconv2d_spectral_medium(gpu, oracle):
    # the pseudospectral workload shape at 1/16 size: kx in [-127,127], ky in [0,127]
    f, g = oracle.inputs_h2(255, 128)
    h, plan = run(gpu, "hermitian", f, g, lengths=[255, 255])
    M = 3 * 128 - 2
    assert oracle.rel_l2(h, oracle.explicit_hconv(f, g, (M, M))) < TOL64


@pytest.mark.parametrize("shape", [(7, 7, 7), (8, 8, 8), (5, 6, 7)])
def test_hconv3d_reference_cases(gpu, oracle, shape):
    Lx, Ly, Lz = shape
    Hz = (Lz + 1) // 2
    f, g = oracle.ramph3(Lx, Ly, Hz, 0), oracle.ramph3(Lx, Ly, Hz, 1)
    h, _ = run(gpu, "hermitian", f, g, lengths=list(shape))
    assert oracle.rel_l2(h, oracle.direct_convh3(f, g)) < TOL64


# ---------------------------------------------------------------- FP32 path

def test_fp32_3d_error_vs_fp64(gpu, oracle):
    import torch
    rng = np.random.default_rng(41)
    f, g = rnd(rng, 32, 32, 32), rnd(rng, 32, 32, 32)
    ref = oracle.explicit_conv(f, g, 64)
    h32, _ = run(gpu, "complex", f, g, dtype=torch.complex64)
    err = oracle.rel_l2(h32, ref)
    print(f"fp32 3D 32^3 rel L2 vs FP64 oracle: {err:.3e}", flush=True)
    assert err < TOL32


def test_fp32_hermitian_2d(gpu, oracle):
    import torch
    f, g = oracle.inputs_h2(127, 64)
    ref = oracle.explicit_hconv(f, g, (190, 190))
    h32, _ = run(gpu, "hermitian", f, g, dtype=torch.complex64, lengths=[127, 127])
    assert oracle.rel_l2(h32, ref) < TOL32


# ---------------------------------------------------------------- host pointers

def test_reference_python_wrapper_host_arrays(gpu, oracle):
    """The reference's Python interface (wrappers/fftwpp.py) on numpy arrays."""
    from fftwpp_amd import fftwpp
    f = fftwpp.complex_align([8])
    g = fftwpp.complex_align([8])
    f[:] = oracle.ramp1(8, 0)
    g[:] = oracle.ramp1(8, 1)
    want = oracle.direct_conv(f.copy(), g.copy())
    gcopy = g.copy()
    fftwpp.Convolution(f.shape).convolve(f, g)
    assert oracle.rel_l2(f, want) < TOL64
    assert np.array_equal(g, gcopy)
    f = fftwpp.complex_align([4, 5, 6])
    g = fftwpp.complex_align([4, 5, 6])
    f[:] = oracle.ramp3(4, 5, 6, 0)
    g[:] = oracle.ramp3(4, 5, 6, 1)
    want = oracle.direct_conv3(f.copy(), g.copy())
    fftwpp.Convolution(f.shape).convolve(f, g)
    assert oracle.rel_l2(f, want) < TOL64
    # HConvolution symmetrises first (wrappers/fftwpp.py:329-351)
    f = np.ascontiguousarray(oracle.ramph2(7, 4, 0))
    g = np.ascontiguousarray(oracle.ramph2(7, 4, 1))
    want = oracle.direct_convh2(f.copy(), g.copy())
    fftwpp.HConvolution(f.shape).convolve(f, g)
    assert oracle.rel_l2(f, want) < TOL64


def test_symmetrize_device_matches_oracle(gpu, oracle):
    import torch
    from fftwpp_amd import _lib
    rng = np.random.default_rng(51)
    for Lx, Ly, Hz in [(7, 7, 4), (8, 6, 3)]:
        f = np.ascontiguousarray(rnd(rng, Lx, Ly, Hz))
        want = oracle.symmetrize_xy(f.copy())
        t = torch.tensor(f, device=gpu)
        _lib.lib().fftwpp_HermitianSymmetrizeXY((Lx + 1) // 2, (Ly + 1) // 2, Hz, Lx // 2,
                                                Ly // 2, t.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy(), want)


# ---------------------------------------------------------------- kernel layer

def test_forward_residues_definition(gpu, oracle):
    """fftwpp_amd_forward_many residue r equals the padded DFT at q*s+r
    (the transform identity of tests/hybrid.cc:60-110)."""
    import ctypes
    import torch
    from fftwpp_amd import _lib
    L_ = _lib.lib()
    rng = np.random.default_rng(61)
    for L, m, q, kind in [(8, 8, 2, 0), (12, 4, 6, 0), (7, 4, 3, 1), (30, 10, 6, 1)]:
        C = 5
        f = rnd(rng, L, C)
        d = ctypes.c_void_p()
        _lib.check(L_.fftwpp_amd_dim_create(ctypes.byref(d), L, m, q, kind))
        fd = torch.tensor(f, device=gpu)
        Fd = torch.zeros((q * m, C), dtype=torch.complex128, device=gpu)
        _lib.check(L_.fftwpp_amd_forward_many(d, 0, fd.data_ptr(), C, C, 1, 0,
                                              Fd.data_ptr(), C, 0, 0, q, None))
        torch.cuda.synchronize()
        F = Fd.cpu().numpy()
        k = np.arange(L) - (L // 2 if kind == 1 else 0)
        N = m * q
        for r in range(q):
            ell = q * np.arange(m) + r
            want = np.exp(2j * np.pi * np.outer(ell, k) / N) @ f
            assert oracle.rel_l2(F[r * m:(r + 1) * m], want) < 1e-13
        # backward o forward = N * identity (tests/hybrid.cc:112-167)
        gd = torch.zeros_like(fd)
        _lib.check(L_.fftwpp_amd_backward_many(d, 0, Fd.data_ptr(), C, 0, gd.data_ptr(), C, C,
                                               1, 0, 0, q, 0, 1.0 / N, None))
        torch.cuda.synchronize()
        assert oracle.rel_l2(gd.cpu().numpy(), f) < 1e-13
        L_.fftwpp_amd_dim_destroy(d)


# ---------------------------------------------------------------- adjoint

def test_adjoint_matches_direct_correlation(gpu, oracle):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(71)
    for shape in [(33,), (9, 11), (5, 6, 7)]:
        a, b, w = rnd(rng, *shape), rnd(rng, *shape), rnd(rng, *shape)
        plan = DeviceConvolution(shape)
        T = lambda x: torch.tensor(x, device=gpu)
        ga, gb = plan.adjoint(T(w), T(a), T(b))
        torch.cuda.synchronize()
        # grad_a[i] = sum_k w[k] conj(b[k-i]) over the stored ranges
        want_a = _correlate(w, b)
        want_b = _correlate(w, a)
        assert oracle.rel_l2(ga.cpu().numpy(), want_a) < TOL64
        assert oracle.rel_l2(gb.cpu().numpy(), want_b) < TOL64


def _correlate(w, b):
    import scipy.signal as ss
    # c[i] = sum_j w[i+j] conj(b[j]) for i, j, i+j inside the box
    full = ss.correlate(w, b, mode="full")  # full[n] = sum_j w[n-(len-1)+j] conj(b[j])
    idx = tuple(slice(s - 1, 2 * s - 1) for s in w.shape)
    return full[idx]


def test_autograd_finite_differences(gpu, oracle):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution, convolve
    rng = np.random.default_rng(81)
    shape = (6, 7)
    plan = DeviceConvolution(shape)
    a0, b0, t0 = rnd(rng, *shape), rnd(rng, *shape), rnd(rng, *shape)
    T = lambda x: torch.tensor(x, device=gpu)
    a = T(a0).requires_grad_(True)
    b = T(b0).requires_grad_(True)
    loss = (convolve(a, b, plan) - T(t0)).abs().pow(2).sum()
    loss.backward()

    def L(av, bv):
        h = oracle.explicit_conv(av, bv, (11, 13))
        return float(np.sum(np.abs(h - t0) ** 2))

    eps = 1e-6
    ga = a.grad.cpu().numpy()
    for idx in [(0, 0), (3, 4), (5, 6)]:
        for unit in [1.0, 1j]:
            ap, am = a0.copy(), a0.copy()
            ap[idx] += eps * unit
            am[idx] -= eps * unit
            fd = (L(ap, b0) - L(am, b0)) / (2 * eps)
            got = ga[idx].real if unit == 1.0 else ga[idx].imag
            assert abs(fd - got) < 1e-5 * max(1.0, abs(fd)), (idx, unit, fd, got)


# ---------------------------------------------------------------- full size

def _workers():
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def test_conv3d_256_full_size(gpu, oracle):
    """The 3D workload at full size, 256^3 -> 512^3 from seed 20261202: FP64
    over the whole array against explicit zero padding to 512^3 (pocketfft),
    and against direct sums at sampled outputs (an FFT-free check); FP32 over
    the whole array against the FP64 oracle; plus f (*) g == g (*) f."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    f, g = oracle.inputs_3d(256)
    fd = torch.tensor(f, device=gpu)
    gd = torch.tensor(g, device=gpu)
    plan = DeviceConvolution((256, 256, 256), M=[512] * 3)
    assert plan.geometry() == ((512,) * 3, (256,) * 3, (2,) * 3)
    h = plan(fd, gd)
    h2 = plan(gd, fd)
    torch.cuda.synchronize()
    H = h.cpu().numpy()
    assert oracle.rel_l2(h2.cpu().numpy(), H) < TOL64
    del h2
    rng = np.random.default_rng(91)
    pts = [(255, 255, 255), (0, 0, 0), (128, 3, 250)] + \
        [tuple(rng.integers(0, 256, 3)) for _ in range(5)]
    num = den = 0.0
    for (i, j, k) in pts:
        direct = np.sum(f[:i + 1, :j + 1, :k + 1] * g[i::-1, j::-1, k::-1])
        num += abs(H[i, j, k] - direct) ** 2
        den += abs(direct) ** 2
    assert math.sqrt(num / den) < TOL64
    ref = oracle.explicit_conv(f, g, 512, workers=_workers())
    err64 = oracle.rel_l2(H, ref)
    print(f"fp64 256^3 rel L2 vs explicit padding (whole array): {err64:.3e}", flush=True)
    assert err64 < TOL64
    del H
    p32 = DeviceConvolution((256, 256, 256), dtype=torch.complex64, M=[512] * 3)
    h32 = p32(fd.to(torch.complex64), gd.to(torch.complex64))
    torch.cuda.synchronize()
    err = oracle.rel_l2(h32.cpu().numpy(), ref)
    print(f"fp32 256^3 rel L2 vs FP64 oracle (whole array): {err:.3e}", flush=True)
    assert err < TOL32


def test_hconv2d_config_full_size(gpu, oracle):
    """The 2D centred Hermitian workload at full size: 2047 x 1024 stored
    modes (kx in [-1023, 1023], ky in [0, 1023]), amplitude |k|^-1 with
    uniform phase from seed 20261202, HermitianSymmetrizeX applied, M = 3070
    per dimension (x: m = 1024, p = 2, q = 3; y: Hermitian, D = 2).  FP64
    against explicit zero padding to 3070^2 (pocketfft) over the whole array,
    FP32 against the same FP64 oracle; the drop-in handle of cfftw++.h
    (fftwpp_create_hconv2d) on device arrays gives the same FP64 result."""
    import torch
    from fftwpp_amd import _lib
    f, g = oracle.inputs_h2(2047, 1024)
    ref = oracle.explicit_hconv(f, g, (3070, 3070), workers=_workers())
    h, plan = run(gpu, "hermitian", f, g, M=[3070, 3070], lengths=[2047, 2047])
    N, m, q = plan.geometry()
    assert m == (1024, 1024) and q == (3, 3), (N, m, q)
    err64 = oracle.rel_l2(h, ref)
    print(f"2D-H 2047x1024 fp64 rel L2 vs explicit padding: {err64:.3e}", flush=True)
    assert err64 < TOL64
    h32, _ = run(gpu, "hermitian", f, g, M=[3070, 3070], lengths=[2047, 2047],
                 dtype=torch.complex64)
    err32 = oracle.rel_l2(h32, ref)
    print(f"2D-H 2047x1024 fp32 rel L2 vs FP64 oracle: {err32:.3e}", flush=True)
    assert err32 < TOL32
    L = _lib.lib()
    hd = L.fftwpp_create_hconv2d(2047, 2047)
    a = torch.tensor(f, device=gpu)
    b = torch.tensor(g, device=gpu)
    assert L.fftwpp_hconv2d_convolve_device(hd, a.data_ptr(), b.data_ptr(), None) == 0
    torch.cuda.synchronize()
    L.fftwpp_hconv2d_delete(hd)
    assert oracle.rel_l2(a.cpu().numpy(), ref) < TOL64


# ---------------------------------------------------------------- m=256 kernels

@pytest.mark.parametrize("L,M", [(256, 511), (200, 399), (256, 700), (131, 600)])
def test_register_fft256_paths(gpu, oracle, L, M):
    """Sub-FFT size 256 (register-resident kernels) with L < 256 and q = 2..3,
    as rows (1D), and as strided columns (2D/3D outer dimensions)."""
    import torch
    rng = np.random.default_rng(L + M)
    _, _, q = oracle.parameters(L, M, 256, False)
    N = 256 * q
    f, g = rnd(rng, L), rnd(rng, L)
    for dtype, tol in [(torch.complex128, TOL64), (torch.complex64, TOL32)]:
        h, plan = run(gpu, "complex", f, g, M=[M], m=[256], dtype=dtype)
        assert plan.geometry()[1][0] == 256
        assert oracle.rel_l2(h, oracle.explicit_conv(f, g, N)) < tol
    f2, g2 = rnd(rng, L, 40), rnd(rng, L, 40)
    h2, _ = run(gpu, "complex", f2, g2, M=[M, 79], m=[256, 0])
    assert oracle.rel_l2(h2, oracle.explicit_conv(f2, g2, (N, 79))) < TOL64
    f3, g3 = rnd(rng, 6, L, 9), rnd(rng, 6, L, 9)
    h3, _ = run(gpu, "complex", f3, g3, M=[11, M, 17], m=[0, 256, 0])
    _, _, q3 = oracle.parameters(6, 11, oracle.parameters(6, 11, 6, False)[0] and 6, False)
    want = oracle.explicit_conv(f3, g3, (11, N, 17))
    assert oracle.rel_l2(h3, want) < TOL64


@pytest.mark.parametrize("L,m", [(4097, None), (6000, None), (12289, None), (16384, None),
                                 (5000, 3000), (9000, 2048)])
def test_conv1d_long_rows(gpu, oracle, L, m):
    """Rows longer than one power-of-two block of m <= 4096 run on the batched
    register kernel as a batch of one item (one block per residue); a forced
    non-power-of-two m whose outputs do not fit beside the sub-FFTs in LDS
    takes the unfused residue path (forward, multiply, backward).  Two calls
    on one plan, the second on other data."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(L)
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, 1e-5)):
        plan = DeviceConvolution((L,), dtype=dtype, M=[2 * L], m=None if m is None else [m])
        if m is None:
            assert plan.geometry()[1][0] == 4096
        for _ in range(2):
            f, g = rnd(rng, L), rnd(rng, L)
            want = oracle.explicit_conv(f, g, 1 << (2 * L - 1).bit_length())
            out = torch.empty(L, dtype=dtype, device=gpu)
            plan.convolve_out(out, torch.tensor(f, dtype=dtype, device=gpu),
                              torch.tensor(g, dtype=dtype, device=gpu))
            torch.cuda.synchronize()
            assert oracle.rel_l2(out.cpu().numpy(), want) < tol


@pytest.mark.parametrize("kind,L", [("centered", 5001), ("centered", 9000), ("hermitian", 6001),
                                    ("hermitian", 8191)])
def test_long_centred_and_hermitian_rows(gpu, oracle, kind, L):
    """Default plans of long centred rows (unfused residue path) and Hermitian
    rows up to H = 4096 (register row kernel at m = 4096)."""
    import torch
    rng = np.random.default_rng(L + 1)
    if kind == "centered":
        f, g = rnd(rng, L), rnd(rng, L)
        want = oracle.explicit_conv(f, g, 2 * L, centered=True)
        h, _ = run(gpu, "centered", f, g)
    else:
        H = (L + 1) // 2
        f, g = rnd(rng, H), rnd(rng, H)
        f[0], g[0] = f[0].real, g[0].real
        want = oracle.explicit_hconv(f, g, oracle.default_M(L, True))
        h, plan = run(gpu, "hermitian", f, g, lengths=[L])
    assert oracle.rel_l2(h, want) < TOL64


@pytest.mark.parametrize("lengths,m", [([8193], None), ([10001], None), ([20001], None),
                                       ([3001], [256]), ([15, 10001], None),
                                       ([5, 7, 9001], None)])
def test_long_hermitian_rows(gpu, oracle, lengths, m):
    """Hermitian rows longer than one block of a power-of-two m: the residue
    pairs of each row run as parallel work units whose partial sums a second
    launch adds (1D, and the inner dimension of 2D/3D), FP64 and FP32."""
    import torch
    rng = np.random.default_rng(sum(lengths))
    if len(lengths) == 1:
        H = (lengths[0] + 1) // 2
        f, g = rnd(rng, H), rnd(rng, H)
        f[0], g[0] = f[0].real, g[0].real
    elif len(lengths) == 2:
        f, g = oracle.inputs_h2(lengths[0], (lengths[1] + 1) // 2, seed=lengths[1])
    else:
        Hz = (lengths[2] + 1) // 2
        f = oracle.ramph3(lengths[0], lengths[1], Hz, 0).astype(complex) / Hz
        g = oracle.ramph3(lengths[0], lengths[1], Hz, 1).astype(complex) / Hz
    want = oracle.explicit_hconv(f, g, tuple(oracle.default_M(L, True) for L in lengths))
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, TOL32)):
        h, plan = run(gpu, "hermitian", f, g, dtype=dtype, m=m, lengths=lengths)
        mm = plan.geometry()[1][-1]
        assert mm & (mm - 1) == 0 and (lengths[-1] + 1) // 2 > mm
        assert oracle.rel_l2(h, want) < tol, dtype


@pytest.mark.parametrize("shape", [(8, 5000), (3, 4, 6001), (600, 4500)])
def test_long_complex_inner_rows(gpu, oracle, shape):
    """Complex 2D/3D convolutions whose inner rows are longer than 4096 run the
    inner stage on the batched register kernel with the rows as items (one
    block per residue for few rows, one block per row for many)."""
    import torch
    rng = np.random.default_rng(sum(shape))
    f, g = rnd(rng, *shape), rnd(rng, *shape)
    want = oracle.explicit_conv(f, g, tuple(2 * L for L in shape))
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, TOL32)):
        h, plan = run(gpu, "complex", f, g, dtype=dtype)
        assert plan.geometry()[1][-1] == 4096
        assert oracle.rel_l2(h, want) < tol, dtype


def test_adjoint_long_row(gpu, oracle):
    """The adjoint (correlations) of a row longer than 4096 on the batched path."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(5001)
    L = 5001
    a, b, w = rnd(rng, L), rnd(rng, L), rnd(rng, L)
    plan = DeviceConvolution((L,))
    T = lambda x: torch.tensor(x, device=gpu)
    ga, gb = plan.adjoint(T(w), T(a), T(b))
    torch.cuda.synchronize()
    assert oracle.rel_l2(ga.cpu().numpy(), _correlate(w, b)) < TOL64
    assert oracle.rel_l2(gb.cpu().numpy(), _correlate(w, a)) < TOL64


def test_cfftw_device_entry_points(gpu, oracle):
    """fftwpp_*_convolve_device on device arrays equals *_convolve on host
    arrays for every handle type, and rejects host pointers."""
    import torch
    from fftwpp_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(123)
    cases = [("conv1d", (37,), (37,)), ("hconv1d", (19,), (37,)), ("conv2d", (9, 11), (9, 11)),
             ("hconv2d", (9, 6), (9, 11)), ("conv3d", (5, 6, 7), (5, 6, 7)),
             ("hconv3d", (5, 7, 4), (5, 7, 7))]
    for name, shape, lengths in cases:
        a, b = rnd(rng, *shape), rnd(rng, *shape)
        if name.startswith("h"):
            plane = oracle.hplane_of(lengths)
            a, b = oracle.hplane(a, 0, plane), oracle.hplane(b, 0, plane)
        h = getattr(L, f"fftwpp_create_{name}")(*lengths)
        host = np.ascontiguousarray(a.copy())
        hb = np.ascontiguousarray(b.copy())
        getattr(L, f"fftwpp_{name}_convolve")(h, host.ctypes.data, hb.ctypes.data)
        da = torch.tensor(a, device=gpu)
        db = torch.tensor(b, device=gpu)
        rc = getattr(L, f"fftwpp_{name}_convolve_device")(h, da.data_ptr(), db.data_ptr(), None)
        torch.cuda.synchronize()
        assert rc == 0, (name, L.fftwpp_amd_last_error())
        assert oracle.rel_l2(da.cpu().numpy(), host) < TOL64, name
        assert getattr(L, f"fftwpp_{name}_convolve_device")(h, host.ctypes.data, hb.ctypes.data,
                                                           None) != 0
        getattr(L, f"fftwpp_{name}_delete")(h)
