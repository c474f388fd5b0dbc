"""Standalone FFTs (the fft1d / mfft1d / fft2d / fft3d / rcfft1d / crfft1d /
mrcfft1d / mcrfft1d transforms of fftw++.h:556-1444) on the GPU against
numpy's pocketfft (a library routine).  FFTW's sign convention: sign -1 is
numpy.fft.fft, sign +1 is n * numpy.fft.ifft; r2c is numpy.fft.rfft, c2r is
n * numpy.fft.irfft.  The reference's own FFT tests only print
(tests/fft1.cc, mfft1.cc), so the library routine is the check.  Bar:
relative L2 error <= 1e-12 in FP64, <= 1e-5 in FP32 against FP64 numpy."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
TESTLIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "support", "libtestmult.so")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    nb = np.linalg.norm(b.ravel())
    return d / nb if nb > 0 else d


def rnd(rng, shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def dft(x, sign, axes=None):
    if sign < 0:
        return np.fft.fftn(x, axes=axes)
    n = np.prod([x.shape[a] for a in (axes or range(x.ndim))])
    return np.fft.ifftn(x, axes=axes) * n


def run(gpu, plan, x, out_shape, out_dtype, inplace=False):
    import torch
    t = torch.tensor(x, device=gpu)
    o = t if inplace else torch.empty(out_shape, dtype=out_dtype, device=gpu)
    plan(t, o)
    torch.cuda.synchronize()
    return o.cpu().numpy()


PRECISIONS = [("f64", TOL64), ("f32", TOL32)]


def types(p):
    import torch
    return (torch.float64, torch.complex128, np.complex128, np.float64) if p == "f64" else \
        (torch.float32, torch.complex64, np.complex64, np.float32)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 64, 100, 256, 1000, 1024, 4093, 4096, 4097,
                               5000, 8192, 12000, 65536, 1 << 20])
def test_c2c_1d(gpu, n):
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(n)
    x = rnd(rng, n)
    for p, tol in PRECISIONS:
        real, cplx, npc, npr = types(p)
        for sign in (-1, 1):
            plan = DeviceFFT(n, sign, dtype=real)
            got = run(gpu, plan, x.astype(npc), (n,), cplx)
            assert rel(got, dft(x, sign)) < tol, (n, p, sign)


@pytest.mark.parametrize("n", [4099, 5003, 8198, 10007, 65537, 1000003])
def test_c2c_bluestein(gpu, n):
    # lengths with no divisor in [n/4096, 4096] (primes above 4096, 2 x 4099):
    # Bluestein's algorithm over a power-of-two convolution
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(n)
    x = rnd(rng, n)
    for p, tol in PRECISIONS:
        real, cplx, npc, npr = types(p)
        for sign in (-1, 1):
            plan = DeviceFFT(n, sign, dtype=real)
            got = run(gpu, plan, x.astype(npc), (n,), cplx)
            assert rel(got, dft(x, sign)) < tol, (n, p, sign)


def test_bluestein_layouts_and_real(gpu):
    # strided transforms, in place, and r2c / c2r through Bluestein
    import torch
    from fftwpp_amd import FFTWPPError
    from fftwpp_amd.torch_fft import DeviceFFT
    n, M, stride, dist = 5003, 3, 2, 10011
    rng = np.random.default_rng(1)
    extent = (M - 1) * dist + (n - 1) * stride + 1
    buf = rnd(rng, extent)
    idx = np.arange(M)[:, None] * dist + np.arange(n)[None, :] * stride
    plan = DeviceFFT(n, -1, dtype=torch.float64, howmany=M, istride=stride, ostride=stride,
                     idist=dist, odist=dist)
    t = torch.tensor(buf, device=gpu)
    plan(t, t)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy()[idx], np.fft.fft(buf[idx], axis=1)) < TOL64
    xr = rng.uniform(-1, 1, n)
    got = run(gpu, DeviceFFT(n, -1, kind="r2c", dtype=torch.float64), xr, (n // 2 + 1,),
              torch.complex128)
    assert rel(got, np.fft.rfft(xr)) < TOL64
    back = run(gpu, DeviceFFT(n, 1, kind="c2r", dtype=torch.float64), np.fft.rfft(xr), (n,),
               torch.float64)
    assert rel(back, xr * n) < TOL64
    with pytest.raises(FFTWPPError):
        DeviceFFT((5003, 4), -1)   # an axis of a multidimensional transform


@pytest.mark.parametrize("n,M,stride,dist", [(48, 5, 1, 48), (48, 5, 5, 1), (30, 3, 2, 70),
                                             (1024, 64, 1, 1024), (256, 16, 16, 1)])
def test_mfft1d_layouts(gpu, n, M, stride, dist):
    import torch
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(n + M)
    extent = (M - 1) * dist + (n - 1) * stride + 1
    buf = rnd(rng, extent)
    idx = np.arange(M)[:, None] * dist + np.arange(n)[None, :] * stride
    for sign in (-1, 1):
        plan = DeviceFFT(n, sign, dtype=torch.float64, howmany=M, istride=stride,
                         ostride=stride, idist=dist, odist=dist)
        t = torch.tensor(buf, device=gpu)
        o = torch.zeros(extent, dtype=torch.complex128, device=gpu)
        plan(t, o)
        torch.cuda.synchronize()
        got = o.cpu().numpy()[idx]
        assert rel(got, dft(buf[idx], sign, axes=(1,))) < TOL64, sign


@pytest.mark.parametrize("shape", [(64, 48), (256, 256), (7, 9, 11), (64, 64, 64), (8, 4096),
                                   (300, 256, 5)])
def test_c2c_multi(gpu, shape):
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(sum(shape))
    x = rnd(rng, shape)
    for p, tol in PRECISIONS:
        real, cplx, npc, npr = types(p)
        for sign in (-1, 1):
            plan = DeviceFFT(shape, sign, dtype=real)
            got = run(gpu, plan, x.astype(npc), shape, cplx)
            assert rel(got, dft(x, sign)) < tol, (shape, p, sign)


@pytest.mark.parametrize("shape", [(100,), (4096,), (5000,), (32, 20), (6, 7, 8)])
def test_c2c_in_place_and_normalized(gpu, shape):
    import torch
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(3 + sum(shape))
    x = rnd(rng, shape)
    fwd = DeviceFFT(shape, -1, dtype=torch.float64)
    got = run(gpu, fwd, x, shape, torch.complex128, inplace=True)
    assert rel(got, dft(x, -1)) < TOL64
    t = torch.tensor(x, device=gpu)
    bwd = DeviceFFT(shape, 1, dtype=torch.float64)
    fwd(t, t)
    bwd(t, t, scale=1.0 / bwd.normalization())
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy(), x) < TOL64


@pytest.mark.parametrize("n", [1, 2, 3, 8, 9, 100, 1024, 4096, 10000, 16384])
def test_r2c_c2r(gpu, n):
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(n + 11)
    x = rng.uniform(-1, 1, n)
    X = np.fft.rfft(x)
    for p, tol in PRECISIONS:
        real, cplx, npc, npr = types(p)
        r2c = DeviceFFT(n, -1, kind="r2c", dtype=real)
        got = run(gpu, r2c, x.astype(npr), (n // 2 + 1,), cplx)
        assert rel(got, X) < tol, (n, p, "r2c")
        c2r = DeviceFFT(n, 1, kind="c2r", dtype=real)
        back = run(gpu, c2r, X.astype(npc), (n,), real)
        assert rel(back, np.fft.irfft(X, n) * n) < tol, (n, p, "c2r")


def test_mrcfft1d_mcrfft1d(gpu):
    # M = 4 transforms of n = 30: reals with stride 2 and distance 70, spectra
    # interleaved (stride 4, distance 1)
    import torch
    from fftwpp_amd.torch_fft import DeviceFFT
    n, M = 30, 4
    h = n // 2 + 1
    rng = np.random.default_rng(5)
    rbuf = rng.uniform(-1, 1, (M - 1) * 70 + (n - 1) * 2 + 1)
    ridx = np.arange(M)[:, None] * 70 + np.arange(n)[None, :] * 2
    cidx = np.arange(M)[:, None] + np.arange(h)[None, :] * M
    r2c = DeviceFFT(n, -1, kind="r2c", dtype=torch.float64, howmany=M, istride=2, ostride=M,
                    idist=70, odist=1)
    cout = torch.zeros(M * h, dtype=torch.complex128, device=gpu)
    r2c(torch.tensor(rbuf, device=gpu), cout)
    torch.cuda.synchronize()
    X = np.fft.rfft(rbuf[ridx], axis=1)
    assert rel(cout.cpu().numpy()[cidx], X) < TOL64
    c2r = DeviceFFT(n, 1, kind="c2r", dtype=torch.float64, howmany=M, istride=M, ostride=2,
                    idist=1, odist=70)
    rback = torch.zeros_like(torch.tensor(rbuf, device=gpu))
    c2r(cout, rback, scale=1.0 / n)
    torch.cuda.synchronize()
    assert rel(rback.cpu().numpy()[ridx], rbuf[ridx]) < TOL64


def test_cxx_face_host_arrays(gpu):
    # fft1d / rcfft1d / crfft1d / fft2d / fft3d of include/fftw++.h on host arrays
    lib = ctypes.CDLL(TESTLIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(17)
    for n in [60, 1024, 6000]:
        cin = rnd(rng, n)
        cout = np.zeros(n, complex)
        cback = np.zeros(n, complex)
        rin = rng.uniform(-1, 1, n)
        rout = np.zeros(n // 2 + 1, complex)
        rback = np.zeros(n)
        args = [a.ctypes.data_as(dp) for a in (cin, cout, cback, rin, rout, rback)]
        assert lib.testfft_cxx(ctypes.c_size_t(n), *args) == 0
        assert rel(cout, np.fft.fft(cin)) < TOL64
        assert rel(cback, cin) < TOL64
        assert rel(rout, np.fft.rfft(rin)) < TOL64
        assert rel(rback, rin) < TOL64
    for shape in [(12, 20), (5, 6, 7)]:
        x = rnd(rng, shape)
        out = np.zeros(shape, complex)
        ns = (ctypes.c_size_t * 3)(*shape)
        assert lib.testfft_multi_cxx(len(shape), ns, x.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
        assert rel(out, np.fft.fftn(x)) < TOL64
        y = x.copy()
        assert lib.testfft_multi_cxx(len(shape), ns, None, y.ctypes.data_as(dp)) == 0
        assert rel(y, np.fft.fftn(x)) < TOL64


@pytest.mark.parametrize("shape", [(8, 10), (7, 9), (64, 48), (256, 256), (5, 6, 7), (64, 64, 64),
                                   (3, 5003), (40, 30, 20)])
def test_real_multi(gpu, shape):
    # rcfft2d / rcfft3d against numpy.fft.rfftn, crfft2d / crfft3d against
    # irfftn (unnormalised), out of place and in place (padded real rows)
    import torch
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1, 1, shape)
    X = np.fft.rfftn(x)
    N = int(np.prod(shape))
    hl = shape[-1] // 2 + 1
    for p, tol in PRECISIONS:
        real, cplx, npc, npr = types(p)
        r2c = DeviceFFT(shape, -1, kind="r2c", dtype=real)
        got = run(gpu, r2c, x.astype(npr), X.shape, cplx)
        assert rel(got, X) < tol, (shape, p, "r2c")
        c2r = DeviceFFT(shape, 1, kind="c2r", dtype=real)
        back = run(gpu, c2r, X.astype(npc), shape, real)
        assert rel(back, np.fft.irfftn(X, shape, axes=tuple(range(len(shape)))) * N) < tol, (shape, p, "c2r")
        # in place: rows of 2 hl reals hold the data, then the spectrum
        pad = np.zeros(shape[:-1] + (2 * hl,))
        pad[..., :shape[-1]] = x
        t = torch.tensor(pad.astype(npr), device=gpu)
        tc = t.view(cplx)
        r2c(t, tc)
        torch.cuda.synchronize()
        assert rel(tc.cpu().numpy(), X) < tol, (shape, p, "r2c in place")
        c2r(tc, t, scale=1.0 / N)
        torch.cuda.synchronize()
        assert rel(t.cpu().numpy()[..., :shape[-1]], x) < tol, (shape, p, "c2r in place")


def test_cxx_face_real_multi(gpu):
    lib = ctypes.CDLL(TESTLIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(23)
    for shape in [(12, 20), (5, 6, 7)]:
        x = rng.uniform(-1, 1, shape)
        out = np.zeros(shape[:-1] + (shape[-1] // 2 + 1,), complex)
        back = np.zeros(shape)
        ns = (ctypes.c_size_t * 3)(*shape)
        assert lib.testfft_real_multi_cxx(len(shape), ns, x.ctypes.data_as(dp),
                                          out.ctypes.data_as(dp), back.ctypes.data_as(dp)) == 0
        assert rel(out, np.fft.rfftn(x)) < TOL64
        assert rel(back, x) < TOL64


@pytest.mark.parametrize("rows,cols,w", [(1, 1, 1), (33, 70, 1), (256, 256, 2), (1000, 3, 2),
                                         (17, 45, 3), (64, 31, 5)])
def test_transpose(gpu, rows, cols, w):
    # fftwpp_amd_transpose (the Transpose class's plan): out[c][r] = in[r][c]
    import torch
    from fftwpp_amd._lib import check, lib
    rng = np.random.default_rng(rows * cols + w)
    x = rng.uniform(-1, 1, (rows, cols, w))
    t = torch.tensor(x, device=gpu)
    o = torch.zeros((cols, rows, w), dtype=torch.float64, device=gpu)
    check(lib().fftwpp_amd_transpose(rows, cols, w, t.data_ptr(), o.data_ptr(), None))
    torch.cuda.synchronize()
    assert np.array_equal(o.cpu().numpy(), x.transpose(1, 0, 2))


def test_cxx_transpose_host_arrays(gpu):
    lib = ctypes.CDLL(TESTLIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(29)
    for rows, cols, cplx, w in [(20, 37, 1, 2), (9, 4, 0, 3)]:
        x = rng.uniform(-1, 1, (rows, cols, w))
        out = np.zeros((cols, rows, w))
        assert lib.testtranspose_cxx(ctypes.c_size_t(rows), ctypes.c_size_t(cols), cplx,
                                     x.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
        # out was transposed back in place: equal to the input again
        assert np.array_equal(out.reshape(rows, cols, w), x)


@pytest.mark.parametrize("shape", [(8, 10), (12, 7), (6, 4, 10), (4, 8, 5)])
def test_cxx_fft0_denyquist(gpu, shape):
    # rcfft2d/3d fft0: the transform of x (-1)^i (2D) or x (-1)^(i+j) (3D)
    # (fftw++.h:144-222); crfft fft0Normalized inverts it; deNyquist zeroes
    # the shifted Nyquist modes (fftw++.h:1342-1360, 1466-1495)
    lib = ctypes.CDLL(TESTLIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1, 1, shape)
    ix = np.indices(shape)
    sign = (-1.0) ** (ix[0] if len(shape) == 2 else ix[0] + ix[1])
    half = shape[:-1] + (shape[-1] // 2 + 1,)
    out = np.zeros(half, complex)
    back = np.zeros(shape)
    nyq = np.zeros(half, complex)
    ns = (ctypes.c_size_t * 3)(*shape)
    assert lib.testfft0_cxx(len(shape), ns, x.ctypes.data_as(dp), out.ctypes.data_as(dp),
                            back.ctypes.data_as(dp), nyq.ctypes.data_as(dp)) == 0
    X = np.fft.rfftn(x * sign)
    assert rel(out, X) < TOL64
    assert rel(back, x) < TOL64
    want = X.copy()
    if shape[0] % 2 == 0:
        want[0] = 0
    if len(shape) == 3 and shape[1] % 2 == 0:
        want[:, 0] = 0
    if shape[-1] % 2 == 0:
        want[..., -1] = 0
    assert np.array_equal(nyq == 0, want == 0)
    assert rel(nyq, want) < TOL64


def test_normalize(gpu):
    # fftwpp_amd_fft_normalize (fftw::Normalize): out / N over the plan's output
    # layout, for strided complex, r2c, c2r 2D and c2c 3D plans
    import torch
    from fftwpp_amd._lib import check, lib
    from fftwpp_amd.torch_fft import DeviceFFT
    rng = np.random.default_rng(41)
    cases = [(DeviceFFT(48, -1, dtype=torch.float64, howmany=5, istride=5, ostride=5, idist=1,
                        odist=1), torch.complex128, 240, 48),
             (DeviceFFT(30, -1, kind="r2c", dtype=torch.float64), torch.complex128, 16, 30),
             (DeviceFFT((6, 10), 1, kind="c2r", dtype=torch.float64), torch.float64, 60, 60),
             (DeviceFFT((4, 5, 6), 1, dtype=torch.float32), torch.complex64, 120, 120)]
    for plan, dt, count, N in cases:
        v = rng.uniform(-1, 1, count) + (1j * rng.uniform(-1, 1, count) if dt.is_complex else 0)
        t = torch.tensor(v, dtype=dt, device=gpu)
        check(lib().fftwpp_amd_fft_normalize(plan._h, t.data_ptr(), None))
        torch.cuda.synchronize()
        assert rel(t.cpu().numpy(), v / N) < (1e-6 if dt == torch.complex64 else 1e-15)


@pytest.mark.parametrize("shape", [(6, 10), (5, 7), (4, 6, 8), (3, 5, 7)])
def test_cxx_inplace_real_normalize(gpu, shape):
    """In-place rcfft{2,3}d then crfft{2,3}d on FFTW's padded rows, then
    Normalize (fftw++.h:357-370) over the padded array: every row, the last
    included, returns to the input reals."""
    lib = ctypes.CDLL(TESTLIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1, 1, shape)
    padded = np.zeros(shape[:-1] + (2 * (shape[-1] // 2 + 1),))
    padded[..., :shape[-1]] = x
    ns = (ctypes.c_size_t * 3)(*shape)
    assert lib.testfft_inplace_normalize_cxx(len(shape), ns, padded.ctypes.data_as(dp)) == 0
    assert rel(padded[..., :shape[-1]], x) < TOL64
