"""Real residue transforms (fftPadReal, convolve.h:805-980; its forward and
backward families, convolve.cc:5449-7483) through the C ABI, and the real
convolutions that run on them.

For real data the padded transform at l and N - l are conjugates, so a real
dimension keeps N/2 + 1 complex planes (residue 0 at s <= m/2, residues
0 < r < q/2 whole, residue q/2 at s < m/2 for even q).  Checks:
  * forward: every kept plane equals sum_j f_j exp(2 pi i j l/N), l = q s + r
    (numpy's inverse FFT of the zero-padded column times N);
  * backward: sum over kept l of c_l Re(F_l exp(-2 pi i j l/N)) with c_l = 1
    on the self-conjugate planes l = 0 and N/2 and 2 elsewhere (numpy's FFT
    of the spectrum holding c_l F_l at l), and backward(forward(f)) = N f;
  * convolutions whose outer dimension takes these transforms (the timeline
    names the stage) against explicit zero padding (scipy/numpy real FFTs),
    up to the full 256^3 -> 512^3 workload.
Bar: relative L2 <= 1e-12 (FP64), <= 1e-5 (FP32)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def kept(m, q):
    """(r, s) of the kept planes in their stored order."""
    out = []
    for r in range(q // 2 + 1):
        cnt = m // 2 + 1 if r == 0 else (m // 2 if 2 * r == q else m)
        out += [(r, s) for s in range(cnt)]
    return out


def weights(m, q):
    """1 on the self-conjugate planes l = 0 and l = N/2, else 2."""
    return np.array([1.0 if (r == 0 and (s == 0 or 2 * s == m)) else 2.0 for r, s in kept(m, q)])


def eligible(L, m):
    return 16 <= m <= 4096 and m & (m - 1) == 0 and L <= 2 * m


# (L, m, q): p = 1 and p = 2, odd and even q, q = 1
CASES = [(9, 16, 2), (16, 16, 1), (20, 16, 3), (31, 16, 4), (100, 64, 5),
         (129, 128, 2), (256, 256, 2), (300, 256, 3), (700, 512, 3), (1024, 1024, 2),
         (2047, 2048, 2), (3000, 4096, 2), (8000, 4096, 2)]


def _dim(L, m, q):
    from fftwpp_amd import _lib
    d = ctypes.c_void_p()
    _lib.check(_lib.lib().fftwpp_amd_dim_create(ctypes.byref(d), L, m, q, 0))
    return d


@pytest.mark.parametrize("L,m,q", CASES)
def test_real_planes_definition(gpu, oracle, L, m, q):
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    N = m * q
    d = _dim(L, m, q)
    try:
        K = lib.fftwpp_amd_real_planes(d)
        assert K == N // 2 + 1 == len(kept(m, q))
        ell = np.array([q * s + r for r, s in kept(m, q)])
        rng = np.random.default_rng(L * 7 + q)
        for C in (1, 3, 37):
            f = rng.uniform(-1, 1, (L, C))
            full = N * np.fft.ifft(np.concatenate([f, np.zeros((N - L, C))]), axis=0)
            want = full[ell]
            spec = rng.uniform(-1, 1, (K, C)) + 1j * rng.uniform(-1, 1, (K, C))
            G = np.zeros((N, C), complex)
            G[ell] = weights(m, q)[:, None] * spec
            want_b = np.fft.fft(G, axis=0).real[:L]
            for key, rt, ct, dcode in (("f64", torch.float64, torch.complex128, 0),
                                       ("f32", torch.float32, torch.complex64, 1)):
                fd = torch.tensor(f, dtype=rt, device=gpu)
                Fd = torch.zeros((K, C), dtype=ct, device=gpu)
                _lib.check(lib.fftwpp_amd_forward_real_many(d, dcode, fd.data_ptr(), C, C, 1, 0,
                                                            Fd.data_ptr(), C, 0, None))
                sd = torch.tensor(spec, dtype=ct, device=gpu)
                bd = torch.zeros((L, C), dtype=rt, device=gpu)
                _lib.check(lib.fftwpp_amd_backward_real_many(d, dcode, sd.data_ptr(), C, 0,
                                                             bd.data_ptr(), C, C, 1, 0, 1.0, None))
                rd = torch.zeros((L, C), dtype=rt, device=gpu)
                _lib.check(lib.fftwpp_amd_backward_real_many(d, dcode, Fd.data_ptr(), C, 0,
                                                             rd.data_ptr(), C, C, 1, 0, 1.0 / N,
                                                             None))
                torch.cuda.synchronize()
                tol = TOL[key]
                assert oracle.rel_l2(Fd.cpu().numpy().astype(complex), want) < tol, (key, C)
                assert oracle.rel_l2(bd.cpu().numpy().astype(float), want_b) < tol, (key, C)
                assert oracle.rel_l2(rd.cpu().numpy().astype(float), f) < tol, (key, C)
    finally:
        lib.fftwpp_amd_dim_destroy(d)


def test_real_planes_batches(gpu, oracle):
    """Batches: single columns of many entries (the entries become the
    columns of one launch) and strided batches of several columns."""
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    L, m, q = 40, 32, 3
    N = m * q
    K = N // 2 + 1
    ell = np.array([q * s + r for r, s in kept(m, q)])
    d = _dim(L, m, q)
    try:
        rng = np.random.default_rng(5)
        # 9 entries of one column, entry stride 50 real words, plane stride 1
        # within an entry of K words
        f = rng.uniform(-1, 1, (9, 50))
        fd = torch.tensor(f, device=gpu)
        Fd = torch.zeros((9, K), dtype=torch.complex128, device=gpu)
        _lib.check(lib.fftwpp_amd_forward_real_many(d, 0, fd.data_ptr(), 1, 1, 9, 50,
                                                    Fd.data_ptr(), 1, K, None))
        back = torch.zeros((9, 50), dtype=torch.float64, device=gpu)
        _lib.check(lib.fftwpp_amd_backward_real_many(d, 0, Fd.data_ptr(), 1, K, back.data_ptr(),
                                                     1, 1, 9, 50, 1.0 / N, None))
        torch.cuda.synchronize()
        x = np.concatenate([f[:, :L], np.zeros((9, N - L))], axis=1)
        want = (N * np.fft.ifft(x, axis=1))[:, ell]
        assert oracle.rel_l2(Fd.cpu().numpy(), want) < 1e-12
        got = back.cpu().numpy()
        assert oracle.rel_l2(got[:, :L], f[:, :L]) < 1e-12
        assert not np.any(got[:, L:])          # words past L untouched
        # 4 batches of 6 columns: column stride 1, element stride 8, batch
        # stride 8 L; planes of 6 at stride 8, batch stride 8 K
        C, S, nb = 6, 8, 4
        f = rng.uniform(-1, 1, (nb, L, S))
        fd = torch.tensor(f, device=gpu)
        Fd = torch.zeros((nb, K, S), dtype=torch.complex128, device=gpu)
        _lib.check(lib.fftwpp_amd_forward_real_many(d, 0, fd.data_ptr(), S, C, nb, L * S,
                                                    Fd.data_ptr(), S, K * S, None))
        torch.cuda.synchronize()
        x = np.concatenate([f[:, :, :C], np.zeros((nb, N - L, C))], axis=1)
        want = (N * np.fft.ifft(x, axis=1))[:, ell, :]
        F = Fd.cpu().numpy()
        assert oracle.rel_l2(F[:, :, :C], want) < 1e-12
        assert not np.any(F[:, :, C:])
    finally:
        lib.fftwpp_amd_dim_destroy(d)


def test_real_planes_reject(gpu):
    """Geometries outside the kernels report 0 planes and refuse."""
    from fftwpp_amd import FFTWPPError, _lib
    lib = _lib.lib()
    for L, m, q in [(12, 12, 2), (40, 16, 3), (8, 8, 2)]:
        d = _dim(L, m, q)
        try:
            assert lib.fftwpp_amd_real_planes(d) == 0
            with pytest.raises(FFTWPPError):
                _lib.check(lib.fftwpp_amd_forward_real_many(d, 0, 1, 1, 1, 1, 0, 1, 1, 0, None))
        finally:
            lib.fftwpp_amd_dim_destroy(d)


def _stages(plan, out, a, b):
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    ms = (ctypes.c_float * 32)()
    n = ctypes.c_int()
    _lib.check(lib.fftwpp_amd_conv_execute_timed(plan._h, out.data_ptr(), a.data_ptr(),
                                                 b.data_ptr(), None, ms, 32, ctypes.byref(n)))
    torch.cuda.synchronize()
    return [lib.fftwpp_amd_conv_stage_name(plan._h, i).decode() for i in range(n.value)]


def _explicit_real(f, g):
    import scipy.fft as sf
    s = tuple(2 * L for L in f.shape)
    H = sf.rfftn(f, s=s, workers=8) * sf.rfftn(g, s=s, workers=8)
    h = sf.irfftn(H, s=s, workers=8)
    return h[tuple(slice(0, L) for L in f.shape)]


@pytest.mark.parametrize("shape", [(4096,), (1000,), (300, 256), (64, 48), (33, 64, 40),
                                   (64, 64, 64)])
def test_real_convolution_planes(gpu, oracle, shape):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(sum(shape))
    f, g = rng.uniform(-1, 1, shape), rng.uniform(-1, 1, shape) * 1e-6
    want = _explicit_real(f, g)
    for dtype, key in ((torch.float64, "f64"), (torch.float32, "f32")):
        plan = DeviceConvolution(shape, kind="real", dtype=dtype)
        (N, m, q) = plan.geometry()
        a = torch.tensor(f, dtype=dtype, device=gpu)
        b = torch.tensor(g, dtype=dtype, device=gpu)
        out = torch.empty_like(a)
        names = _stages(plan, out, a, b)
        if eligible(shape[0], m[0]):
            assert any("real" in n for n in names), names
        got = out.cpu().numpy().astype(np.float64)
        assert oracle.rel_l2(got, want) < TOL[key], (shape, key, names)
        plan.convolve_(a, b)
        torch.cuda.synchronize()
        assert np.array_equal(a.cpu().numpy().astype(np.float64), got)


def test_real3d_full_size(gpu, oracle):
    """The 3D real workload 256^3 -> 512^3 (M = 2L) in FP64 against explicit
    zero padding with scipy's real FFTs."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    L = 256
    rng = np.random.default_rng(20261207)
    f, g = rng.uniform(-1, 1, (L, L, L)), rng.uniform(-1, 1, (L, L, L))
    plan = DeviceConvolution((L, L, L), kind="real", dtype=torch.float64)
    assert plan.geometry()[1][0] == 256 and plan.geometry()[2][0] == 2
    a = torch.tensor(f, device=gpu)
    b = torch.tensor(g, device=gpu)
    out = torch.empty_like(a)
    names = _stages(plan, out, a, b)
    assert "x real forward f" in names and "x real backward" in names, names
    got = out.cpu().numpy()
    del a, b, out
    want = _explicit_real(f, g)
    err = oracle.rel_l2(got, want)
    print(f"real 3D 256^3 FP64 rel L2 {err:.3e}")
    assert err < 1e-12
