"""The y-z plane kernel of 3D convolutions (fftwpp_amd_plane_conv): every
x-plane of Ly x lz = 256 x 256 words convolved over y and z (m = 256, q = 2
each, padded to 512 x 512) in one launch whose intermediates stay in the L2
cache -- the inner Convolution2 of Convolution3::convolveRaw
(convolve.h:1812-1879).

The unscaled plane result is checked against the explicit zero-padded 2D
transform pair with the reference's sign convention (forward e^{+2 pi i},
backward e^{-2 pi i}, no normalisation): for multBinary, 512^2 times the
linear convolution of the plane pair; for multcorrelation, the correlation;
for multNone, 512^2 times f.  FP64 within 1e-12, FP32 within 1e-5 relative
L2, plane counts that do not divide among the 8 XCD teams, in place and out
of place.  The launch's abort code must be 0 (every team wait completed)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N2 = 512 * 512


def dims(L):
    from fftwpp_amd import _lib
    lib = _lib.lib()
    out = []
    for _ in range(2):
        d = ctypes.c_void_p()
        _lib.check(lib.fftwpp_amd_dim_create(ctypes.byref(d), L, 256, 2, 0))
        out.append(d)
    return lib, out


def explicit(f, g, mult):
    """Unscaled padded pair per plane, reference signs: F = sum f zeta^{+jk}."""
    pad = lambda a: np.pad(a, ((0, 0), (0, 256), (0, 256)))
    F = np.fft.ifft2(pad(f), axes=(1, 2)) * N2
    G = np.fft.ifft2(pad(g), axes=(1, 2)) * N2
    if mult == 1:
        P = F * G
    elif mult == 3:
        P = F * np.conj(G)
    else:
        P = F
    return np.fft.fft2(P, axes=(1, 2))[:, :256, :256]


def planes(n, seed, dtype):
    rng = np.random.default_rng(seed)
    shape = (n, 256, 256)
    f = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    g = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return f.astype(dtype), g.astype(dtype)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_plane_kernel_available(gpu):
    """On an MI355X (8 XCDs of 32 CUs) the 3D path must take the plane kernel."""
    lib, (dy, dz) = dims(256)
    try:
        ok = lib.fftwpp_amd_plane_supported(dy, dz)
        assert ok == 1, lib.fftwpp_amd_last_error().decode()
    finally:
        for d in (dy, dz):
            lib.fftwpp_amd_dim_destroy(d)


@pytest.mark.parametrize("nplanes,mult,inplace,prec", [
    (13, 1, False, "f64"), (16, 1, True, "f64"), (9, 3, False, "f64"), (8, 0, True, "f64"),
    (13, 1, False, "f32"), (24, 3, True, "f32"),
])
def test_plane_kernel_vs_explicit(gpu, nplanes, mult, inplace, prec):
    import torch
    from fftwpp_amd import _lib
    lib, (dy, dz) = dims(256)
    try:
        npdt = np.complex128 if prec == "f64" else np.complex64
        tdt = torch.complex128 if prec == "f64" else torch.complex64
        dtype = 0 if prec == "f64" else 1
        f, g = planes(nplanes, 100 + nplanes + mult, npdt)
        print(f"plane kernel: {nplanes} planes, mult {mult}, inplace {inplace}, {prec}", flush=True)
        a = torch.tensor(f, device=gpu, dtype=tdt)
        b = torch.tensor(g, device=gpu, dtype=tdt)
        out = a if inplace else torch.empty_like(a)
        work = torch.empty(lib.fftwpp_amd_plane_work(dtype), dtype=torch.uint8, device=gpu)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.fftwpp_amd_plane_conv(dy, dz, dtype, mult, a.data_ptr(), b.data_ptr(),
                                             out.data_ptr(), nplanes, 256, 256,
                                             work.data_ptr(), stream))
        st = lib.fftwpp_amd_plane_status(work.data_ptr(), stream)
        assert st == 0, st
        h = out.cpu().numpy().astype(np.complex128)
        ref = explicit(f.astype(np.complex128), g.astype(np.complex128), mult)
        err = rel(h, ref)
        print(f"  rel L2 vs explicit padding {err:.3e}", flush=True)
        assert err < (1e-12 if prec == "f64" else 1e-5)
        # a second launch on the same work area (counters cleared per launch)
        out2 = torch.empty_like(b)
        _lib.check(lib.fftwpp_amd_plane_conv(dy, dz, dtype, mult, b.data_ptr(), b.data_ptr(),
                                             out2.data_ptr(), nplanes, 256, 256,
                                             work.data_ptr(), stream))
        assert lib.fftwpp_amd_plane_status(work.data_ptr(), stream) == 0
        ref2 = explicit(g.astype(np.complex128), g.astype(np.complex128), mult)
        assert rel(out2.cpu().numpy().astype(np.complex128), ref2) < (1e-12 if prec == "f64" else 1e-5)
    finally:
        for d in (dy, dz):
            lib.fftwpp_amd_dim_destroy(d)


def test_plane_kernel_rejects_other_shapes(gpu):
    from fftwpp_amd import _lib
    lib, (dy, dz) = dims(200)
    try:
        assert lib.fftwpp_amd_plane_supported(dy, dz) == 0
        rc = lib.fftwpp_amd_plane_conv(dy, dz, 0, 1, 1, 1, 1, 4, 200, 200, 1, None)
        assert rc == -3
    finally:
        for d in (dy, dz):
            lib.fftwpp_amd_dim_destroy(d)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_memory_capped_3d_uses_planes(gpu, oracle, prec):
    """A 256^3 -> 512^3 plan under a memory limit that excludes the separate
    y, z and y stages takes the plane kernel (about a third of the work
    memory) and gives the same result as the unlimited plan, which the
    full-size parity test checks against explicit padding; a limit below the
    plane kernel's need is refused."""
    import torch
    from fftwpp_amd._lib import FFTWPPError
    from fftwpp_amd.torch_conv import DeviceConvolution
    tdt = torch.complex128 if prec == "f64" else torch.complex64
    f, g = oracle.inputs_3d(256)
    a = torch.tensor(f, device=gpu, dtype=tdt)
    b = torch.tensor(g, device=gpu, dtype=tdt)
    free = DeviceConvolution((256,) * 3, dtype=tdt, M=[512] * 3)
    full, plane = free.work_bytes()
    assert not plane
    ref = torch.empty_like(a)
    free.convolve_out(ref, a, b)
    torch.cuda.synchronize()
    del free
    capped = DeviceConvolution((256,) * 3, dtype=tdt, M=[512] * 3)
    E = 16 if prec == "f64" else 8
    X = 512 * 256 * 256 * E
    capped.set_memory_limit(3 * X)
    need, plane = capped.work_bytes()
    print(f"{prec}: work bytes {full / 2**30:.2f} GiB separate stages, {need / 2**30:.2f} GiB plane kernel",
          flush=True)
    assert plane and need <= 3 * X < full
    out = torch.empty_like(a)
    capped.convolve_out(out, a, b)
    torch.cuda.synchronize()
    h, r = out.cpu().numpy(), ref.cpu().numpy()
    err = float(np.linalg.norm((h - r).ravel()) / np.linalg.norm(r.ravel()))
    print(f"  plane vs separate stages rel L2 {err:.3e}", flush=True)
    assert err < (1e-14 if prec == "f64" else 1e-6)
    capped.set_memory_limit(X)
    with pytest.raises(FFTWPPError):
        capped.convolve_out(out, a, b)
