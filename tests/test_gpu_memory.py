"""Memory-bounded execution by outer residue blocks (the reference's residue
loop, convolve.h:1475-1535 and 1812-1879, processing D residues per pass,
convolve.h:445-451): under a memory limit that excludes the whole expansion,
2D and 3D plans run their x residues in blocks, each block's backward
transform added to the output, and report work bytes within the limit.
Results equal the unlimited plan's (which the parity tests check against
explicit padding) to 1e-14 (FP64) / 1e-6 (FP32), in and out of place."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rnd(rng, shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _run(gpu, shape, kind, dtype, limit_frac, lengths=None, M=None, m=None, seed=1):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(seed)
    f, g = _rnd(rng, shape), _rnd(rng, shape)
    free = DeviceConvolution(shape, kind=kind, dtype=dtype, lengths=lengths, M=M, m=m)
    assert free.geometry()[2][0] > 1, free.geometry()
    full, _ = free.work_bytes()
    a = torch.tensor(f, dtype=dtype, device=gpu)
    b = torch.tensor(g, dtype=dtype, device=gpu)
    ref = torch.empty_like(a)
    free.convolve_out(ref, a, b)
    capped = DeviceConvolution(shape, kind=kind, dtype=dtype, lengths=lengths, M=M, m=m)
    limit = int(full * limit_frac)
    capped.set_memory_limit(limit)
    need, plane = capped.work_bytes()
    assert not plane and need <= limit < full, (need, limit, full)
    out = torch.empty_like(a)
    capped.convolve_out(out, a, b)
    inplace = a.clone()
    capped.convolve_(inplace, b)
    torch.cuda.synchronize()
    r = ref.cpu().numpy()
    tol = 1e-14 if dtype == torch.complex128 else 1e-6
    for got in (out, inplace):
        h = got.cpu().numpy()
        err = float(np.linalg.norm((h - r).ravel()) / np.linalg.norm(r.ravel()))
        assert err < tol, (shape, kind, err)
    return full, need


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_residue_blocks_3d(gpu, oracle, prec):
    import torch
    dt = torch.complex128 if prec == "f64" else torch.complex64
    # x: m = 32, q = 3 (M = 96): blocks of one or two residues
    full, need = _run(gpu, (32, 24, 40), "complex", dt, 0.75, M=[96, 48, 80], m=[32, 0, 0])
    print(f"3D {prec}: {full} bytes unlimited, {need} in residue blocks", flush=True)


def test_residue_blocks_2d_and_hermitian(gpu, oracle):
    import torch
    _run(gpu, (40, 33), "complex", torch.complex128, 0.75, M=[120, 66], m=[40, 0])
    # centred Hermitian 2D: x centred with p = 2
    _run(gpu, (63, 32), "hermitian", torch.complex128, 0.75, lengths=[63, 63], m=[32, 0])
    # 3D centred Hermitian
    _run(gpu, (15, 13, 8), "hermitian", torch.complex128, 0.75, lengths=[15, 13, 15],
         m=[8, 0, 0])


def test_memory_limit_too_small_is_refused(gpu, oracle):
    import torch
    from fftwpp_amd._lib import FFTWPPError
    from fftwpp_amd.torch_conv import DeviceConvolution
    plan = DeviceConvolution((40, 33), M=[120, 66])
    full, _ = plan.work_bytes()
    plan.set_memory_limit(full // 50)
    a = torch.ones((40, 33), dtype=torch.complex128, device=gpu)
    with pytest.raises(FFTWPPError):
        plan.convolve_out(torch.empty_like(a), a, a)
