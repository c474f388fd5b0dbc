"""Hand-written adjoint of the centred Hermitian convolutions
(HybridConvolutionHermitian{,2,3}) and its torch.autograd.Function: against
the oracle's adjoint formula on direct Hermitian sums (itself pinned by finite
differences in tests/test_oracle.py), against central finite differences of
the GPU forward, and, at the size of the differentiable 2D workload
(1023 x 512 stored modes, M = 1534), through the adjoint identity
Re <grad_f, d> = Re <w, C(S d, S g)>."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _rnd(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _shape(lengths):
    return tuple(lengths[:-1]) + ((lengths[-1] + 1) // 2,)


@pytest.mark.parametrize("lengths", [(31,), (15, 13), (7, 9, 7)])
def test_hermitian_adjoint_matches_oracle(gpu, oracle, lengths):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.default_rng(90 + len(lengths))
    shape = _shape(lengths)
    a, b, w = _rnd(rng, shape), _rnd(rng, shape), _rnd(rng, shape)
    want_a, want_b = oracle.hconv_adjoint(w, a, b, lengths)
    for dtype, tol in ((torch.complex128, TOL64), (torch.complex64, TOL32)):
        plan = DeviceConvolution(shape, kind="hermitian", dtype=dtype, lengths=lengths)
        T = lambda x: torch.tensor(x, device=gpu, dtype=dtype)
        ga, gb = plan.adjoint(T(w), T(a), T(b))
        torch.cuda.synchronize()
        assert oracle.rel_l2(ga.cpu().numpy(), want_a) < tol
        assert oracle.rel_l2(gb.cpu().numpy(), want_b) < tol


def test_hermitian_forward_of_symmetrized_inputs(gpu, oracle):
    """The differentiable forward convolves the symmetrized inputs."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution, convolve
    rng = np.random.default_rng(95)
    lengths = (15, 13)
    shape = _shape(lengths)
    a, b = _rnd(rng, shape), _rnd(rng, shape)
    plane = oracle.hplane_of(lengths)
    want = oracle.direct_convh2(oracle.hplane(a, 0, plane), oracle.hplane(b, 0, plane))
    plan = DeviceConvolution(shape, kind="hermitian", lengths=lengths)
    h = convolve(torch.tensor(a, device=gpu), torch.tensor(b, device=gpu), plan)
    torch.cuda.synchronize()
    assert oracle.rel_l2(h.cpu().numpy(), want) < TOL64


def test_hermitian_autograd_finite_differences(gpu, oracle):
    """torch.autograd gradients of L = ||C(S f, S g) - t||^2 against central
    finite differences of the same GPU forward (FP64)."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution, convolve
    rng = np.random.default_rng(96)
    lengths = (11, 9)
    shape = _shape(lengths)
    plan = DeviceConvolution(shape, kind="hermitian", lengths=lengths)
    f0, g0, t0 = _rnd(rng, shape), _rnd(rng, shape), _rnd(rng, shape)
    T = lambda x: torch.tensor(x, device=gpu)
    t = T(t0)
    f = T(f0).requires_grad_(True)
    g = T(g0).requires_grad_(True)
    loss = (convolve(f, g, plan) - t).abs().pow(2).sum()
    loss.backward()

    def L(fv, gv):
        with torch.no_grad():
            return float((convolve(T(fv), T(gv), plan) - t).abs().pow(2).sum())
    eps = 1e-6
    for grad, which in ((f.grad.cpu().numpy(), 0), (g.grad.cpu().numpy(), 1)):
        for _ in range(4):
            d = _rnd(rng, shape)
            if which == 0:
                fd = (L(f0 + eps * d, g0) - L(f0 - eps * d, g0)) / (2 * eps)
            else:
                fd = (L(f0, g0 + eps * d) - L(f0, g0 - eps * d)) / (2 * eps)
            an = float(np.real(np.sum(np.conj(grad) * d)))
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, fd, an)


def _h2_inputs(oracle, Lx, Hy, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [oracle._spectral(rng, Lx, Hy) for _ in range(n)]


def test_hermitian_adjoint_identity_workload_size(gpu, oracle):
    """1023 x 512 stored modes, M = 1534 (the differentiable 2D workload):
    Re <grad_f, d> equals Re <w, C(S d, S g)> and Re <grad_g, d> equals
    Re <w, C(S f, S d)> in FP64; FP32 gradients agree with FP64 to 1e-5."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    Lx, Hy = 1023, 512
    lengths = (Lx, 2 * Hy - 1)
    f0, g0, t0, d0 = _h2_inputs(oracle, Lx, Hy, 4, oracle.SEED)
    grads = {}
    for dtype in (torch.complex128, torch.complex64):
        plan = DeviceConvolution((Lx, Hy), kind="hermitian", dtype=dtype, M=[1534, 1534],
                                 lengths=lengths)
        T = lambda x: torch.tensor(x, device=gpu, dtype=dtype)
        f, g, t, d = T(f0), T(g0), T(t0), T(d0)
        h = plan(plan.symmetrized(f), plan.symmetrized(g))
        w = 2 * (h - t)
        gf, gg = plan.adjoint(w, f, g)
        grads[dtype] = (gf, gg)
        if dtype == torch.complex128:
            dh_f = plan(plan.symmetrized(d), plan.symmetrized(g))
            dh_g = plan(plan.symmetrized(f), plan.symmetrized(d))
            lhs_f = torch.sum(torch.conj(gf) * d).real
            rhs_f = torch.sum(torch.conj(w) * dh_f).real
            lhs_g = torch.sum(torch.conj(gg) * d).real
            rhs_g = torch.sum(torch.conj(w) * dh_g).real
            torch.cuda.synchronize()
            assert abs(float(lhs_f - rhs_f)) <= TOL64 * abs(float(rhs_f))
            assert abs(float(lhs_g - rhs_g)) <= TOL64 * abs(float(rhs_g))
    torch.cuda.synchronize()
    for k in range(2):
        a = grads[torch.complex128][k].cpu().numpy()
        b = grads[torch.complex64][k].cpu().numpy().astype(np.complex128)
        assert oracle.rel_l2(b, a) < TOL32


def test_hermitian_adjoint_rejects_even_centred_length(gpu):
    import torch
    from fftwpp_amd._lib import FFTWPPError
    from fftwpp_amd.torch_conv import DeviceConvolution
    plan = DeviceConvolution((8, 5), kind="hermitian", lengths=(8, 9))
    z = torch.zeros((8, 5), dtype=torch.complex128, device=gpu)
    with pytest.raises(FFTWPPError):
        plan.adjoint(z, z, z)
