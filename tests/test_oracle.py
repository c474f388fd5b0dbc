"""Pins the CPU oracle (oracle/oracle.c, oracle/oracle.py) before anything is
checked against it.  Every oracle function is compared with something it does
not share code with: numpy/scipy library routines, closed forms from the
reference's accuracy test, brute-force definitions, or invariants."""
import math

import numpy as np
import pytest
import scipy.signal as ss

TOL = 1e-12


def rnd(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


# ---------------------------------------------------------------- direct sums

def test_direct_conv_vs_numpy(oracle):
    rng = np.random.default_rng(1)
    for L in [1, 2, 7, 8, 33]:
        f, g = rnd(rng, L), rnd(rng, L)
        assert oracle.rel_l2(oracle.direct_conv(f, g), np.convolve(f, g)[:L]) < TOL


def test_direct_conv_closed_form(oracle):
    # tests/hybridconv.cc:13-14,50-69: f=iF e^{ij}, g=iG e^{ij} => h_j=iF iG (j+1) e^{ij}
    iF, iG = complex(math.sqrt(3), math.sqrt(7)), complex(math.sqrt(5), math.sqrt(11))
    for L in [64, 256]:
        j = np.arange(L)
        e = np.exp(1j * j)
        h = oracle.direct_conv(iF * e, iG * e)
        assert oracle.rel_l2(h, iF * iG * (j + 1) * e) < 1e-13


def test_direct_convC_vs_numpy(oracle):
    rng = np.random.default_rng(2)
    for L in [1, 2, 7, 8, 20]:
        f, g = rnd(rng, L), rnd(rng, L)
        H = L // 2
        full = np.convolve(f, g)  # index n <-> k = n - 2H
        want = full[H:H + L]      # k in [-H, L-H)
        assert oracle.rel_l2(oracle.direct_convC(f, g), want) < TOL


def _hermitian_full(f):
    H = f.size
    return np.concatenate([np.conj(f[1:][::-1]), f])  # k = -(H-1)..H-1


def test_direct_convh_vs_expansion(oracle):
    rng = np.random.default_rng(3)
    for H in [1, 2, 4, 9]:
        f, g = rnd(rng, H), rnd(rng, H)
        f[0] = f[0].real
        g[0] = g[0].real
        full = np.convolve(_hermitian_full(f), _hermitian_full(g))  # k=-(2H-2)..
        want = full[2 * H - 2:2 * H - 2 + H]
        assert oracle.rel_l2(oracle.direct_convh(f, g), want) < TOL


def test_direct_conv2_conv3_vs_scipy(oracle):
    rng = np.random.default_rng(4)
    f, g = rnd(rng, 5, 7), rnd(rng, 5, 7)
    assert oracle.rel_l2(oracle.direct_conv2(f, g), ss.convolve(f, g)[:5, :7]) < TOL
    f, g = rnd(rng, 3, 4, 5), rnd(rng, 3, 4, 5)
    assert oracle.rel_l2(oracle.direct_conv3(f, g), ss.convolve(f, g)[:3, :4, :5]) < TOL


def _expand2(f):
    """[Lx][Hy] centred-x, non-negative-y -> full [Lx][2Hy-1] grid (kx centred,
    ky = -(Hy-1)..Hy-1), mirror rows outside the stored range are zero."""
    Lx, Hy = f.shape
    off = Lx // 2
    full = np.zeros((Lx, 2 * Hy - 1), dtype=complex)
    full[:, Hy - 1:] = f
    for i in range(Lx):
        kx = i - off
        mi = -kx + off
        if 0 <= mi < Lx:
            full[i, :Hy - 1] = np.conj(f[mi, 1:][::-1])
    return full


@pytest.mark.parametrize("Lx,Ly", [(7, 7), (8, 8), (7, 8), (8, 7), (9, 5)])
def test_direct_convh2_vs_scipy(oracle, Lx, Ly):
    Hy = (Ly + 1) // 2
    f = oracle.ramph2(Lx, Hy, 0)
    g = oracle.ramph2(Lx, Hy, 1)
    h = oracle.direct_convh2(f, g)
    F, G = _expand2(f), _expand2(g)
    full = ss.convolve(F, G)  # x index n <-> kx = n - 2*off ; y index <-> ky = n - 2(Hy-1)
    off = Lx // 2
    want = full[off:off + Lx, 2 * (Hy - 1):2 * (Hy - 1) + Hy]
    if Lx % 2 == 0:
        # even Lx: the kx=-Lx/2 source row is excluded (xcompact=0), which the
        # zeroed Nyquist row of the symmetrised input already guarantees
        pass
    assert oracle.rel_l2(h, want) < TOL


@pytest.mark.parametrize("shape", [(7, 7, 7), (8, 8, 8), (5, 6, 7)])
def test_direct_convh3_vs_explicit_and_expansion(oracle, shape):
    Lx, Ly, Lz = shape
    Hz = (Lz + 1) // 2
    f = oracle.ramph3(Lx, Ly, Hz, 0)
    g = oracle.ramph3(Lx, Ly, Hz, 1)
    h = oracle.direct_convh3(f, g)
    # brute-force full-spectrum expansion + scipy linear convolution
    F = oracle.expand_hermitian(f, None)
    G = oracle.expand_hermitian(g, None)
    full = ss.convolve(F, G)
    ox, oy = Lx // 2, Ly // 2
    want = full[ox:ox + Lx, oy:oy + Ly, 2 * (Hz - 1):2 * (Hz - 1) + Hz]
    assert oracle.rel_l2(h, want) < TOL


def test_symmetrize_x_properties(oracle):
    rng = np.random.default_rng(5)
    for Lx in [7, 8]:
        f = np.ascontiguousarray(rnd(rng, Lx, 4))
        oracle.symmetrize_x(f)
        Hx, x0 = (Lx + 1) // 2, Lx // 2
        for i in range(1, Hx):
            assert f[x0 - i, 0] == np.conj(f[x0 + i, 0])
        assert f[x0, 0].imag == 0.0
        if x0 == Hx:
            assert np.all(f[0] == 0)
        g = f.copy()
        oracle.symmetrize_x(g)
        assert np.array_equal(f, g)  # idempotent


def test_symmetrize_xy_properties(oracle):
    rng = np.random.default_rng(6)
    for Lx, Ly in [(7, 7), (8, 6), (5, 8)]:
        f = np.ascontiguousarray(rnd(rng, Lx, Ly, 3))
        oracle.symmetrize_xy(f)
        x0, y0 = Lx // 2, Ly // 2
        Hx, Hy = (Lx + 1) // 2, (Ly + 1) // 2
        for i in range(-Hx + 1, Hx):
            for j in range(-Hy + 1, Hy):
                a, b = x0 + i, y0 + j
                ma, mb = x0 - i, y0 - j
                if (0 <= a < Lx and 0 <= b < Ly and 0 <= ma < Lx and 0 <= mb < Ly):
                    assert abs(f[a, b, 0] - np.conj(f[ma, mb, 0])) < 1e-15
        g = f.copy()
        oracle.symmetrize_xy(g)
        assert np.array_equal(f, g)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 11, 12, 13, 60, 64, 210, 256, 1024])
def test_oracle_fft_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rnd(rng, n)
    assert oracle.rel_l2(oracle.fft(x, 1), np.fft.ifft(x) * n) < 1e-13
    assert oracle.rel_l2(oracle.fft(x, -1), np.fft.fft(x)) < 1e-13


# ---------------------------------------------------------------- hybrid port

def complex_grid():
    # tests/tests.py:561-573: L=8, M in {16,20}, m in {M, L+1, L, L/2, L/4}
    L = 8
    out = []
    for M in [16, 20]:
        for m in [M, L + 1, L, 4, 2]:
            p, n, q = __import__("oracle.oracle", fromlist=["x"]).parameters(L, M, m, False)
            out.append((L, M, m, q))
    return out


@pytest.mark.parametrize("L,M,m,q", complex_grid())
def test_hybrid_conv1_grid(oracle, L, M, m, q):
    f, g = oracle.ramp1(L, 0), oracle.ramp1(L, 1)
    assert oracle.rel_l2(oracle.hybrid_conv1(f, g, m, q), oracle.direct_conv(f, g)) < TOL


def test_hybrid_conv1_closed_form(oracle):
    iF, iG = complex(math.sqrt(3), math.sqrt(7)), complex(math.sqrt(5), math.sqrt(11))
    for L in [64, 512, 2048]:
        j = np.arange(L)
        e = np.exp(1j * j)
        h = oracle.hybrid_conv1(iF * e, iG * e, L, 2)
        assert oracle.rel_l2(h, iF * iG * (j + 1) * e) < 1e-12


@pytest.mark.parametrize("L", [7, 8])
def test_hybrid_centered_and_hermitian(oracle, L):
    H = (L + 1) // 2
    M = 3 * H - 2 * (L % 2)
    f, g = oracle.ramp1(L, 0), oracle.ramp1(L, 1)
    d = oracle.direct_convC(f, g)
    for m in [M, 4, 2]:
        _, _, q = oracle.parameters(L, M, m, True)
        assert oracle.rel_l2(oracle.hybrid_conv1(f, g, m, q, True), d) < TOL
    assert oracle.rel_l2(oracle.explicit_conv(f, g, M, True), d) < TOL
    fh, gh = oracle.rampH1(H, 0), oracle.rampH1(H, 1)
    dh = oracle.direct_convh(fh, gh)
    for m in [M, 4, 2]:
        _, _, q = oracle.parameters(L, M, m, True)
        assert oracle.rel_l2(oracle.hybrid_hconv1(fh, gh, L, m, q), dh) < TOL
    assert oracle.rel_l2(oracle.explicit_hconv(fh, gh, M), dh) < TOL


def test_hybrid_multid(oracle):
    rng = np.random.default_rng(7)
    f, g = rnd(rng, 6, 5), rnd(rng, 6, 5)
    d = oracle.direct_conv2(f, g)
    assert oracle.rel_l2(oracle.hybrid_conv2(f, g, 5, 8, 2, 3, 4), d) < TOL
    assert oracle.rel_l2(oracle.explicit_conv(f, g, (11, 9)), d) < TOL
    f, g = rnd(rng, 4, 3, 5), rnd(rng, 4, 3, 5)
    d = oracle.direct_conv3(f, g)
    assert oracle.rel_l2(oracle.hybrid_conv3(f, g, 5, 4, 2, 2, 3, 8, 2), d) < TOL
    for Lx, Ly in [(7, 7), (8, 8)]:
        Hy = (Ly + 1) // 2
        f, g = oracle.ramph2(Lx, Hy, 0), oracle.ramph2(Lx, Hy, 1)
        d = oracle.direct_convh2(f, g)
        Mx = 3 * ((Lx + 1) // 2) - 2 * (Lx % 2)
        My = 3 * Hy - 2 * (Ly % 2)
        assert oracle.rel_l2(oracle.hybrid_conv2(f, g, Ly, Mx, 1, 4, 3, True), d) < TOL
        assert oracle.rel_l2(oracle.explicit_hconv(f, g, (Mx, My)), d) < TOL


def test_explicit_independent_of_padding(oracle):
    # M >= 2L-1: the result does not depend on the padded length
    rng = np.random.default_rng(8)
    f, g = rnd(rng, 40), rnd(rng, 40)
    a = oracle.explicit_conv(f, g, 79)
    for N in [80, 96, 127]:
        assert oracle.rel_l2(oracle.explicit_conv(f, g, N), a) < TOL


def test_explicit_multithreaded_path(oracle):
    """The in-place multithreaded path used for full-size arrays gives the
    direct sums too (2D, 3D and the centred Hermitian 2D case)."""
    rng = np.random.default_rng(9)
    f, g = rnd(rng, 6, 5, 7), rnd(rng, 6, 5, 7)
    d = oracle.direct_conv3(f, g)
    assert oracle.rel_l2(oracle.explicit_conv(f, g, (11, 9, 13), workers=2), d) < TOL
    f, g = rnd(rng, 9, 8), rnd(rng, 9, 8)
    assert oracle.rel_l2(oracle.explicit_conv(f, g, (17, 16), workers=2),
                         oracle.direct_conv2(f, g)) < TOL
    for Lx, Ly in [(7, 7), (9, 5)]:
        Hy = (Ly + 1) // 2
        fh, gh = oracle.ramph2(Lx, Hy, 0), oracle.ramph2(Lx, Hy, 1)
        Mx = 3 * ((Lx + 1) // 2) - 2 * (Lx % 2)
        My = 3 * Hy - 2 * (Ly % 2)
        assert oracle.rel_l2(oracle.explicit_hconv(fh, gh, (Mx, My), workers=2),
                             oracle.direct_convh2(fh, gh)) < TOL


def _cyclic_brute(f, g, N):
    # h[n] = sum over i + j = n (mod N) of f_i g_j, by definition
    h = np.zeros(N, dtype=complex)
    for i, fi in enumerate(f):
        for j, gj in enumerate(g):
            h[(i + j) % N] += fi * gj
    return h


def test_explicit_general_products(oracle):
    # user multipliers by explicit padding: binary, ternary, square and
    # two-output products against direct sums and definitions
    rng = np.random.default_rng(31)
    L = 23
    f, g, h = rnd(rng, L), rnd(rng, L), rnd(rng, L)
    mul2 = lambda F, x: [F[0] * F[1]]
    assert oracle.rel_l2(oracle.explicit_general([f, g], 2 * L, mul2)[0],
                         oracle.direct_conv(f, g)) < TOL
    ternary = lambda F, x: [F[0] * F[1] * F[2]]
    want3 = oracle.direct_conv(oracle.direct_conv(f, g), h)
    for N in [3 * L - 2, 80]:
        assert oracle.rel_l2(oracle.explicit_general([f, g, h], N, ternary)[0], want3) < TOL
    square = lambda F, x: [F[0] * F[0]]
    assert oracle.rel_l2(oracle.explicit_general([f], 50, square)[0],
                         oracle.direct_conv(f, f)) < TOL
    pair = lambda F, x: [F[0] * F[1], F[0] - F[1]]
    c, d = oracle.explicit_general([f, g], 48, pair)
    assert oracle.rel_l2(c, oracle.direct_conv(f, g)) < TOL
    assert oracle.rel_l2(d, f - g) < TOL
    # centred data against the centred direct sums
    assert oracle.rel_l2(oracle.explicit_general([f, g], 36, mul2, centered=True)[0],
                         oracle.direct_convC(f, g)) < TOL


def test_explicit_general_grid_point_sign(oracle):
    # a weight exp(2 pi i x/N) at padded grid point x shifts the circular
    # convolution up by one under the forward sign +1: out_k = h_{k-1 mod N}
    rng = np.random.default_rng(32)
    L, N = 9, 20
    f, g = rnd(rng, L), rnd(rng, L)
    h = _cyclic_brute(f, g, N)
    up = lambda F, x: [F[0] * F[1] * np.exp(2j * np.pi * x[0] / N)]
    got = oracle.explicit_general([f, g], N, up)[0]
    assert oracle.rel_l2(got, h[(np.arange(L) - 1) % N]) < TOL
    # 2D: a weight along the last axis shifts that axis only
    Lx, Ly, Nx, Ny = 4, 5, 8, 11
    f2, g2 = rnd(rng, Lx, Ly), rnd(rng, Lx, Ly)
    H = np.zeros((Nx, Ny), dtype=complex)
    for i in range(Lx):
        for j in range(Ly):
            for a in range(Lx):
                for b in range(Ly):
                    H[(i + a) % Nx, (j + b) % Ny] += f2[i, j] * g2[a, b]
    up2 = lambda F, x: [F[0] * F[1] * np.exp(2j * np.pi * x[1] / Ny)]
    got2 = oracle.explicit_general([f2, g2], (Nx, Ny), up2)[0]
    want2 = H[np.ix_(np.arange(Lx), (np.arange(Ly) - 1) % Ny)]
    assert oracle.rel_l2(got2, want2) < TOL


def test_pad_forward_definition(oracle):
    # residue r of the padded transform: F_{qs+r} = sum_k f_k e^{2pi i k(qs+r)/N}
    rng = np.random.default_rng(9)
    for L, m, q, centered in [(8, 8, 2, False), (8, 4, 4, False), (7, 4, 3, True),
                              (8, 2, 6, True), (9, 3, 4, False)]:
        f = rnd(rng, L)
        N = m * q
        k = np.arange(L) - (L // 2 if centered else 0)
        for r in range(q):
            ell = q * np.arange(m) + r
            want = np.exp(2j * np.pi * np.outer(ell, k) / N) @ f
            got = oracle.pad_forward(f, m, q, r, centered)
            assert oracle.rel_l2(got, want) < 1e-13


# ---------------------------------------------------------------- parameters

def test_parameters_worked_geometries(oracle):
    # convolve.cc:403-410 on the measured configurations
    assert oracle.parameters(1024, 2048, 1024, False) == (1, 2, 2)
    assert oracle.parameters(256, 512, 256, False) == (1, 2, 2)
    assert oracle.parameters(256, 512, 128, False)[2] == 4
    assert oracle.parameters(2047, 3070, 1024, True) == (2, 3, 3)
    assert oracle.parameters(2047, 3070, 512, True)[2] == 6
    assert oracle.default_M(2047, hermitian=True) == 3070
    assert oracle.default_M(1023, hermitian=True) == 1534
    assert oracle.default_M(256) == 511


@pytest.mark.parametrize("lengths", [(11,), (9, 9), (5, 7, 5)])
def test_hermitian_adjoint_formula_fd(oracle, lengths):
    """Central finite differences of L = sum |C(S f, S g) - t|^2 (direct
    Hermitian sums of symmetrized inputs) pin the adjoint formula that the
    GPU implements: dL = 2 Re <grad, delta> with grad = dL/d conj f."""
    rng = np.random.default_rng(17 + len(lengths))
    shape = tuple(lengths[:-1]) + ((lengths[-1] + 1) // 2,)
    plane = oracle.hplane_of(lengths)
    conv = {1: oracle.direct_convh, 2: oracle.direct_convh2, 3: oracle.direct_convh3}[len(lengths)]

    def rnd():
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f, g, t = rnd(), rnd(), rnd()

    def loss(f, g):
        h = conv(oracle.hplane(f, 0, plane), oracle.hplane(g, 0, plane))
        return np.sum(np.abs(h - t) ** 2)
    h = conv(oracle.hplane(f, 0, plane), oracle.hplane(g, 0, plane))
    gf, gg = oracle.hconv_adjoint(h - t, f, g, lengths)
    eps = 1e-6
    for _ in range(4):
        d = rnd()
        for grad, lp, lm in ((gf, loss(f + eps * d, g), loss(f - eps * d, g)),
                             (gg, loss(f, g + eps * d), loss(f, g - eps * d))):
            fd = (lp - lm) / (2 * eps)
            an = 2 * np.real(np.sum(np.conj(grad) * d))
            assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_hplane_symmetrize_matches_reference_rule(oracle):
    """hplane op 0 is HermitianSymmetrizeX / XY (odd lengths) as restated in
    oracle.c."""
    rng = np.random.default_rng(4)
    f = rng.standard_normal((9, 5)) + 1j * rng.standard_normal((9, 5))
    assert np.array_equal(oracle.hplane(f, 0, oracle.hplane_of((9, 9))),
                          oracle.symmetrize_x(f.copy()))
    f3 = rng.standard_normal((5, 7, 4)) + 1j * rng.standard_normal((5, 7, 4))
    assert np.array_equal(oracle.hplane(f3, 0, oracle.hplane_of((5, 7, 7))),
                          oracle.symmetrize_xy(f3.copy()))
