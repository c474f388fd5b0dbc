"""2D and 3D convolutions whose innermost transform runs outside the fused row
kernels (tests/support/innerc.cc through the C++ classes of convolve.h):

  * inner transforms over C > 1 interleaved columns, the reference's
    fftPad(L, M, app, C, C) as the y (2D) or z (3D) dimension of
    Convolution2/3 (convolve.h:1475-1535): every column c convolves
    independently over the outer dimensions;
  * a forced inner m that no row engine takes (m = 3000: not a power of two,
    and its sequences overflow LDS), which runs unfused (every residue
    forward, the multiplier, every residue backward).

Against the oracle's direct sums (tests/direct.h) per column, FP64 <= 1e-12."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TESTLIB = os.path.join(HERE, "support", "libtestmult.so")


def _fn(name, nsizes):
    f = getattr(ctypes.CDLL(TESTLIB), name)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_size_t] * nsizes + [ctypes.c_void_p] * 3
    return f


def _rnd(rng, shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


@pytest.mark.parametrize("Lx,Ly,C,my", [(8, 7, 3, 0), (20, 30, 2, 0), (5, 40, 4, 64),
                                        (6, 2000, 1, 3000), (3, 1500, 2, 3000)])
def test_conv2_inner(gpu, oracle, Lx, Ly, C, my):
    rng = np.random.default_rng(Lx * Ly + C)
    f, g = _rnd(rng, (Lx, Ly, C)), _rnd(rng, (Lx, Ly, C))
    h = np.zeros_like(f)
    rc = _fn("testface_conv2_inner", 4)(Lx, Ly, C, my, f.ctypes.data, g.ctypes.data,
                                        h.ctypes.data)
    assert rc == 0, rc
    for c in range(C):
        want = oracle.direct_conv2(f[:, :, c], g[:, :, c])
        assert oracle.rel_l2(h[:, :, c], want) < 1e-12, (Lx, Ly, C, my, c)


@pytest.mark.parametrize("Lx,Ly,Lz,C,mz", [(4, 5, 6, 3, 0), (3, 4, 20, 2, 32),
                                           (2, 3, 2000, 1, 3000)])
def test_conv3_inner(gpu, oracle, Lx, Ly, Lz, C, mz):
    rng = np.random.default_rng(Lx * Ly * Lz + C)
    f, g = _rnd(rng, (Lx, Ly, Lz, C)), _rnd(rng, (Lx, Ly, Lz, C))
    h = np.zeros_like(f)
    rc = _fn("testface_conv3_inner", 5)(Lx, Ly, Lz, C, mz, f.ctypes.data, g.ctypes.data,
                                        h.ctypes.data)
    assert rc == 0, rc
    for c in range(C):
        want = oracle.explicit_conv(f[..., c], g[..., c], (2 * Lx, 2 * Ly, 2 * Lz))
        assert oracle.rel_l2(h[..., c], want) < 1e-12, (Lx, Ly, Lz, C, mz, c)


@pytest.mark.parametrize("shape,A,B,code", [((40,), 3, 2, 1), ((33,), 2, 2, 0), ((30,), 4, 1, 3),
                                            ((8, 7), 3, 1, 3), ((6, 9), 3, 3, 1),
                                            ((4, 5, 6), 4, 3, 1)])
def test_library_multiplier_any_AB(gpu, oracle, shape, A, B, code):
    """multNone / multBinary / multcorrelation with A != 2 or B != 1 through the
    C++ classes (the reference's residue loop takes any A and B,
    convolve.cc:7513-7575): all A inputs transformed, the multiplier on F[0]
    and F[1] (convolve.cc:26-110), the first B arrays transformed back; against
    explicit padding with the same products, FP64 <= 1e-12."""
    rng = np.random.default_rng(sum(shape) + 10 * A + B)
    na = max(A, B)
    arrays = [_rnd(rng, shape) for _ in range(na)]
    buf = np.ascontiguousarray(np.stack(arrays))
    fn = getattr(ctypes.CDLL(TESTLIB), "testface_conv_ab")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t,
                   ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    L = (ctypes.c_size_t * 3)(*(list(shape) + [0] * (3 - len(shape))))
    assert fn(len(shape), L, A, B, code, buf.ctypes.data) == 0

    def op(Fs, x):
        P = list(Fs)
        if code == 1:
            P[0] = Fs[0] * Fs[1]
        elif code == 3:
            P[0] = Fs[0] * np.conj(Fs[1])
        return P[:B]
    want = oracle.explicit_general(arrays[:A] + [np.zeros(shape)] * (B - A if B > A else 0),
                                   tuple(2 * n for n in shape), op)
    for b in range(B):
        assert oracle.rel_l2(buf[b], want[b]) < 1e-12, (shape, A, B, code, b)
    for a in range(B, na):
        assert np.array_equal(buf[a], arrays[a])
