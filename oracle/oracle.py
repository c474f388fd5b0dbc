"""CPU oracle for the hybrid dealiased convolution path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker (or as the timed CPU
port for the baseline), never as part of the product path in fftwpp_amd/.

Three independent families live here:

* ``direct_*``: ctypes bindings of oracle.c's O(N^2) restatement of the
  reference test oracle (tests/direct.h:11-156, tests/direct.cc:5-98).
* ``explicit_*``: explicitly zero-padded FFT convolutions through numpy's
  pocketfft (a library routine).  For M >= the dealiasing bound these equal the
  reference's normalised ``convolve()`` output; for any (L, N=m*q) they give the
  length-N circular convolution that the implicit residue method computes.
* ``hybrid_*``: ctypes bindings of oracle.c's port of the implicitly padded
  residue algorithm (convolve.cc), used as the CPU baseline.

Plus the reference's parameter algebra (fftBase::parameters, index) and the
seeded input generators for the measured workloads.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force=False):
    """Compile oracle.c into liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIBPATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []),
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = ctypes.CDLL(LIBPATH)
        P = ctypes.c_void_p
        Z = ctypes.c_size_t
        Iv = ctypes.c_int
        sigs = {
            "oracle_direct_conv": [P, P, P, Z],
            "oracle_direct_convC": [P, P, P, Z],
            "oracle_direct_convh": [P, P, P, Z],
            "oracle_direct_conv2": [P, P, P, Z, Z, Z],
            "oracle_direct_conv3": [P, P, P, Z, Z, Z, Z, Z],
            "oracle_direct_convh2": [P, P, P, Z, Z, Iv, Z, Iv],
            "oracle_direct_convh3": [P, P, P, Z, Z, Z, Iv, Iv, Z, Z, Iv],
            "oracle_symmetrize_x": [Z, Z, Z, P, Z],
            "oracle_symmetrize_xy": [Z, Z, Z, Z, Z, P, Z, Z],
            "oracle_fft": [P, Z, Iv],
            "oracle_hybrid_conv1": [P, P, Z, Z, Z, Iv],
            "oracle_hybrid_hconv1": [P, P, Z, Z, Z],
            "oracle_hybrid_conv2": [P, P, Z, Z, Z, Z, Z, Z, Iv],
            "oracle_hybrid_conv3": [P, P, Z, Z, Z, Z, Z, Z, Z, Z, Z, Iv],
            "oracle_pad_forward": [P, Z, Z, Z, Iv, Z, P],
            "oracle_set_threads": [Iv],
            "oracle_get_threads": [],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = Iv if name == "oracle_get_threads" else None
        _lib = L
    return _lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Reference parameter algebra (convolve.cc:403-410, convolve.h:297-326)
# --------------------------------------------------------------------------

def ceilquotient(a, b):
    return -(-a // b)


def parameters(L, M, m, centered):
    """fftBase::parameters (convolve.cc:403-410): returns (p, n, q)."""
    p = ceilquotient(L, m)
    P = p // 2 if ((centered and p % 2 == 0) or p == 2) else p
    n = ceilquotient(M, P * m)
    return p, n, P * n


def default_M(L, hermitian=False, A=2):
    """Default padded size of wrappers/HybridConvolution.h:19,39."""
    if hermitian:
        return 3 * ceilquotient(L, 2) - 2 * (L % 2)
    return A * L - A + 1


# --------------------------------------------------------------------------
# Direct (O(N^2)) oracle -- restatement of tests/direct.{h,cc}
# --------------------------------------------------------------------------

def direct_conv(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    h = np.empty_like(f)
    lib().oracle_direct_conv(h.ctypes.data_as(ctypes.c_void_p), pf, pg, f.size)
    return h


def direct_convC(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    h = np.empty_like(f)
    lib().oracle_direct_convC(h.ctypes.data_as(ctypes.c_void_p), pf, pg, f.size)
    return h


def direct_convh(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    h = np.empty_like(f)
    lib().oracle_direct_convh(h.ctypes.data_as(ctypes.c_void_p), pf, pg, f.size)
    return h


def direct_conv2(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    Lx, Ly = f.shape
    h = np.empty_like(f)
    lib().oracle_direct_conv2(h.ctypes.data_as(ctypes.c_void_p), pf, pg, Lx, Ly, Ly)
    return h


def direct_conv3(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    Lx, Ly, Lz = f.shape
    h = np.empty_like(f)
    lib().oracle_direct_conv3(h.ctypes.data_as(ctypes.c_void_p), pf, pg,
                              Lx, Ly, Lz, Ly * Lz, Lz)
    return h


def symmetrize_x(f, Lx=None):
    """HermitianSymmetrizeX on an [Lx][Hy] array in place (the Python wrapper's
    call pattern, wrappers/fftwpp.py:337-343)."""
    assert f.dtype == np.complex128 and f.flags.c_contiguous
    Lx, Hy = f.shape
    Hx = (Lx + 1) // 2
    lib().oracle_symmetrize_x(Hx, Hy, Lx // 2, f.ctypes.data_as(ctypes.c_void_p), Hy)
    return f


def symmetrize_xy(f):
    """HermitianSymmetrizeXY on an [Lx][Ly][Hz] array in place
    (wrappers/fftwpp.py:344-351)."""
    assert f.dtype == np.complex128 and f.flags.c_contiguous
    Lx, Ly, Hz = f.shape
    Hx, Hy = (Lx + 1) // 2, (Ly + 1) // 2
    lib().oracle_symmetrize_xy(Hx, Hy, Hz, Lx // 2, Ly // 2,
                               f.ctypes.data_as(ctypes.c_void_p), Ly * Hz, Hz)
    return f


def direct_convh2(f, g):
    """[Lx][Hy] centred-Hermitian direct convolution of already symmetrised
    inputs (tests/hybridconvh2.cc:65-69 calls it with symmetrize=false)."""
    f, pf = _c(f)
    g, pg = _c(g)
    Lx, Hy = f.shape
    Hx = (Lx + 1) // 2
    h = np.empty_like(f)
    lib().oracle_direct_convh2(h.ctypes.data_as(ctypes.c_void_p), pf, pg,
                               Hx, Hy, Lx % 2, Hy, 0)
    return h


def direct_convh3(f, g):
    f, pf = _c(f)
    g, pg = _c(g)
    Lx, Ly, Hz = f.shape
    Hx, Hy = (Lx + 1) // 2, (Ly + 1) // 2
    h = np.empty_like(f)
    lib().oracle_direct_convh3(h.ctypes.data_as(ctypes.c_void_p), pf, pg,
                               Hx, Hy, Hz, Lx % 2, Ly % 2, Ly * Hz, Hz, 0)
    return h


def fft(x, sign):
    x, px = _c(np.array(x, dtype=np.complex128))
    lib().oracle_fft(px, x.size, sign)
    return x


# --------------------------------------------------------------------------
# Hybrid (implicit residue) port -- CPU baseline and second oracle
# --------------------------------------------------------------------------

def hybrid_conv1(f, g, m, q, centered=False):
    f, pf = _c(np.array(f))
    g, pg = _c(np.array(g))
    lib().oracle_hybrid_conv1(pf, pg, f.size, m, q, int(centered))
    return f


def hybrid_hconv1(f, g, L, m, q):
    f, pf = _c(np.array(f))
    g, pg = _c(np.array(g))
    lib().oracle_hybrid_hconv1(pf, pg, L, m, q)
    return f


def hybrid_conv2(f, g, Ly, mx, qx, my, qy, hermitian=False):
    f, pf = _c(np.array(f))
    g, pg = _c(np.array(g))
    lib().oracle_hybrid_conv2(pf, pg, f.shape[0], Ly, mx, qx, my, qy, int(hermitian))
    return f


def hybrid_conv3(f, g, Lz, mx, qx, my, qy, mz, qz, hermitian=False):
    f, pf = _c(np.array(f))
    g, pg = _c(np.array(g))
    Lx, Ly = f.shape[0], f.shape[1]
    lib().oracle_hybrid_conv3(pf, pg, Lx, Ly, Lz, mx, qx, my, qy, mz, qz, int(hermitian))
    return f


def pad_forward(f, m, q, r, centered=False):
    f, pf = _c(f)
    out = np.empty(m, dtype=np.complex128)
    lib().oracle_pad_forward(pf, f.size, m, q, int(centered), r,
                             out.ctypes.data_as(ctypes.c_void_p))
    return out


# --------------------------------------------------------------------------
# Explicit zero padding via numpy FFT (library routine, independent method)
# --------------------------------------------------------------------------

def _offsets(L, centered):
    return -(L // 2) if centered else 0


def explicit_conv(f, g, N, centered=False, workers=None):
    """Normalised length-N circular convolution of zero-padded (optionally
    centred) complex data, truncated back to the input shape.  ``N`` is an int
    or a tuple with one padded length per dimension.  ``workers`` (an int)
    runs the transforms through scipy.fft's multithreaded pocketfft in place
    (the same library routine as numpy.fft), for full-size arrays."""
    f = np.asarray(f, dtype=np.complex128)
    g = np.asarray(g, dtype=np.complex128)
    d = f.ndim
    Ns = (N,) * d if np.isscalar(N) else tuple(N)
    cs = (centered,) * d if isinstance(centered, bool) else tuple(centered)
    idx = []
    for ax in range(d):
        L = f.shape[ax]
        k = np.arange(L) + _offsets(L, cs[ax])
        idx.append(np.mod(k, Ns[ax]))
    F = np.zeros(Ns, dtype=np.complex128)
    G = np.zeros(Ns, dtype=np.complex128)
    ix = np.ix_(*idx)
    F[ix] = f
    G[ix] = g
    if workers is None:
        H = np.fft.ifftn(np.fft.fftn(F) * np.fft.fftn(G))
        return H[ix]
    import scipy.fft as sf
    F = sf.fftn(F, overwrite_x=True, workers=workers)
    G = sf.fftn(G, overwrite_x=True, workers=workers)
    F *= G
    del G
    F = sf.ifftn(F, overwrite_x=True, workers=workers)
    return F[ix]


def explicit_general(arrays, N, op, centered=False):
    """A convolution with a user multiplier by explicit padding (Convolution
    with A inputs and B outputs, the multiplier applied pointwise to the padded
    transforms, convolve.h:1120-1135).  Each input is zero-padded (optionally
    centred) to N and transformed with sign +1, F[x] = sum_k f_k
    exp(2 pi i k x / N) per dimension (the reference's forward sign,
    convolve.h:297-326 gives the padded index x of each transformed value);
    ``op(Fs, x)`` maps the list of A transforms and the tuple of grid-point
    arrays x[d] (broadcastable against F) to the list of B products, each
    transformed back with sign -1, divided by prod(N) and truncated to the
    input shape."""
    arrays = [np.asarray(a, dtype=np.complex128) for a in arrays]
    d = arrays[0].ndim
    Ns = (N,) * d if np.isscalar(N) else tuple(N)
    cs = (centered,) * d if isinstance(centered, bool) else tuple(centered)
    idx = []
    for ax in range(d):
        L = arrays[0].shape[ax]
        idx.append(np.mod(np.arange(L) + _offsets(L, cs[ax]), Ns[ax]))
    ix = np.ix_(*idx)
    total = float(np.prod(Ns))
    Fs = []
    for a in arrays:
        P = np.zeros(Ns, dtype=np.complex128)
        P[ix] = a
        Fs.append(np.fft.ifftn(P) * total)
    x = np.ix_(*[np.arange(n) for n in Ns])
    return [np.fft.fftn(G)[ix] / total for G in op(Fs, x)]


def expand_hermitian(f, compact_axes_even):
    """Return (full, kranges) for a centred Hermitian array whose last axis
    stores the non-negative modes.  Leading axes are centred (index i <-> k =
    i - floor(L/2)); the last axis is extended to k in (-H, H) by
    f(-k) = conj f(k).  Entries whose mirror lies outside the stored range
    (the Nyquist row of an even leading length) extend as zero."""
    f = np.asarray(f, dtype=np.complex128)
    lead = f.shape[:-1]
    H = f.shape[-1]
    full = np.zeros(lead + (2 * H - 1,), dtype=np.complex128)
    full[..., H - 1:] = f
    # negative last-axis modes
    neg = np.conj(f[..., 1:][..., ::-1])  # k=-(H-1)..-1 from +(H-1)..1
    # mirror the leading axes: k -> -k
    for ax, L in enumerate(lead):
        off = L // 2
        k = np.arange(L) - off
        mk = -k
        valid = (mk >= -off) & (mk < L - off)
        src = np.where(valid, mk + off, 0)
        neg = np.take(neg, src, axis=ax)
        shape = [1] * neg.ndim
        shape[ax] = L
        neg = neg * valid.reshape(shape)
    full[..., :H - 1] = neg
    return full


def explicit_hconv(f, g, N, workers=None):
    """Centred Hermitian convolution by explicit padding: the stored last axis
    holds H non-negative modes, leading axes are centred.  ``N``: padded length
    per dimension (last entry for the Hermitian axis).  ``workers`` as for
    explicit_conv."""
    f = np.asarray(f, dtype=np.complex128)
    d = f.ndim
    Ns = (N,) * d if np.isscalar(N) else tuple(N)
    F = expand_hermitian(f, None)
    G = expand_hermitian(g, None)
    H = f.shape[-1]
    idx = []
    for ax in range(d - 1):
        L = f.shape[ax]
        idx.append(np.mod(np.arange(L) - L // 2, Ns[ax]))
    idx.append(np.mod(np.arange(2 * H - 1) - (H - 1), Ns[-1]))
    FF = np.zeros(Ns, dtype=np.complex128)
    GG = np.zeros(Ns, dtype=np.complex128)
    ix = np.ix_(*idx)
    FF[ix] = F
    GG[ix] = G
    del F, G
    out_idx = np.ix_(*(idx[:-1] + [np.arange(H)]))
    if workers is None:
        HH = np.fft.ifftn(np.fft.fftn(FF) * np.fft.fftn(GG))
        return HH[out_idx]
    import scipy.fft as sf
    FF = sf.fftn(FF, overwrite_x=True, workers=workers)
    GG = sf.fftn(GG, overwrite_x=True, workers=workers)
    FF *= GG
    del GG
    FF = sf.ifftn(FF, overwrite_x=True, workers=workers)
    return FF[out_idx]


def hplane_of(lengths):
    """Self-conjugate plane (Hx, Hy, x0, y0, Sx, Sy) of the stored array of a
    centred Hermitian convolution with full lengths ``lengths``: k = 0 in 1D,
    the ky = 0 line in 2D, the kz = 0 plane in 3D (odd centred lengths)."""
    d = len(lengths)
    if d == 1:
        return (1, 1, 0, 0, 1, 1)
    Lx, Ly = lengths[0], lengths[1]
    if d == 2:
        return ((Lx + 1) // 2, 1, Lx // 2, 0, (Ly + 1) // 2, 1)
    Hz = (lengths[2] + 1) // 2
    return ((Lx + 1) // 2, (Ly + 1) // 2, Lx // 2, Ly // 2, Ly * Hz, Hz)


def hplane(a, op, plane):
    """Operations on the self-conjugate plane (element (i,j) at
    (x0+i) Sx + (y0+j) Sy pairs with (-i,-j); upper = j > 0 or j = 0, i > 0):
    op 0 symmetrize (lower <- conj upper, origin real, as HermitianSymmetrizeX/XY,
    convolve.h:1177-1256), op 1 w + conj w(-k) on both members and 2 Re at the
    origin, op 2 keep upper, zero lower, halve the origin (adjoint of op 0)."""
    Hx, Hy, x0, y0, Sx, Sy = plane
    f = np.array(a, dtype=np.complex128, copy=True).reshape(-1)
    for i in range(-(Hx - 1), Hx):
        for j in range(Hy):
            if j == 0 and i < 0:
                continue
            u = (x0 + i) * Sx + (y0 + j) * Sy
            if i == 0 and j == 0:
                v = f[u]
                f[u] = v.real if op == 0 else (2 * v.real if op == 1 else v / 2)
                continue
            lo = (x0 - i) * Sx + (y0 - j) * Sy
            if op == 0:
                f[lo] = np.conj(f[u])
            elif op == 1:
                a_, b_ = f[u], f[lo]
                f[u] = a_ + np.conj(b_)
                f[lo] = b_ + np.conj(a_)
            else:
                f[lo] = 0
    return f.reshape(np.shape(a))


def hconv_adjoint(w, f, g, lengths):
    """Cotangents (d/d conj f, d/d conj g) of h = C(S f, S g) for a real loss
    with dL/d conj h = w, C the centred Hermitian convolution of stored modes
    (direct sums) and S the symmetrizer of the self-conjugate plane:
    grad_f = P(C(w_H, S g)), grad_g = P(C(w_H, S f)) with w_H = hplane op 1 and
    P = hplane op 2 (the Hermitian G makes the correlation with G a
    convolution with G)."""
    plane = hplane_of(lengths)
    conv = {1: direct_convh, 2: direct_convh2, 3: direct_convh3}[len(lengths)]
    wh = hplane(w, 1, plane)
    gf = hplane(conv(wh, hplane(g, 0, plane)), 2, plane)
    gg = hplane(conv(wh, hplane(f, 0, plane)), 2, plane)
    return gf, gg


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


# --------------------------------------------------------------------------
# Reference test inputs (integer ramps) and seeded workload generators
# --------------------------------------------------------------------------

def ramp1(L, a):
    """tests/hybridconv.cc:60: Complex(j+a+1,(a+1)*j+3)."""
    j = np.arange(L)
    return (j + a + 1) + 1j * ((a + 1) * j + 3)


def ramp2(Lx, Ly, a):
    """tests/hybridconv2.cc input pattern: Complex(i+a+1, (a+1)*j+3)... the
    2D variant used by the reference's 2D test programs."""
    i = np.arange(Lx)[:, None]
    j = np.arange(Ly)[None, :]
    return (i + a + 1) + 1j * ((a + 1) * j + 3) + 0 * (i + j)


def ramp3(Lx, Ly, Lz, a):
    """tests/hybridconv3.cc:56-66: Complex(i+a*k+1,(a+1)*j+3+k)."""
    i = np.arange(Lx)[:, None, None]
    j = np.arange(Ly)[None, :, None]
    k = np.arange(Lz)[None, None, :]
    return (i + a * k + 1) + 1j * ((a + 1) * j + 3 + k) + 0 * (i + j + k)


def ramph2(Lx, Hy, a):
    """tests/hybridconvh2.cc:54-63 input, then HermitianSymmetrizeX."""
    i = np.arange(Lx)[:, None]
    I = i if Lx % 2 else i - 1
    j = np.arange(Hy)[None, :]
    f = np.ascontiguousarray((I + a + 1) + 1j * ((a + 1) * j + 3) + 0 * (i + j),
                             dtype=np.complex128)
    return symmetrize_x(f)


def ramph3(Lx, Ly, Hz, a):
    """tests/hybridconvh3.cc:63-76 input, then HermitianSymmetrizeXY."""
    i = np.arange(Lx)[:, None, None]
    I = i if Lx % 2 else i - 1
    j = np.arange(Ly)[None, :, None]
    J = j if Ly % 2 else j - 1
    k = np.arange(Hz)[None, None, :]
    f = np.ascontiguousarray((I + a * k + 1) + 1j * ((a + 1) * J + 3 + k)
                             + 0 * (i + j + k), dtype=np.complex128)
    return symmetrize_xy(f)


def rampH1(H, a):
    """tests/hybridconvh.cc: Complex(j+a+1,(a+1)*j+3), then HermitianSymmetrize."""
    f = ramp1(H, a).astype(np.complex128)
    f[0] = f[0].real
    return f


SEED = 20261202


def _uc(rng, shape):
    re = rng.uniform(-1.0, 1.0, size=shape)
    im = rng.uniform(-1.0, 1.0, size=shape)
    return re + 1j * im


def inputs_1d(L=1024, seed=SEED):
    rng = np.random.Generator(np.random.PCG64(seed))
    return _uc(rng, L), _uc(rng, L)


def inputs_3d(L=256, seed=SEED):
    rng = np.random.Generator(np.random.PCG64(seed))
    return _uc(rng, (L, L, L)), _uc(rng, (L, L, L))


def _spectral(rng, Lx, Hy):
    kx = (np.arange(Lx) - Lx // 2)[:, None]
    ky = np.arange(Hy)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.where(k > 0, 1.0 / np.where(k > 0, k, 1.0), 0.0)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=(Lx, Hy))
    return np.ascontiguousarray(amp * np.exp(1j * phase), dtype=np.complex128)


def inputs_h2(Lx=2047, Hy=1024, n=2, seed=SEED):
    """|k|^-1 amplitude, uniform phase, HermitianSymmetrizeX applied."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        out.append(symmetrize_x(_spectral(rng, Lx, Hy)))
    return out


def batched_lengths(count=4096, lo=64, hi=16384, seed=SEED):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(np.int64)
