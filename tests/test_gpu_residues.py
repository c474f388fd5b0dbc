"""The C++ residue-transform face (fftPad, fftPadCentered, fftPadHermitian
forward/backward in the reference's output layout, convolve.h:297-326) on
the reference's transform test grids (tests/tests.py:468-613, the programs
tests/hybrid.cc and tests/hybridh.cc): for every (L, M, m, C, S, D, I) the
residue blocks of forward() must equal the explicit padded transform at
index(r, k) ("Forward Error") and backward() of the explicit values over all
blocks must return N times the input ("Backward Error"), both <= 1e-12.
Parameter combinations the reference rejects must be rejected."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

TESTLIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "support", "libtestmult.so")


def cq(a, b):
    return -(-a // b)


def pqn(kind, L, M, m):
    # fftBase::parameters (convolve.cc:403-410)
    p = cq(L, m)
    P = p // 2 if ((kind != 0 and p % 2 == 0) or p == 2) else p
    n = cq(M, m * P)
    return p, n, P * n


def valid(kind, m, p, q, n, D, C, S):
    # fftPad::valid, fftPadCentered::Opt::valid, fftPadHermitian::Opt::valid
    # (convolve.h:482-485, 605-607, 721-723)
    if kind == 2:
        return (D == 1 and q == 1) or (D == 2 and p % 2 == 0 and (p == 2 or C == 1))
    ok = D == 1 if q == 1 else (D == 1 or (S == 1 and ((D < n and D % 2 == 0) or D == n)))
    return ok and (kind == 0 or q == 1 or p % 2 == 0)


def get_ds(start, stop):
    out = [start] + list(range(start + 2 - start % 2, stop, 2))
    if stop > start:
        out.append(stop)
    return out


def collect(kind, L, M, m, C):
    p, n, q = pqn(kind, L, M, m)
    if q == 1:
        ds = [1]
    elif C > 1:
        ds = [2 if kind == 2 else 1]
    elif kind == 2:
        ds = [2]
    else:
        ds = get_ds(1, max(n, 1))
    for I in (0, 1):
        for D in ds:
            yield (L, M, m, D, I)


def grid(kind):
    cases = []
    for C, S in ((1, 1), (2, 2), (2, 3)):
        if kind == 2 and S != C:
            continue
        if kind == 0:
            L = 8
            for M in (2 * L, cq(5 * L, 2)):
                for m in (M, L + 1, L, cq(L, 2), cq(L, 4)):
                    cases += [(C, S) + c for c in collect(kind, L, M, m, C)]
        elif kind == 1:
            for L in (8, 7):
                L2 = cq(L, 2)
                for M in (3 * L2 - 2 * (L % 2), 3 * L2 - 2 * (L % 2) + 1, 2 * L, 5 * L2):
                    for m in (cq(L, 4), L2, cq(L2 + L, 2), M):
                        cases += [(C, S) + c for c in collect(kind, L, M, m, C)]
        else:
            for L in (8, 7):
                L2 = cq(L, 2)
                for M in (3 * L2 - 2 * (L % 2), 2 * L, 5 * L2):
                    ms = (M, L2, cq(L, 4)) if C == 1 else (L2,)
                    for m in ms:
                        cases += [(C, S) + c for c in collect(kind, L, M, m, C)]
    return sorted(set(cases))


CASES = [(k,) + c for k in (0, 1, 2) for c in grid(k)]


@pytest.mark.parametrize("kind,C,S,L,M,m,D,I", CASES)
def test_reference_transform_grid(gpu, kind, C, S, L, M, m, D, I):
    lib = ctypes.CDLL(TESTLIB)
    lib.testface_residues.argtypes = [ctypes.c_int] + [ctypes.c_size_t] * 6 + \
        [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_size_t)]
    errs = (ctypes.c_double * 2)()
    info = (ctypes.c_size_t * 7)()
    rc = lib.testface_residues(kind, L, M, C, S, m, D, I, errs, info)
    p, n, q = pqn(kind, L, M, m)
    ok = q == 1 and D == 1 or valid(kind, m, p, q, n, D, C, S)
    if not ok:
        assert rc == 1, "the reference rejects these parameters"
        return
    if rc == 2:
        from fftwpp_amd._lib import lib as flib
        pytest.fail(flib().fftwpp_amd_last_error().decode())
    assert rc == 0
    assert (info[0], info[2]) == (m, q)
    assert errs[0] < 1e-12, ("Forward Error", errs[0], list(info))
    assert errs[1] < 1e-12, ("Backward Error", errs[1], list(info))
