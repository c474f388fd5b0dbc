"""The C-ABI library loads and exports every function the C headers declare;
host-only parameter logic of the Python layer.  No GPU work is launched."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fftwpp_amd", "libfftwpp_amd.so")
C_HEADERS = ["fftwpp_amd.h", "cfftw++.h"]


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"([A-Za-z_]\w*)\s*\(([^;{}()]*)\)\s*;", text):
        start = text.rfind("\n", 0, m.start()) + 1
        line = text[start:m.end()]
        if "typedef" in line:
            continue
        names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        from fftwpp_amd import build
        build.build()
    return LIB


@pytest.mark.parametrize("header", C_HEADERS)
def test_header_symbols_exported(built, header):
    names = declared_functions(header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, f"{header}: not exported: {missing}"


def test_library_loads_and_binds(built):
    from fftwpp_amd import _lib
    L = _lib.lib()
    # every binding in the Python table resolves
    for name in _lib.SIGNATURES:
        assert hasattr(L, name)
    # the Python table covers every C declaration
    for header in C_HEADERS:
        for name in declared_functions(header):
            assert name in _lib.SIGNATURES, name


def test_cfftw_symbols_match_reference_python_wrapper():
    # the names wrappers/fftwpp.py:33-82 binds through ctypes
    needed = ["fftwpp_create_conv1d", "fftwpp_conv1d_delete", "fftwpp_conv1d_convolve",
              "fftwpp_create_hconv1d", "fftwpp_hconv1d_delete", "fftwpp_hconv1d_convolve",
              "fftwpp_create_conv2d", "fftwpp_conv2d_delete", "fftwpp_conv2d_convolve",
              "fftwpp_HermitianSymmetrize", "fftwpp_create_hconv2d", "fftwpp_hconv2d_delete",
              "fftwpp_hconv2d_convolve", "fftwpp_HermitianSymmetrizeX", "fftwpp_create_conv3d",
              "fftwpp_conv3d_delete", "fftwpp_conv3d_convolve", "fftwpp_create_hconv3d",
              "fftwpp_hconv3d_delete", "fftwpp_hconv3d_convolve",
              "fftwpp_HermitianSymmetrizeXY", "set_fftwpp_maxthreads",
              "get_fftwpp_maxthreads"]
    declared = declared_functions("cfftw++.h")
    for n in needed:
        assert n in declared, n


def test_maxthreads_roundtrip(built):
    from fftwpp_amd import fftwpp
    old = fftwpp.fftwpp_get_maxthreads()
    fftwpp.fftwpp_set_maxthreads(7)
    assert fftwpp.fftwpp_get_maxthreads() == 7
    fftwpp.fftwpp_set_maxthreads(old)


def test_aligned_allocators(built):
    from fftwpp_amd import _lib
    L = _lib.lib()
    p = L.create_complexAlign(33)
    assert p and p % 16 == 0
    L.delete_complexAlign(p)
    p = L.create_doubleAlign(9)
    assert p and p % 16 == 0
    L.delete_doubleAlign(p)


def test_complex_align_python():
    from fftwpp_amd.fftwpp import complex_align
    a = complex_align((5, 3))
    assert a.shape == (5, 3) and a.ctypes.data % 16 == 0


def test_host_pointers_rejected_by_device_entry(built):
    """fftwpp_amd_conv_execute refuses non-device pointers before any launch
    (the guard runs on the host; no kernel is started)."""
    from fftwpp_amd import _lib
    L = _lib.lib()
    if L.fftwpp_amd_device_count() > 0:
        pytest.skip("checked on GPU-less hosts only")
    h = ctypes.c_void_p()
    rc = L.fftwpp_amd_conv_create(ctypes.byref(h), 1, 0, _lib.sizes([8]), None, None, 0)
    # without a device the twiddle upload fails with a HIP status, not a crash
    assert rc != 0


def test_cxx_headers_compile_and_link(built, tmp_path):
    """A C++ user of convolve.h and fftw++.h (the reference's class names,
    a user device multiplier with the reference signature) compiles with g++
    and links against the library; nothing is run."""
    src = tmp_path / "user.cc"
    src.write_text(r'''
#include "convolve.h"
#include "fftw++.h"
using namespace fftwpp;
static void mult(Complex **F, size_t n, Indices *indices, size_t) {
  (void) F; (void) n; (void) indices->layout; (void) indices->stream;
}
int run(size_t L, Complex **f, Complex *g) {
  Application app(3,1,mult);
  fftPad fft(L,3*L-2,app);
  Convolution C(&fft);
  C.convolve(f);
  fft1d Forward(L,-1);
  Forward.fft(g);
  mfft1d Many(L,1,4,1,L);
  rcfft1d R(2*L);
  crfft1d Cr(2*L);
  fft2d F2(L,L,-1);
  fft3d F3(L,L,L,1);
  rcfft2d R2(L,L,(double *) g);
  crfft3d C3(L,L,L,g);
  R2.fft0((double *) g,g);
  C3.deNyquist(g);
  Transpose T(L,L,1,g);
  T.transpose(g);
  Forward.Normalize(g);
  Cr.Normalize((double *) g);
  return (int) (R.normalization()+Cr.normalization()+F2.normalization()+F3.normalization()+
                 Many.normalization()+R2.normalization()+C3.normalization());
}
int main() { return 0; }
''')
    exe = tmp_path / "user"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", os.path.dirname(LIB), "-lfftwpp_amd", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_fft_and_general_plan_argument_checks(built):
    """Host-side validation of the standalone FFT and user-multiplier plans:
    rejected before any device memory is touched."""
    from fftwpp_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    P = ctypes.c_void_p
    # bad dimension count, kind, sign, dtype
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 0, 0, _lib.sizes([8]), -1, 0, 1, 0, 0, 0, 0) == -1
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 3, 1, _lib.sizes([8]), -1, 0, 1, 0, 0, 0, 0) == -1
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 0, 1, _lib.sizes([8]), 2, 0, 1, 0, 0, 0, 0) == -1
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 0, 1, _lib.sizes([8]), -1, 5, 1, 0, 0, 0, 0) == -1
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 0, 1, _lib.sizes([0]), -1, 0, 1, 0, 0, 0, 0) == -1
    # multidimensional transforms are single transforms
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 1, 2, _lib.sizes([8, 8]), -1, 0, 2, 0, 0, 0, 0) == -3
    # a prime above the largest sub-FFT as an axis of a multidimensional
    # transform (1D lengths take Bluestein's algorithm instead)
    assert L.fftwpp_amd_fft_create(ctypes.byref(h), 0, 2, _lib.sizes([5003, 4]), -1, 0, 1, 0, 0, 0,
                                   0) == -3
    assert h.value is None
    # user-multiplier plans: A, B in [1, 8], a multiplier is required
    dummy = P(1)
    for A, B, mult in [(0, 1, dummy), (2, 9, dummy), (2, 1, None)]:
        rc = L.fftwpp_amd_conv_create_general(ctypes.byref(h), 1, 0, _lib.sizes([16]), None, None,
                                              0, A, B, mult, None)
        assert rc == -1, (A, B)
