// fftface.cc -- the C++ face of the standalone FFTs (include/fftw++.h) on
// host arrays, as a user of FFTW++'s fftw++.h calls it (examples/example0.cc,
// example0r.cc, example1.cc): exported for the tests through ctypes.
#include <cstring>

#include "fftw++.h"

using namespace fftwpp;

extern "C" {

// cout = FFT_{-1}(cin); cback = normalised FFT_{+1}(cout) in place (= cin);
// rout = rcfft1d(rin); rback = normalised crfft1d(rout) (= rin).
// Returns 0, or 1 when the library threw.
int testfft_cxx(size_t n, double *cin, double *cout, double *cback, double *rin,
                double *rout, double *rback)
{
  try {
    fft1d Forward(n,-1);
    Forward.fft((Complex *) cin,(Complex *) cout);
    memcpy(cback,cout,n*sizeof(Complex));
    fft1d Backward(n,1);
    Backward.fftNormalized((Complex *) cback);
    rcfft1d R(n);
    R.fft(rin,(Complex *) rout);
    crfft1d C(n);
    C.fftNormalized((Complex *) rout,rback);
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

// out = fft2d / fft3d (sign -1) of in, or of out in place when in is NULL.
int testfft_multi_cxx(int dims, const size_t *n, double *in, double *out)
{
  try {
    if(dims == 2) {
      fft2d F(n[0],n[1],-1);
      if(in) F.fft((Complex *) in,(Complex *) out);
      else F.fft((Complex *) out);
    } else {
      fft3d F(n[0],n[1],n[2],-1);
      if(in) F.fft((Complex *) in,(Complex *) out);
      else F.fft((Complex *) out);
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

// rcfft2d / rcfft3d of the reals `in` (n[0] x ... x n[dims-1]) into `out`
// (... x (n[dims-1]/2+1) complex), then crfft2d / crfft3d normalised of out
// into `back`.
int testfft_real_multi_cxx(int dims, const size_t *n, double *in, double *out, double *back)
{
  try {
    if(dims == 2) {
      rcfft2d F(n[0],n[1],in,(Complex *) out);
      F.fft(in,(Complex *) out);
      crfft2d B(n[0],n[1],(Complex *) out,back);
      B.fftNormalized((Complex *) out,back);
    } else {
      rcfft3d F(n[0],n[1],n[2],in,(Complex *) out);
      F.fft(in,(Complex *) out);
      crfft3d B(n[0],n[1],n[2],(Complex *) out,back);
      B.fftNormalized((Complex *) out,back);
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

// In place on a host array of n[0] x ... x 2 (n[dims-1]/2+1) reals (FFTW's
// padded rows) holding the input reals in the first n[dims-1] of each row:
// rcfft{2,3}d in place, crfft{2,3}d in place (unnormalised), then Normalize
// over the padded array (fftw++.h:357-370), which leaves the input reals.
int testfft_inplace_normalize_cxx(int dims, const size_t *n, double *data)
{
  try {
    if(dims == 2) {
      rcfft2d F(n[0],n[1],(Complex *) data);
      F.fft(data,(Complex *) data);
      crfft2d B(n[0],n[1],(Complex *) data);
      B.fft((Complex *) data,data);
      B.Normalize(data);
    } else {
      rcfft3d F(n[0],n[1],n[2],(Complex *) data);
      F.fft(data,(Complex *) data);
      crfft3d B(n[0],n[1],n[2],(Complex *) data);
      B.fft((Complex *) data,data);
      B.Normalize(data);
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

// Origin-shifted real transforms on host arrays: out = rcfft{2,3}d fft0 of
// in; back = crfft{2,3}d fft0Normalized of out (= in); nyq = deNyquist of a
// copy of out.
int testfft0_cxx(int dims, const size_t *n, double *in, double *out, double *back, double *nyq)
{
  try {
    const size_t words=(dims == 2 ? n[0] : n[0]*n[1])*(n[dims-1]/2+1);
    if(dims == 2) {
      rcfft2d F(n[0],n[1],in,(Complex *) out);
      F.fft0(in,(Complex *) out);
      crfft2d B(n[0],n[1],(Complex *) out,back);
      B.fft0Normalized((Complex *) out,back);
      memcpy(nyq,out,words*sizeof(Complex));
      F.deNyquist((Complex *) nyq);
    } else {
      rcfft3d F(n[0],n[1],n[2],in,(Complex *) out);
      F.fft0(in,(Complex *) out);
      crfft3d B(n[0],n[1],n[2],(Complex *) out,back);
      B.fft0Normalized((Complex *) out,back);
      memcpy(nyq,out,words*sizeof(Complex));
      F.deNyquist((Complex *) nyq);
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

// Transpose of a rows x cols matrix of Complex (length 1) or of 3-double
// elements: out-of-place from in to out, then in place on out (back to in).
int testtranspose_cxx(size_t rows, size_t cols, int complexElements, double *in, double *out)
{
  try {
    if(complexElements) {
      Transpose T(rows,cols,1,(Complex *) in,(Complex *) out);
      T.transpose((Complex *) in,(Complex *) out);
      Transpose B(cols,rows,1,(Complex *) out);
      B.transpose((Complex *) out);
    } else {
      Transpose T(rows,cols,3,in,out);
      T.transpose(in,out);
      Transpose B(cols,rows,3,out);
      B.transpose(out);
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

} // extern "C"
