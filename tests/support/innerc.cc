// innerc.cc -- 2D and 3D convolutions through the C++ classes of convolve.h
// whose innermost transform runs outside the fused row kernels: C > 1
// interleaved columns (the reference's fftPad(L, M, app, C, C), a row of C
// sequences with element stride C), or a forced inner m that no row engine
// takes.  On host arrays (the classes stage them), exported for the GPU tests
// through ctypes.
#include <cstring>
#include <vector>

#include "convolve.h"

using namespace fftwpp;

extern "C" {

// f, g: Lx x Ly x C complex words (interleaved doubles), x stride Ly C; the
// normalised convolution over (x, y) of every column c is written to h.
// my = 0 lets the library choose the inner m.  Returns 0, or the error code.
int testface_conv2_inner(size_t Lx, size_t Ly, size_t C, size_t my, const double *f,
                         const double *g, double *h)
{
  try {
    const size_t n=Lx*Ly*C;
    Complex *a=utils::ComplexAlign(n), *b=utils::ComplexAlign(n);
    memcpy(a,f,n*sizeof(Complex));
    memcpy(b,g,n*sizeof(Complex));
    Application appx(2,1,multNone,1);
    fftPad fftx(Lx,2*Lx-1,appx,Ly*C,Ly*C);
    Application appy(2,1,multBinary,appx,my);
    fftPad *ffty=my ? new fftPad(Ly,2*Ly-1,appy,C,C,my,1,false) :
      new fftPad(Ly,2*Ly-1,appy,C,C);
    {
      Convolution2 conv(&fftx,ffty);
      Complex *F[2]={a,b};
      conv.convolve(F);
    }
    delete ffty;
    memcpy(h,a,n*sizeof(Complex));
    utils::deleteAlign(a);
    utils::deleteAlign(b);
  } catch(const Error &e) {
    return e.code;
  }
  return 0;
}

// The same in 3D: Lx x Ly x Lz x C words, contiguous (y stride Lz C, x stride
// Ly Lz C), the inner z transform over C interleaved columns with forced
// mz (0: chosen).
int testface_conv3_inner(size_t Lx, size_t Ly, size_t Lz, size_t C, size_t mz,
                         const double *f, const double *g, double *h)
{
  try {
    const size_t n=Lx*Ly*Lz*C;
    Complex *a=utils::ComplexAlign(n), *b=utils::ComplexAlign(n);
    memcpy(a,f,n*sizeof(Complex));
    memcpy(b,g,n*sizeof(Complex));
    Application appx(2,1,multNone,1);
    fftPad fftx(Lx,2*Lx-1,appx,Ly*Lz*C,Ly*Lz*C);
    Application appy(2,1,multNone,appx);
    fftPad ffty(Ly,2*Ly-1,appy,Lz*C,Lz*C);
    Application appz(2,1,multBinary,appx,mz);
    fftPad *fftz=mz ? new fftPad(Lz,2*Lz-1,appz,C,C,mz,1,false) :
      new fftPad(Lz,2*Lz-1,appz,C,C);
    {
      Convolution3 conv(&fftx,&ffty,fftz);
      Complex *F[2]={a,b};
      conv.convolve(F);
    }
    delete fftz;
    memcpy(h,a,n*sizeof(Complex));
    utils::deleteAlign(a);
    utils::deleteAlign(b);
  } catch(const Error &e) {
    return e.code;
  }
  return 0;
}

// A library multiplier (code 0 multNone, 1 multBinary, 3 multcorrelation)
// with A inputs and B outputs through the C++ classes: dim = 1..3, lengths
// L[0..dim-1] (contiguous, non-centred, M = 2L - 1 per dimension); arrays
// holds max(A, B) arrays of prod(L) complex words back to back, the first A
// the inputs; the first B are overwritten with the outputs.
int testface_conv_ab(int dim, const size_t *L, size_t A, size_t B, int code, double *arrays)
{
  try {
    multiplier *mult=code == 0 ? multNone : (code == 1 ? multBinary : multcorrelation);
    size_t n=1;
    for(int d=0; d < dim; ++d) n *= L[d];
    const size_t na=A > B ? A : B;
    std::vector<Complex *> F(na);
    for(size_t a=0; a < na; ++a) {
      F[a]=utils::ComplexAlign(n);
      memcpy(F[a],arrays+2*a*n,n*sizeof(Complex));
    }
    if(dim == 1) {
      Application app(A,B,mult,1);
      fftPad fft(L[0],2*L[0]-1,app);
      Convolution conv(&fft);
      conv.convolve(F.data());
    } else if(dim == 2) {
      Application appx(A,B,multNone,1);
      fftPad fftx(L[0],2*L[0]-1,appx,L[1],L[1]);
      Application appy(A,B,mult,appx);
      fftPad ffty(L[1],2*L[1]-1,appy);
      Convolution2 conv(&fftx,&ffty);
      conv.convolve(F.data());
    } else {
      Application appx(A,B,multNone,1);
      fftPad fftx(L[0],2*L[0]-1,appx,L[1]*L[2],L[1]*L[2]);
      Application appy(A,B,multNone,appx);
      fftPad ffty(L[1],2*L[1]-1,appy,L[2],L[2]);
      Application appz(A,B,mult,appx);
      fftPad fftz(L[2],2*L[2]-1,appz);
      Convolution3 conv(&fftx,&ffty,&fftz);
      conv.convolve(F.data());
    }
    for(size_t a=0; a < na; ++a) {
      if(a < B) memcpy(arrays+2*a*n,F[a],n*sizeof(Complex));
      utils::deleteAlign(F[a]);
    }
  } catch(const Error &e) {
    return e.code;
  }
  return 0;
}

} // extern "C"
