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

} // extern "C"
