// residues.cc -- the reference's transform tests tests/hybrid.cc:60-167
// (fftPad, fftPadCentered) and tests/hybridh.cc:56-131 (fftPadHermitian)
// through the C++ classes of convolve.h on device arrays: the residue blocks
// of forward() against the explicit padded transform re-indexed by
// index(r,k) ("Forward Error"), and backward() of the explicit values over
// every block against N times the input ("Backward Error").  Exported for
// the tests through ctypes.
#include <cmath>
#include <vector>

#include "convolve.h"

using namespace fftwpp;

extern "C" void fftwpp_amd_internal_set_error(const char *msg);

namespace {

struct Dev {
  void *p=NULL;
  size_t bytes=0;
  explicit Dev(size_t n) : bytes(n*sizeof(Complex)) {
    if(fftwpp_amd_malloc(&p,bytes)) throw Error(FFTWPP_AMD_ERR_HIP,"malloc");
    fftwpp_amd_memset(p,bytes,NULL);
  }
  ~Dev() { fftwpp_amd_free(p); }
  Complex *c() { return (Complex *) p; }
  void put(const std::vector<Complex> &h) { fftwpp_amd_memcpy(p,h.data(),h.size()*sizeof(Complex),NULL); sync(); }
  std::vector<Complex> get(size_t n) {
    std::vector<Complex> h(n);
    fftwpp_amd_memcpy(h.data(),p,n*sizeof(Complex),NULL);
    sync();
    return h;
  }
  static void sync() { fftwpp_amd_synchronize(NULL); }
};

double sq(Complex z) { return z.re*z.re+z.im*z.im; }

}

extern "C" {

// kind 0 fftPad, 1 fftPadCentered, 2 fftPadHermitian, with forced m, D and
// in-place flag I.  errs[0] = Forward Error, errs[1] = Backward Error,
// info = {m, p, q, n, D, R, overwrite}.  Returns 0, 1 when the parameters
// are rejected (the constructor threw), 2 on another failure.
int testface_residues(int kind, size_t L, size_t M, size_t C, size_t S, size_t m,
                      size_t D, int I, double *errs, size_t *info)
{
  fftBase *fft=NULL, *fft2=NULL;
  Application app(1,1,multNone,1,false,m,D,I);
  try {
    if(kind == 2) fft=new fftPadHermitian(L,M,app,C,m,D,I != 0);
    else if(kind == 1) fft=new fftPadCentered(L,M,app,C,S,m,D,I != 0);
    else fft=new fftPad(L,M,app,C,S,m,D,I != 0);
  } catch(const std::exception &) {
    return 1;
  }
  int rc=0;
  try {
    if(S == 0) S=C;
    if(kind == 2) S=C;
    const size_t Mp=fft->M;
    if(kind == 2) fft2=new fftPadHermitian(L,Mp,app,C,Mp,1,true);
    else if(kind == 1) fft2=new fftPadCentered(L,Mp,app,C,S,Mp,1,true);
    else fft2=new fftPad(L,Mp,app,C,S,Mp,1,true);
    const bool herm=kind == 2;
    const size_t H=fft->inputLength();
    const size_t nin=herm ? C*H : S*L;
    std::vector<Complex> f(nin,Complex(0,0));
    if(herm) {
      for(size_t c=0; c < C; ++c) f[c]=Complex(1.0+c,0.0);
      for(size_t j=1; j < H; ++j)
        for(size_t c=0; c < C; ++c) f[C*j+c]=Complex(C*j+c+1,C*j+c+2);
    } else {
      for(size_t j=0; j < L; ++j)
        for(size_t c=0; c < C; ++c) f[S*j+c]=Complex(C*j+c+1,C*j+c+2);
    }
    Dev df(nin), dh(nin), dF(fft->outputSize()), dF2(fft2->outputSize());
    df.put(f);
    fft2->forward(df.c(),dF2.c());
    std::vector<Complex> F2=dF2.get(fft2->outputSize());
    const double *F2r=(const double *) F2.data();
    const bool Overwrite=fft->overwrite;
    double error=0, norm=0, error2=0, norm2=0;
    std::vector<Complex> hh;
    if(Overwrite) {
      dh.put(f);
      fft->forward(dh.c(),dF.c(),0);
      hh=dh.get(nin);
    }
    const size_t stride=fft->noutputs(0);
    for(size_t r=0; r < fft->R; r += fft->increment(r)) {
      if(!Overwrite) fft->forward(df.c(),dF.c(),r);
      std::vector<Complex> F=dF.get(fft->outputSize());
      if(herm) {
        const size_t D1=r == 0 ? fft->D0 : fft->D;
        for(size_t d=0; d < D1; ++d) {
          const double *Fr=(const double *) (F.data()+fft->b*d);
          for(size_t k=0; k < stride; ++k) {
            const size_t i=fft->index(r,k+stride*d);
            for(size_t c=0; c < C; ++c) {
              const double val=F2r[C*i+c], d2=Fr[C*k+c]-val;
              error += d2*d2;
              norm += val*val;
            }
          }
        }
      } else {
        for(size_t k=0; k < fft->noutputs(r); ++k) {
          const size_t i=fft->index(r,k);
          for(size_t c=0; c < C; ++c) {
            const Complex val=F2[S*i+c];
            const Complex out=!Overwrite || r == fft->n-1 ? F[S*k+c] :
              (r == 0 ? hh[S*k+c] : hh[S*k+c+fft->b]);
            error += sq(Complex(out.re-val.re,out.im-val.im));
            norm += sq(val);
          }
        }
      }
    }
    // backward of the explicit values, block by block
    std::vector<Complex> F(fft->outputSize(),Complex(0,0));
    std::vector<Complex> hb(nin,Complex(0,0));
    for(size_t r=0; r < fft->R; r += fft->increment(r)) {
      if(herm) {
        const size_t D1=r == 0 ? fft->D0 : fft->D;
        for(size_t d=0; d < D1; ++d) {
          double *Fr=(double *) (F.data()+fft->b*d);
          for(size_t k=0; k < stride; ++k) {
            const size_t i=fft->index(r,k+stride*d);
            for(size_t c=0; c < C; ++c) Fr[C*k+c]=F2r[C*i+c];
          }
        }
      } else {
        for(size_t k=0; k < fft->noutputs(r); ++k) {
          const size_t i=fft->index(r,k);
          for(size_t c=0; c < C; ++c) {
            if(!Overwrite || r == fft->n-1) F[S*k+c]=F2[S*i+c];
            else if(r == 0) hb[S*k+c]=F2[S*i+c];
            else hb[S*k+c+fft->b]=F2[S*i+c];
          }
        }
      }
      if(!Overwrite) {
        dF.put(F);
        fft->backward(dF.c(),dh.c(),r);
      }
    }
    if(Overwrite) {
      dF.put(F);
      dh.put(hb);
      fft->backward(dF.c(),dh.c(),0);
    }
    std::vector<Complex> h=dh.get(nin);
    const double scale=1.0/fft->normalization();
    for(size_t j=0; j < (herm ? H : L); ++j)
      for(size_t c=0; c < C; ++c) {
        const size_t J=herm ? C*j+c : S*j+c;
        error2 += sq(Complex(h[J].re*scale-f[J].re,h[J].im*scale-f[J].im));
        norm2 += sq(f[J]);
      }
    errs[0]=norm > 0 ? std::sqrt(error/norm) : error;
    errs[1]=norm2 > 0 ? std::sqrt(error2/norm2) : error2;
    const size_t v[7]={fft->m,fft->p,fft->q,fft->n,fft->D,fft->R,(size_t) fft->overwrite};
    for(int i=0; i < 7; ++i) info[i]=v[i];
  } catch(const std::exception &e) {
    fftwpp_amd_internal_set_error(e.what());
    rc=2;
  }
  delete fft2;
  delete fft;
  return rc;
}

} // extern "C"
