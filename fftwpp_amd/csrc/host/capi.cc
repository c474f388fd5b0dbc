// capi.cc -- the drop-in C wrapper (include/cfftw++.h) and the convolution
// layer of include/fftwpp_amd.h, both built on the C++ classes of
// include/convolve.h.
#include "convolve.h"
#include "cfftw++.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>

using namespace fftwpp;

namespace fftwpp {

// The compositions of wrappers/HybridConvolution.h:12-182: default padded
// sizes, outer transforms with multNone and C = product of the inner stored
// lengths, the inner transform carrying the binary multiplier.
struct HybridConvolution {
  Application *app;
  fftPad *fft;
  Convolution *convolve;
  // real: real data (fftPadReal, as tests/hybridconvr.cc builds it)
  HybridConvolution(size_t L, size_t M=0, size_t m=0, multiplier mult=multBinary,
                    bool real=false) {
    if(M == 0) M=2*L-1;
    app=new Application(2,1,mult,fftw::maxthreads,false,m);
    fft=real ? new fftPadReal(L,M,*app) : new fftPad(L,M,*app);
    convolve=new Convolution(fft);
  }
  ~HybridConvolution() { delete convolve; delete fft; delete app; }
};

static size_t hdefault(size_t L) { return 3*utils::ceilquotient(L,2)-2*(L % 2); }

// Centred complex 1D convolution: fftPadCentered + Convolution as in
// tests/hybridconv.cc:44-46 with -c; default M is the 3/2 rule.
struct CenteredConvolution {
  Application *app;
  fftPadCentered *fft;
  Convolution *convolve;
  CenteredConvolution(size_t L, size_t M=0, size_t m=0) {
    if(M == 0) M=hdefault(L);
    app=new Application(2,1,multBinary,fftw::maxthreads,false,m);
    fft=new fftPadCentered(L,M,*app);
    convolve=new Convolution(fft);
  }
  ~CenteredConvolution() { delete convolve; delete fft; delete app; }
};

struct HybridConvolutionHermitian {
  Application *app;
  fftPadHermitian *fft;
  Convolution *convolve;
  HybridConvolutionHermitian(size_t L, size_t M=0, size_t m=0) {
    if(M == 0) M=hdefault(L);
    app=new Application(2,1,realMultBinary,fftw::maxthreads,false,m);
    fft=new fftPadHermitian(L,M,*app);
    convolve=new Convolution(fft);
  }
  ~HybridConvolutionHermitian() { delete convolve; delete fft; delete app; }
};

struct HybridConvolution2 {
  Application *appx, *appy;
  fftPad *fftx, *ffty;
  Convolution2 *convolve2;
  // real: real data, the outer dimension an fftPadReal (tests/hybridconvr2.cc)
  HybridConvolution2(size_t Lx, size_t Ly, size_t Mx=0, size_t My=0,
                     size_t mx=0, size_t my=0, multiplier mult=multBinary,
                     bool real=false) {
    if(Mx == 0) Mx=2*Lx-1;
    if(My == 0) My=2*Ly-1;
    appx=new Application(2,1,multNone,fftw::maxthreads,false,mx);
    fftx=real ? new fftPadReal(Lx,Mx,*appx,Ly) : new fftPad(Lx,Mx,*appx,Ly);
    appy=new Application(2,1,mult,*appx,my);
    ffty=new fftPad(Ly,My,*appy);
    convolve2=new Convolution2(fftx,ffty);
  }
  ~HybridConvolution2() {
    delete convolve2; delete ffty; delete appy; delete fftx; delete appx;
  }
};

struct HybridConvolutionHermitian2 {
  Application *appx, *appy;
  fftPadCentered *fftx;
  fftPadHermitian *ffty;
  Convolution2 *convolve2;
  HybridConvolutionHermitian2(size_t Lx, size_t Ly, size_t Mx=0, size_t My=0,
                              size_t mx=0, size_t my=0) {
    if(Mx == 0) Mx=hdefault(Lx);
    if(My == 0) My=hdefault(Ly);
    size_t Hy=utils::ceilquotient(Ly,2);
    appx=new Application(2,1,multNone,fftw::maxthreads,false,mx);
    fftx=new fftPadCentered(Lx,Mx,*appx,Hy,Hy);
    appy=new Application(2,1,realMultBinary,*appx,my);
    ffty=new fftPadHermitian(Ly,My,*appy);
    convolve2=new Convolution2(fftx,ffty);
  }
  ~HybridConvolutionHermitian2() {
    delete convolve2; delete ffty; delete appy; delete fftx; delete appx;
  }
};

struct HybridConvolution3 {
  Application *appx, *appy, *appz;
  fftPad *fftx, *ffty, *fftz;
  Convolution3 *convolve3;
  // real: real data, the outer dimension an fftPadReal (tests/hybridconvr3.cc)
  HybridConvolution3(size_t Lx, size_t Ly, size_t Lz, size_t Mx=0,
                     size_t My=0, size_t Mz=0, size_t mx=0, size_t my=0,
                     size_t mz=0, multiplier mult=multBinary, bool real=false) {
    if(Mx == 0) Mx=2*Lx-1;
    if(My == 0) My=2*Ly-1;
    if(Mz == 0) Mz=2*Lz-1;
    appx=new Application(2,1,multNone,fftw::maxthreads,false,mx);
    fftx=real ? new fftPadReal(Lx,Mx,*appx,Ly*Lz) : new fftPad(Lx,Mx,*appx,Ly*Lz);
    appy=new Application(2,1,multNone,*appx,my);
    ffty=new fftPad(Ly,My,*appy,Lz);
    appz=new Application(2,1,mult,*appy,mz);
    fftz=new fftPad(Lz,Mz,*appz);
    convolve3=new Convolution3(fftx,ffty,fftz);
  }
  ~HybridConvolution3() {
    delete convolve3; delete fftz; delete appz; delete ffty; delete appy;
    delete fftx; delete appx;
  }
};

struct HybridConvolutionHermitian3 {
  Application *appx, *appy, *appz;
  fftPadCentered *fftx, *ffty;
  fftPadHermitian *fftz;
  Convolution3 *convolve3;
  HybridConvolutionHermitian3(size_t Lx, size_t Ly, size_t Lz, size_t Mx=0,
                              size_t My=0, size_t Mz=0, size_t mx=0,
                              size_t my=0, size_t mz=0) {
    if(Mx == 0) Mx=hdefault(Lx);
    if(My == 0) My=hdefault(Ly);
    if(Mz == 0) Mz=hdefault(Lz);
    size_t Hz=utils::ceilquotient(Lz,2);
    appx=new Application(2,1,multNone,fftw::maxthreads,false,mx);
    fftx=new fftPadCentered(Lx,Mx,*appx,Ly*Hz);
    appy=new Application(2,1,multNone,*appx,my);
    ffty=new fftPadCentered(Ly,My,*appy,Hz);
    appz=new Application(2,1,realMultBinary,*appy,mz);
    fftz=new fftPadHermitian(Lz,Mz,*appz);
    convolve3=new Convolution3(fftx,ffty,fftz);
  }
  ~HybridConvolutionHermitian3() {
    delete convolve3; delete fftz; delete appz; delete ffty; delete appy;
    delete fftx; delete appx;
  }
};

} // namespace fftwpp

static thread_local char capi_error[512];

extern "C" void fftwpp_amd_internal_set_error(const char *msg);

static void report_error(const char *where, const std::exception &e)
{
  snprintf(capi_error,sizeof(capi_error),"%s: %s",where,e.what());
  fftwpp_amd_internal_set_error(capi_error);
  std::cerr << capi_error << std::endl;
}

template<class F>
static auto guard(const char *where, F f) -> decltype(f())
{
  try {
    return f();
  } catch(const std::exception &e) {
    report_error(where,e);
    return decltype(f())();
  }
}

// Device-pointer variants (extensions): status codes instead of messages.
template<typename Conv>
static int convolve_device(const char *where, Conv *conv, void *a, const void *b,
                           void *stream)
{
  if(!conv || !a || !b) return FFTWPP_AMD_ERR_ARG;
  if(!fftwpp_amd_is_device_pointer(a) || !fftwpp_amd_is_device_pointer(b)) {
    snprintf(capi_error,sizeof(capi_error),"%s: a and b must be device pointers",where);
    fftwpp_amd_internal_set_error(capi_error);
    return FFTWPP_AMD_ERR_ARG;
  }
  try {
    conv->execDevice(a,a,b,true,FFTWPP_AMD_F64,stream);
  } catch(const std::exception &e) {
    const fftwpp::Error *fe=dynamic_cast<const fftwpp::Error *>(&e);
    snprintf(capi_error,sizeof(capi_error),"%s: %s",where,e.what());
    fftwpp_amd_internal_set_error(capi_error);
    return fe ? fe->code : FFTWPP_AMD_ERR_ARG;
  }
  return FFTWPP_AMD_OK;
}

extern "C" {
namespace fftwpp {

double *create_doubleAlign(size_t n) { return utils::doubleAlign(n); }
void delete_doubleAlign(double *p) { utils::deleteAlign(p); }
double *create_complexAlign(size_t n) { return (double *) utils::ComplexAlign(n); }
void delete_complexAlign(double *p) { utils::deleteAlign(p); }

size_t get_fftwpp_maxthreads(void) { return fftw::maxthreads; }
void set_fftwpp_maxthreads(size_t nthreads) { fftw::maxthreads=nthreads; }

HybridConvolution *fftwpp_create_conv1d(size_t L)
{
  return guard("fftwpp_create_conv1d",[&]() { return new HybridConvolution(L); });
}
void fftwpp_conv1d_delete(HybridConvolution *conv) { delete conv; }
void fftwpp_conv1d_convolve(HybridConvolution *conv, double *a, double *b)
{
  guard("fftwpp_conv1d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve->convolve(F);
    return 0;
  });
}

HybridConvolutionHermitian *fftwpp_create_hconv1d(size_t L)
{
  return guard("fftwpp_create_hconv1d",[&]() { return new HybridConvolutionHermitian(L); });
}
void fftwpp_hconv1d_delete(HybridConvolutionHermitian *conv) { delete conv; }
void fftwpp_hconv1d_convolve(HybridConvolutionHermitian *conv, double *a,
                             double *b)
{
  guard("fftwpp_hconv1d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve->convolve(F);
    return 0;
  });
}

HybridConvolution2 *fftwpp_create_conv2d(size_t Lx, size_t Ly)
{
  return guard("fftwpp_create_conv2d",[&]() { return new HybridConvolution2(Lx,Ly); });
}
void fftwpp_conv2d_delete(HybridConvolution2 *conv) { delete conv; }
void fftwpp_conv2d_convolve(HybridConvolution2 *conv, double *a, double *b)
{
  guard("fftwpp_conv2d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve2->convolve(F);
    return 0;
  });
}

HybridConvolutionHermitian2 *fftwpp_create_hconv2d(size_t Lx, size_t Ly)
{
  return guard("fftwpp_create_hconv2d",[&]() { return new HybridConvolutionHermitian2(Lx,Ly); });
}
void fftwpp_hconv2d_delete(HybridConvolutionHermitian2 *conv) { delete conv; }
void fftwpp_hconv2d_convolve(HybridConvolutionHermitian2 *conv, double *a,
                             double *b)
{
  guard("fftwpp_hconv2d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve2->convolve(F);
    return 0;
  });
}

HybridConvolution3 *fftwpp_create_conv3d(size_t Lx, size_t Ly, size_t Lz)
{
  return guard("fftwpp_create_conv3d",[&]() { return new HybridConvolution3(Lx,Ly,Lz); });
}
void fftwpp_conv3d_delete(HybridConvolution3 *conv) { delete conv; }
void fftwpp_conv3d_convolve(HybridConvolution3 *conv, double *a, double *b)
{
  guard("fftwpp_conv3d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve3->convolve(F);
    return 0;
  });
}

HybridConvolutionHermitian3 *fftwpp_create_hconv3d(size_t Lx, size_t Ly,
                                                   size_t Lz)
{
  return guard("fftwpp_create_hconv3d",[&]() { return new HybridConvolutionHermitian3(Lx,Ly,Lz); });
}
void fftwpp_hconv3d_delete(HybridConvolutionHermitian3 *conv) { delete conv; }
void fftwpp_hconv3d_convolve(HybridConvolutionHermitian3 *conv, double *a,
                             double *b)
{
  guard("fftwpp_hconv3d_convolve",[&]() {
    Complex *F[]={(Complex *) a,(Complex *) b};
    conv->convolve3->convolve(F);
    return 0;
  });
}

int fftwpp_conv1d_convolve_device(HybridConvolution *conv, double *a,
                                  const double *b, void *stream)
{
  return convolve_device("fftwpp_conv1d_convolve_device",conv ? conv->convolve : NULL,a,b,stream);
}
int fftwpp_hconv1d_convolve_device(HybridConvolutionHermitian *conv, double *a,
                                   const double *b, void *stream)
{
  return convolve_device("fftwpp_hconv1d_convolve_device",conv ? conv->convolve : NULL,a,b,stream);
}
int fftwpp_conv2d_convolve_device(HybridConvolution2 *conv, double *a,
                                  const double *b, void *stream)
{
  return convolve_device("fftwpp_conv2d_convolve_device",conv ? conv->convolve2 : NULL,a,b,stream);
}
int fftwpp_hconv2d_convolve_device(HybridConvolutionHermitian2 *conv, double *a,
                                   const double *b, void *stream)
{
  return convolve_device("fftwpp_hconv2d_convolve_device",conv ? conv->convolve2 : NULL,a,b,stream);
}
int fftwpp_conv3d_convolve_device(HybridConvolution3 *conv, double *a,
                                  const double *b, void *stream)
{
  return convolve_device("fftwpp_conv3d_convolve_device",conv ? conv->convolve3 : NULL,a,b,stream);
}
int fftwpp_hconv3d_convolve_device(HybridConvolutionHermitian3 *conv, double *a,
                                   const double *b, void *stream)
{
  return convolve_device("fftwpp_hconv3d_convolve_device",conv ? conv->convolve3 : NULL,a,b,stream);
}

void fftwpp_HermitianSymmetrize(double *f)
{
  guard("fftwpp_HermitianSymmetrize",[&]() { HermitianSymmetrize((Complex *) f); return 0; });
}

void fftwpp_HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0, double *f)
{
  guard("fftwpp_HermitianSymmetrizeX",[&]() {
    HermitianSymmetrizeX(Hx,Hy,x0,(Complex *) f);
    return 0;
  });
}

void fftwpp_HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                                  size_t y0, double *f)
{
  guard("fftwpp_HermitianSymmetrizeXY",[&]() {
    HermitianSymmetrizeXY(Hx,Hy,Hz,x0,y0,(Complex *) f);
    return 0;
  });
}

} // namespace fftwpp
} // extern "C"

// ------------------------------------------------------------------------
// Convolution layer of fftwpp_amd.h
// ------------------------------------------------------------------------
struct fftwpp_amd_conv {
  int dim, hermitian, dtype;
  HybridConvolution *c1;
  HybridConvolutionHermitian *h1;
  CenteredConvolution *k1;
  HybridConvolution2 *c2;
  HybridConvolutionHermitian2 *h2;
  HybridConvolution3 *c3;
  HybridConvolutionHermitian3 *h3;
  size_t words;   // complex words per array
  fftwpp::DeviceWork work;
  void *adj[3]={nullptr,nullptr,nullptr};   // Hermitian adjoint temporaries

  fftBase *dims[3];
  // user-multiplier plans (conv_create_general): one Application and one
  // complex or centred transform per dimension
  Application *gapp[3]={nullptr,nullptr,nullptr};
  fftBase *gdims[3]={nullptr,nullptr,nullptr};
  size_t A=2, B=1;
  fftwpp_amd_multiplier umult=nullptr;
  void *uuser=nullptr;
  fftwpp::Timeline timeline;
  void setTimeline(fftwpp::Timeline *t) {
    if(c1) c1->convolve->timeline=t;
    if(h1) h1->convolve->timeline=t;
    if(k1) k1->convolve->timeline=t;
    if(c2) c2->convolve2->timeline=t;
    if(h2) h2->convolve2->timeline=t;
    if(c3) c3->convolve3->timeline=t;
    if(h3) h3->convolve3->timeline=t;
  }
  void setRowHook(const fftwpp::RowHook *r) {
    if(c1) c1->convolve->rowHook=r;
    if(h1) h1->convolve->rowHook=r;
    if(k1) k1->convolve->rowHook=r;
    if(c2) c2->convolve2->rowHook=r;
    if(h2) h2->convolve2->rowHook=r;
    if(c3) c3->convolve3->rowHook=r;
    if(h3) h3->convolve3->rowHook=r;
  }
  void exec(void *out, const void *a, const void *b, void *stream, int mult) {
    bool norm=true;
    if(c1) c1->convolve->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(h1) h1->convolve->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(k1) k1->convolve->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(c2) c2->convolve2->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(h2) h2->convolve2->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(c3) c3->convolve3->execDevice(out,a,b,norm,dtype,stream,mult);
    else if(h3) h3->convolve3->execDevice(out,a,b,norm,dtype,stream,mult);
    else throw fftwpp::Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a user-multiplier plan runs through fftwpp_amd_conv_execute_general");
  }
};

static int conv_status(const char *where, const std::exception &e)
{
  const fftwpp::Error *fe=dynamic_cast<const fftwpp::Error *>(&e);
  snprintf(capi_error,sizeof(capi_error),"%s: %s",where,e.what());
  fftwpp_amd_internal_set_error(capi_error);
  return fe ? fe->code : FFTWPP_AMD_ERR_ARG;
}

// Self-conjugate plane of the stored array of a Hermitian plan (see
// fftwpp_amd_hermitian_plane); false when a centred dimension has even
// length (its Nyquist mode has no stored partner).
static bool hermitian_plane_of(const fftwpp_amd_conv *c, size_t p[6])
{
  if(c->dim == 1) {
    size_t q[6]={1,1,0,0,1,1};
    for(int i=0; i < 6; ++i) p[i]=q[i];
    return true;
  }
  size_t Lx=c->dims[0]->L, Ly=c->dims[1]->L;
  if(Lx % 2 == 0) return false;
  if(c->dim == 2) {
    size_t Hy=(Ly+1)/2;
    size_t q[6]={(Lx+1)/2,1,Lx/2,0,Hy,1};
    for(int i=0; i < 6; ++i) p[i]=q[i];
    return true;
  }
  if(Ly % 2 == 0) return false;
  size_t Hz=(c->dims[2]->L+1)/2;
  size_t q[6]={(Lx+1)/2,(Ly+1)/2,Lx/2,Ly/2,Ly*Hz,Hz};
  for(int i=0; i < 6; ++i) p[i]=q[i];
  return true;
}

// Adjoint of the Hermitian convolution h = C(S a, S b), S the symmetrizer of
// the self-conjugate plane (HermitianSymmetrize{,X,XY}), with respect to the
// stored modes of a and b.  With W the stored cotangent w extended by zero,
// dL/d conj U = W (*) conj-reflected G = W (*) G for Hermitian G = E(S b), and
// folding the implied modes back onto the stored ones gives
//   grad_a = P( C(w_H, S b) ),  grad_b = P( C(w_H, S a) ),
// where w_H = w + conj w(-k) on the self-conjugate plane (plane op 1) and P
// is the adjoint of S (plane op 2).  The multiplier is the plan's own
// realMultBinary: the forward residue transforms are reused unchanged.
static int hermitian_adjoint(fftwpp_amd_conv *c, const void *w, const void *a,
                             const void *b, void *grad_a, void *grad_b,
                             void *stream, size_t bytes)
{
  size_t p[6];
  if(!hermitian_plane_of(c,p)) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_adjoint: Hermitian adjoint needs odd lengths in the centred dimensions");
    return FFTWPP_AMD_ERR_UNSUPPORTED;
  }
  if((grad_a && !b) || (grad_b && !a)) return FFTWPP_AMD_ERR_ARG;
  for(void *&q : c->adj)
    if(!q) {
      int rc=fftwpp_amd_malloc(&q,bytes);
      if(rc) return rc;
    }
  void *wh=c->adj[0], *sa=c->adj[1], *sb=c->adj[2];
  const int dt=c->dtype;
  auto plane=[&](int op, void *f) {
    return fftwpp_amd_hermitian_plane(dt,op,p[0],p[1],p[2],p[3],f,p[4],p[5],stream);
  };
  int rc=fftwpp_amd_memcpy(wh,w,bytes,stream);
  if(!rc) rc=plane(1,wh);
  try {
    if(!rc && grad_a) {
      rc=fftwpp_amd_memcpy(sb,b,bytes,stream);
      if(!rc) rc=plane(0,sb);
      if(!rc) c->exec(grad_a,wh,sb,stream,-1);
      if(!rc) rc=plane(2,grad_a);
    }
    if(!rc && grad_b) {
      rc=fftwpp_amd_memcpy(sa,a,bytes,stream);
      if(!rc) rc=plane(0,sa);
      if(!rc) c->exec(grad_b,wh,sa,stream,-1);
      if(!rc) rc=plane(2,grad_b);
    }
  } catch(const std::exception &e) {
    return conv_status("fftwpp_amd_conv_adjoint",e);
  }
  return rc;
}

extern "C" {

int fftwpp_amd_conv_create(fftwpp_amd_conv **h, int dim, int hermitian,
                           const size_t *L, const size_t *M, const size_t *m,
                           int dtype)
{
  if(!h || !L) return FFTWPP_AMD_ERR_ARG;
  *h=NULL;
  if(dim < 1 || dim > 3) return FFTWPP_AMD_ERR_ARG;
  if(dtype != FFTWPP_AMD_F64 && dtype != FFTWPP_AMD_F32) return FFTWPP_AMD_ERR_ARG;
  size_t Mz[3]={0,0,0}, mz[3]={0,0,0};
  for(int d=0; d < dim; ++d) {
    if(L[d] == 0) return FFTWPP_AMD_ERR_ARG;
    if(M) Mz[d]=M[d];
    if(m) mz[d]=m[d];
  }
  fftwpp_amd_conv *c=new fftwpp_amd_conv();
  c->dim=dim; c->hermitian=hermitian; c->dtype=dtype;
  try {
    if(hermitian == 2) {
      if(dim != 1) throw fftwpp::Error(FFTWPP_AMD_ERR_UNSUPPORTED,"centred complex convolutions are one-dimensional");
      c->k1=new CenteredConvolution(L[0],Mz[0],mz[0]);
      c->words=L[0];
      c->dims[0]=c->k1->fft;
    } else if(hermitian == 3) {
      // real data: the complex compositions with an fftPadReal outer dimension
      const bool r=true;
      if(dim == 1) {
        c->c1=new HybridConvolution(L[0],Mz[0],mz[0],multBinary,r);
        c->dims[0]=c->c1->fft;
        c->words=L[0];
      } else if(dim == 2) {
        c->c2=new HybridConvolution2(L[0],L[1],Mz[0],Mz[1],mz[0],mz[1],multBinary,r);
        c->dims[0]=c->c2->fftx; c->dims[1]=c->c2->ffty;
        c->words=L[0]*L[1];
      } else {
        c->c3=new HybridConvolution3(L[0],L[1],L[2],Mz[0],Mz[1],Mz[2],mz[0],mz[1],mz[2],
                                     multBinary,r);
        c->dims[0]=c->c3->fftx; c->dims[1]=c->c3->ffty; c->dims[2]=c->c3->fftz;
        c->words=L[0]*L[1]*L[2];
      }
    } else if(dim == 1) {
      if(hermitian) {
        c->h1=new HybridConvolutionHermitian(L[0],Mz[0],mz[0]);
        c->words=(L[0]+1)/2;
        c->dims[0]=c->h1->fft;
      } else {
        c->c1=new HybridConvolution(L[0],Mz[0],mz[0]);
        c->words=L[0];
        c->dims[0]=c->c1->fft;
      }
    } else if(dim == 2) {
      if(hermitian) {
        c->h2=new HybridConvolutionHermitian2(L[0],L[1],Mz[0],Mz[1],mz[0],mz[1]);
        c->words=L[0]*((L[1]+1)/2);
        c->dims[0]=c->h2->fftx; c->dims[1]=c->h2->ffty;
      } else {
        c->c2=new HybridConvolution2(L[0],L[1],Mz[0],Mz[1],mz[0],mz[1]);
        c->words=L[0]*L[1];
        c->dims[0]=c->c2->fftx; c->dims[1]=c->c2->ffty;
      }
    } else {
      if(hermitian) {
        c->h3=new HybridConvolutionHermitian3(L[0],L[1],L[2],Mz[0],Mz[1],Mz[2],
                                              mz[0],mz[1],mz[2]);
        c->words=L[0]*L[1]*((L[2]+1)/2);
        c->dims[0]=c->h3->fftx; c->dims[1]=c->h3->ffty; c->dims[2]=c->h3->fftz;
      } else {
        c->c3=new HybridConvolution3(L[0],L[1],L[2],Mz[0],Mz[1],Mz[2],
                                     mz[0],mz[1],mz[2]);
        c->words=L[0]*L[1]*L[2];
        c->dims[0]=c->c3->fftx; c->dims[1]=c->c3->ffty; c->dims[2]=c->c3->fftz;
      }
    }
  } catch(const std::exception &e) {
    int rc=conv_status("fftwpp_amd_conv_create",e);
    fftwpp_amd_conv_destroy(c);
    return rc;
  }
  *h=c;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_set_memory_limit(fftwpp_amd_conv *c, size_t bytes)
{
  if(!c) return FFTWPP_AMD_ERR_ARG;
  if(c->c3) c->c3->convolve3->memoryLimit=bytes;
  else if(c->h3) c->h3->convolve3->memoryLimit=bytes;
  else if(c->c2) c->c2->convolve2->memoryLimit=bytes;
  else if(c->h2) c->h2->convolve2->memoryLimit=bytes;
  else {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_set_memory_limit: 2D and 3D plans only");
    return FFTWPP_AMD_ERR_UNSUPPORTED;
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_work_bytes(const fftwpp_amd_conv *c, size_t *bytes, int *plane)
{
  if(!c || !bytes) return FFTWPP_AMD_ERR_ARG;
  const bool real=c->hermitian == 3;
  if(c->c3 || c->h3) {
    const Convolution3 *k=c->c3 ? c->c3->convolve3 : c->h3->convolve3;
    const bool p=k->usesPlanes(c->dtype);
    const size_t D=p ? 0 : k->residueBlock(c->dtype,real);
    *bytes=p ? k->workBytes(c->dtype,true) : k->workBytes(c->dtype,false,D);
    if(plane) *plane=p ? 1 : 0;
  } else if(c->c2 || c->h2) {
    const Convolution2 *k=c->c2 ? c->c2->convolve2 : c->h2->convolve2;
    *bytes=k->workBytes(c->dtype,k->residueBlock(c->dtype,real));
    if(plane) *plane=0;
  } else {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_work_bytes: 2D and 3D plans only");
    return FFTWPP_AMD_ERR_UNSUPPORTED;
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_destroy(fftwpp_amd_conv *c)
{
  if(!c) return FFTWPP_AMD_OK;
  delete c->c1; delete c->h1; delete c->k1; delete c->c2; delete c->h2; delete c->c3;
  delete c->h3;
  for(int d=0; d < 3; ++d) { delete c->gdims[d]; delete c->gapp[d]; }
  for(void *p : c->adj)
    if(p) fftwpp_amd_free(p);
  delete c;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_execute(fftwpp_amd_conv *c, void *a, const void *b,
                            void *stream)
{
  if(!c || !a || !b) return FFTWPP_AMD_ERR_ARG;
  if(!fftwpp_amd_is_device_pointer(a) || !fftwpp_amd_is_device_pointer(b)) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute: a and b must be device pointers");
    return FFTWPP_AMD_ERR_ARG;
  }
  try {
    c->exec(a,a,b,stream,-1);
  } catch(const std::exception &e) {
    return conv_status("fftwpp_amd_conv_execute",e);
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_execute_out(fftwpp_amd_conv *c, void *out, const void *a,
                                const void *b, void *stream)
{
  if(!c || !out || !a || !b) return FFTWPP_AMD_ERR_ARG;
  if(!fftwpp_amd_is_device_pointer(out) || !fftwpp_amd_is_device_pointer(a) ||
     !fftwpp_amd_is_device_pointer(b)) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute_out: arguments must be device pointers");
    return FFTWPP_AMD_ERR_ARG;
  }
  try {
    c->exec(out,a,b,stream,-1);
  } catch(const std::exception &e) {
    return conv_status("fftwpp_amd_conv_execute_out",e);
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_execute_fused(fftwpp_amd_conv *c, void *out, const void *a,
                                  const void *b, fftwpp_amd_rows_hook hook,
                                  void *user, void *stream)
{
  if(!c || !out || !a || !b || !hook) return FFTWPP_AMD_ERR_ARG;
  if(!fftwpp_amd_is_device_pointer(out) || !fftwpp_amd_is_device_pointer(a) ||
     !fftwpp_amd_is_device_pointer(b)) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute_fused: arguments must be device pointers");
    return FFTWPP_AMD_ERR_ARG;
  }
  const fftwpp::RowHook rh={hook,user};
  try {
    c->setRowHook(&rh);
    c->exec(out,a,b,stream,-1);
    c->setRowHook(NULL);
  } catch(const std::exception &e) {
    c->setRowHook(NULL);
    return conv_status("fftwpp_amd_conv_execute_fused",e);
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_execute_timed(fftwpp_amd_conv *c, void *out, const void *a,
                                  const void *b, void *stream,
                                  float *stage_ms, int max_stages,
                                  int *nstages)
{
  if(!c || !out || !a || !b) return FFTWPP_AMD_ERR_ARG;
  if(!fftwpp_amd_is_device_pointer(out) || !fftwpp_amd_is_device_pointer(a) ||
     !fftwpp_amd_is_device_pointer(b)) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute_timed: a and b must be device pointers");
    return FFTWPP_AMD_ERR_ARG;
  }
  try {
    c->setTimeline(&c->timeline);
    c->exec(out,a,b,stream,-1);
    c->setTimeline(NULL);
    int n=(int) c->timeline.stages();
    if(nstages) *nstages=n;
    for(int i=0; i < n && i < max_stages; ++i)
      if(stage_ms) stage_ms[i]=(float) c->timeline.stage(i);
  } catch(const std::exception &e) {
    c->setTimeline(NULL);
    return conv_status("fftwpp_amd_conv_execute_timed",e);
  }
  return FFTWPP_AMD_OK;
}

const char *fftwpp_amd_conv_stage_name(const fftwpp_amd_conv *c, int i)
{
  if(!c || i < 0 || i >= (int) c->timeline.names.size()) return "";
  return c->timeline.names[i].c_str();
}

int fftwpp_amd_conv_adjoint(fftwpp_amd_conv *c, const void *w, const void *a,
                            const void *b, void *grad_a, void *grad_b,
                            void *stream)
{
  if(!c || !w) return FFTWPP_AMD_ERR_ARG;
  // real plans hold real words, the others complex words
  size_t bytes=c->words*(c->dtype == FFTWPP_AMD_F32 ? 8 : 16)/(c->hermitian == 3 ? 2 : 1);
  const void *ptrs[5]={w,a,b,grad_a,grad_b};
  for(const void *p : ptrs)
    if(p && !fftwpp_amd_is_device_pointer(p)) {
      fftwpp_amd_internal_set_error("fftwpp_amd_conv_adjoint: arguments must be device pointers");
      return FFTWPP_AMD_ERR_ARG;
    }
  if(c->hermitian == 1) return hermitian_adjoint(c,w,a,b,grad_a,grad_b,stream,bytes);
  try {
    if(grad_a) {
      if(!b) return FFTWPP_AMD_ERR_ARG;
      int rc=fftwpp_amd_memcpy(grad_a,w,bytes,stream);
      if(rc) return rc;
      c->exec(grad_a,grad_a,b,stream,FFTWPP_AMD_MULT_CORRELATION);
    }
    if(grad_b) {
      if(!a) return FFTWPP_AMD_ERR_ARG;
      int rc=fftwpp_amd_memcpy(grad_b,w,bytes,stream);
      if(rc) return rc;
      c->exec(grad_b,grad_b,a,stream,FFTWPP_AMD_MULT_CORRELATION);
    }
  } catch(const std::exception &e) {
    return conv_status("fftwpp_amd_conv_adjoint",e);
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_create_general(fftwpp_amd_conv **h, int dim, int centred,
                                   const size_t *L, const size_t *M,
                                   const size_t *m, int dtype, size_t A,
                                   size_t B, fftwpp_amd_multiplier mult,
                                   void *user)
{
  if(!h || !L || !mult) return FFTWPP_AMD_ERR_ARG;
  *h=NULL;
  if(dim < 1 || dim > 3) return FFTWPP_AMD_ERR_ARG;
  if(dtype != FFTWPP_AMD_F64 && dtype != FFTWPP_AMD_F32) return FFTWPP_AMD_ERR_ARG;
  if(A < 1 || B < 1 || A > 8 || B > 8) return FFTWPP_AMD_ERR_ARG;
  for(int d=0; d < dim; ++d)
    if(L[d] == 0) return FFTWPP_AMD_ERR_ARG;
  fftwpp_amd_conv *c=new fftwpp_amd_conv();
  c->dim=dim; c->hermitian=centred ? 2 : 0; c->dtype=dtype;
  c->A=A; c->B=B; c->umult=mult; c->uuser=user;
  c->words=1;
  try {
    for(int d=0; d < dim; ++d) {
      size_t Md=M && M[d] ? M[d] : centred ? hdefault(L[d]) : 2*L[d]-1;
      c->gapp[d]=new Application(A,B,multNone,fftw::maxthreads,false,m ? m[d] : 0);
      if(centred)
        c->gdims[d]=new fftPadCentered(L[d],Md,*c->gapp[d]);
      else
        c->gdims[d]=new fftPad(L[d],Md,*c->gapp[d]);
      c->dims[d]=c->gdims[d];
      c->words *= L[d];
    }
  } catch(const std::exception &e) {
    fftwpp_amd_conv_destroy(c);
    return conv_status("fftwpp_amd_conv_create_general",e);
  }
  *h=c;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_execute_general(fftwpp_amd_conv *c, void *const *f,
                                    void *stream)
{
  if(!c || !f) return FFTWPP_AMD_ERR_ARG;
  if(!c->umult) {
    fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute_general: the plan has no user multiplier");
    return FFTWPP_AMD_ERR_ARG;
  }
  const size_t na=std::max(c->A,c->B);
  for(size_t a=0; a < na; ++a)
    if(!f[a] || !fftwpp_amd_is_device_pointer(f[a])) {
      fftwpp_amd_internal_set_error("fftwpp_amd_conv_execute_general: f[a] must be device pointers");
      return FFTWPP_AMD_ERR_ARG;
    }
  double N=1.0;
  for(int d=0; d < c->dim; ++d) N *= (double) (c->gdims[d]->m*c->gdims[d]->q);
  try {
    fftwpp::generalConvolution(c->gdims,c->dim,c->A,c->B,c->work,c->dtype,f,1.0/N,stream,
                               [&](void **F, size_t n, const fftwpp_amd_layout &layout) {
                                 c->umult(F,n,&layout,stream,c->uuser);
                               });
  } catch(const std::exception &e) {
    return conv_status("fftwpp_amd_conv_execute_general",e);
  }
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_conv_geometry(const fftwpp_amd_conv *c, size_t *N, size_t *m,
                             size_t *q)
{
  if(!c) return FFTWPP_AMD_ERR_ARG;
  for(int d=0; d < c->dim; ++d) {
    if(N) N[d]=c->dims[d]->m*c->dims[d]->q;
    if(m) m[d]=c->dims[d]->m;
    if(q) q[d]=c->dims[d]->q;
  }
  return FFTWPP_AMD_OK;
}

} // extern "C"
