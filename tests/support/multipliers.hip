// multipliers.hip -- user device multipliers for the tests of the
// multiplier/Indices contract (convolve.h:48-82): what a user of the library
// writes in place of a host multiplier.  Each exists as a C callback of the
// C ABI (fftwpp_amd_multiplier) and, for the C++ face, as a multiplier with
// the reference signature driving a Convolution/Convolution2/Convolution3 on
// host arrays.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>

#include "convolve.h"

namespace {

template<typename T> struct cx { T x, y; };

template<typename T>
__device__ inline cx<T> mul(cx<T> a, cx<T> b)
{
  return {a.x*b.x-a.y*b.y,a.x*b.y+a.y*b.x};
}

constexpr int TPB=256;

inline unsigned blocks(size_t n) { return (unsigned) ((n+TPB-1)/TPB); }

// F0 <- F0 F1 F2: a ternary product (A = 3, B = 1)
template<typename T>
__global__ void ternary_kernel(cx<T> *F0, const cx<T> *F1, const cx<T> *F2, size_t n)
{
  size_t i=(size_t) blockIdx.x*TPB+threadIdx.x;
  if(i < n) F0[i]=mul(mul(F0[i],F1[i]),F2[i]);
}

// F0 <- F0 F1: the binary product of multBinary, as a user multiplier
template<typename T>
__global__ void binary_kernel(cx<T> *F0, const cx<T> *F1, size_t n)
{
  size_t i=(size_t) blockIdx.x*TPB+threadIdx.x;
  if(i < n) F0[i]=mul(F0[i],F1[i]);
}

// F0 <- F0 F0: autoconvolution (A = 1, B = 1)
template<typename T>
__global__ void square_kernel(cx<T> *F0, size_t n)
{
  size_t i=(size_t) blockIdx.x*TPB+threadIdx.x;
  if(i < n) F0[i]=mul(F0[i],F0[i]);
}

// (F0, F1) <- (F0 F1, F0 - F1): two outputs (A = 2, B = 2)
template<typename T>
__global__ void pair_kernel(cx<T> *F0, cx<T> *F1, size_t n)
{
  size_t i=(size_t) blockIdx.x*TPB+threadIdx.x;
  if(i < n) {
    cx<T> a=F0[i], b=F1[i];
    F0[i]=mul(a,b);
    F1[i]={a.x-b.x,a.y-b.y};
  }
}

// F0 <- F0 F1 w(x) with a weight that depends on the padded grid point x:
// w = 1 + cos(2 pi x_0/N_0)/2 + sin(2 pi sum_d (d+1) x_d/N_d)/4.
template<typename T>
__global__ void weighted_kernel(cx<T> *F0, const cx<T> *F1, size_t n,
                                fftwpp_amd_layout l)
{
  size_t i=(size_t) blockIdx.x*TPB+threadIdx.x;
  if(i >= n) return;
  double x0=0.0, t=0.0;
  for(int d=0; d < l.dims; ++d) {
    size_t j=(i/l.stride[d]) % l.N[d];
    size_t x=FFTWPP_AMD_GRID_POINT(l.m[d],l.q[d],j);
    double u=(double) x/(double) l.N[d];
    if(d == 0) x0=u;
    t += (d+1)*u;
  }
  t -= floor(t);
  double w=1.0+0.5*cospi(2.0*x0)+0.25*sinpi(2.0*t);
  cx<T> p=mul(F0[i],F1[i]);
  F0[i]={(T) (p.x*w),(T) (p.y*w)};
}

template<typename T>
void launch(int which, void **F, size_t n, const fftwpp_amd_layout *l, hipStream_t s)
{
  cx<T> **G=(cx<T> **) F;
  switch(which) {
    case 0: hipLaunchKernelGGL(ternary_kernel<T>,dim3(blocks(n)),dim3(TPB),0,s,G[0],G[1],G[2],n); break;
    case 1: hipLaunchKernelGGL(square_kernel<T>,dim3(blocks(n)),dim3(TPB),0,s,G[0],n); break;
    case 2: hipLaunchKernelGGL(pair_kernel<T>,dim3(blocks(n)),dim3(TPB),0,s,G[0],G[1],n); break;
    case 4: hipLaunchKernelGGL(binary_kernel<T>,dim3(blocks(n)),dim3(TPB),0,s,G[0],G[1],n); break;
    default: hipLaunchKernelGGL(weighted_kernel<T>,dim3(blocks(n)),dim3(TPB),0,s,G[0],G[1],n,*l); break;
  }
}

void dispatch(int which, void **F, size_t n, const fftwpp_amd_layout *l, void *stream)
{
  if(l->dtype == FFTWPP_AMD_F32) launch<float>(which,F,n,l,(hipStream_t) stream);
  else launch<double>(which,F,n,l,(hipStream_t) stream);
}

// The reference-signature multipliers (C++ face): F are device arrays of the
// whole padded transforms, their layout and stream in indices.
void cxxTernary(Complex **F, size_t n, fftwpp::Indices *indices, size_t)
{
  dispatch(0,(void **) F,n,indices->layout,indices->stream);
}

void cxxWeighted(Complex **F, size_t n, fftwpp::Indices *indices, size_t)
{
  dispatch(3,(void **) F,n,indices->layout,indices->stream);
}

} // namespace

extern "C" {

// fftwpp_amd_multiplier callbacks
void testmult_ternary(void **F, size_t n, const fftwpp_amd_layout *l, void *s, void *)
{ dispatch(0,F,n,l,s); }
void testmult_square(void **F, size_t n, const fftwpp_amd_layout *l, void *s, void *)
{ dispatch(1,F,n,l,s); }
void testmult_pair(void **F, size_t n, const fftwpp_amd_layout *l, void *s, void *)
{ dispatch(2,F,n,l,s); }
void testmult_weighted(void **F, size_t n, const fftwpp_amd_layout *l, void *s, void *)
{ dispatch(3,F,n,l,s); }
void testmult_binary(void **F, size_t n, const fftwpp_amd_layout *l, void *s, void *)
{ dispatch(4,F,n,l,s); }

// The C++ face on host arrays (interleaved complex doubles), as a user of
// convolve.h writes it (tests/hybridconv.cc:44-46 with A = 3 or a weighted
// binary multiplier): which = 0 ternary (f[0..2]), 3 weighted (f[0..1]).
// N[d] receives the padded lengths.  Returns 0, or 1 when the library threw.
int testmult_cxx(int which, int dim, const size_t *L, const size_t *M, double **f,
                 size_t *N)
{
  using namespace fftwpp;
  try {
    multiplier *mult=which == 0 ? cxxTernary : cxxWeighted;
    const size_t A=which == 0 ? 3 : 2;
    Complex **F=(Complex **) f;
    if(dim == 1) {
      Application app(A,1,mult,1,false);
      fftPad fft(L[0],M[0],app);
      Convolution C(&fft);
      C.convolve(F);
      N[0]=fft.m*fft.q;
    } else if(dim == 2) {
      Application appx(A,1,multNone,1,false);
      fftPad fftx(L[0],M[0],appx,L[1]);
      Application appy(A,1,mult,appx);
      fftPad ffty(L[1],M[1],appy);
      Convolution2 C(&fftx,&ffty);
      C.convolve(F);
      N[0]=fftx.m*fftx.q; N[1]=ffty.m*ffty.q;
    } else {
      Application appx(A,1,multNone,1,false);
      fftPad fftx(L[0],M[0],appx,L[1]*L[2]);
      Application appy(A,1,multNone,appx);
      fftPad ffty(L[1],M[1],appy,L[2]);
      Application appz(A,1,mult,appy);
      fftPad fftz(L[2],M[2],appz);
      Convolution3 C(&fftx,&ffty,&fftz);
      C.convolve(F);
      N[0]=fftx.m*fftx.q; N[1]=ffty.m*ffty.q; N[2]=fftz.m*fftz.q;
    }
  } catch(const std::exception &) {
    return 1;
  }
  return 0;
}

} // extern "C"
