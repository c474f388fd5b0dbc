// kernels256.h -- residue transforms and fused row convolution for the
// sub-FFT size m = 256 with p = 1 (L <= 256), non-centred data: the shape of
// every dimension of the 256^3 -> 512^3 convolution.  Same contracts as the
// generic kernels in kernels.hip, with the FFT held in registers (fft256.h).
//
// Addressing: element e = j + 16k of a lane lives at
//     (uniform base + 16 k * stride) + (lane offset)
// so each unrolled access uses a scalar base and one 32-bit vector offset
// shared by all k; the host guarantees offsets below 2^31 elements.
#pragma once
#include "fft256.h"

namespace fftwpp_amd {

template<typename T>
__device__ __forceinline__ cplx<T> zero() { return mk((T) 0,(T) 0); }

// Columns: lane (c, j) = (threadIdx.x % TC, threadIdx.x / TC) so consecutive
// lanes read consecutive columns; the 16 lanes j of column c share an LDS row.
template<typename T, int TC>
__global__ void __launch_bounds__(16*TC,sizeof(T) == 4 ? 3 : 2)
colfwd256_kernel(const cplx<T> *f, int64_t S, int64_t C, int64_t bsi,
                 cplx<T> *F, int64_t SF, int64_t bso, int L, int64_t r0,
                 int64_t nres, const cplx<T> *zN, int64_t q)
{
  __shared__ cplx<T> lds[TC*16*17];
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  const int64_t c0=(int64_t) blockIdx.x*TC;
  const bool ok=c0+c < C;
  const cplx<T> *fb=f+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *Fb=F+(int64_t) blockIdx.y*bso+c0;
  const unsigned lin=(unsigned) (c+j*S);
  const unsigned lout=(unsigned) (c+j*SF);
  cplx<T> *row=lds+c*16*17;
  for(int64_t rr=0; rr < nres; ++rr) {
    const int r=(int) (r0+rr);
    cplx<T> v[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      const cplx<T> *bk=fb+(int64_t) 16*k*S;
      cplx<T> x=(ok && j+16*k < L) ? bk[lin] : zero<T>();
      v[k]=r == 0 ? x : x*(zN+16*k*r)[j*r];
    }
    fft256<T,1>(v,row,j,zN,q);
    cplx<T> *Fr=Fb+rr*256*SF;
    if(ok) {
#pragma unroll
      for(int u=0; u < 16; ++u) (Fr+(int64_t) 16*u*SF)[lout]=v[u];
    }
  }
}

template<typename T, int TC>
__global__ void __launch_bounds__(16*TC,sizeof(T) == 4 ? 3 : 2)
colbwd256_kernel(const cplx<T> *F, int64_t SF, int64_t bsi, cplx<T> *f,
                 int64_t S, int64_t C, int64_t bso, int L, int64_t r0,
                 int64_t nres, int accumulate, T scale, const cplx<T> *zN,
                 int64_t q)
{
  __shared__ cplx<T> lds[TC*16*17];
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  const int64_t c0=(int64_t) blockIdx.x*TC;
  const bool ok=c0+c < C;
  const cplx<T> *Fb=F+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *fb=f+(int64_t) blockIdx.y*bso+c0;
  const unsigned lin=(unsigned) (c+j*SF);
  const unsigned lout=(unsigned) (c+j*S);
  cplx<T> *row=lds+c*16*17;
  cplx<T> acc[16];
#pragma unroll
  for(int k=0; k < 16; ++k) acc[k]=zero<T>();
  for(int64_t rr=0; rr < nres; ++rr) {
    const int r=(int) (r0+rr);
    const cplx<T> *Fr=Fb+rr*256*SF;
    cplx<T> v[16];
#pragma unroll
    for(int u=0; u < 16; ++u)
      v[u]=ok ? (Fr+(int64_t) 16*u*SF)[lin] : zero<T>();
    fft256<T,-1>(v,row,j,zN,q);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? v[k] : mulc(v[k],(zN+16*k*r)[j*r]));
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < 16; ++k) {
      if(j+16*k < L) {
        cplx<T> *o=fb+(int64_t) 16*k*S;
        cplx<T> v=acc[k];
        if(accumulate) v=v+o[lout];
        o[lout]=scale*v;
      }
    }
  }
}

// Rows: NR rows per block of 16*NR lanes; lane j of row i = threadIdx.x % 16.
// The spectrum of f is parked in LDS while g is transformed.
template<typename T, int MULT, int NR>
__global__ void __launch_bounds__(16*NR,sizeof(T) == 4 ? 4 : 2)
rowconv256_kernel(const cplx<T> *f, const cplx<T> *g, int64_t rows, int64_t rs,
                  cplx<T> *out, int64_t os, int L, T scale, const cplx<T> *zN,
                  int64_t q)
{
  __shared__ cplx<T> lds[NR*16*17];
  __shared__ cplx<T> spec[NR*256];
  const int i=threadIdx.x/16, j=threadIdx.x % 16;
  const int64_t row0=(int64_t) blockIdx.x*NR;
  const bool ok=row0+i < rows;
  const cplx<T> *fr=f+row0*rs;
  const cplx<T> *gr=g+row0*rs;
  cplx<T> *orow=out+row0*os;
  const unsigned lin=(unsigned) (i*rs+j);
  const unsigned lo=(unsigned) (i*os+j);
  cplx<T> *row=lds+i*16*17;
  cplx<T> *sp=spec+i*256+j;
  cplx<T> acc[16];
#pragma unroll
  for(int k=0; k < 16; ++k) acc[k]=zero<T>();
  for(int r=0; r < (int) q; ++r) {
    cplx<T> a[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      cplx<T> x=(ok && j+16*k < L) ? (fr+16*k)[lin] : zero<T>();
      a[k]=r == 0 ? x : x*(zN+16*k*r)[j*r];
    }
    fft256<T,1>(a,row,j,zN,q);
#pragma unroll
    for(int k=0; k < 16; ++k) sp[16*k]=a[k];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      cplx<T> x=(ok && j+16*k < L) ? (gr+16*k)[lin] : zero<T>();
      a[k]=r == 0 ? x : x*(zN+16*k*r)[j*r];
    }
    fft256<T,1>(a,row,j,zN,q);
#pragma unroll
    for(int k=0; k < 16; ++k) {
      cplx<T> u=sp[16*k];
      if(MULT == 1) a[k]=u*a[k];
      else if(MULT == 3) a[k]=mulc(u,a[k]);
      else a[k]=u;
    }
    fft256<T,-1>(a,row,j,zN,q);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? a[k] : mulc(a[k],(zN+16*k*r)[j*r]));
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < 16; ++k)
      if(j+16*k < L) (orow+16*k)[lo]=scale*acc[k];
  }
}

} // namespace fftwpp_amd
