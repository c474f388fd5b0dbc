// fft256.h -- register-resident 256-point FFT for the sub-FFT size m=256.
//
// Sixteen cooperating lanes hold one sequence; lane j keeps x[j + 16k] in
// v[k], k < 16 (the same natural layout before and after the transform).
// 256 = 16 x 16 Stockham: a radix-16 DFT in registers, one transpose through
// LDS (rows padded to 17 entries), a twiddle by w_256^{t j} and a second
// radix-16 DFT.  Twiddles are copies of the dimension's FP64-generated table
// zN (N = 256 q), w_256^n = zN[n q], staged in LDS.
#pragma once
#include "common.h"

namespace fftwpp_amd {

template<typename T> struct K16;
template<> struct K16<double> {
  static constexpr double c1=0.92387953251128675612818318939678829;  // cos(pi/8)
  static constexpr double s1=0.38268343236508977172845998403039887;  // sin(pi/8)
  static constexpr double r2=0.70710678118654752440084436210484904;
};
template<> struct K16<float> {
  static constexpr float c1=0.92387953251128675612818318939678829f;
  static constexpr float s1=0.38268343236508977172845998403039887f;
  static constexpr float r2=0.70710678118654752440084436210484904f;
};

// In-place DFT4 of (a,b,c,d) with kernel e^{S 2 pi i ut/4}.
template<typename T, int S>
__device__ __forceinline__ void dft4(cplx<T> &a, cplx<T> &b, cplx<T> &c,
                                     cplx<T> &d)
{
  cplx<T> s02=a+c, d02=a-c, s13=b+d, d13=muli(b-d,S);
  a=s02+s13; c=s02-s13; b=d02+d13; d=d02-d13;
}

// v <- DFT16(v) with kernel e^{S 2 pi i ut/16}; t = t2 + 4 t1, u = u1 + 4 u2.
template<typename T, int S>
__device__ __forceinline__ void dft16(cplx<T> *v)
{
  // DFT4 over t1 for each t2: a[t2][u1] stored back at v[t2 + 4 u1]
#pragma unroll
  for(int t2=0; t2 < 4; ++t2)
    dft4<T,S>(v[t2],v[t2+4],v[t2+8],v[t2+12]);
  // twiddles w16^{u1 t2}
  const T c1=K16<T>::c1, s1=K16<T>::s1, r2=K16<T>::r2;
  const T s=(T) S;
  auto tw=[&](cplx<T> &x, T wr, T wi) { x=x*mk(wr,wi); };
  tw(v[1+4*1],c1,s*s1);   // t2=1,u1=1: w^1
  tw(v[1+4*2],r2,s*r2);   // w^2
  tw(v[1+4*3],s1,s*c1);   // w^3
  tw(v[2+4*1],r2,s*r2);   // t2=2,u1=1: w^2
  v[2+4*2]=muli(v[2+4*2],S);  // w^4
  tw(v[2+4*3],-r2,s*r2);  // w^6
  tw(v[3+4*1],s1,s*c1);   // t2=3,u1=1: w^3
  tw(v[3+4*2],-r2,s*r2);  // w^6
  tw(v[3+4*3],-c1,-s*s1); // w^9
  // DFT4 over t2 for each u1: inputs v[t2 + 4 u1], output index u1 + 4 u2
  cplx<T> y[16];
#pragma unroll
  for(int u1=0; u1 < 4; ++u1) {
    cplx<T> a=v[4*u1], b=v[4*u1+1], c=v[4*u1+2], d=v[4*u1+3];
    dft4<T,S>(a,b,c,d);
    y[u1]=a; y[u1+4]=b; y[u1+8]=c; y[u1+12]=d;
  }
#pragma unroll
  for(int k=0; k < 16; ++k) v[k]=y[k];
}

// Synchronisation of the lanes that share an LDS row: the whole block
// (block barrier) or, when the 16 lanes of a sequence sit in one wave, the
// wave (LDS operations of a wave complete in order; the fences keep the
// compiler from moving LDS accesses across the exchange).
template<int WAVE>
__device__ __forceinline__ void lanes_sync()
{
  if(WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE,"wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE,"wavefront");
  } else
    __syncthreads();
}

// 256-point transform with kernel e^{S 2 pi i jk/256} of the sequence held by
// the 16 lanes j of a group, exchanging real and imaginary parts in two
// rounds through the group's scalar LDS row area `row` (at least 16*RS
// words, RS >= 17).  Lane j writes its stage-1 output u to row[RS j + (u ^ h)]
// and reads stage-2 input t from row[RS t + (j ^ h)]: the swizzle h (0 or 2, the
// same for the 16 lanes of a group) lets groups whose areas start on the same
// bank use disjoint banks.  Stage twiddles w_256^{t j} come from the 16 x 16
// table tw16[16 t + j], so the 16 lanes read consecutive words.  With WAVE=0
// every thread of the block must call this (block barriers).
template<typename T, int S, int WAVE, int RS=17>
__device__ __forceinline__ void fft256s(cplx<T> *v, T *row, int j, int h,
                                        const cplx<T> *tw16)
{
  dft16<T,S>(v);
  // u ^ h = u + h when bit 1 of u is clear, u - h when it is set (h = 0, 2)
  T *wa=row+RS*j+h, *wb=row+RS*j-h;
  const T *ra=row+(j ^ h);
#pragma unroll
  for(int u=0; u < 16; ++u) ((u & 2) ? wb : wa)[u]=v[u].x;
  lanes_sync<WAVE>();
#pragma unroll
  for(int t=0; t < 16; ++t) v[t].x=ra[RS*t];
  lanes_sync<WAVE>();
#pragma unroll
  for(int u=0; u < 16; ++u) ((u & 2) ? wb : wa)[u]=v[u].y;
  lanes_sync<WAVE>();
#pragma unroll
  for(int t=0; t < 16; ++t) v[t].y=ra[RS*t];
  lanes_sync<WAVE>();
#pragma unroll
  for(int t=1; t < 16; ++t) {
    cplx<T> w=(tw16+16*t)[j];
    v[t]=S > 0 ? v[t]*w : mulc(v[t],w);
  }
  dft16<T,S>(v);
}

// fft256s of two sequences v and u sharing the exchange row: the 16-point
// DFTs and twiddles of both are independent work for the same lanes, the
// exchanges run one after the other (the same synchronisations as two calls).
template<typename T, int S, int WAVE, int RS>
__device__ __forceinline__ void fft256s_pair(cplx<T> *v, cplx<T> *u, T *row, int j, int h,
                                             const cplx<T> *tw16)
{
  dft16<T,S>(v);
  dft16<T,S>(u);
  T *wa=row+RS*j+h, *wb=row+RS*j-h;
  const T *ra=row+(j ^ h);
#pragma unroll
  for(int part=0; part < 4; ++part) {
    cplx<T> *z=part < 2 ? v : u;
    const bool re=(part & 1) == 0;
#pragma unroll
    for(int e=0; e < 16; ++e) ((e & 2) ? wb : wa)[e]=re ? z[e].x : z[e].y;
    lanes_sync<WAVE>();
#pragma unroll
    for(int t=0; t < 16; ++t) {
      if(re) z[t].x=ra[RS*t];
      else z[t].y=ra[RS*t];
    }
    lanes_sync<WAVE>();
  }
#pragma unroll
  for(int t=1; t < 16; ++t) {
    const cplx<T> w=(tw16+16*t)[j];
    v[t]=S > 0 ? v[t]*w : mulc(v[t],w);
    u[t]=S > 0 ? u[t]*w : mulc(u[t],w);
  }
  dft16<T,S>(v);
  dft16<T,S>(u);
}

// The 16 x 16 stage-twiddle table tw16[16 t + j] = w_256^{t j} from the
// dimension's table zN (N = 256 q, zN[n] = w_N^n): w_256^n = zN[q n].
// Cooperative fill by `nthreads` threads of the block.
template<typename T>
__device__ __forceinline__ void fill_tw16(cplx<T> *tw16, const cplx<T> *zN,
                                          int q, int tid, int nthreads)
{
  for(int n=tid; n < 256; n += nthreads)
    tw16[n]=zN[q*((n >> 4)*(n & 15))];
}

// Residue twiddles rtab[256 rr + n] = w_N^{(r0 + rr) n}, n < 256, rr < nres.
template<typename T>
__device__ __forceinline__ void fill_rtab(cplx<T> *rtab, const cplx<T> *zN,
                                          int r0, int nres, int tid,
                                          int nthreads)
{
  for(int e=tid; e < 256*nres; e += nthreads)
    rtab[e]=zN[(r0+(e >> 8))*(e & 255)];
}

} // namespace fftwpp_amd
