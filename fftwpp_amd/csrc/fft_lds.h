// fft_lds.h -- batched Stockham FFT of length-m sequences held in LDS.
//
// This is the sub-FFT layer of the hybrid convolution (the role FFTW's
// mfft1d/Mfft1d plays in the reference, fftw++.h:576-896).  Each stage of
// radix R maps butterfly j (0 <= j < m/R, k = j mod Ns) as
//     v_t = in[j + t m/R] * w^{t k},  w = e^{sign 2 pi i/(Ns R)}
//     out[(j/Ns) Ns R + k + u Ns] = sum_t v_t e^{sign 2 pi i u t/R}
// so the result is in natural order after the last stage (autosort).  Two
// LDS buffers alternate between stages; one barrier separates stages.
// Twiddles come from the dimension's FP64-generated table zN (N=m*q):
// e^{2 pi i/m} = zN[q].
#pragma once
#include "common.h"

namespace fftwpp_amd {

template<typename T> struct K;
template<> struct K<double> {
  static constexpr double r2=0.70710678118654752440084436210484904;
  static constexpr double c3=-0.5;
  static constexpr double s3=0.86602540378443864676372317075293618;
  static constexpr double c51=0.30901699437494742410229341718281906;
  static constexpr double c52=-0.80901699437494742410229341718281906;
  static constexpr double s51=0.95105651629515357211643933337938214;
  static constexpr double s52=0.58778525229247312916870595463907277;
};
template<> struct K<float> {
  static constexpr float r2=0.70710678118654752440084436210484904f;
  static constexpr float c3=-0.5f;
  static constexpr float s3=0.86602540378443864676372317075293618f;
  static constexpr float c51=0.30901699437494742410229341718281906f;
  static constexpr float c52=-0.80901699437494742410229341718281906f;
  static constexpr float s51=0.95105651629515357211643933337938214f;
  static constexpr float s52=0.58778525229247312916870595463907277f;
};

// In-register DFT of R points with kernel e^{sign 2 pi i ut/R}.
template<typename T, int R> struct DFT;

template<typename T> struct DFT<T,2> {
  __device__ __forceinline__ static void run(cplx<T> *v, int) {
    cplx<T> a=v[0], b=v[1];
    v[0]=a+b; v[1]=a-b;
  }
};

template<typename T> struct DFT<T,4> {
  __device__ __forceinline__ static void run(cplx<T> *v, int sign) {
    cplx<T> a=v[0]+v[2], b=v[0]-v[2], c=v[1]+v[3], d=muli(v[1]-v[3],sign);
    v[0]=a+c; v[2]=a-c; v[1]=b+d; v[3]=b-d;
  }
};

template<typename T> struct DFT<T,8> {
  __device__ __forceinline__ static void run(cplx<T> *v, int sign) {
    cplx<T> e[4]={v[0],v[2],v[4],v[6]}, o[4]={v[1],v[3],v[5],v[7]};
    DFT<T,4>::run(e,sign);
    DFT<T,4>::run(o,sign);
    const T r=K<T>::r2;
    T s=(T) sign;
    // w^1=(1+s i)/sqrt2, w^2=s i, w^3=(-1+s i)/sqrt2
    cplx<T> o1=mk(r*(o[1].x-s*o[1].y),r*(o[1].y+s*o[1].x));
    cplx<T> o2=muli(o[2],sign);
    cplx<T> o3=mk(r*(-o[3].x-s*o[3].y),r*(-o[3].y+s*o[3].x));
    v[0]=e[0]+o[0]; v[4]=e[0]-o[0];
    v[1]=e[1]+o1;   v[5]=e[1]-o1;
    v[2]=e[2]+o2;   v[6]=e[2]-o2;
    v[3]=e[3]+o3;   v[7]=e[3]-o3;
  }
};

template<typename T> struct DFT<T,3> {
  __device__ __forceinline__ static void run(cplx<T> *v, int sign) {
    cplx<T> t1=v[1]+v[2], t2=v[1]-v[2];
    cplx<T> a=v[0]+K<T>::c3*t1;
    cplx<T> b=muli(((T) sign*K<T>::s3)*t2,1);
    v[0]=v[0]+t1; v[1]=a+b; v[2]=a-b;
  }
};

template<typename T> struct DFT<T,5> {
  __device__ __forceinline__ static void run(cplx<T> *v, int sign) {
    T s=(T) sign;
    cplx<T> a1=v[1]+v[4], b1=v[1]-v[4], a2=v[2]+v[3], b2=v[2]-v[3];
    cplx<T> p1=v[0]+K<T>::c51*a1+K<T>::c52*a2;
    cplx<T> p2=v[0]+K<T>::c52*a1+K<T>::c51*a2;
    cplx<T> q1=muli((s*K<T>::s51)*b1+(s*K<T>::s52)*b2,1);
    cplx<T> q2=muli((s*K<T>::s52)*b1-(s*K<T>::s51)*b2,1);
    v[0]=v[0]+a1+a2;
    v[1]=p1+q1; v[4]=p1-q1;
    v[2]=p2+q2; v[3]=p2-q2;
  }
};

// One Stockham stage for a compile-time radix.
template<typename T, int R>
__device__ void stage(const cplx<T> *in, cplx<T> *out, int nb, int ld,
                      int m, int Ns, int sign, const DimArgs &d,
                      int tid, int nt)
{
  const int mR=m/R;
  const int total=nb*mR;
  const int64_t twstep=(int64_t) (m/(Ns*R))*d.q; // zN index of e^{2pi i/(Ns R)}
  for(int idx=tid; idx < total; idx += nt) {
    int b=idx/mR;
    int j=idx-b*mR;
    int k=j % Ns;
    const cplx<T> *ib=in+(int64_t) b*ld;
    cplx<T> *ob=out+(int64_t) b*ld;
    cplx<T> v[R];
#pragma unroll
    for(int t=0; t < R; ++t) v[t]=ib[j+t*mR];
    if(k != 0) {
#pragma unroll
      for(int t=1; t < R; ++t) {
        cplx<T> w=ztab<T>(d,(int64_t) t*k*twstep);
        v[t]=sign > 0 ? v[t]*w : mulc(v[t],w);
      }
    }
    DFT<T,R>::run(v,sign);
    int o0=(j/Ns)*Ns*R+k;
#pragma unroll
    for(int u=0; u < R; ++u) ob[o0+u*Ns]=v[u];
  }
}

// Stage for any other radix: O(R^2) butterflies straight from LDS, one
// output u of butterfly j per thread (a single large prime radix still
// spreads over the block).  The twiddle of term t is w^{tk} e^{2pi i ut/R}
// = zN[t E1 mod N] with E1 = k twstep + u N/R, stepped without division.
template<typename T>
__device__ void stage_generic(const cplx<T> *in, cplx<T> *out, int nb,
                              int ld, int m, int Ns, int R, int sign,
                              const DimArgs &d, int tid, int nt)
{
  const int mR=m/R;
  const int total=nb*m;
  const int64_t twstep=(int64_t) (m/(Ns*R))*d.q;
  const int64_t wR=d.N/R; // zN index of e^{2pi i/R}
  for(int idx=tid; idx < total; idx += nt) {
    int b=idx/m;
    int rem=idx-b*m;
    int u=rem/mR;
    int j=rem-u*mR;
    int k=j % Ns;
    const cplx<T> *ib=in+(int64_t) b*ld;
    const int64_t E1=((int64_t) k*twstep+(int64_t) u*wR) % d.N;
    int64_t e=0;
    cplx<T> s=mk((T) 0,(T) 0);
    for(int t=0; t < R; ++t) {
      cplx<T> x=ib[j+t*mR];
      cplx<T> w=ztab<T>(d,e);
      s=s+(sign > 0 ? x*w : mulc(x,w));
      e += E1;
      if(e >= d.N) e -= d.N;
    }
    out[(int64_t) b*ld+(j/Ns)*Ns*R+k+u*Ns]=s;
  }
}

// Transform nb sequences (sequence b at A[b*ld .. b*ld+m-1]).  All threads of
// the block must call this.  Returns the buffer holding the result.
template<typename T>
__device__ cplx<T> *lds_fft(cplx<T> *A, cplx<T> *B, int nb, int ld,
                            const DimArgs &d, int sign)
{
  const int tid=threadIdx.x, nt=blockDim.x;
  const int m=(int) d.m;
  int Ns=1;
  for(int st=0; st < d.nstages; ++st) {
    int R=d.radix[st];
    switch(R) {
    case 2: stage<T,2>(A,B,nb,ld,m,Ns,sign,d,tid,nt); break;
    case 3: stage<T,3>(A,B,nb,ld,m,Ns,sign,d,tid,nt); break;
    case 4: stage<T,4>(A,B,nb,ld,m,Ns,sign,d,tid,nt); break;
    case 5: stage<T,5>(A,B,nb,ld,m,Ns,sign,d,tid,nt); break;
    case 8: stage<T,8>(A,B,nb,ld,m,Ns,sign,d,tid,nt); break;
    default: stage_generic<T>(A,B,nb,ld,m,Ns,R,sign,d,tid,nt); break;
    }
    __syncthreads();
    cplx<T> *tmp=A; A=B; B=tmp;
    Ns *= R;
  }
  return A;
}

} // namespace fftwpp_amd
