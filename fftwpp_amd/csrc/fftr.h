// fftr.h -- register-resident Stockham FFT of a compile-time power-of-two
// length M (16 <= M <= 4096): the sub-FFT layer (the role of FFTW's
// mfft1d in the reference, fftw++.h:576-896) for the fused kernels whose
// sub-FFT size is not 256.
//
// TPS = M/V threads hold one sequence, thread t keeping x[t + TPS k] in
// v[k], k < V (V = 16, or 8 to halve the registers of FP64 kernels), before
// and after the transform.  The stages are radix V (and one final smaller
// power of two when log2 M is not a multiple of log2 V); stage s with radix
// R and Ns = (product of the earlier radices) does, for each butterfly
// j = t + TPS b (b < V/R) with k' = j mod Ns,
//     u_t' = v[b + t' V/R] w^{t' k'},   w = e^{S 2 pi i/(Ns R)},
//     out[(j/Ns) Ns R + k' + u Ns] = sum_t' u_t' e^{S 2 pi i u t'/R},
// in place in registers; between stages the outputs go through LDS (real
// and imaginary parts in two rounds, scalar words at e + e/16, which keeps
// every access of the gfx950 LDS bank-conflict free) and come back as
// v[k] = out[t + TPS k].  After the last stage v[k] holds X[t + TPS k].
//
// Twiddles: one LDS table of M-1 complex words, stage s at offset off_s
// holding e^{2 pi i t' k'/(Ns R)} at off_s + (t'-1) Ns + k', so the threads
// of a stage read consecutive words; fill_twr() builds it from the
// dimension's FP64-generated table zN (N = M q): e^{2 pi i n/M} = zN[q n].
#pragma once
#include "common.h"
#include "fft_lds.h"
#include "fft256.h"

namespace fftwpp_amd {

template<int M, int V=16> struct RPlan {
  static_assert(M >= 16 && M <= 4096 && (M & (M-1)) == 0, "M must be 2^k, 16..4096");
  static_assert(V == 8 || V == 16, "8 or 16 values per thread");
  static constexpr int TPS=M/V;
  static constexpr int LOG=__builtin_ctz(M);
  static constexpr int LV=__builtin_ctz(V);
  static constexpr int NV=LOG/LV;                       // radix-V stages
  static constexpr int LAST=1 << (LOG-LV*NV);           // remaining factor
  static constexpr int NST=NV+(LAST > 1 ? 1 : 0);
  static constexpr int PAD=M+M/16;                      // exchange words
  __host__ __device__ static constexpr int radix(int s) { return s < NV ? V : LAST; }
  // loops rather than recursion, so that calls with a run-time s inline
  __host__ __device__ static constexpr int ns(int s)
  {
    int n=1;
    for(int i=0; i < s; ++i) n *= radix(i);
    return n;
  }
  __host__ __device__ static constexpr int off(int s)
  {
    int o=0;
    for(int i=1; i < s; ++i) o += (radix(i)-1)*ns(i);
    return o;
  }
};

__device__ __forceinline__ int padr(int e) { return e+(e >> 4); }

// Word of sequence element e in a one-round (CX) exchange area: an XOR
// swizzle of the low bits by higher bits that makes every exchange of the
// transform conflict-free on gfx950's LDS (64-bit words: writes served 16
// lanes per pass over 32 banks, reads 32 lanes over 64 banks; 32-bit words:
// 32 lanes over 32 banks) for M >= 128 (FP64, V = 8), M >= 256 (FP32, V = 8)
// or M >= 512 (FP32, V = 16); e + e/16 below that.  Indices stay below
// M + M/16.
template<typename T, int M, int V>
__device__ __forceinline__ int xidx(int e)
{
  if constexpr(sizeof(T) == 8 && V == 8 && M >= 128)
    return e ^ ((e >> 4) & 7) ^ (((e >> 6) & 1) << 3);
  else if constexpr(sizeof(T) == 4 && V == 8 && M >= 256)
    return e ^ ((e >> 5) & 7) ^ (((e >> 6) & 3) << 3);
  else if constexpr(sizeof(T) == 4 && V == 16 && M >= 512)
    return e ^ ((e >> 4) & 7) ^ (((e >> 7) & 3) << 3);
  else
    return padr(e);
}

// Twiddle table (M - 1 complex words) from zN of a dimension with N = M q.
template<typename T, int M, int V=16>
__device__ __forceinline__ void fill_twr(cplx<T> *tw, const cplx<T> *zN, int q,
                                         int tid, int nthreads)
{
  typedef RPlan<M,V> P;
  for(int e=tid; e < M-1; e += nthreads) {
    // locate the stage of entry e (constants per unrolled k)
    int Ns=P::ns(1), R=P::radix(1), off=0;
#pragma unroll
    for(int k=2; k < P::NST; ++k)
      if(e >= P::off(k)) { Ns=P::ns(k); R=P::radix(k); off=P::off(k); }
    const int rel=e-off;
    const int tp=rel/Ns+1, kp=rel-(tp-1)*Ns;
    // e^{2 pi i tp kp/(Ns R)} = zN[q tp kp M/(Ns R)]
    tw[e]=zN[(int64_t) q*tp*kp*(M/(Ns*R))];
  }
}

// Radix-R butterfly (R = 2, 4, 8, 16) with kernel e^{S 2 pi i ut/R}.
template<typename T, int R, int S>
__device__ __forceinline__ void bfly(cplx<T> *u)
{
  if constexpr(R == 16) dft16<T,S>(u);
  else DFT<T,R>::run(u,S);
}

// Synchronisation of the threads of a sequence: the wave when a sequence
// does not span waves (its TPS threads are consecutive lanes), else the block.
template<int TPS, bool BLK=false>
__device__ __forceinline__ void rsync()
{
  if constexpr(TPS <= 64 && !BLK) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE,"wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE,"wavefront");
  } else
    __syncthreads();
}

// Word of element e in a one-round exchange area (the default index map).
template<typename T, int M, int V> struct RowXI {
  static __device__ __forceinline__ int at(int e) { return xidx<T,M,V>(e); }
};

// Stage s of the transform (see the top of this file); recurses into s+1.
// CX: xb holds 2 PAD words, real parts in the first PAD and imaginary parts in
// the second, and each exchange is one round (two synchronisations instead
// of four); element e sits at word XI::at(e) ^ xm of each half.
template<typename T, int M, int V, int S, int s, bool BLK, bool CX=false,
         class XI=RowXI<T,M,V>>
__device__ __forceinline__ void fftr_stage(cplx<T> *v, T *xb, const cplx<T> *tw,
                                           int t, int xm=0)
{
  typedef RPlan<M,V> P;
  constexpr int TPS=P::TPS;
  constexpr int R=P::radix(s), NB=V/R, Ns=P::ns(s), OFF=P::off(s);
  // butterflies in place: slot b + t' NB holds input t' / output u = t'
#pragma unroll
  for(int b=0; b < NB; ++b) {
    const int j=t+TPS*b;
    const int kp=j & (Ns-1);
    cplx<T> u[R];
#pragma unroll
    for(int tp=0; tp < R; ++tp) u[tp]=v[b+tp*NB];
    if constexpr(s > 0) {
      const cplx<T> *w=tw+OFF+kp;
#pragma unroll
      for(int tp=1; tp < R; ++tp) {
        const cplx<T> z=w[(tp-1)*Ns];
        u[tp]=S > 0 ? u[tp]*z : mulc(u[tp],z);
      }
    }
    bfly<T,R,S>(u);
#pragma unroll
    for(int tp=0; tp < R; ++tp) v[b+tp*NB]=u[tp];
  }
  if constexpr(s+1 < P::NST && CX) {
    // exchange: output u of butterfly b goes to (j/Ns) Ns R + k' + u Ns
#pragma unroll
    for(int b=0; b < NB; ++b) {
      const int j=t+TPS*b;
      const int kp=j & (Ns-1);
      const int base=(j-kp)*R+kp;
#pragma unroll
      for(int u=0; u < R; ++u) {
        const int e=XI::at(base+u*Ns) ^ xm;
        xb[e]=v[b+u*NB].x;
        xb[P::PAD+e]=v[b+u*NB].y;
      }
    }
    rsync<TPS,BLK>();
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int e=XI::at(t+TPS*k) ^ xm;
      v[k]=mk(xb[e],xb[P::PAD+e]);
    }
    rsync<TPS,BLK>();
    fftr_stage<T,M,V,S,s+1,BLK,CX,XI>(v,xb,tw,t,xm);
  } else if constexpr(s+1 < P::NST) {
    // exchange: output u of butterfly b goes to (j/Ns) Ns R + k' + u Ns
#pragma unroll
    for(int part=0; part < 2; ++part) {
#pragma unroll
      for(int b=0; b < NB; ++b) {
        const int j=t+TPS*b;
        const int kp=j & (Ns-1);
        const int base=(j-kp)*R+kp;
#pragma unroll
        for(int u=0; u < R; ++u)
          xb[padr(base+u*Ns)]=part == 0 ? v[b+u*NB].x : v[b+u*NB].y;
      }
      rsync<TPS,BLK>();
#pragma unroll
      for(int k=0; k < V; ++k) {
        const T val=xb[padr(t+TPS*k)];
        if(part == 0) v[k].x=val;
        else v[k].y=val;
      }
      rsync<TPS,BLK>();
    }
    fftr_stage<T,M,V,S,s+1,BLK>(v,xb,tw,t);
  }
}

// Stage s of two transforms of the same size and sign at once (one-round
// exchanges, the default index map): the butterflies of both sequences are
// independent work for the same waves, then each is exchanged in turn through
// the one area (the same barriers as two fftr calls).
template<typename T, int M, int V, int S, int s>
__device__ __forceinline__ void fftr_pair_stage(cplx<T> *X, cplx<T> *Y, T *xb,
                                                const cplx<T> *tw, int t)
{
  typedef RPlan<M,V> P;
  typedef RowXI<T,M,V> XI;
  constexpr int TPS=P::TPS;
  constexpr int R=P::radix(s), NB=V/R, Ns=P::ns(s), OFF=P::off(s);
#pragma unroll
  for(int b=0; b < NB; ++b) {
    const int j=t+TPS*b;
    const int kp=j & (Ns-1);
    cplx<T> u[R], v[R];
#pragma unroll
    for(int tp=0; tp < R; ++tp) {
      u[tp]=X[b+tp*NB];
      v[tp]=Y[b+tp*NB];
    }
    if constexpr(s > 0) {
      const cplx<T> *w=tw+OFF+kp;
#pragma unroll
      for(int tp=1; tp < R; ++tp) {
        const cplx<T> z=w[(tp-1)*Ns];
        u[tp]=S > 0 ? u[tp]*z : mulc(u[tp],z);
        v[tp]=S > 0 ? v[tp]*z : mulc(v[tp],z);
      }
    }
    bfly<T,R,S>(u);
    bfly<T,R,S>(v);
#pragma unroll
    for(int tp=0; tp < R; ++tp) {
      X[b+tp*NB]=u[tp];
      Y[b+tp*NB]=v[tp];
    }
  }
  if constexpr(s+1 < P::NST) {
#pragma unroll
    for(int pass=0; pass < 2; ++pass) {
      cplx<T> *Z=pass == 0 ? X : Y;
#pragma unroll
      for(int b=0; b < NB; ++b) {
        const int j=t+TPS*b;
        const int kp=j & (Ns-1);
        const int base=(j-kp)*R+kp;
#pragma unroll
        for(int u=0; u < R; ++u) {
          const int e=XI::at(base+u*Ns);
          xb[e]=Z[b+u*NB].x;
          xb[P::PAD+e]=Z[b+u*NB].y;
        }
      }
      rsync<TPS>();
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int e=XI::at(t+TPS*k);
        Z[k]=mk(xb[e],xb[P::PAD+e]);
      }
      rsync<TPS>();
    }
    fftr_pair_stage<T,M,V,S,s+1>(X,Y,xb,tw,t);
  }
}

// fftr<T,M,V,S,true> of X and of Y, their butterflies interleaved stage by
// stage; same results, same exchange area and barriers.
template<typename T, int M, int V, int S>
__device__ __forceinline__ void fftr_pair(cplx<T> *X, cplx<T> *Y, T *xb, const cplx<T> *tw,
                                          int t)
{
  fftr_pair_stage<T,M,V,S,0>(X,Y,xb,tw,t);
}

// In-place transform of the sequence held by threads t < TPS (v[V] each)
// with kernel e^{S 2 pi i jk/M}; xb is the sequence's exchange area
// (RPlan<M,V>::PAD scalar words), tw the twiddle table of RPlan<M,V>.  When
// TPS > 64 every thread of the block must call this.
template<typename T, int M, int V, int S, bool CX=false>
__device__ __forceinline__ void fftr(cplx<T> *v, T *xb, const cplx<T> *tw, int t)
{
  fftr_stage<T,M,V,S,0,false,CX>(v,xb,tw,t);
}

// The same transform when the threads of a sequence are not consecutive
// lanes (block barriers; every thread of the block must call it).
template<typename T, int M, int V, int S, bool CX=false>
__device__ __forceinline__ void fftr_blk(cplx<T> *v, T *xb, const cplx<T> *tw,
                                         int t)
{
  fftr_stage<T,M,V,S,0,true,CX>(v,xb,tw,t);
}

// fftr_blk with one-round exchanges through the index map XI, XORed by xm.
template<typename T, int M, int V, int S, class XI>
__device__ __forceinline__ void fftr_blkx(cplx<T> *v, T *xb, const cplx<T> *tw,
                                          int t, int xm)
{
  fftr_stage<T,M,V,S,0,true,true,XI>(v,xb,tw,t,xm);
}

} // namespace fftwpp_amd
