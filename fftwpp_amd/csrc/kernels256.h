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

__device__ __forceinline__ void wsync_rows()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE,"wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE,"wavefront");
}

// Dynamic LDS bytes of the column kernels for TC columns and q residues.
template<typename T>
inline size_t col256_lds(int TC, size_t q)
{
  return sizeof(T) == 4 ? 256*q*sizeof(cplx<T>)+TC*16*17*sizeof(T)
    : TC*16*17*sizeof(cplx<T>);
}

// Columns: lane (c, j) = (threadIdx.x % TC, threadIdx.x / TC) so consecutive
// lanes read consecutive columns; the 16 lanes j of column c share an LDS row
// (block barriers, since they span waves).  FP32: the block copies the
// twiddle table zN (256 q entries) into LDS and exchanges real and imaginary
// parts separately; FP64: twiddles come from the global table and complex
// words are exchanged (fewer barriers).  Dynamic LDS: see col256_lds().
// Column transforms for one block; `tab` holds the twiddles (LDS or global,
// fixed per instantiation) and `row` is this column's exchange area.
template<typename T, int TC, bool SPLIT, typename TAB>
__device__ __forceinline__ void colfwd_body(const cplx<T> *fb, cplx<T> *Fb,
                                            int64_t S, int64_t SF, bool ok,
                                            int c, int j, int L, int64_t r0,
                                            int64_t nres, TAB tab, int qi,
                                            void *row)
{
  const unsigned lin=(unsigned) (c+j*S);
  const unsigned lout=(unsigned) (c+j*SF);
  for(int64_t rr=0; rr < nres; ++rr) {
    const int r=(int) (r0+rr);
    cplx<T> v[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      const cplx<T> *bk=fb+(int64_t) 16*k*S;
      cplx<T> x=(ok && j+16*k < L) ? bk[lin] : zero<T>();
      v[k]=r == 0 ? x : x*(tab+16*k*r)[j*r];
    }
    if constexpr(SPLIT) fft256s<T,1,0>(v,(T *) row,j,tab,qi);
    else fft256<T,1,0>(v,(cplx<T> *) row,j,tab,qi);
    cplx<T> *Fr=Fb+rr*256*SF;
    if(ok) {
#pragma unroll
      for(int u=0; u < 16; ++u) (Fr+(int64_t) 16*u*SF)[lout]=v[u];
    }
  }
}

template<typename T, int TC, bool SPLIT, typename TAB>
__device__ __forceinline__ void colbwd_body(const cplx<T> *Fb, cplx<T> *fb,
                                            int64_t S, int64_t SF, bool ok,
                                            int c, int j, int L, int64_t r0,
                                            int64_t nres, int accumulate,
                                            T scale, TAB tab, int qi,
                                            void *row)
{
  const unsigned lin=(unsigned) (c+j*SF);
  const unsigned lout=(unsigned) (c+j*S);
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
    if constexpr(SPLIT) fft256s<T,-1,0>(v,(T *) row,j,tab,qi);
    else fft256<T,-1,0>(v,(cplx<T> *) row,j,tab,qi);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? v[k] : mulc(v[k],(tab+16*k*r)[j*r]));
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

template<typename T, int TC>
__global__ void __launch_bounds__(16*TC,sizeof(T) == 4 ? 3 : 2)
colfwd256_kernel(const cplx<T> *f, int64_t S, int64_t C, int64_t bsi,
                 cplx<T> *F, int64_t SF, int64_t bso, int L, int64_t r0,
                 int64_t nres, const cplx<T> *zN, int64_t q)
{
  extern __shared__ char smem_raw[];
  const int qi=(int) q;
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  const int64_t c0=(int64_t) blockIdx.x*TC;
  const bool ok=c0+c < C;
  const cplx<T> *fb=f+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *Fb=F+(int64_t) blockIdx.y*bso+c0;
  if constexpr(sizeof(T) == 4) {
    cplx<T> *tab=reinterpret_cast<cplx<T> *>(smem_raw);
    for(int n=threadIdx.x; n < 256*qi; n += 16*TC) tab[n]=zN[n];
    __syncthreads();
    T *row=reinterpret_cast<T *>(tab+256*qi)+c*16*17;
    colfwd_body<T,TC,true>(fb,Fb,S,SF,ok,c,j,L,r0,nres,(const cplx<T> *) tab,qi,row);
  } else {
    cplx<T> *row=reinterpret_cast<cplx<T> *>(smem_raw)+c*16*17;
    colfwd_body<T,TC,false>(fb,Fb,S,SF,ok,c,j,L,r0,nres,zN,qi,row);
  }
}

template<typename T, int TC>
__global__ void __launch_bounds__(16*TC,sizeof(T) == 4 ? 3 : 2)
colbwd256_kernel(const cplx<T> *F, int64_t SF, int64_t bsi, cplx<T> *f,
                 int64_t S, int64_t C, int64_t bso, int L, int64_t r0,
                 int64_t nres, int accumulate, T scale, const cplx<T> *zN,
                 int64_t q)
{
  extern __shared__ char smem_raw[];
  const int qi=(int) q;
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  const int64_t c0=(int64_t) blockIdx.x*TC;
  const bool ok=c0+c < C;
  const cplx<T> *Fb=F+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *fb=f+(int64_t) blockIdx.y*bso+c0;
  if constexpr(sizeof(T) == 4) {
    cplx<T> *tab=reinterpret_cast<cplx<T> *>(smem_raw);
    for(int n=threadIdx.x; n < 256*qi; n += 16*TC) tab[n]=zN[n];
    __syncthreads();
    T *row=reinterpret_cast<T *>(tab+256*qi)+c*16*17;
    colbwd_body<T,TC,true>(Fb,fb,S,SF,ok,c,j,L,r0,nres,accumulate,scale,
                           (const cplx<T> *) tab,qi,row);
  } else {
    cplx<T> *row=reinterpret_cast<cplx<T> *>(smem_raw)+c*16*17;
    colbwd_body<T,TC,false>(Fb,fb,S,SF,ok,c,j,L,r0,nres,accumulate,scale,zN,qi,row);
  }
}

// Rows: NR rows per block of 16*NR lanes; lane j of row i = threadIdx.x % 16,
// so the 16 lanes of a row share a wave and exchange through LDS without
// block barriers.  The block first copies the twiddle table zN (N = 256 q
// entries) into LDS.  Dynamic LDS: 256 q complex + NR*16*17 scalar words.
template<typename T, int MULT, int NR>
__global__ void __launch_bounds__(16*NR,sizeof(T) == 4 ? 3 : 2)
rowconv256_kernel(const cplx<T> *f, const cplx<T> *g, int64_t rows, int64_t rs,
                  cplx<T> *out, int64_t os, int L, T scale, const cplx<T> *zN,
                  int64_t q)
{
  extern __shared__ char smem_raw[];
  cplx<T> *tab=reinterpret_cast<cplx<T> *>(smem_raw);   // 256 q twiddles
  const int qi=(int) q;
  T *lds=reinterpret_cast<T *>(tab+256*qi);
  for(int n=threadIdx.x; n < 256*qi; n += 16*NR) tab[n]=zN[n];
  __syncthreads();
  const int i=threadIdx.x/16, j=threadIdx.x % 16;
  const int64_t row0=(int64_t) blockIdx.x*NR;
  const bool ok=row0+i < rows;
  const cplx<T> *fr=f+row0*rs;
  const cplx<T> *gr=g+row0*rs;
  cplx<T> *orow=out+row0*os;
  const unsigned lin=(unsigned) (i*rs+j);
  const unsigned lo=(unsigned) (i*os+j);
  T *row=lds+i*16*17;
  cplx<T> acc[16];
#pragma unroll
  for(int k=0; k < 16; ++k) acc[k]=zero<T>();
  for(int r=0; r < qi; ++r) {
    cplx<T> a[16], b[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      bool in=ok && j+16*k < L;
      cplx<T> x=in ? (fr+16*k)[lin] : zero<T>();
      cplx<T> y=in ? (gr+16*k)[lin] : zero<T>();
      if(r != 0) {
        cplx<T> w=(tab+16*k*r)[j*r];
        x=x*w;
        y=y*w;
      }
      a[k]=x;
      b[k]=y;
    }
    fft256s<T,1,1>(a,row,j,tab,qi);
    fft256s<T,1,1>(b,row,j,tab,qi);
#pragma unroll
    for(int k=0; k < 16; ++k) {
      if(MULT == 1) a[k]=a[k]*b[k];
      else if(MULT == 3) a[k]=mulc(a[k],b[k]);
    }
    fft256s<T,-1,1>(a,row,j,tab,qi);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? a[k] : mulc(a[k],(tab+16*k*r)[j*r]));
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < 16; ++k)
      if(j+16*k < L) (orow+16*k)[lo]=scale*acc[k];
  }
}

} // namespace fftwpp_amd
