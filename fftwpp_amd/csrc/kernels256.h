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

// Column kernels: lane (c, j) = (threadIdx.x % TC, threadIdx.x / TC), so
// consecutive lanes read consecutive columns; the 16 lanes j of column c
// share an exchange area of CS scalar words (block barriers, since they span
// waves).  TC, CS and the swizzle h(c) make every LDS exchange access
// bank-conflict free on gfx950 (64-bit writes are served 16 lanes per pass
// over 32 banks, 64-bit reads 32 lanes per pass over 64 banks, 32-bit
// accesses 32 lanes per pass over 32 banks): FP64 uses 8 columns, areas 276
// words apart and h = 2 for columns 4..7; FP32 uses 16 columns, areas 274
// words apart and no swizzle.
template<typename T> struct Col256 {
  static constexpr int TC=sizeof(T) == 8 ? 8 : 16;
  static constexpr int CS=sizeof(T) == 8 ? 276 : 274;
  __device__ static int swz(int c) { return sizeof(T) == 8 ? ((c >> 1) & 2) : 0; }
};

// Dynamic LDS bytes of the column kernels for nres residues: residue
// twiddles (256 nres complex), stage twiddles (256 complex), exchange areas.
template<typename T>
inline size_t col256_lds(size_t nres)
{
  return (256*nres+256)*sizeof(cplx<T>)+Col256<T>::TC*Col256<T>::CS*sizeof(T);
}

// Dynamic LDS bytes of the row kernel with NR rows per block and q residues.
template<typename T>
inline size_t row256_lds(int NR, size_t q)
{
  return (256*q+256)*sizeof(cplx<T>)+NR*16*17*sizeof(T);
}

// Forward residue transforms of one column: for each residue r = r0 + rr,
// F_r(s) = DFT_256[w_N^{r k} f(k)](s), written at Fb + rr*256*SF.
template<typename T>
__device__ __forceinline__ void colfwd_body(const cplx<T> *fb, cplx<T> *Fb,
                                            int64_t S, int64_t SF, bool ok,
                                            int c, int j, int L, int64_t r0,
                                            int64_t nres, const cplx<T> *rtab,
                                            const cplx<T> *tw16, T *row, int h)
{
  const unsigned lin=(unsigned) (c+j*S);
  const unsigned lout=(unsigned) (c+j*SF);
  // the column is read once and kept in registers for all residues
  cplx<T> x0[16];
#pragma unroll
  for(int k=0; k < 16; ++k) {
    const cplx<T> *bk=fb+(int64_t) 16*k*S;
    x0[k]=(ok && j+16*k < L) ? bk[lin] : zero<T>();
  }
  for(int64_t rr=0; rr < nres; ++rr) {
    const int64_t r=r0+rr;
    const cplx<T> *rt=rtab+256*rr;
    cplx<T> v[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      v[k]=r == 0 ? x0[k] : x0[k]*(rt+16*k)[j];
    }
    fft256s<T,1,0>(v,row,j,h,tw16);
    cplx<T> *Fr=Fb+rr*256*SF;
    if(ok) {
#pragma unroll
      for(int u=0; u < 16; ++u) (Fr+(int64_t) 16*u*SF)[lout]=v[u];
    }
  }
}

// Backward residue transforms of one column, accumulated over the residues:
// f(k) = scale * sum_rr w_N^{-r k} DFT_256^{-1}[F_r](k) (+ f(k) if accumulate).
template<typename T>
__device__ __forceinline__ void colbwd_body(const cplx<T> *Fb, cplx<T> *fb,
                                            int64_t S, int64_t SF, bool ok,
                                            int c, int j, int L, int64_t r0,
                                            int64_t nres, int accumulate,
                                            T scale, const cplx<T> *rtab,
                                            const cplx<T> *tw16, T *row, int h)
{
  const unsigned lin=(unsigned) (c+j*SF);
  const unsigned lout=(unsigned) (c+j*S);
  cplx<T> acc[16];
#pragma unroll
  for(int k=0; k < 16; ++k) acc[k]=zero<T>();
  for(int64_t rr=0; rr < nres; ++rr) {
    const int64_t r=r0+rr;
    const cplx<T> *rt=rtab+256*rr;
    const cplx<T> *Fr=Fb+rr*256*SF;
    cplx<T> v[16];
#pragma unroll
    for(int u=0; u < 16; ++u)
      v[u]=ok ? (Fr+(int64_t) 16*u*SF)[lin] : zero<T>();
    fft256s<T,-1,0>(v,row,j,h,tw16);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? v[k] : mulc(v[k],(rt+16*k)[j]));
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

template<typename T>
__global__ void __launch_bounds__(16*Col256<T>::TC,sizeof(T) == 4 ? 3 : 2)
colfwd256_kernel(const cplx<T> *f, int64_t S, int64_t C, int64_t bsi,
                 cplx<T> *F, int64_t SF, int64_t bso, int L, int64_t r0,
                 int64_t nres, const cplx<T> *zN, int64_t q)
{
  constexpr int TC=Col256<T>::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *rtab=reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tw16=rtab+256*nres;
  fill_rtab(rtab,zN,(int) r0,(int) nres,threadIdx.x,16*TC);
  fill_tw16(tw16,zN,(int) q,threadIdx.x,16*TC);
  __syncthreads();
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  T *row=reinterpret_cast<T *>(tw16+256)+c*Col256<T>::CS;
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const cplx<T> *fb=f+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *Fb=F+(int64_t) blockIdx.y*bso+c0;
  colfwd_body<T>(fb,Fb,S,SF,ok,c,j,L,r0,nres,rtab,tw16,row,Col256<T>::swz(c));
}

template<typename T>
__global__ void __launch_bounds__(16*Col256<T>::TC,sizeof(T) == 4 ? 3 : 2)
colbwd256_kernel(const cplx<T> *F, int64_t SF, int64_t bsi, cplx<T> *f,
                 int64_t S, int64_t C, int64_t bso, int L, int64_t r0,
                 int64_t nres, int accumulate, T scale, const cplx<T> *zN,
                 int64_t q)
{
  constexpr int TC=Col256<T>::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *rtab=reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tw16=rtab+256*nres;
  fill_rtab(rtab,zN,(int) r0,(int) nres,threadIdx.x,16*TC);
  fill_tw16(tw16,zN,(int) q,threadIdx.x,16*TC);
  __syncthreads();
  const int c=threadIdx.x % TC, j=threadIdx.x/TC;
  T *row=reinterpret_cast<T *>(tw16+256)+c*Col256<T>::CS;
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const cplx<T> *Fb=F+(int64_t) blockIdx.y*bsi+c0;
  cplx<T> *fb=f+(int64_t) blockIdx.y*bso+c0;
  colbwd_body<T>(Fb,fb,S,SF,ok,c,j,L,r0,nres,accumulate,scale,rtab,tw16,row,
                 Col256<T>::swz(c));
}

// Rows: NR rows per block of 16*NR lanes; lane j of row i = threadIdx.x % 16,
// so the 16 lanes of a row share a wave and exchange through LDS without
// block barriers (row areas 16*17 words apart are conflict free).  The block
// first stages the residue twiddles (256 q) and the stage twiddles (256) in
// LDS.  Dynamic LDS: row256_lds().
// KEEP: the rows of f and g are read once and kept in registers for all
// residues (FP32); otherwise each residue reads them again (FP64, where
// keeping them would leave one wave per SIMD).
template<typename T> struct Row256 {
  static constexpr bool KEEP=sizeof(T) == 4;
  static constexpr int MINW=2;
};
template<typename T, class Op, int NR>
__global__ void __launch_bounds__(16*NR,Row256<T>::MINW)
rowconv256_kernel(const cplx<T> *f, const cplx<T> *g, int64_t rows, int64_t rs,
                  cplx<T> *out, int64_t os, int L, T scale, const cplx<T> *zN,
                  int64_t q, Op op)
{
  extern __shared__ char smem_raw[];
  const int qi=(int) q;
  cplx<T> *rtab=reinterpret_cast<cplx<T> *>(smem_raw);   // 256 q
  cplx<T> *tw16=rtab+256*qi;                              // 256
  T *lds=reinterpret_cast<T *>(tw16+256);
  fill_rtab(rtab,zN,0,qi,threadIdx.x,16*NR);
  fill_tw16(tw16,zN,qi,threadIdx.x,16*NR);
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
  constexpr bool KEEP=Row256<T>::KEEP;
  cplx<T> a0[KEEP ? 16 : 1], b0[KEEP ? 16 : 1];
  if constexpr(KEEP) {
#pragma unroll
    for(int k=0; k < 16; ++k) {
      bool in=ok && j+16*k < L;
      a0[k]=in ? (fr+16*k)[lin] : zero<T>();
      b0[k]=in ? (gr+16*k)[lin] : zero<T>();
    }
  }
  for(int r=0; r < qi; ++r) {
    const cplx<T> *rt=rtab+256*r;
    cplx<T> a[16], b[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      cplx<T> x, y;
      if constexpr(KEEP) {
        x=a0[k];
        y=b0[k];
      } else {
        bool in=ok && j+16*k < L;
        x=in ? (fr+16*k)[lin] : zero<T>();
        y=in ? (gr+16*k)[lin] : zero<T>();
      }
      if(r != 0) {
        cplx<T> w=(rt+16*k)[j];
        x=x*w;
        y=y*w;
      }
      a[k]=x;
      b[k]=y;
    }
    fft256s<T,1,1>(a,row,j,0,tw16);
    fft256s<T,1,1>(b,row,j,0,tw16);
    {
      const auto rop=bind_row(op,row0+i);
#pragma unroll
      for(int k=0; k < 16; ++k) a[k]=rop(a[k],b[k],qi*(j+16*k)+r);
    }
    fft256s<T,-1,1>(a,row,j,0,tw16);
#pragma unroll
    for(int k=0; k < 16; ++k)
      acc[k]=acc[k]+(r == 0 ? a[k] : mulc(a[k],(rt+16*k)[j]));
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < 16; ++k)
      if(j+16*k < L) (orow+16*k)[lo]=scale*acc[k];
  }
}

// Sum of a value over the two 16-lane groups of each 32-lane half of the
// wave (lanes l and l ^ 16): v_permlane16_swap exchanges the odd 16-lane
// rows of its first operand with the even rows of its second, so with both
// operands v the two results hold v of the group and v of its partner.
__device__ __forceinline__ float pair16_sum(float v)
{
  auto s=__builtin_amdgcn_permlane16_swap(__float_as_uint(v),__float_as_uint(v),
                                          false,false);
  return __uint_as_float(s[0])+__uint_as_float(s[1]);
}
__device__ __forceinline__ double pair16_sum(double v)
{
  const unsigned long long u=(unsigned long long) __double_as_longlong(v);
  auto lo=__builtin_amdgcn_permlane16_swap((unsigned) u,(unsigned) u,false,false);
  auto hi=__builtin_amdgcn_permlane16_swap((unsigned) (u >> 32),(unsigned) (u >> 32),
                                           false,false);
  const double a=__longlong_as_double((long long) (((unsigned long long) hi[0] << 32) | lo[0]));
  const double b=__longlong_as_double((long long) (((unsigned long long) hi[1] << 32) | lo[1]));
  return a+b;
}

// Rows with q = 2 (N = 512): a wave holds 2 rows x 2 residues x 16 lanes,
// lane = j + 16 (2 i + r) for row i of the wave and residue r.  The two
// residue groups of a row load the same elements in the same instruction, so
// f and g are read from memory once; each group transforms its residue and
// the contributions are summed across the group pair in registers.  Group r
// then stores the outputs j + 16 k with k in [8 r, 8 r + 8).  The grid is
// persistent: each block fills its twiddle tables once and its waves, which
// exchange only within themselves, loop over row pairs independently.
// Dynamic LDS: row256_lds(4 WPB, 2).
template<typename T, class Op, int WPB>
__global__ void __launch_bounds__(64*WPB,2)
rowconv256p_kernel(const cplx<T> *f, const cplx<T> *g, int64_t rows,
                   int64_t rs, cplx<T> *out, int64_t os, int L, T scale,
                   const cplx<T> *zN, Op op)
{
  extern __shared__ char smem_raw[];
  cplx<T> *rtab=reinterpret_cast<cplx<T> *>(smem_raw);   // 2 x 256
  cplx<T> *tw16=rtab+512;                                 // 256
  T *lds=reinterpret_cast<T *>(tw16+256);
  fill_rtab(rtab,zN,0,2,threadIdx.x,64*WPB);
  fill_tw16(tw16,zN,2,threadIdx.x,64*WPB);
  __syncthreads();
  const int j=threadIdx.x & 15;
  const int grp=threadIdx.x >> 4;          // 4 groups per wave
  const int r=grp & 1;
  const int wave=threadIdx.x >> 6;
  const cplx<T> *rt=rtab+256*r;
  constexpr int RS=17;                     // row stride of the exchange areas
  T *xrow=lds+grp*16*RS;
  const int64_t pairs=(rows+1)/2;
  for(int64_t pr=(int64_t) blockIdx.x*WPB+wave; pr < pairs; pr += (int64_t) gridDim.x*WPB) {
    const int64_t row=2*pr+((grp & 3) >> 1);
    const bool ok=row < rows;
    const cplx<T> *fr=f+(ok ? row : 0)*rs;
    const cplx<T> *gr=g+(ok ? row : 0)*rs;
    cplx<T> a[16], b[16];
#pragma unroll
    for(int k=0; k < 16; ++k) {
      bool in=ok && j+16*k < L;
      cplx<T> x=in ? (fr+16*k)[j] : zero<T>();
      cplx<T> y=in ? (gr+16*k)[j] : zero<T>();
      cplx<T> w=(rt+16*k)[j];                 // w_512^{r (j + 16 k)}
      a[k]=x*w;
      b[k]=y*w;
    }
    fft256s_pair<T,1,1,RS>(a,b,xrow,j,0,tw16);
    {
      const auto rop=bind_row(op,row);
#pragma unroll
      for(int k=0; k < 16; ++k) a[k]=rop(a[k],b[k],2*(j+16*k)+r);
    }
    fft256s<T,-1,1,RS>(a,xrow,j,0,tw16);
#pragma unroll
    for(int k=0; k < 16; ++k) {
      cplx<T> y=mulc(a[k],(rt+16*k)[j]);
      a[k]=mk(pair16_sum(y.x),pair16_sum(y.y));
    }
    if(ok) {
      cplx<T> *orow=out+row*os;
#pragma unroll
      for(int k=0; k < 16; ++k)
        if((k >> 3) == r && j+16*k < L) (orow+16*k)[j]=scale*a[k];
    }
  }
}

} // namespace fftwpp_amd
