// batch.h -- fused kernel for a batch of independent 1D complex dealiased
// convolutions of uneven lengths that share one padded geometry (m, q).
//
// A convolution i of the batch is an item {off, L}: its inputs are
// f[off .. off+L) and g[off .. off+L) of packed arrays and its result goes
// to out[off .. off+L).  Every item of one launch has the same sub-FFT size
// m and residue count q, hence the same padded length N = m q >= 2L - 1 and
// the same twiddle table; the fold over the p = ceil(L/m) blocks of each item
// uses the item's own L.  This is the residue loop of the reference's
// Convolution::convolveRaw (convolve.cc:7513-7575) for one item, with the
// forward (convolve.cc:836-1030, 1032-1225: fold by w_N^{j r}, m-point DFT),
// the multiplier (convolve.cc:26-110) and the backward transform
// (convolve.cc:1468-1763: inverse m-point DFT, conjugate twiddle, accumulate
// over residues) fused in one workgroup.
//
// Work decomposition: a tile is TR consecutive items; blocks take tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... (a persistent grid that drains
// the class).  The sums over residues accumulate in LDS when the tile's
// outputs fit there (LACC) and otherwise in the output itself, which each
// tile owns exclusively.
#pragma once
#include "common.h"
#include "fft_lds.h"

namespace fftwpp_amd {

struct BatchItem {
  int64_t off;   // first element of the item in the packed arrays
  int64_t L;     // item length
};

// sum_{j = sigma mod m, j < L} w_N^{j r} f[j]
template<typename T>
__device__ __forceinline__ cplx<T> fold_len(const DimArgs &d,
                                            const cplx<T> *f, int64_t L,
                                            int64_t sigma, int64_t r)
{
  cplx<T> acc=mk((T) 0,(T) 0);
  if(r == 0) {
    for(int64_t j=sigma; j < L; j += d.m) acc=acc+f[j];
  } else {
    for(int64_t j=sigma; j < L; j += d.m) acc=acc+zeta<T>(d,j*r)*f[j];
  }
  return acc;
}

// Dynamic LDS: TR items, 4 TR sequences of ld complex words (two ping-pong
// buffers of the f and g transforms), and with LACC TR rows of lacc words.
template<typename T, int MULT, bool LACC>
__global__ void __launch_bounds__(1024)
batch_conv_kernel(DimArgs d, const BatchItem *items, int64_t count,
                  const cplx<T> *f, const cplx<T> *g, cplx<T> *out, T scale,
                  int TR, int ld, int lacc)
{
  extern __shared__ char smem_raw[];
  BatchItem *its=reinterpret_cast<BatchItem *>(smem_raw);
  cplx<T> *A=reinterpret_cast<cplx<T> *>(its+TR);
  cplx<T> *B=A+(int64_t) 2*TR*ld;
  cplx<T> *acc=B+(int64_t) 2*TR*ld;
  const int m=(int) d.m;
  const int tid=threadIdx.x, nt=blockDim.x;
  const int64_t ntiles=(count+TR-1)/TR;
  for(int64_t tile=blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i0=tile*TR;
    const int nr=(int) min((int64_t) TR,count-i0);
    for(int i=tid; i < nr; i += nt) its[i]=items[i0+i];
    __syncthreads();
    int Lmax=0;
    for(int i=0; i < nr; ++i) Lmax=max(Lmax,(int) its[i].L);
    for(int64_t r=0; r < d.q; ++r) {
      // folded, twiddled sequences: f of item i at row i, g at row nr + i
      for(int idx=tid; idx < 2*nr*m; idx += nt) {
        const int seq=idx/m, sigma=idx-seq*m;
        const int i=seq < nr ? seq : seq-nr;
        const BatchItem it=its[i];
        const cplx<T> *src=(seq < nr ? f : g)+it.off;
        A[(int64_t) seq*ld+sigma]=fold_len<T>(d,src,it.L,sigma,r);
      }
      __syncthreads();
      cplx<T> *X=lds_fft<T>(A,B,2*nr,ld,d,1);
      cplx<T> *Y=X == A ? B : A;
      if(MULT != 0) {
        for(int idx=tid; idx < nr*m; idx += nt) {
          const int i=idx/m, s=idx-i*m;
          const cplx<T> u=X[(int64_t) i*ld+s], v=X[(int64_t) (nr+i)*ld+s];
          X[(int64_t) i*ld+s]=MULT == 3 ? mulc(u,v) : u*v;
        }
        __syncthreads();
      }
      cplx<T> *Z=lds_fft<T>(X,Y,nr,ld,d,-1);
      const bool first=r == 0, last=r == d.q-1;
      for(int idx=tid; idx < nr*Lmax; idx += nt) {
        const int i=idx/Lmax, j=idx-i*Lmax;
        const BatchItem it=its[i];
        if(j >= it.L) continue;
        cplx<T> v=Z[(int64_t) i*ld+(j % m)];
        if(r != 0) v=mulc(v,zeta<T>(d,(int64_t) j*r));
        if(LACC) {
          cplx<T> *a=acc+(int64_t) i*lacc+j;
          const cplx<T> s=first ? v : *a+v;
          if(last) out[it.off+j]=scale*s;
          else *a=s;
        } else {
          cplx<T> *o=out+it.off+j;
          const cplx<T> s=first ? v : *o+v;
          *o=last ? scale*s : s;
        }
      }
      __syncthreads();
    }
  }
}

} // namespace fftwpp_amd
