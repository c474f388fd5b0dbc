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
// Work decomposition: a tile is TR consecutive items and a residue group is
// D consecutive residues; a work unit is (tile, group) and blocks take units
// blockIdx.x, blockIdx.x + gridDim.x, ... (a persistent grid that drains
// the class).  The sum over a unit's residues accumulates in LDS when the
// tile's outputs fit there (LACC) and otherwise in its destination, which
// the unit owns exclusively: the output itself when one group covers all q
// residues (G = 1), else the group's slice of the partial sums, which
// batch_reduce_kernel adds in group order (deterministic).
#pragma once
#include "common.h"
#include "fft_lds.h"
#include "fftr.h"

namespace fftwpp_amd {

struct BatchItem {
  int64_t off;   // first element of the item in the packed arrays
  int64_t L;     // item length
  int64_t poff;  // first element of the item in its class's partial sums
};

// LDS bytes of the item table of a tile, rounded up to 16-byte alignment.
__host__ __device__ inline size_t batch_items_bytes(size_t TR)
{
  return (TR*sizeof(BatchItem)+15) & ~(size_t) 15;
}

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
// part: G = ceil(q/D) slices of ptotal words (unused when G = 1).
template<typename T, int MULT, bool LACC>
__global__ void __launch_bounds__(1024)
batch_conv_kernel(DimArgs d, const BatchItem *items, int64_t count,
                  const cplx<T> *f, const cplx<T> *g, cplx<T> *out, T scale,
                  int TR, int ld, int lacc, int D, cplx<T> *part,
                  int64_t ptotal)
{
  extern __shared__ char smem_raw[];
  BatchItem *its=reinterpret_cast<BatchItem *>(smem_raw);
  cplx<T> *A=reinterpret_cast<cplx<T> *>(smem_raw+batch_items_bytes(TR));
  cplx<T> *B=A+(int64_t) 2*TR*ld;
  cplx<T> *acc=B+(int64_t) 2*TR*ld;
  const int m=(int) d.m;
  const int tid=threadIdx.x, nt=blockDim.x;
  const int64_t ntiles=(count+TR-1)/TR;
  const int64_t G=(d.q+D-1)/D;
  for(int64_t unit=blockIdx.x; unit < ntiles*G; unit += gridDim.x) {
    const int64_t tile=unit/G, grp=unit-tile*G;
    const int64_t rbeg=grp*D, rend=min(d.q,rbeg+D);
    const T sc=G == 1 ? scale : (T) 1;
    const int64_t i0=tile*TR;
    const int nr=(int) min((int64_t) TR,count-i0);
    for(int i=tid; i < nr; i += nt) its[i]=items[i0+i];
    __syncthreads();
    int Lmax=0;
    for(int i=0; i < nr; ++i) Lmax=max(Lmax,(int) its[i].L);
    for(int64_t r=rbeg; r < rend; ++r) {
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
      const bool first=r == rbeg, last=r == rend-1;
      for(int idx=tid; idx < nr*Lmax; idx += nt) {
        const int i=idx/Lmax, j=idx-i*Lmax;
        const BatchItem it=its[i];
        if(j >= it.L) continue;
        cplx<T> v=Z[(int64_t) i*ld+(j % m)];
        if(r != 0) v=mulc(v,zeta<T>(d,(int64_t) j*r));
        cplx<T> *dst=G == 1 ? out+it.off+j : part+grp*ptotal+it.poff+j;
        if(LACC) {
          cplx<T> *a=acc+(int64_t) i*lacc+j;
          const cplx<T> sum=first ? v : *a+v;
          if(last) *dst=sc*sum;
          else *a=sum;
        } else {
          const cplx<T> sum=first ? v : *dst+v;
          *dst=last ? sc*sum : sum;
        }
      }
      __syncthreads();
    }
  }
}

// a[k] = f[j_k], b[k] = g[j_k] at j_k = min(j0 + TPS k, L - 1), every load
// issued before any is used.  FP64 reads through buffer resources with 32-bit
// offsets (one address register per load instead of two: the kernel then
// stays below 256 VGPRs without spills); FP32 through plain pointers, which
// measured faster there.  Clamped offsets stay inside the item, so no range
// check is relied on.  Items of 2^27 words or more (byte offsets beyond 31
// bits) take the pointer loads in FP64 too.
template<typename T, int V, int TPS>
__device__ __forceinline__ void batch_fold_loads(const cplx<T> *f, const cplx<T> *g, int L,
                                                 int j0, cplx<T> *a, cplx<T> *b)
{
  if constexpr(sizeof(T) == 8) {
    if(L < (1 << 27)) {
      const __amdgpu_buffer_rsrc_t rf=buf_rsrc<T>(f,(size_t) L), rg=buf_rsrc<T>(g,(size_t) L);
#pragma unroll
      for(int k=0; k < V; ++k) {
        const unsigned o=(unsigned) min(j0+TPS*k,L-1)*(unsigned) sizeof(cplx<T>);
        a[k]=ld_buf<T>(rf,o,0);
        b[k]=ld_buf<T>(rg,o,0);
      }
      return;
    }
  }
#pragma unroll
  for(int k=0; k < V; ++k) {
    const int j=min(j0+TPS*k,L-1);
    a[k]=f[j];
    b[k]=g[j];
  }
}

// Values per thread of the register FFT: 16 in FP32, 8 in FP64 (two
// transforms and the FFT's temporaries then fit two waves per SIMD).
template<typename T> struct BatchV { static constexpr int V=sizeof(T) == 8 ? 8 : 16; };

// One work unit of the FP64 sequence mode: residue r of the item at fi, gi
// of length L, its sequence Z_r written to z (see batch_reg_kernel).  With a
// slot of whole waves the fold reads through buffer resources over each block
// of M words (the base advanced per block, so offsets stay below M words for
// any L; fi must be wave-uniform); otherwise through pointers.  The forward
// transforms of f and g run with their butterflies interleaved.
template<typename T, int M>
__device__ __forceinline__ void batch_zm_unit(const cplx<T> *zN, const cplx<T> *fi,
                                              const cplx<T> *gi, int L, int r, int q,
                                              const cplx<T> *tw, const cplx<T> *tq, T *xb,
                                              int t, int mult, cplx<T> *z)
{
  constexpr int V=BatchV<T>::V;
  constexpr int TPS=RPlan<M,V>::TPS;
  constexpr bool BUF=sizeof(T) == 8 && TPS >= 64;
  cplx<T> X[V], Y[V];
#pragma unroll
  for(int k=0; k < V; ++k) X[k]=Y[k]=mk((T) 0,(T) 0);
  for(int jb=0, eb=0; jb < L; jb += M) {
    const cplx<T> wb=zN[eb*M];                 // w_q^{b r}, eb = b r mod q
    const int rem=min(L-jb,M);                 // words of this block
    cplx<T> a[V], b[V];
    if constexpr(BUF) {
      const __amdgpu_buffer_rsrc_t rf=buf_rsrc<T>(fi+jb,(size_t) rem);
      const __amdgpu_buffer_rsrc_t rg=buf_rsrc<T>(gi+jb,(size_t) rem);
#pragma unroll
      for(int k=0; k < V; ++k) {
        const unsigned o=(unsigned) min(t+TPS*k,rem-1)*(unsigned) sizeof(cplx<T>);
        a[k]=ld_buf<T>(rf,o,0);
        b[k]=ld_buf<T>(rg,o,0);
      }
    } else {
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int j=jb+min(t+TPS*k,rem-1);
        a[k]=fi[j];
        b[k]=gi[j];
      }
    }
#pragma unroll
    for(int k=0; k < V; ++k) {
      const cplx<T> w=t+TPS*k < rem ? wb : mk((T) 0,(T) 0);
      X[k]=X[k]+w*a[k];
      Y[k]=Y[k]+w*b[k];
    }
    eb += r;
    if(eb >= q) eb -= q;
  }
  const cplx<T> bt=zN[t*r];
  const cplx<T> *tr=tq+r*V;
  if(r) {
#pragma unroll
    for(int k=0; k < V; ++k) {
      const cplx<T> w=bt*tr[k];               // w_N^{sigma r}
      X[k]=w*X[k];
      Y[k]=w*Y[k];
    }
  }
  fftr_pair<T,M,V,1>(X,Y,xb,tw,t);
#pragma unroll
  for(int k=0; k < V; ++k)
    X[k]=mult == 1 ? X[k]*Y[k] : (mult == 3 ? mulc(X[k],Y[k]) : X[k]);
  fftr<T,M,V,-1,true>(X,xb,tw,t);
  if(r) {
#pragma unroll
    for(int k=0; k < V; ++k) X[k]=mulc(X[k],bt*tr[k]);
  }
#pragma unroll
  for(int k=0; k < V; ++k) z[t+TPS*k]=X[k];
}

// The same work units with the register-resident FFT (fftr.h) of a
// power-of-two sub-FFT size M: a block has blockDim.x/TPS sequence slots of
// TPS = M/V threads and each slot convolves one work unit (item, residue
// group) at a time; the sums over residues accumulate in the unit's
// destination (thread-private words: thread t owns j = t + TPS k + M b).
// When TPS > 64 the host launches one slot per block, so the block barriers
// inside fftr() see the same calls from every thread.  Dynamic LDS: M
// complex twiddles + slots * 2 RPlan<M,V>::PAD scalar words (one-round
// exchanges) + q V complex residue twiddles.
//
// D = 0 selects the sequence mode: a work unit is one residue r of an item
// and writes its inverse-transformed, conjugate-twiddled sequence
// Z_r(sigma) = w_N^{-sigma r} IFFT_r(sigma), sigma < M, to part[i q M + r M]
// (N words per item instead of partial sums of the unfolded outputs);
// batch_unfold_kernel then forms out(sigma + M b) = sum_r w_q^{-b r} Z_r(sigma),
// the backward fold over the blocks b of the reference's backward transforms
// (convolve.cc:1468-1613, 1765-1965), so no output word is read and written
// once per residue.
template<typename T, int M, bool ZM=false>
__global__ void __launch_bounds__(512)
batch_reg_kernel(DimArgs d, const BatchItem *items, int64_t count,
                 const cplx<T> *f, const cplx<T> *g, cplx<T> *out, T scale,
                 int mult, int D, cplx<T> *part, int64_t ptotal)
{
  constexpr int V=BatchV<T>::V;
  typedef RPlan<M,V> P;
  constexpr int TPS=P::TPS;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,(int) d.q,threadIdx.x,blockDim.x);
  const int slots=blockDim.x/TPS;
  // residue twiddles w_N^{(t + TPS k) r} = w_N^{t r} w_N^{TPS k r}: the
  // uniform factor from an LDS table tq[r V + k] after the exchange areas
  cplx<T> *tq=reinterpret_cast<cplx<T> *>(reinterpret_cast<T *>(tw+M)+slots*2*P::PAD);
  for(int e=threadIdx.x; e < (int) d.q*V; e += blockDim.x) {
    const int r=e/V, k=e-r*V;
    tq[e]=zN[TPS*k*r];                   // TPS k r < M r < N
  }
  __syncthreads();
  const int slot=threadIdx.x/TPS, t=threadIdx.x-slot*TPS;
  T *xb=reinterpret_cast<T *>(tw+M)+slot*2*P::PAD;
  if constexpr(ZM && sizeof(T) == 8) {
    // the FP64 sequence mode (the host passes D = 0): units (item,
    // residue), dealt to XCDs as the residue groups below
    const int qi=(int) d.q;
    const int64_t units=count*qi;
    for(int64_t ub=(int64_t) blockIdx.x*slots; ub < units; ub += (int64_t) gridDim.x*slots) {
      const int64_t unit=ub+slot;
      if(unit >= units) break;          // only when slots > 1 (TPS <= 64)
      int64_t ii;
      int r;
      if(qi > 1 && slots == 1) {
        const int64_t chunk=8*(int64_t) qi, c=unit/chunk, full=count/8;
        const int64_t u=unit-c*chunk;
        const int64_t W=c < full ? 8 : count-8*full;
        ii=8*c+u % W;
        r=(int) (u/W);
      } else {
        ii=unit/qi;
        r=(int) (unit-ii*qi);
      }
      if constexpr(TPS >= 64) {
        // a slot spans whole waves: its unit, and the item's base, are
        // uniform over the wave
        ii=__builtin_amdgcn_readfirstlane((int) ii);
        r=__builtin_amdgcn_readfirstlane(r);
      }
      const BatchItem it=items[ii];
      cplx<T> *z=part+(ii*qi+r)*(int64_t) M;
      batch_zm_unit<T,M>(zN,f+it.off,g+it.off,(int) it.L,r,qi,tw,tq,xb,t,mult,z);
    }
    return;
  }
  constexpr bool zm=ZM;                  // the FP32 sequence mode (the host passes D = 0)
  if(zm) D=1;
  const int64_t G=(d.q+D-1)/D;
  const int64_t units=count*G;
  for(int64_t ub=(int64_t) blockIdx.x*slots; ub < units; ub += (int64_t) gridDim.x*slots) {
    const int64_t unit=ub+slot;
    if(unit >= units) break;            // only when slots > 1 (TPS <= 64)
    int64_t ii, grp;
    if(G > 1 && slots == 1) {
      // Blocks are dealt to the 8 XCDs round robin (block b on XCD b mod 8)
      // and the host makes the grid a multiple of 8, so unit u runs on XCD
      // u mod 8.  Units c 8G + g 8 + w (g < G) are the G residue groups of
      // item 8c + w: they run together on one XCD and share its L2 for
      // their reads of the item.  The last, partial chunk of rem < 8 items
      // maps u' = g rem + w.
      const int64_t chunk=8*G, c=unit/chunk, full=count/8;
      const int64_t u=unit-c*chunk;
      const int64_t W=c < full ? 8 : count-8*full;
      ii=8*c+u % W;
      grp=u/W;
    } else {
      ii=unit/G;
      grp=unit-ii*G;
    }
    const BatchItem it=items[ii];
    const T sc=G == 1 ? scale : (T) 1;
    cplx<T> *dst=G == 1 ? out+it.off : part+grp*ptotal+it.poff;
    const cplx<T> *fi=f+it.off, *gi=g+it.off;
    const int L=(int) it.L;
    const int qi=(int) d.q;
    const int rbeg=(int) (grp*D), rend=(int) min(d.q,(int64_t) rbeg+D);
    for(int r=rbeg; r < rend; ++r) {
      // w_N^{(sigma + M b) r} = w_N^{sigma r} w_q^{b r}: one per-thread factor
      // and one factor shared by the slot per block b
      cplx<T> X[V], Y[V];
#pragma unroll
      for(int k=0; k < V; ++k) X[k]=Y[k]=mk((T) 0,(T) 0);
      for(int jb=0, eb=0; jb < L; jb += M) {
        const cplx<T> wb=zN[eb*M];               // w_q^{b r}, eb = b r mod q
        // all 2 V loads of a block issued before the first use: indices
        // clamped into the item, the terms beyond it zeroed by their factor
        // (per-element branches made the loads wait one by one)
        cplx<T> a[V], b[V];
        batch_fold_loads<T,V,TPS>(fi,gi,L,jb+t,a,b);
#pragma unroll
        for(int k=0; k < V; ++k) {
          const cplx<T> w=jb+t+TPS*k < L ? wb : mk((T) 0,(T) 0);
          X[k]=X[k]+w*a[k];
          Y[k]=Y[k]+w*b[k];
        }
        eb += r;
        if(eb >= qi) eb -= qi;
      }
      const cplx<T> bt=zN[t*r];
      const cplx<T> *tr=tq+r*V;
      if(r) {
#pragma unroll
        for(int k=0; k < V; ++k) {
          const cplx<T> w=bt*tr[k];             // w_N^{sigma r}
          X[k]=w*X[k];
          Y[k]=w*Y[k];
        }
      }
      fftr<T,M,V,1,true>(X,xb,tw,t);
      fftr<T,M,V,1,true>(Y,xb,tw,t);
#pragma unroll
      for(int k=0; k < V; ++k)
        X[k]=mult == 1 ? X[k]*Y[k] : (mult == 3 ? mulc(X[k],Y[k]) : X[k]);
      fftr<T,M,V,-1,true>(X,xb,tw,t);
      if(r) {
#pragma unroll
        for(int k=0; k < V; ++k) X[k]=mulc(X[k],bt*tr[k]);
      }
      if constexpr(zm) {
        cplx<T> *z=part+(ii*qi+r)*(int64_t) M;
#pragma unroll
        for(int k=0; k < V; ++k) z[t+TPS*k]=X[k];
        continue;
      }
      const bool first=r == rbeg, last=r == rend-1;
      for(int jb=0, eb=0; jb < L; jb += M) {
        const cplx<T> wb=zN[eb*M];
        // earlier residues' sums read unconditionally (clamped), as the fold
        cplx<T> s[V];
#pragma unroll
        for(int k=0; k < V; ++k) s[k]=mk((T) 0,(T) 0);
        if(!first) {
#pragma unroll
          for(int k=0; k < V; ++k) s[k]=dst[min(jb+t+TPS*k,L-1)];
        }
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int j=jb+t+TPS*k;
          if(j < L) {
            const cplx<T> v=mulc(X[k],wb)+s[k];
            dst[j]=last ? sc*v : v;
          }
        }
        eb += r;
        if(eb >= qi) eb -= qi;
      }
    }
  }
}

// out[off + sigma + M b] = scale * sum_{r < q} w_q^{-b r} Z[i q M + r M + sigma]
// for the blocks b of each item (sequence mode of batch_reg_kernel); one
// block per item (grid-stride), threads over sigma.  With QMAX > 0 (q <= QMAX)
// a thread loads its q values Z_r(sigma) once, before the sums over the
// blocks (indices clamped to r < q, the extra terms given a zero factor);
// QMAX = 0 takes any q and re-reads them per block.
template<typename T, int QMAX>
__global__ void __launch_bounds__(256)
batch_unfold_kernel(const BatchItem *items, int64_t count, const cplx<T> *Z,
                    int q, int M, const cplx<T> *zN, cplx<T> *out, T scale)
{
  for(int64_t i=blockIdx.x; i < count; i += gridDim.x) {
    const BatchItem it=items[i];
    const cplx<T> *zi=Z+i*(int64_t) q*M;
    const int L=(int) it.L;
    for(int sg=threadIdx.x; sg < M; sg += blockDim.x) {
      if constexpr(QMAX > 0) {
        cplx<T> z[QMAX];
#pragma unroll
        for(int r=0; r < QMAX; ++r) z[r]=zi[min(r,q-1)*M+sg];
        for(int b=0, jb=0; jb+sg < L; ++b, jb += M) {
          cplx<T> v=z[0];
#pragma unroll
          for(int r=1, e=b; r < QMAX; ++r) {
            e=e >= q ? e-q : e;                    // e = b r mod q (b < q)
            const cplx<T> w=r < q ? zN[(int64_t) e*M] : mk((T) 0,(T) 0);
            v=v+mulc(z[r],w);
            e += b;
          }
          out[it.off+jb+sg]=scale*v;
        }
      } else {
        for(int b=0, jb=0; jb < L; ++b, jb += M) {
          if(jb+sg >= L) break;
          cplx<T> v=zi[sg];
          for(int r=1, e=b; r < q; ++r) {
            v=v+mulc(zi[r*M+sg],zN[(int64_t) e*M]);    // w_q^{-b r}, e = b r mod q
            e += b;
            if(e >= q) e -= q;
          }
          out[it.off+jb+sg]=scale*v;
        }
      }
    }
  }
}

// out[off + j] = scale * sum_{g < G} part[g ptotal + poff + j], one block
// per item (grid-stride), groups added in order.
template<typename T>
__global__ void __launch_bounds__(256)
batch_reduce_kernel(const BatchItem *items, int64_t count,
                    const cplx<T> *part, int64_t ptotal, int G,
                    cplx<T> *out, T scale)
{
  for(int64_t i=blockIdx.x; i < count; i += gridDim.x) {
    const BatchItem it=items[i];
    for(int64_t j=threadIdx.x; j < it.L; j += blockDim.x) {
      cplx<T> sum=part[it.poff+j];
      for(int g=1; g < G; ++g) sum=sum+part[g*ptotal+it.poff+j];
      out[it.off+j]=scale*sum;
    }
  }
}

} // namespace fftwpp_amd
