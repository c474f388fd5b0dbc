// rowsreg.h -- fused row convolutions on the register-resident FFT (fftr.h)
// for a power-of-two sub-FFT size M (16..4096) when one block of M covers the
// row: complex or centred rows with L <= M, Hermitian rows with H <= M.
//
// A slot of TPS = M/V threads owns one row (thread t holds sigma = t + TPS k,
// k < V), loops over the residues with the sums kept in registers and writes
// the row once at the end, so out may equal f.  Rows are contiguous.
//
// Complex rows (Convolution::convolveRaw, convolve.cc:7513-7575, with the
// forward1/backward1 residue transforms, convolve.cc:836-1030, 1468-1613, and
// the centred folds of fftPadCentered): element j has wavenumber k = j + k0
// and sits at sigma = k mod M; residue r is DFT_M[w_N^{k r} f](s).
//
// Hermitian rows (fftPadHermitian, convolve.cc:4309-5447): residues r1, r2 =
// r1 + 1 are packed as W1 + i W2 with W_sigma = sum over k = sigma mod M,
// |k| < H, of w_N^{k r} f(k), f(-k) = conj f(k) (k = sigma and sigma - M),
// multiplied as realMultBinary (real parts, imaginary parts) and unpacked
// after the inverse with the mirror element Z(M - sigma), fetched through the
// slot's LDS exchange area.
#pragma once
#include "batch.h"

namespace fftwpp_amd {

// e mod N for -N <= e < 2N: the twiddle exponents k r of these kernels have
// |k| < M (rows) or -M <= k < 2M (columns with p <= 2) and 0 <= r < q, so
// they lie in [-N, 2N) with N = M q (kept below 2^30 by the dispatchers).
__device__ __forceinline__ int wrapN(int e, int N)
{
  return e < 0 ? e+N : (e >= N ? e-N : e);
}

// Values per thread: 8 in both precisions (the register sums over residues
// sit beside the two transforms).
struct RowV { static constexpr int V=8; };

// acc of residue lane 0 <- sum of acc over the RP residue lanes of the row,
// through the lanes' exchange areas (2 PAD words each: real parts, then
// imaginary parts).  All threads of the block call it when RP > 1.
template<typename T, int M, int RP>
__device__ __forceinline__ void rowreg_combine(cplx<T> *acc, T *xbase, T *xb,
                                               int rp, int t)
{
  if constexpr(RP > 1) {
    constexpr int V=RowV::V;
    typedef RPlan<M,V> P;
    constexpr int TPS=P::TPS;
    if(rp > 0) {
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int e=xidx<T,M,V>(t+TPS*k);
        xb[e]=acc[k].x;
        xb[P::PAD+e]=acc[k].y;
      }
    }
    __syncthreads();
    if(rp == 0) {
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int e=xidx<T,M,V>(t+TPS*k);
#pragma unroll
        for(int o=1; o < RP; ++o)
          acc[k]=acc[k]+mk(xbase[2*o*P::PAD+e],xbase[(2*o+1)*P::PAD+e]);
      }
    }
    __syncthreads();
  }
}

// Dynamic LDS: M complex twiddles + slots * 2 RPlan<M,V>::PAD scalar words
// (each slot's exchange area holds a whole complex sequence, so every
// exchange is one round).
// RP = 1: each slot owns a row (one slot per block when TPS > 64).  RP > 1:
// the block holds one row and its RP slots take every RP-th residue (pair);
// their sums are added through LDS before the row is written.
template<typename T, int M, bool HERM, int RP, class Op>
__global__ void __launch_bounds__(512)
rowreg_kernel(DimArgs d, const cplx<T> *f, const cplx<T> *g, int64_t rows,
              int64_t rs, cplx<T> *out, int64_t os, T scale, Op op)
{
  constexpr int V=RowV::V;
  typedef RPlan<M,V> P;
  constexpr int TPS=P::TPS;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,(int) d.q,threadIdx.x,blockDim.x);
  __syncthreads();
  const int slots=RP > 1 ? 1 : blockDim.x/TPS;      // rows per block
  const int slot=threadIdx.x/TPS, t=threadIdx.x-slot*TPS;
  const int rp=RP > 1 ? slot : 0;                     // residue lane
  T *xbase=reinterpret_cast<T *>(tw+M);
  T *xb=xbase+slot*2*P::PAD;
  const int q=(int) d.q, L=(int) d.L, H=(int) d.H, k0=(int) d.k0;
  for(int64_t rb=(int64_t) blockIdx.x*slots; rb < rows; rb += (int64_t) gridDim.x*slots) {
    const int64_t row=rb+(RP > 1 ? 0 : slot);
    if(row >= rows) break;               // uniform over the block when RP > 1
    const cplx<T> *fr=f+row*rs, *gr=g+row*rs;
    cplx<T> acc[V];
#pragma unroll
    for(int k=0; k < V; ++k) acc[k]=mk((T) 0,(T) 0);
    if constexpr(!HERM) {
      for(int r=rp; r < q; r += RP) {
        // w_N^{(j+k0) r}: j + k0 is sigma or sigma - M
        const cplx<T> bt=zN[t*r], zm=conjg(zN[M*r]);
        cplx<T> X[V], Y[V];
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          const int j=(sigma-k0) & (M-1);      // k0 <= 0, M a power of two
          cplx<T> a=mk((T) 0,(T) 0), b=a;
          if(j < L) {
            a=fr[j];
            b=gr[j];
            if(r) {
              cplx<T> w=bt*zN[TPS*k*r];
              if(j+k0 < 0) w=w*zm;
              a=w*a;
              b=w*b;
            }
          }
          X[k]=a;
          Y[k]=b;
        }
        fftr_pair<T,M,V,1>(X,Y,xb,tw,t);
        {
          const auto rop=bind_row(op,row);
#pragma unroll
          for(int k=0; k < V; ++k) X[k]=rop(X[k],Y[k],q*(t+TPS*k)+r);
        }
        fftr<T,M,V,-1,true>(X,xb,tw,t);
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          const int j=(sigma-k0) & (M-1);
          if(j < L) {
            if(r) {
              cplx<T> w=bt*zN[TPS*k*r];
              if(j+k0 < 0) w=w*zm;
              acc[k]=acc[k]+mulc(X[k],w);
            } else
              acc[k]=acc[k]+X[k];
          }
        }
      }
      rowreg_combine<T,M,RP>(acc,xbase,xb,rp,t);
      if(rp == 0) {
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          const int j=(sigma-k0) & (M-1);
          if(j < L) out[row*os+j]=scale*acc[k];
        }
      }
    } else {
      for(int r1=2*rp; r1 < q; r1 += 2*RP) {
        const int r2=r1+1 < q ? r1+1 : -1;
        // w_N^{sigma r} = w_N^{t r} w_N^{TPS k r}; w_N^{(sigma - M) r} adds
        // the factor conj w_N^{M r}
        const cplx<T> zero=mk((T) 0,(T) 0);
        const cplx<T> b1=zN[t*r1], m1=conjg(zN[M*r1]);
        const cplx<T> b2=r2 >= 0 ? zN[t*r2] : zero, m2=r2 >= 0 ? conjg(zN[M*r2]) : zero;
        cplx<T> X[V], Y[V];
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          const cplx<T> w1=b1*zN[TPS*k*r1];
          const cplx<T> w2=r2 >= 0 ? b2*zN[TPS*k*r2] : zero;
          cplx<T> a=mk((T) 0,(T) 0), b=a;
          // k = sigma (when < H) and k = sigma - M (when > -H)
          if(sigma < H) {
            const cplx<T> w=w1+muli(w2,1);
            a=w*fr[sigma];
            b=w*gr[sigma];
          }
          if(M-sigma < H) {
            const cplx<T> w=w1*m1+muli(w2*m2,1);
            a=a+w*conjg(fr[M-sigma]);
            b=b+w*conjg(gr[M-sigma]);
          }
          X[k]=a;
          Y[k]=b;
        }
        fftr<T,M,V,1,true>(X,xb,tw,t);
        fftr<T,M,V,1,true>(Y,xb,tw,t);
        {
          const auto rop=bind_row(op,row);
#pragma unroll
          for(int k=0; k < V; ++k) {
            const int l=q*(t+TPS*k);
            X[k]=mk(rop(X[k].x,Y[k].x,l+r1),r2 >= 0 ? rop(X[k].y,Y[k].y,l+r2) : (T) 0);
          }
        }
        fftr<T,M,V,-1,true>(X,xb,tw,t);
        // mirror elements Z(M - sigma) through LDS
        cplx<T> Zm[V];
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int e=xidx<T,M,V>(t+TPS*k);
          xb[e]=X[k].x;
          xb[P::PAD+e]=X[k].y;
        }
        rsync<TPS>();
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int e=xidx<T,M,V>((M-(t+TPS*k)) & (M-1));
          Zm[k]=mk(xb[e],xb[P::PAD+e]);
        }
        rsync<TPS>();
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          if(sigma < H) {
            const cplx<T> z=X[k], zc=conjg(Zm[k]);
            const cplx<T> y1=(T) 0.5*(z+zc);
            cplx<T> v=mulc(y1,b1*zN[TPS*k*r1]);
            if(r2 >= 0) {
              const cplx<T> y2=muli((T) 0.5*(z-zc),-1);
              v=v+mulc(y2,b2*zN[TPS*k*r2]);
            }
            acc[k]=acc[k]+v;
          }
        }
      }
      rowreg_combine<T,M,RP>(acc,xbase,xb,rp,t);
      if(rp == 0) {
#pragma unroll
        for(int k=0; k < V; ++k) {
          const int sigma=t+TPS*k;
          if(sigma < H) out[row*os+sigma]=scale*acc[k];
        }
      }
    }
    if(RP > 1) __syncthreads();          // exchange areas are reused by the next row
  }
}


// Residues r1 and r2 of the f + i g pair of a Hermitian row (H <= M), folded
// for the forward transform from one read of the row: X1[k] = U_r1(sigma),
// X2[k] = U_r2(sigma) at sigma = t + TPS k, with
//   U_r(sigma) = w_N^{sigma r} [u(sigma) + w_q^{-r} v(sigma)],
//   u(sigma) = f(sigma) + i g(sigma)                    (sigma < H),
//   v(sigma) = conj f(M - sigma) + i conj g(M - sigma)  (M - sigma < H),
// the wavenumbers k = sigma and sigma - M of the block (f(-k) = conj f(k),
// w_N^{(sigma - M) r} = w_N^{sigma r} w_q^{-r}).  rf, rg: buffers over the
// row's H words; tr[k] = w_N^{TPS k r}, tr[V] = w_N^{M r} (LDS); b = w_N^{t r}.
// Out-of-range words are read at index 0 and then zeroed (no branches).
template<typename T, int M>
__device__ __forceinline__ void hpk_fold2(cplx<T> *X1, cplx<T> *X2, __amdgpu_buffer_rsrc_t rf,
                                          __amdgpu_buffer_rsrc_t rg, const cplx<T> *tr1,
                                          const cplx<T> *tr2, cplx<T> b1, cplx<T> b2, int t,
                                          int r1, int r2, int H)
{
  constexpr int V=RowV::V, TPS=M/V;
  constexpr unsigned E=sizeof(cplx<T>);
  const cplx<T> c1=conjg(tr1[V]), c2=conjg(tr2[V]);
  const cplx<T> zero=mk((T) 0,(T) 0);
#pragma unroll
  for(int k=0; k < V; ++k) {
    const int sigma=t+TPS*k;
    const bool iu=sigma < H, iv=M-sigma < H;
    const unsigned ou=(unsigned) (iu ? sigma : 0)*E, ov=(unsigned) (iv ? M-sigma : 0)*E;
    const cplx<T> fu=ld_buf<T>(rf,ou,0), gu=ld_buf<T>(rg,ou,0);
    const cplx<T> fv=ld_buf<T>(rf,ov,0), gv=ld_buf<T>(rg,ov,0);
    const cplx<T> u=iu ? mk(fu.x-gu.y,fu.y+gu.x) : zero;
    const cplx<T> v=iv ? mk(fv.x+gv.y,gv.x-fv.y) : zero;
    cplx<T> a=u+c1*v, b=u+c2*v;
    if(r1) a=(b1*tr1[k])*a;
    if(r2) b=(b2*tr2[k])*b;
    X1[k]=a;
    X2[k]=b;
  }
}

// Hermitian rows with H <= M and TPS = M/8 >= 64 (M >= 512), SLOTS = 256/TPS
// rows per block of 256 threads (M = 1024: two rows of two waves).  The
// forward transforms of f and g share one complex FFT per residue: the fold
// U_r of f + i g (hpk_fold2) transforms to F_r + i G_r with F_r, G_r real
// (the folds of Hermitian data are Hermitian), so P_r = F_r G_r is
// realMultBinary's product (fftPadHermitian::forward2 and the multiplier,
// convolve.cc:4517-4609, 26-110).  Residues are paired for the inverse,
// P_{r1} + i P_{r2}, and unpacked with the mirror element Z(M - sigma) as in
// rowreg_kernel (backward2, convolve.cc:4749-4843): q forward and ceil(q/2)
// inverse transforms per row instead of 3 ceil(q/2).  Rows are read through
// buffer resources (a 32-bit offset per element) and the uniform residue
// twiddles w_N^{TPS k r} come from an LDS table, which keeps addresses and
// twiddles out of the registers that the residue loop carries.  Slots past
// the last row run on row rows - 1 without storing (the block barriers of the
// transforms need every thread).  Dynamic LDS: hrowpk_lds().
template<typename T, int M>
struct HRowPk {
  static constexpr int V=RowV::V;
  static constexpr int TPS=M/V;
  static constexpr int SLOTS=256/TPS > 0 ? 256/TPS : 1;
  static constexpr int NT=SLOTS*TPS;
};

template<typename T, int M>
inline size_t hrowpk_lds(int q)
{
  typedef HRowPk<T,M> K;
  return (size_t) (M+q*(K::V+1))*sizeof(cplx<T>)+(size_t) K::SLOTS*2*RPlan<M,K::V>::PAD*sizeof(T);
}

template<typename T, int M, class Op>
__global__ void __launch_bounds__((HRowPk<T,M>::NT))
hrowpk_kernel(DimArgs d, const cplx<T> *f, const cplx<T> *g, int64_t rows, int64_t rs,
              cplx<T> *out, int64_t os, T scale, Op op)
{
  typedef HRowPk<T,M> K;
  constexpr int V=K::V, TPS=K::TPS, SLOTS=K::SLOTS;
  static_assert(TPS >= 64, "a slot spans whole waves");
  typedef RPlan<M,V> P;
  constexpr unsigned E=sizeof(cplx<T>);
  extern __shared__ char smem_raw[];
  const int q=(int) d.q, H=(int) d.H;
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tq=tw+M;                           // tq[r (V + 1) + k] = w_N^{TPS k r}
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,q,threadIdx.x,blockDim.x);
  for(int e=threadIdx.x; e < q*(V+1); e += blockDim.x) {
    const int r=e/(V+1), k=e-r*(V+1);
    tq[e]=zN[TPS*k*r];                        // TPS k r <= M r < N
  }
  __syncthreads();
  // a slot spans whole waves: its index is uniform over the wave
  const int slot=__builtin_amdgcn_readfirstlane((int) threadIdx.x)/TPS;
  const int t=threadIdx.x-slot*TPS;
  T *xb=reinterpret_cast<T *>(tq+q*(V+1))+slot*2*P::PAD;
  for(int64_t rb=(int64_t) blockIdx.x*SLOTS; rb < rows; rb += (int64_t) gridDim.x*SLOTS) {
    const int64_t row=rb+slot;
    const bool live=row < rows;
    const int64_t rr=live ? row : rows-1;
    const __amdgpu_buffer_rsrc_t rf=buf_rsrc(f+rr*rs,H), rg=buf_rsrc(g+rr*rs,H);
    cplx<T> acc[V];
#pragma unroll
    for(int k=0; k < V; ++k) acc[k]=mk((T) 0,(T) 0);
    const auto rop=bind_row(op,row);
    for(int r1=0; r1 < q; r1 += 2) {
      const int r2=r1+1 < q ? r1+1 : -1;
      const cplx<T> b1=zN[t*r1];
      // both residues folded from one read of the row
      cplx<T> X[V], X2[V];
      T p1[V];
      hpk_fold2<T,M>(X,X2,rf,rg,tq+r1*(V+1),tq+(r2 >= 0 ? r2 : r1)*(V+1),b1,
                     r2 >= 0 ? zN[t*r2] : b1,t,r1,r2 >= 0 ? r2 : r1,H);
      if(r2 >= 0) {
        // the two residues' transforms with their butterflies interleaved
        fftr_pair<T,M,V,1>(X,X2,xb,tw,t);
#pragma unroll
        for(int k=0; k < V; ++k) {
          p1[k]=rop(X[k].x,X[k].y,q*(t+TPS*k)+r1);
          X[k]=mk(p1[k],rop(X2[k].x,X2[k].y,q*(t+TPS*k)+r2));
        }
      } else {
        fftr<T,M,V,1,true>(X,xb,tw,t);
#pragma unroll
        for(int k=0; k < V; ++k) X[k]=mk(rop(X[k].x,X[k].y,q*(t+TPS*k)+r1),(T) 0);
      }
      fftr<T,M,V,-1,true>(X,xb,tw,t);
      // mirror elements Z(M - sigma) through LDS
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int e=xidx<T,M,V>(t+TPS*k);
        xb[e]=X[k].x;
        xb[P::PAD+e]=X[k].y;
      }
      rsync<TPS>();
      const cplx<T> *t1=tq+r1*(V+1), *t2=tq+(r2 >= 0 ? r2 : 0)*(V+1);
      const cplx<T> b2=r2 >= 0 ? zN[t*r2] : mk((T) 0,(T) 0);
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int e=xidx<T,M,V>((M-(t+TPS*k)) & (M-1));
        const cplx<T> z=X[k], zc=mk(xb[e],-xb[P::PAD+e]);
        cplx<T> v=mulc((T) 0.5*(z+zc),b1*t1[k]);
        if(r2 >= 0) v=v+mulc(muli((T) 0.5*(z-zc),-1),b2*t2[k]);
        acc[k]=acc[k]+v;
      }
      rsync<TPS>();
    }
    if(live) {
      const __amdgpu_buffer_rsrc_t ro=buf_rsrc(out+row*os,H);
#pragma unroll
      for(int k=0; k < V; ++k)
        if(t+TPS*k < H) st_buf<T>(ro,(unsigned) (t+TPS*k)*E,0,scale*acc[k]);
    }
  }
}


// Hermitian rows longer than one block (H > M): the residue-pair passes of a
// row (fftPadHermitian's D = 2 pairs, convolve.cc:4517-5012) run as
// independent work units (row, pass), each on one slot of TPS threads.  A
// unit folds every block b of the row (wavenumbers k = sigma + b M, |k| < H,
// f(-k) = conj f(k)) into the packed pair W1 + i W2, transforms f and g,
// multiplies as realMultBinary, inverts, unpacks with the mirror element
// Z(M - sigma) and writes its contribution to every stored mode k < H of the
// row into its slice of `part` (slice p = rows x H words); hrows_reduce_kernel
// adds the slices in pass order (deterministic).  Twiddles
// w_N^{(sigma + b M) r} = w_N^{sigma r} w_q^{b r} with w_q^{b r} = zN[(b r mod q) M].
// Dynamic LDS: M complex twiddles + slots * 2 PAD scalar words.
template<typename T, int M>
__global__ void __launch_bounds__(512)
hrows_long_kernel(DimArgs d, const cplx<T> *f, const cplx<T> *g, int64_t rows,
                  int64_t rs, cplx<T> *part)
{
  constexpr int V=RowV::V;
  typedef RPlan<M,V> P;
  constexpr int TPS=P::TPS;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,(int) d.q,threadIdx.x,blockDim.x);
  __syncthreads();
  const int slots=blockDim.x/TPS;
  const int slot=threadIdx.x/TPS, t=threadIdx.x-slot*TPS;
  T *xb=reinterpret_cast<T *>(tw+M)+slot*2*P::PAD;
  const int q=(int) d.q, H=(int) d.H;
  const int passes=(q+1)/2;
  const int nb=(H-1)/M+1;                 // blocks b in [0, nb) hold the modes k < H
  const int64_t units=rows*passes;
  const cplx<T> zero=mk((T) 0,(T) 0);
  for(int64_t ub=(int64_t) blockIdx.x*slots; ub < units; ub += (int64_t) gridDim.x*slots) {
    const int64_t unit=ub+slot;
    if(unit >= units) break;              // only when slots > 1 (TPS <= 64)
    const int64_t row=unit/passes;
    const int pass=(int) (unit-row*passes);
    const int r1=2*pass, r2=r1+1 < q ? r1+1 : -1;
    const cplx<T> *fr=f+row*rs, *gr=g+row*rs;
    cplx<T> *pr=part+((int64_t) pass*rows+row)*H;
    cplx<T> X[V], Y[V];
#pragma unroll
    for(int k=0; k < V; ++k) X[k]=Y[k]=zero;
    // b r mod q from b = -nb on, advanced by r per block
    int e1=(int) ((q-((int64_t) nb*r1) % q) % q);
    int e2=r2 >= 0 ? (int) ((q-((int64_t) nb*r2) % q) % q) : 0;
    for(int b=-nb; b < nb; ++b) {
      const cplx<T> c1=zN[e1*M], c2=r2 >= 0 ? zN[e2*M] : zero;
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int sigma=t+TPS*k;
        const int kk=sigma+b*M;
        if(kk > -H && kk < H) {
          cplx<T> w=zN[sigma*r1]*c1;
          if(r2 >= 0) w=w+muli(zN[sigma*r2]*c2,1);
          const cplx<T> fv=kk >= 0 ? fr[kk] : conjg(fr[-kk]);
          const cplx<T> gv=kk >= 0 ? gr[kk] : conjg(gr[-kk]);
          X[k]=X[k]+w*fv;
          Y[k]=Y[k]+w*gv;
        }
      }
      e1 += r1;
      if(e1 >= q) e1 -= q;
      if(r2 >= 0) {
        e2 += r2;
        if(e2 >= q) e2 -= q;
      }
    }
    fftr<T,M,V,1,true>(X,xb,tw,t);
    fftr<T,M,V,1,true>(Y,xb,tw,t);
#pragma unroll
    for(int k=0; k < V; ++k) X[k]=mk(X[k].x*Y[k].x,X[k].y*Y[k].y);
    fftr<T,M,V,-1,true>(X,xb,tw,t);
    // mirror elements Z(M - sigma) through LDS
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int e=xidx<T,M,V>(t+TPS*k);
      xb[e]=X[k].x;
      xb[P::PAD+e]=X[k].y;
    }
    rsync<TPS>();
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int e=xidx<T,M,V>((M-(t+TPS*k)) & (M-1));
      const cplx<T> z=X[k], zc=conjg(mk(xb[e],xb[P::PAD+e]));
      // this thread's unpacked pair, with the conjugate twiddles of sigma
      Y[k]=mulc((T) 0.5*(z+zc),zN[(t+TPS*k)*r1]);
      X[k]=r2 >= 0 ? mulc(muli((T) 0.5*(z-zc),-1),zN[(t+TPS*k)*r2]) : zero;
    }
    rsync<TPS>();
    e1=0;
    e2=0;
    for(int b=0; b < nb; ++b) {
      const cplx<T> c1=zN[e1*M], c2=r2 >= 0 ? zN[e2*M] : zero;
#pragma unroll
      for(int k=0; k < V; ++k) {
        const int kk=t+TPS*k+b*M;
        if(kk < H) pr[kk]=mulc(Y[k],c1)+(r2 >= 0 ? mulc(X[k],c2) : zero);
      }
      e1 += r1;
      if(e1 >= q) e1 -= q;
      if(r2 >= 0) {
        e2 += r2;
        if(e2 >= q) e2 -= q;
      }
    }
  }
}

// out[row][k] = scale * sum over passes of part[pass][row][k], k < H, in
// pass order; one block per row (grid-stride).
template<typename T>
__global__ void __launch_bounds__(256)
hrows_reduce_kernel(const cplx<T> *part, int64_t rows, int H, int passes,
                    cplx<T> *out, int64_t os, T scale)
{
  for(int64_t row=blockIdx.x; row < rows; row += gridDim.x)
    for(int k=threadIdx.x; k < H; k += blockDim.x) {
      cplx<T> sum=part[row*H+k];
      for(int p=1; p < passes; ++p) sum=sum+part[((int64_t) p*rows+row)*H+k];
      out[row*os+k]=scale*sum;
    }
}

// ------------------------------------------------------------------------
// Column residue transforms on the register FFT (fftPad forward/backward of
// strided columns, convolve.cc:771-1965) for power-of-two M and at most P2 = 2
// blocks of M per column (p <= 2), complex or centred data.  Lane (c, t) =
// (threadIdx.x % TC, threadIdx.x / TC): consecutive lanes take consecutive
// columns, so every load and store of a wave covers TC adjacent columns; the
// TPS threads of a column span waves and exchange with block barriers, one
// round per stage, through the column's LDS area of CS = 2 PAD + X words
// (real parts, then imaginary parts PAD words on).  Element e of column c
// sits at word base(e) ^ (c BETA mod 32), base being the identity or the
// row kernels' swizzle (both keep words below M): the (base, X, BETA) of each
// size were found
// by an exhaustive search under the gfx950 LDS bank model (64-bit words:
// writes 16 lanes per pass over 32 banks, reads 32 lanes over 64 banks;
// 32-bit words: 32 lanes over 32 banks) for the lane layout (c, t) above,
// and give conflict-free exchanges (FP32 M = 256: 128 extra passes in 1536).
// Dynamic LDS: M complex twiddles + TC * CS scalar words.
// ------------------------------------------------------------------------
template<typename T, int M> struct ColReg {
  static constexpr int V=8;
  static constexpr int TPS=M/V;
  static constexpr int TC=TPS >= 512 ? 1 : (512/TPS < 64 ? 512/TPS : 64);
  static constexpr int PAD=RPlan<M,V>::PAD;
  static constexpr bool F64=sizeof(T) == 8;
  // base: 0 identity, 2 swizzle (xidx)
  static constexpr int BASE=M >= 1024 ? 2 : 0;
  static constexpr int X=M == 16 ? 1 : 0;
  static constexpr int BETA=M == 16 || M == 4096 ? 0 : M == 256 ? (F64 ? 17 : 2) : M == 512 ? 9 :
    M == 1024 ? 10 : M == 2048 ? (F64 ? 28 : 20) : 1;
  static constexpr int CS=2*PAD+X;
  static constexpr int NT=TPS*TC;
  static __device__ __forceinline__ int at(int e)
  {
    if constexpr(BASE == 2) return xidx<T,M,V>(e);
    else return e;
  }
  // the column's XOR of the word index (M = 16: the words stay below 16)
  static __device__ __forceinline__ int xm(int c) { return (c*BETA) & (M < 32 ? M-1 : 31); }
};

// The stage twiddle table of the column kernels (tw) and the residue twiddle
// table they read: zlds = 1 copies the dimension's N twiddles into LDS after
// the exchange areas (the residue twiddles are per-lane gathers, about 10%
// faster from LDS; the host asks for it when the LDS fits without lowering
// occupancy), else the table in global memory.  Ends with a block barrier.
template<typename T, int M>
__device__ __forceinline__ const cplx<T> *col_twiddles(const DimArgs &d, cplx<T> *tw, int zlds)
{
  typedef ColReg<T,M> CR;
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,CR::V>(tw,zN,(int) d.q,threadIdx.x,CR::NT);
  cplx<T> *zl=reinterpret_cast<cplx<T> *>(reinterpret_cast<T *>(tw+M)+CR::TC*CR::CS);
  if(zlds)
    for(int e=threadIdx.x; e < (int) d.N; e += CR::NT) zl[e]=zN[e];
  __syncthreads();
  return zlds ? zl : zN;
}

// F[b][(r-r0) M + s][c] = sum_j f[b][j][c] w_N^{(j+k0)(q s + r)}, j < L;
// column c starts at c*ics of f and c*ocs of F (unit strides for
// interleaved columns, row strides for contiguous rows taken as columns).
template<typename T, int M, int P2>
__global__ void __launch_bounds__(512)
colfwd_reg_kernel(DimArgs d, const cplx<T> *f, int64_t S, int64_t C,
                  int64_t bsi, cplx<T> *F, int64_t SF, int64_t bso, int64_t r0,
                  int64_t nres, int64_t ics, int64_t ocs, int zlds)
{
  typedef ColReg<T,M> CR;
  constexpr int V=CR::V, TPS=CR::TPS, TC=CR::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=col_twiddles<T,M>(d,tw,zlds);
  // interleaved columns: consecutive lanes on consecutive columns; rows
  // (column stride ics > 1): a row's threads on consecutive lanes
  const bool rowmap=ics != 1;
  const int c=rowmap ? threadIdx.x/TPS : threadIdx.x % TC;
  const int t=rowmap ? threadIdx.x % TPS : threadIdx.x/TC;
  T *xb=reinterpret_cast<T *>(tw+M)+c*CR::CS;
  const int xm=CR::xm(c);
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const cplx<T> *fb=f+(int64_t) blockIdx.y*bsi+(c0+c)*ics;
  cplx<T> *Fb=F+(int64_t) blockIdx.y*bso+(c0+c)*ocs;
  const int N=(int) d.N, L=(int) d.L, k0=(int) d.k0;
  // the column, read once: element j = j0 + M i with j0 = (sigma - k0) mod M
  cplx<T> raw[P2][V];
#pragma unroll
  for(int k=0; k < V; ++k) {
    const int j0=(t+TPS*k-k0) & (M-1);
#pragma unroll
    for(int i=0; i < P2; ++i) {
      const int j=j0+M*i;
      raw[i][k]=ok && j < L ? fb[(int64_t) j*S] : mk((T) 0,(T) 0);
    }
  }
  for(int64_t rr=0; rr < nres; ++rr) {
    const int r=(int) (r0+rr);
    // the residue's twiddles loaded before the first use (r = 0 reads
    // w_N^0 = 1): a load per branch made each wait for the previous one
    cplx<T> w[P2][V];
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int j0=(t+TPS*k-k0) & (M-1);
#pragma unroll
      for(int i=0; i < P2; ++i) w[i][k]=zN[wrapN((j0+M*i+k0)*r,N)];
    }
    cplx<T> X[V];
#pragma unroll
    for(int k=0; k < V; ++k) {
      cplx<T> a=raw[0][k]*w[0][k];
#pragma unroll
      for(int i=1; i < P2; ++i) a=a+raw[i][k]*w[i][k];
      X[k]=a;
    }
    fftr_blkx<T,M,V,1,CR>(X,xb,tw,t,xm);
    if(ok) {
      cplx<T> *Fr=Fb+rr*M*SF;
#pragma unroll
      for(int k=0; k < V; ++k) Fr[(int64_t) (t+TPS*k)*SF]=X[k];
    }
  }
}

// f[b][j][c] = scale * ((accumulate ? f : 0) + sum_r w_N^{-(j+k0) r}
// DFT^{-1}[F_r]((j+k0) mod M)), j < L.
template<typename T, int M, int P2>
__global__ void __launch_bounds__(512)
colbwd_reg_kernel(DimArgs d, const cplx<T> *F, int64_t SF, int64_t bsi,
                  cplx<T> *f, int64_t S, int64_t C, int64_t bso, int64_t r0,
                  int64_t nres, int accumulate, T scale, int64_t ics, int64_t ocs,
                  int zlds)
{
  typedef ColReg<T,M> CR;
  constexpr int V=CR::V, TPS=CR::TPS, TC=CR::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=col_twiddles<T,M>(d,tw,zlds);
  // interleaved columns: consecutive lanes on consecutive columns; rows
  // (column stride ics > 1): a row's threads on consecutive lanes
  const bool rowmap=ics != 1;
  const int c=rowmap ? threadIdx.x/TPS : threadIdx.x % TC;
  const int t=rowmap ? threadIdx.x % TPS : threadIdx.x/TC;
  T *xb=reinterpret_cast<T *>(tw+M)+c*CR::CS;
  const int xm=CR::xm(c);
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const cplx<T> *Fb=F+(int64_t) blockIdx.y*bsi+(c0+c)*ics;
  cplx<T> *fb=f+(int64_t) blockIdx.y*bso+(c0+c)*ocs;
  const int N=(int) d.N, L=(int) d.L, k0=(int) d.k0;
  cplx<T> acc[P2][V];
#pragma unroll
  for(int i=0; i < P2; ++i)
#pragma unroll
    for(int k=0; k < V; ++k) acc[i][k]=mk((T) 0,(T) 0);
  for(int64_t rr=0; rr < nres; ++rr) {
    const int r=(int) (r0+rr);
    const cplx<T> *Fr=Fb+rr*M*SF;
    cplx<T> X[V];
#pragma unroll
    for(int k=0; k < V; ++k) X[k]=ok ? Fr[(int64_t) (t+TPS*k)*SF] : mk((T) 0,(T) 0);
    fftr_blkx<T,M,V,-1,CR>(X,xb,tw,t,xm);
    // twiddles loaded before the first use (r = 0 reads 1), as forward
    cplx<T> w[P2][V];
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int j0=(t+TPS*k-k0) & (M-1);
#pragma unroll
      for(int i=0; i < P2; ++i) w[i][k]=zN[wrapN((j0+M*i+k0)*r,N)];
    }
#pragma unroll
    for(int k=0; k < V; ++k) {
#pragma unroll
      for(int i=0; i < P2; ++i) acc[i][k]=acc[i][k]+mulc(X[k],w[i][k]);
    }
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int j0=(t+TPS*k-k0) & (M-1);
#pragma unroll
      for(int i=0; i < P2; ++i) {
        const int j=j0+M*i;
        if(j < L) {
          cplx<T> *o=fb+(int64_t) j*S;
          cplx<T> v=acc[i][k];
          if(accumulate) v=v+*o;
          *o=scale*v;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// Real columns (fftPadReal, convolve.h:805-980; its forward and backward
// residue transforms, convolve.cc:5449-7483) for power-of-two M and p <= 2.
// For real data the padded transform at l = q s + r and at N - l are complex
// conjugates, so a real dimension keeps K = N/2 + 1 planes, residue-major:
//   r = 0:              s in [0, M/2]  (partner of s: M - s)
//   0 < r < q/2:        s in [0, M)    (partner: residue q - r)
//   r = q/2 (q even):   s in [0, M/2)  (partner of s: M - 1 - s)
// at plane RealKept::base(r) + s.  The planes are complex, so the inner
// dimensions of a real convolution are ordinary complex convolutions over
// K planes; the backward transform folds each kept plane with its partner:
//   f(j) = sum over kept l of c_l Re(F_l w_N^{-j l}),
// c_l = 1 for the self-conjugate planes l = 0 and l = N/2 and 2 for the
// others.  Residues 0 and q/2, which hold their own partners, are inverted
// over all M points with the missing half rebuilt as conjugates of the kept
// half (their real part is then the sum above); the other residues add
// 2 Re(w_N^{-j r} IDFT_M[F_r](j)).  Lanes and LDS as in colfwd_reg_kernel.
// ------------------------------------------------------------------------
struct RealKept {
  __host__ __device__ static int64_t residues(int64_t q) { return q/2+1; }
  __host__ __device__ static int64_t planes(int64_t m, int64_t q) { return m*q/2+1; }
  __host__ __device__ static int64_t base(int64_t m, int64_t r) { return r == 0 ? 0 : m/2+1+(r-1)*m; }
  __host__ __device__ static int64_t count(int64_t m, int64_t q, int64_t r)
  {
    return r == 0 ? m/2+1 : (2*r == q ? m/2 : m);
  }
};

// F[b][base(r) + s][c] = sum_{j < L} f[b][j][c] w_N^{j (q s + r)} for the
// kept (r, s); f real with element stride S and column stride ics (real
// words), F complex with plane stride SF and column stride ocs.
template<typename T, int M, int P2>
__global__ void __launch_bounds__(512)
colfwd_real_kernel(DimArgs d, const T *f, int64_t S, int64_t C, int64_t bsi,
                   cplx<T> *F, int64_t SF, int64_t bso, int64_t ics, int64_t ocs)
{
  typedef ColReg<T,M> CR;
  constexpr int V=CR::V, TPS=CR::TPS, TC=CR::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,(int) d.q,threadIdx.x,CR::NT);
  __syncthreads();
  const bool rowmap=ics != 1;
  const int c=rowmap ? threadIdx.x/TPS : threadIdx.x % TC;
  const int t=rowmap ? threadIdx.x % TPS : threadIdx.x/TC;
  T *xb=reinterpret_cast<T *>(tw+M)+c*CR::CS;
  const int xm=CR::xm(c);
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const T *fb=f+(int64_t) blockIdx.y*bsi+(c0+c)*ics;
  cplx<T> *Fb=F+(int64_t) blockIdx.y*bso+(c0+c)*ocs;
  const int N=(int) d.N, L=(int) d.L, q=(int) d.q;
  T raw[P2][V];
#pragma unroll
  for(int k=0; k < V; ++k)
#pragma unroll
    for(int i=0; i < P2; ++i) {
      const int j=t+TPS*k+M*i;
      raw[i][k]=ok && j < L ? fb[(int64_t) j*S] : (T) 0;
    }
  const int R=(int) RealKept::residues(q);
  for(int r=0; r < R; ++r) {
    // twiddles loaded before the first use (r = 0 reads 1)
    cplx<T> w[P2][V];
#pragma unroll
    for(int k=0; k < V; ++k)
#pragma unroll
      for(int i=0; i < P2; ++i) w[i][k]=zN[wrapN((t+TPS*k+M*i)*r,N)];
    cplx<T> X[V];
#pragma unroll
    for(int k=0; k < V; ++k) {
      cplx<T> a=raw[0][k]*w[0][k];
#pragma unroll
      for(int i=1; i < P2; ++i) a=a+raw[i][k]*w[i][k];
      X[k]=a;
    }
    fftr_blkx<T,M,V,1,CR>(X,xb,tw,t,xm);
    if(ok) {
      const int cnt=(int) RealKept::count(M,q,r);
      cplx<T> *Fr=Fb+RealKept::base(M,r)*SF;
#pragma unroll
      for(int k=0; k < V; ++k)
        if(t+TPS*k < cnt) Fr[(int64_t) (t+TPS*k)*SF]=X[k];
    }
  }
}

// f[b][j][c] = scale sum over kept l of c_l Re(F[b][l][c] w_N^{-j l}), j < L.
template<typename T, int M, int P2>
__global__ void __launch_bounds__(512)
colbwd_real_kernel(DimArgs d, const cplx<T> *F, int64_t SF, int64_t bsi,
                   T *f, int64_t S, int64_t C, int64_t bso, T scale,
                   int64_t ics, int64_t ocs)
{
  typedef ColReg<T,M> CR;
  constexpr int V=CR::V, TPS=CR::TPS, TC=CR::TC;
  extern __shared__ char smem_raw[];
  cplx<T> *tw=reinterpret_cast<cplx<T> *>(smem_raw);
  const cplx<T> *zN=(const cplx<T> *) d.zN;
  fill_twr<T,M,V>(tw,zN,(int) d.q,threadIdx.x,CR::NT);
  __syncthreads();
  const bool rowmap=ics != 1;
  const int c=rowmap ? threadIdx.x/TPS : threadIdx.x % TC;
  const int t=rowmap ? threadIdx.x % TPS : threadIdx.x/TC;
  T *xb=reinterpret_cast<T *>(tw+M)+c*CR::CS;
  const int xm=CR::xm(c);
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const bool ok=c0+c < C;
  const cplx<T> *Fb=F+(int64_t) blockIdx.y*bsi+(c0+c)*ics;
  T *fb=f+(int64_t) blockIdx.y*bso+(c0+c)*ocs;
  const int N=(int) d.N, L=(int) d.L, q=(int) d.q;
  T acc[P2][V];
#pragma unroll
  for(int i=0; i < P2; ++i)
#pragma unroll
    for(int k=0; k < V; ++k) acc[i][k]=(T) 0;
  const int R=(int) RealKept::residues(q);
  for(int r=0; r < R; ++r) {
    const cplx<T> *Fr=Fb+RealKept::base(M,r)*SF;
    const bool self=r == 0 || 2*r == q;
    // the full residue: kept s, or the conjugate of the partner's kept word
    const int mirror=r == 0 ? M : M-1;
    cplx<T> X[V];
#pragma unroll
    for(int k=0; k < V; ++k) {
      const int s=t+TPS*k;
      const bool direct=!self || s < (int) RealKept::count(M,q,r);
      cplx<T> v=mk((T) 0,(T) 0);
      if(ok) {
        v=Fr[(int64_t) (direct ? s : mirror-s)*SF];
        if(!direct) v=conjg(v);
      }
      X[k]=v;
    }
    fftr_blkx<T,M,V,-1,CR>(X,xb,tw,t,xm);
    const T cr=self ? (T) 1 : (T) 2;
    // twiddles loaded before their use (r = 0 reads 1), half of the values
    // at a time (all at once costs a wave per SIMD)
    constexpr int KH=V/2;
#pragma unroll
    for(int h=0; h < V; h += KH) {
      cplx<T> w[P2][KH];
#pragma unroll
      for(int k=0; k < KH; ++k)
#pragma unroll
        for(int i=0; i < P2; ++i) w[i][k]=zN[wrapN((t+TPS*(h+k)+M*i)*r,N)];
#pragma unroll
      for(int k=0; k < KH; ++k)
#pragma unroll
        for(int i=0; i < P2; ++i)
          acc[i][h+k] += cr*(X[h+k].x*w[i][k].x+X[h+k].y*w[i][k].y);
    }
  }
  if(ok) {
#pragma unroll
    for(int k=0; k < V; ++k)
#pragma unroll
      for(int i=0; i < P2; ++i) {
        const int j=t+TPS*k+M*i;
        if(j < L) fb[(int64_t) j*S]=scale*acc[i][k];
      }
  }
}

} // namespace fftwpp_amd
