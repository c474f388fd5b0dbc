// rowsfused.h -- launch of the fused innermost-row convolutions on the
// register FFT kernels for a product functor Op (common.h: LibMult, RunMult,
// or a caller's functor), shared by the library (kernels.hip) and by code
// that compiles its own multiplier into these kernels
// (include/fftwpp_amd_fused.h).  The row stage of Convolution::convolveRaw
// (convolve.cc:7513-7575; Convolution::operate calls the multiplier inside
// the residue loop, convolve.h:1120-1135).
//
// rows_fused<T, Op>() returns 0 after a launch, 1 when no register row kernel
// takes the dimension (the caller picks another engine), or -e for a HIP
// error e.  Kernels:
//   m = 256, complex (not centred), L <= 256, q <= 8: rowconv256p_kernel
//     (q = 2: both residues of a row in one wave) or rowconv256_kernel;
//   power-of-two m in [16, 4096] covering the row (L <= m, or H <= m for
//     Hermitian rows), m q < 2^30: rowreg_kernel.
#pragma once
#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>
#include "kernels256.h"
#include "rowsreg.h"

namespace fftwpp_amd {

// Raise a kernel's dynamic LDS limit to 160 KB once (per kernel).
template<typename K>
inline hipError_t fused_allow_lds(K kernel, size_t bytes)
{
  if(bytes <= 64*1024) return hipSuccess;
  static std::mutex mu;
  static std::vector<const void *> raised;
  std::lock_guard<std::mutex> lock(mu);
  const void *fn=(const void *) kernel;
  if(std::find(raised.begin(),raised.end(),fn) != raised.end()) return hipSuccess;
  const hipError_t e=hipFuncSetAttribute(fn,hipFuncAttributeMaxDynamicSharedMemorySize,160*1024);
  if(e == hipSuccess) raised.push_back(fn);
  return e;
}

// Blocks of `kernel` resident on the current device at once (at least 1),
// cached per (kernel, device): the grid of a persistent kernel.
template<typename K>
inline size_t fused_resident(K kernel, int threads, size_t lds)
{
  static std::mutex mu;
  static std::vector<std::pair<std::pair<const void *,int>,size_t>> known;
  int dev=0;
  if(hipGetDevice(&dev) != hipSuccess) return 1024;
  const std::pair<const void *,int> key((const void *) kernel,dev);
  std::lock_guard<std::mutex> lock(mu);
  for(const auto &e : known) if(e.first == key) return e.second;
  int cus=0, per=0;
  if(hipDeviceGetAttribute(&cus,hipDeviceAttributeMultiprocessorCount,dev) != hipSuccess ||
     hipOccupancyMaxActiveBlocksPerMultiprocessor(&per,(const void *) kernel,threads,lds) != hipSuccess)
    return 1024;
  const size_t n=std::max<size_t>(1,(size_t) cus*(size_t) per);
  known.push_back({key,n});
  return n;
}

inline int fused_status(hipError_t e) { return e == hipSuccess ? 0 : -(int) e; }

template<typename T, class Op>
inline int rows256_launch(const DimArgs &a, const void *f, const void *g, size_t rows,
                          size_t rs, void *out, size_t os, double scale, const Op &op,
                          hipStream_t st)
{
  const cplx<T> *zt=(const cplx<T> *) a.zN;
  hipError_t e;
  if(a.q == 2) {
    // both residues of a row in one wave (each row read once); persistent grid
    constexpr int WPB=4;
    const size_t need=(rows+2*WPB-1)/(2*WPB);
    const size_t lds=row256_lds<T>(4*WPB,2);
    if((e=fused_allow_lds(rowconv256p_kernel<T,Op,WPB>,lds)) != hipSuccess) return fused_status(e);
    dim3 grid((unsigned) std::min(need,fused_resident(rowconv256p_kernel<T,Op,WPB>,64*WPB,lds)));
    hipLaunchKernelGGL((rowconv256p_kernel<T,Op,WPB>),grid,dim3(64*WPB),lds,st,
                       (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,(int64_t) rs,
                       (cplx<T> *) out,(int64_t) os,(int) a.L,(T) scale,zt,op);
    return fused_status(hipGetLastError());
  }
  constexpr int NR=sizeof(T) == 8 ? 8 : 16;
  dim3 grid((unsigned) ((rows+NR-1)/NR));
  const size_t lds=row256_lds<T>(NR,(size_t) a.q);
  if((e=fused_allow_lds(rowconv256_kernel<T,Op,NR>,lds)) != hipSuccess) return fused_status(e);
  hipLaunchKernelGGL((rowconv256_kernel<T,Op,NR>),grid,dim3(16*NR),lds,st,
                     (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,(int64_t) rs,
                     (cplx<T> *) out,(int64_t) os,(int) a.L,(T) scale,zt,(int64_t) a.q,op);
  return fused_status(hipGetLastError());
}

// RP residue lanes per row when rows are too few to fill the chip (under 512)
// and the row has at least two residue passes (residues, or residue pairs for
// Hermitian rows).
template<typename T, int M, bool HERM, int RP, class Op>
inline int rowreg_launch_rp(const DimArgs &a, const void *f, const void *g, size_t rows,
                            size_t rs, void *out, size_t os, double scale, const Op &op,
                            hipStream_t st)
{
  constexpr int V=RowV::V;
  constexpr int TPS=M/V;
  const size_t nt=RP > 1 ? RP*TPS : (TPS > 64 ? TPS : 256);
  const size_t slots=nt/TPS;
  const size_t lds=M*sizeof(cplx<T>)+slots*2*RPlan<M,V>::PAD*sizeof(T);
  const hipError_t e=fused_allow_lds(rowreg_kernel<T,M,HERM,RP,Op>,lds);
  if(e != hipSuccess) return fused_status(e);
  const size_t rpb=RP > 1 ? 1 : slots;
  size_t nb=(rows+rpb-1)/rpb;
  // persistent grid: the twiddle table is filled once per resident block
  nb=std::min(nb,fused_resident(rowreg_kernel<T,M,HERM,RP,Op>,(int) nt,lds));
  hipLaunchKernelGGL((rowreg_kernel<T,M,HERM,RP,Op>),dim3((unsigned) nb),dim3((unsigned) nt),
                     lds,st,a,(const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,
                     (int64_t) rs,(cplx<T> *) out,(int64_t) os,(T) scale,op);
  return fused_status(hipGetLastError());
}

// Hermitian rows of M = 512 ... 2048 on the packed kernel (f + i g in one
// transform per residue), persistent grid.
template<typename T, int M, class Op>
inline int hrowpk_launch(const DimArgs &a, const void *f, const void *g, size_t rows,
                         size_t rs, void *out, size_t os, double scale, const Op &op,
                         hipStream_t st)
{
  typedef HRowPk<T,M> K;
  const size_t lds=hrowpk_lds<T,M>((int) a.q);
  const hipError_t e=fused_allow_lds(hrowpk_kernel<T,M,Op>,lds);
  if(e != hipSuccess) return fused_status(e);
  size_t nb=(rows+K::SLOTS-1)/K::SLOTS;
  nb=std::min(nb,fused_resident(hrowpk_kernel<T,M,Op>,K::NT,lds));
  hipLaunchKernelGGL((hrowpk_kernel<T,M,Op>),dim3((unsigned) nb),dim3(K::NT),lds,st,a,
                     (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,(int64_t) rs,
                     (cplx<T> *) out,(int64_t) os,(T) scale,op);
  return fused_status(hipGetLastError());
}

template<typename T, int M, bool HERM, class Op>
inline int rowreg_launch(const DimArgs &a, const void *f, const void *g, size_t rows,
                         size_t rs, void *out, size_t os, double scale, const Op &op,
                         hipStream_t st)
{
  constexpr int TPS=M/RowV::V;
  const int64_t passes=HERM ? (a.q+1)/2 : a.q;
  // Hermitian rows of 512 ... 2048 points with enough rows to fill the chip:
  // the packed kernel (measured 12-28% faster than rowreg_kernel there)
  if constexpr(HERM && M >= 512 && M <= 2048)
    if(rows >= 512) return hrowpk_launch<T,M>(a,f,g,rows,rs,out,os,scale,op,st);
  if(TPS <= 256 && rows < 512 && passes >= 2)
    return rowreg_launch_rp<T,M,HERM,2>(a,f,g,rows,rs,out,os,scale,op,st);
  return rowreg_launch_rp<T,M,HERM,1>(a,f,g,rows,rs,out,os,scale,op,st);
}

// Row sizes compiled: bit i of MMASK stands for m = 16 << i (16 ... 4096);
// sizes left out are reported as not taken (1).
constexpr unsigned ROWS_ALL_SIZES=0x1FFu;

template<typename T, int M, bool HERM, unsigned MMASK, class Op>
inline int rowreg_if(const DimArgs &a, const void *f, const void *g, size_t rows, size_t rs,
                     void *out, size_t os, double scale, const Op &op, hipStream_t st)
{
  if constexpr((MMASK >> (__builtin_ctz(M)-4)) & 1u)
    return rowreg_launch<T,M,HERM>(a,f,g,rows,rs,out,os,scale,op,st);
  else
    return 1;
}

template<typename T, bool HERM, unsigned MMASK, class Op>
inline int rowreg_dispatch(const DimArgs &a, const void *f, const void *g, size_t rows,
                           size_t rs, void *out, size_t os, double scale, const Op &op,
                           hipStream_t st)
{
  switch(a.m) {
  case 16: return rowreg_if<T,16,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 32: return rowreg_if<T,32,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 64: return rowreg_if<T,64,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 128: return rowreg_if<T,128,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 256: return rowreg_if<T,256,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 512: return rowreg_if<T,512,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 1024: return rowreg_if<T,1024,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 2048: return rowreg_if<T,2048,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  case 4096: return rowreg_if<T,4096,HERM,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  default: return 1;
  }
}

// Whether the m = 256 row kernels take the dimension (lane offsets of at most
// 16 strides plus a tile width stay within 31 bits).
inline bool rows256_take(const DimArgs &a, bool herm, size_t rs, size_t os)
{
  const size_t lim=(size_t) 1 << 30;
  return !herm && a.k0 == 0 && a.m == 256 && a.L <= 256 && a.q <= 8 &&
    16*rs+64 < lim && 16*os+64 < lim;
}

// Whether the register row kernel takes the dimension.
inline bool rowreg_take(const DimArgs &a, bool herm)
{
  const int64_t m=a.m;
  if(m < 16 || m > 4096 || (m & (m-1))) return false;
  if(m*a.q >= ((int64_t) 1 << 30)) return false;     // 32-bit twiddle exponents
  return herm ? a.H <= m : a.L <= m;
}

// HERMS: which row kinds are compiled (bit 0 complex or centred, bit 1
// Hermitian).
template<typename T, class Op, unsigned MMASK=ROWS_ALL_SIZES, unsigned HERMS=3u>
inline int rows_fused(const DimArgs &a, bool herm, const void *f, const void *g, size_t rows,
                      size_t rs, void *out, size_t os, double scale, const Op &op,
                      hipStream_t st)
{
  if(rows == 0) return 0;
  if(!((herm ? 2u : 1u) & HERMS)) return 1;
  if constexpr((MMASK >> 4) & 1u)
    if(rows256_take(a,herm,rs,os))
      return rows256_launch<T>(a,f,g,rows,rs,out,os,scale,op,st);
  if(rowreg_take(a,herm)) {
    if constexpr((HERMS & 2u) != 0)
      if(herm) return rowreg_dispatch<T,true,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
    if constexpr((HERMS & 1u) != 0)
      if(!herm) return rowreg_dispatch<T,false,MMASK>(a,f,g,rows,rs,out,os,scale,op,st);
  }
  return 1;
}

} // namespace fftwpp_amd
