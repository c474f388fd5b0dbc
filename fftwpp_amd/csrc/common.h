// common.h -- complex arithmetic, per-dimension geometry and twiddle access
// shared by every HIP kernel of the hybrid dealiased convolution library.
//
// Conventions (from the reference):
//   * the padded "forward" transform uses the kernel e^{+2 pi i jk/N}
//     (convolve.cc:576-577); the backward transform uses e^{-2 pi i jk/N};
//   * a padded dimension has N = m*q >= M >= L (convolve.cc:403-410,493-509);
//   * data index j in [0,L) carries wavenumber k = j + k0, with k0 = 0 for
//     fftPad and k0 = -floor(L/2) for the centred transforms
//     (fftPadCentered::forward2, convolve.cc:2244-2250);
//   * residue r in [0,q) is the set of padded indices l = q*s + r, s < m.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

namespace fftwpp_amd {

template<typename T>
struct cplx {
  T x, y;
};

template<typename T>
__device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> c; c.x=a; c.y=b; return c; }
template<typename T>
__device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x+b.x,a.y+b.y); }
template<typename T>
__device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x-b.x,a.y-b.y); }
template<typename T>
__device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b)
{
  return mk(fma(a.x,b.x,-a.y*b.y),fma(a.x,b.y,a.y*b.x));
}
template<typename T>
__device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return mk(s*a.x,s*a.y); }
template<typename T>
__device__ __forceinline__ cplx<T> conjg(cplx<T> a) { return mk(a.x,-a.y); }
// a * conj(b)
template<typename T>
__device__ __forceinline__ cplx<T> mulc(cplx<T> a, cplx<T> b)
{
  return mk(fma(a.x,b.x,a.y*b.y),fma(a.y,b.x,-a.x*b.y));
}
// multiply by +i (sign>0) or -i (sign<0)
template<typename T>
__device__ __forceinline__ cplx<T> muli(cplx<T> a, int sign)
{
  return sign > 0 ? mk(-a.y,a.x) : mk(a.y,-a.x);
}

// The pointwise products of the fused row kernels.  op(F, G, row, l) returns
// the product of the two transformed inputs at padded index l = q s + r of
// row `row` of the row stage; op(u, v, row, l) on two reals is the product of
// one residue of Hermitian rows, whose transforms are real (realMultBinary).
// The kernels call it through bind_row (below) once per row.
// LibMult<MULT>: multNone (0), multBinary (1), multcorrelation (3) fixed at
// compile time; RunMult: the same chosen at run time.  A caller's functor with
// these two members runs in the same kernels (fused user multipliers).
template<int MULT> struct LibMult {
  template<typename T>
  __device__ __forceinline__ cplx<T> operator()(cplx<T> a, cplx<T> b, int64_t, int) const
  {
    if constexpr(MULT == 1) return a*b;
    else if constexpr(MULT == 3) return mulc(a,b);
    else return a;
  }
  template<typename T>
  __device__ __forceinline__ T operator()(T a, T b, int64_t, int) const { return a*b; }
};

// A product functor bound to one row: op.at(row) when Op provides it (the
// functor's work that depends on the row alone, done once per row), else a
// wrapper that passes the row to every call.  Either way rop(F, G, l).
template<class Op> struct RowBound {
  Op op;
  int64_t row;
  template<typename T>
  __device__ __forceinline__ cplx<T> operator()(cplx<T> a, cplx<T> b, int l) const
  {
    return op(a,b,row,l);
  }
  template<typename T>
  __device__ __forceinline__ T operator()(T a, T b, int l) const { return op(a,b,row,l); }
};

template<class Op, class=void> struct HasAt { static constexpr bool value=false; };
template<class Op>
struct HasAt<Op,decltype((void) std::declval<const Op &>().at((int64_t) 0))> {
  static constexpr bool value=true;
};

template<class Op>
__device__ __forceinline__ auto bind_row(const Op &op, int64_t row)
{
  if constexpr(HasAt<Op>::value) return op.at(row);
  else return RowBound<Op>{op,row};
}

struct RunMult {
  int mult;
  template<typename T>
  __device__ __forceinline__ cplx<T> operator()(cplx<T> a, cplx<T> b, int64_t, int) const
  {
    return mult == 1 ? a*b : (mult == 3 ? mulc(a,b) : a);
  }
  template<typename T>
  __device__ __forceinline__ T operator()(T a, T b, int64_t, int) const { return a*b; }
};

// Geometry of one padded dimension as seen by a kernel.
constexpr int MAXSTAGES=24;

struct DimArgs {
  int64_t L;      // input length (full length for Hermitian dimensions)
  int64_t m;      // sub-FFT size
  int64_t q;      // number of residues
  int64_t N;      // padded length m*q
  int64_t k0;     // wavenumber of data index 0
  int64_t H;      // stored modes: L, or ceil(L/2) for a Hermitian dimension
  int nstages;
  int radix[MAXSTAGES];
  const void *zN; // zN[j] = e^{2 pi i j/N}, j < N, in the kernel's precision
};

// Column tile of block x of a grid row of gx blocks.  Blocks are dealt to
// the 8 XCDs round robin by linear id, so block x runs on XCD x mod 8 (when
// gx is a multiple of 8); runs of gx/8 consecutive tiles are given to one
// XCD, so that neighbouring tiles, which share cache lines when a tile row is
// narrower than a line, share that XCD's L2 instead of fetching the line
// twice.  A bijection of [0, gx).
__device__ __forceinline__ int64_t xcd_tile(int64_t x, int64_t gx)
{
  const int64_t per=gx/8;
  return x < 8*per ? (x % 8)*per+x/8 : x;
}

template<typename T>
__device__ __forceinline__ cplx<T> ztab(const DimArgs &d, int64_t j)
{
  return reinterpret_cast<const cplx<T> *>(d.zN)[j];
}

// e^{2 pi i e/N} for an arbitrary signed exponent e.
template<typename T>
__device__ __forceinline__ cplx<T> zeta(const DimArgs &d, int64_t e)
{
  int64_t r=e % d.N;
  if(r < 0) r += d.N;
  return ztab<T>(d,r);
}

// Buffer access: a complex word at byte offset off (per lane) + soff
// (uniform) of a raw buffer; one offset register serves every load of a
// thread whose words differ by uniform strides.
// Store of one complex word at byte offset off + soff of the buffer r.
template<typename T>
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff,
                                       cplx<T> v);
template<>
__device__ __forceinline__ void st_buf<double>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                               unsigned soff, cplx<double> v)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  d2 w; w.x=v.x; w.y=v.y;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4,w),r,off,soff,0);
}
template<>
__device__ __forceinline__ void st_buf<float>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                              unsigned soff, cplx<float> v)
{
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  f2 w; w.x=v.x; w.y=v.y;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2,w),r,off,soff,0);
}
// Plain (L1-cached) load of one complex word at byte offset off + soff of r.
template<typename T>
__device__ __forceinline__ cplx<T> ld_buf(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff);
template<>
__device__ __forceinline__ cplx<double> ld_buf<double>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                       unsigned soff)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v=__builtin_bit_cast(d2,__builtin_amdgcn_raw_buffer_load_b128(r,off,soff,0));
  return mk(v.x,v.y);
}
template<>
__device__ __forceinline__ cplx<float> ld_buf<float>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                     unsigned soff)
{
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v=__builtin_bit_cast(f2,__builtin_amdgcn_raw_buffer_load_b64(r,off,soff,0));
  return mk(v.x,v.y);
}

// Raw buffer resource over `words` complex words from base (offsets at or
// beyond the end read 0 and drop stores).
template<typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const cplx<T> *base, size_t words)
{
  return __builtin_amdgcn_make_buffer_rsrc((void *) base,0,(int) (words*sizeof(cplx<T>)),
                                           0x00020000);
}

} // namespace fftwpp_amd
