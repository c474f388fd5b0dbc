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

} // namespace fftwpp_amd
