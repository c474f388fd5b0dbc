// kernels.hip -- generic HIP kernels of the hybrid implicitly dealiased
// convolution and the thin C ABI (include/fftwpp_amd.h) through which the
// host C++ layer reaches them.
//
// Kernel families (reference counterparts in convolve.cc):
//   fwd_many    forward residue transforms of C strided columns
//               (fftPad/fftPadCentered forward*Many, :836-1466, :2199-4307)
//   bwd_many    backward residue transforms with residue accumulation
//               (backward*Many, :1468-1965)
//   conv_rows   fused forward -> multiply -> backward -> accumulate over all
//               residues of contiguous rows (Convolution::convolveRaw,
//               :7513-7575, with multBinary/multcorrelation, :33-110)
//   hconv_rows  the same for Hermitian rows (fftPadHermitian forward2/
//               backward2 :4517-5012 with realMultBinary :60-83); two real
//               residues share one complex sub-FFT
//   multiply    the multipliers applied to device arrays
//   symmetrize  HermitianSymmetrizeX/XY (convolve.h:1168-1267)
//   twiddles    FP64 generation of zeta_N^j (Complex.h:284-296 expi)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "fft_lds.h"
#include "kernels256.h"
#include "batch.h"
#include "rowsreg.h"
#include "plane256.h"
#include "rowsfused.h"
#include "../../include/fftwpp_amd.h"

using namespace fftwpp_amd;

namespace {

// ------------------------------------------------------------------------
// Twiddle generation: zN[j] = e^{2 pi i j/N} in FP64, rounded to FP32 for
// the single-precision copy.
// ------------------------------------------------------------------------
__global__ void twiddle_kernel(int64_t N, cplx<double> *z64, cplx<float> *z32)
{
  int64_t j=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(j >= N) return;
  double s, c;
  // reduce to [0,1/2] octant-free: sincospi(2j/N) with exact integer ratio
  sincospi(2.0*(double) j/(double) N,&s,&c);
  z64[j]=mk(c,s);
  z32[j]=mk((float) c,(float) s);
}

template<typename T>
__device__ __forceinline__ int64_t floormod(int64_t a, int64_t m)
{
  int64_t r=a % m;
  return r < 0 ? r+m : r;
}

// W_sigma = sum_{k = sigma mod m, k in data range} zeta_N^{k r} f(k)
// for a complex or centred dimension (k = j + k0, j < L).
template<typename T>
__device__ __forceinline__ cplx<T> fold(const DimArgs &d, const cplx<T> *f,
                                        int64_t stride, int64_t sigma,
                                        int64_t r)
{
  cplx<T> acc=mk((T) 0,(T) 0);
  int64_t j=floormod<T>(sigma-d.k0,d.m);
  for(; j < d.L; j += d.m) {
    int64_t k=j+d.k0;
    cplx<T> v=f[j*stride];
    acc=acc+zeta<T>(d,k*r)*v;
  }
  return acc;
}

// Hermitian dimension: f holds modes k in [0,H); f(-k)=conj f(k).
// Returns sum_{k = sigma mod m, |k| < H} (zeta^{k r1} + i zeta^{k r2}) f(k)
// (the second term only when r2 >= 0).
template<typename T>
__device__ __forceinline__ cplx<T> hfold(const DimArgs &d, const cplx<T> *f,
                                         int64_t stride, int64_t sigma,
                                         int64_t r1, int64_t r2)
{
  cplx<T> acc=mk((T) 0,(T) 0);
  int64_t H=d.H;
  // smallest k > -H with k = sigma (mod m)
  int64_t k=sigma-d.m*((sigma+H-1)/d.m);
  for(; k < H; k += d.m) {
    cplx<T> v=k >= 0 ? f[k*stride] : conjg(f[-k*stride]);
    cplx<T> w=zeta<T>(d,k*r1);
    if(r2 >= 0) w=w+muli(zeta<T>(d,k*r2),1);
    acc=acc+w*v;
  }
  return acc;
}

// ------------------------------------------------------------------------
// fwd_many: F[b][(r-r0)*m+s][c] = sum_k f[b][k][c] zeta_N^{k(qs+r)}
// ------------------------------------------------------------------------
// Column c of the input starts at c*ics and of the output at c*ocs: unit
// strides for interleaved columns; the row strides when contiguous rows are
// taken as the columns (then consecutive threads walk along a row).
template<typename T>
__global__ void fwd_many_kernel(DimArgs d, const cplx<T> *f, int64_t S,
                                int64_t C, int64_t bsi, cplx<T> *F,
                                int64_t SF, int64_t bso, int64_t r0,
                                int64_t nres, int TC, int64_t ics, int64_t ocs)
{
  extern __shared__ char smem_raw[];
  cplx<T> *A=reinterpret_cast<cplx<T> *>(smem_raw);
  const int m=(int) d.m;
  const int ld=m+1;
  cplx<T> *B=A+(int64_t) TC*ld;
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const int64_t b=blockIdx.y;
  const int nc=(int) min((int64_t) TC,C-c0);
  const cplx<T> *fb=f+b*bsi+c0*ics;
  cplx<T> *Fb=F+b*bso+c0*ocs;
  const int tid=threadIdx.x, nt=blockDim.x;
  for(int64_t rr=0; rr < nres; ++rr) {
    int64_t r=r0+rr;
    for(int idx=tid; idx < nc*m; idx += nt) {
      int sigma, cc;
      if(ics == 1) { sigma=idx/nc; cc=idx-sigma*nc; }
      else { cc=idx/m; sigma=idx-cc*m; }
      A[cc*ld+sigma]=fold<T>(d,fb+cc*ics,S,sigma,r);
    }
    __syncthreads();
    cplx<T> *X=lds_fft<T>(A,B,nc,ld,d,1);
    cplx<T> *Fr=Fb+rr*m*SF;
    for(int idx=tid; idx < nc*m; idx += nt) {
      int s, cc;
      if(ocs == 1) { s=idx/nc; cc=idx-s*nc; }
      else { cc=idx/m; s=idx-cc*m; }
      Fr[(int64_t) s*SF+cc*ocs]=X[cc*ld+s];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// bwd_many: f[b][j][c] (+)= sum_{r,s} F[b][(r-r0)*m+s][c] zeta_N^{-k(qs+r)},
// k = j + k0; the total is multiplied by scale after the last residue.
// ------------------------------------------------------------------------
template<typename T>
__global__ void bwd_many_kernel(DimArgs d, const cplx<T> *F, int64_t SF,
                                int64_t bsi, cplx<T> *f, int64_t S,
                                int64_t C, int64_t bso, int64_t r0,
                                int64_t nres, int accumulate, T scale, int TC,
                                int64_t ics, int64_t ocs)
{
  extern __shared__ char smem_raw[];
  cplx<T> *A=reinterpret_cast<cplx<T> *>(smem_raw);
  const int m=(int) d.m;
  const int ld=m+1;
  cplx<T> *B=A+(int64_t) TC*ld;
  const int64_t c0=xcd_tile(blockIdx.x,gridDim.x)*TC;
  const int64_t b=blockIdx.y;
  const int nc=(int) min((int64_t) TC,C-c0);
  const cplx<T> *Fb=F+b*bsi+c0*ics;
  cplx<T> *fb=f+b*bso+c0*ocs;
  const int tid=threadIdx.x, nt=blockDim.x;
  const int L=(int) d.L;
  for(int64_t rr=0; rr < nres; ++rr) {
    int64_t r=r0+rr;
    const cplx<T> *Fr=Fb+rr*m*SF;
    for(int idx=tid; idx < nc*m; idx += nt) {
      int s, cc;
      if(ics == 1) { s=idx/nc; cc=idx-s*nc; }
      else { cc=idx/m; s=idx-cc*m; }
      A[cc*ld+s]=Fr[(int64_t) s*SF+cc*ics];
    }
    __syncthreads();
    cplx<T> *X=lds_fft<T>(A,B,nc,ld,d,-1);
    bool last=rr == nres-1;
    bool add=accumulate || rr > 0;
    for(int idx=tid; idx < nc*L; idx += nt) {
      int j, cc;
      if(ocs == 1) { j=idx/nc; cc=idx-j*nc; }
      else { cc=idx/L; j=idx-cc*L; }
      int64_t k=j+d.k0;
      int sigma=(int) floormod<T>(k,d.m);
      cplx<T> v=mulc(X[cc*ld+sigma],zeta<T>(d,k*r));
      cplx<T> *o=fb+(int64_t) j*S+cc*ocs;
      if(add) v=v+*o;
      if(last) v=scale*v;
      *o=v;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// conv_rows: fused complex/centred row convolution.  Rows are contiguous
// (element stride 1); row i of f, g starts at i*rs, of out at i*os.
// MULT: 1 = multBinary (F0 *= F1), 3 = multcorrelation (F0 *= conj F1).
// ------------------------------------------------------------------------
template<typename T, int MULT>
__global__ void conv_rows_kernel(DimArgs d, const cplx<T> *f,
                                 const cplx<T> *g, int64_t rows, int64_t rs,
                                 cplx<T> *out, int64_t os, T scale, int TR,
                                 int ld)
{
  extern __shared__ char smem_raw[];
  const int m=(int) d.m;
  const int L=(int) d.L;
  cplx<T> *A=reinterpret_cast<cplx<T> *>(smem_raw);   // 2*TR sequences
  cplx<T> *B=A+(int64_t) 2*TR*ld;                     // 2*TR sequences
  cplx<T> *acc=B+(int64_t) 2*TR*ld;                   // TR rows of L
  const int64_t row0=(int64_t) blockIdx.x*TR;
  const int nr=(int) min((int64_t) TR,rows-row0);
  const int tid=threadIdx.x, nt=blockDim.x;
  for(int64_t r=0; r < d.q; ++r) {
    // build the folded, twiddled sequences of f (rows 0..nr-1) and g
    for(int idx=tid; idx < 2*nr*m; idx += nt) {
      int seq=idx/m, sigma=idx-seq*m;
      int a=seq/nr, i=seq-a*nr;
      const cplx<T> *src=(a == 0 ? f : g)+(row0+i)*rs;
      A[(a*TR+i)*ld+sigma]=fold<T>(d,src,1,sigma,r);
    }
    __syncthreads();
    // both inputs transformed together: sequences [0,TR) f, [TR,2TR) g
    cplx<T> *X=lds_fft<T>(A,B,2*TR,ld,d,1);
    cplx<T> *Y=X == A ? B : A;
    for(int idx=tid; idx < nr*m; idx += nt) {
      int i=idx/m, s=idx-i*m;
      cplx<T> u=X[i*ld+s], v=X[(TR+i)*ld+s];
      if(MULT != 0) X[i*ld+s]=MULT == 3 ? mulc(u,v) : u*v;
    }
    __syncthreads();
    cplx<T> *Z=lds_fft<T>(X,Y,TR,ld,d,-1);
    for(int idx=tid; idx < nr*L; idx += nt) {
      int i=idx/L, j=idx-i*L;
      int64_t k=j+d.k0;
      int sigma=(int) floormod<T>(k,d.m);
      cplx<T> v=mulc(Z[i*ld+sigma],zeta<T>(d,k*r));
      acc[i*L+j]=r == 0 ? v : acc[i*L+j]+v;
    }
    __syncthreads();
  }
  for(int idx=tid; idx < nr*L; idx += nt) {
    int i=idx/L, j=idx-i*L;
    out[(row0+i)*os+j]=scale*acc[i*L+j];
  }
}

// ------------------------------------------------------------------------
// hconv_rows: Hermitian rows (H stored modes).  Residues r1, r2=r1+1 are
// packed as W1 + i W2; both transforms are real, so one complex sub-FFT
// yields u_{r1} + i u_{r2}.  realMultBinary multiplies real parts and
// imaginary parts separately; the inverse is unpacked by Hermitian symmetry.
// ------------------------------------------------------------------------
template<typename T>
__global__ void hconv_rows_kernel(DimArgs d, const cplx<T> *f,
                                  const cplx<T> *g, int64_t rows, int64_t rs,
                                  cplx<T> *out, int64_t os, T scale, int TR,
                                  int ld)
{
  extern __shared__ char smem_raw[];
  const int m=(int) d.m;
  const int H=(int) d.H;
  cplx<T> *A=reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *B=A+(int64_t) 2*TR*ld;
  cplx<T> *acc=B+(int64_t) 2*TR*ld;
  const int64_t row0=(int64_t) blockIdx.x*TR;
  const int nr=(int) min((int64_t) TR,rows-row0);
  const int tid=threadIdx.x, nt=blockDim.x;
  for(int64_t r1=0; r1 < d.q; r1 += 2) {
    int64_t r2=r1+1 < d.q ? r1+1 : -1;
    for(int idx=tid; idx < 2*nr*m; idx += nt) {
      int seq=idx/m, sigma=idx-seq*m;
      int a=seq/nr, i=seq-a*nr;
      const cplx<T> *src=(a == 0 ? f : g)+(row0+i)*rs;
      A[(a*TR+i)*ld+sigma]=hfold<T>(d,src,1,sigma,r1,r2);
    }
    __syncthreads();
    cplx<T> *X=lds_fft<T>(A,B,2*TR,ld,d,1);
    cplx<T> *Y=X == A ? B : A;
    for(int idx=tid; idx < nr*m; idx += nt) {
      int i=idx/m, s=idx-i*m;
      cplx<T> u=X[i*ld+s], v=X[(TR+i)*ld+s];
      X[i*ld+s]=mk(u.x*v.x,u.y*v.y);
    }
    __syncthreads();
    cplx<T> *Z=lds_fft<T>(X,Y,TR,ld,d,-1);
    for(int idx=tid; idx < nr*H; idx += nt) {
      int i=idx/H, j=idx-i*H;
      int sigma=j % m;
      int msig=sigma == 0 ? 0 : m-sigma;
      cplx<T> z=Z[i*ld+sigma], zc=conjg(Z[i*ld+msig]);
      cplx<T> y1=(T) 0.5*(z+zc);
      cplx<T> v=mulc(y1,zeta<T>(d,(int64_t) j*r1));
      if(r2 >= 0) {
        cplx<T> y2=muli((T) 0.5*(z-zc),-1);
        v=v+mulc(y2,zeta<T>(d,(int64_t) j*r2));
      }
      acc[i*H+j]=r1 == 0 ? v : acc[i*H+j]+v;
    }
    __syncthreads();
  }
  for(int idx=tid; idx < nr*H; idx += nt) {
    int i=idx/H, j=idx-i*H;
    out[(row0+i)*os+j]=scale*acc[i*H+j];
  }
}

// ------------------------------------------------------------------------
// Elementwise multipliers on device arrays (convolve.cc:26-110).
// ------------------------------------------------------------------------
template<typename T>
__global__ void multiply_kernel(int mult, cplx<T> *F0, const cplx<T> *F1,
                                int64_t n)
{
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= n) return;
  if(mult == FFTWPP_AMD_MULT_BINARY) F0[i]=F0[i]*F1[i];
  else if(mult == FFTWPP_AMD_MULT_CORRELATION) F0[i]=mulc(F0[i],F1[i]);
  else if(mult == FFTWPP_AMD_MULT_REAL_BINARY) {
    // realMultBinary: n counts real words (doubles or floats)
    T *a=reinterpret_cast<T *>(F0);
    const T *b=reinterpret_cast<const T *>(F1);
    a[i]=a[i]*b[i];
  }
}

// Real convolutions: z = f0 + i 2^e f1, and out = s 2^-e Im z (the halved
// imaginary part of the self-convolution of z is 2^e f0 * f1).  The exponent
// e balances the two norms, ||f0|| ~ 2^e ||f1||, so that the rounding error
// of the self-convolution, of order eps ||z||^2, stays of order
// eps ||f0|| ||f1|| 2^e whatever the ratio of the input norms; scaling by a
// power of two is exact.  e = 0 when either input is zero or no exponent is
// given.
constexpr int REAL_NORM_BLOCKS=256;

// Partial sums of squares of f0 and f1 (in FP64) per block.
template<typename T>
__global__ void __launch_bounds__(256) real_norms_kernel(const T *f0, const T *f1, int64_t n,
                                                          double *part)
{
  __shared__ double s0[256], s1[256];
  double a=0.0, b=0.0;
  for(int64_t i=(int64_t) blockIdx.x*256+threadIdx.x; i < n; i += (int64_t) gridDim.x*256) {
    const double x=f0[i], y=f1[i];
    a=fma(x,x,a);
    b=fma(y,y,b);
  }
  s0[threadIdx.x]=a;
  s1[threadIdx.x]=b;
  __syncthreads();
  for(int w=128; w > 0; w >>= 1) {
    if((int) threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x+w];
      s1[threadIdx.x] += s1[threadIdx.x+w];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    part[blockIdx.x]=s0[0];
    part[REAL_NORM_BLOCKS+blockIdx.x]=s1[0];
  }
}

// One block: the block sums in a fixed order, then the balancing exponent
// e = round(log2(||f0|| / ||f1||)) written as an int after the partials.
__global__ void __launch_bounds__(256) real_balance_kernel(double *part)
{
  __shared__ double s0[256], s1[256];
  s0[threadIdx.x]=part[threadIdx.x];
  s1[threadIdx.x]=part[REAL_NORM_BLOCKS+threadIdx.x];
  __syncthreads();
  for(int w=128; w > 0; w >>= 1) {
    if((int) threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x+w];
      s1[threadIdx.x] += s1[threadIdx.x+w];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    int e=0;
    if(s0[0] > 0.0 && s1[0] > 0.0) {
      const double l=0.5*(log2(s0[0])-log2(s1[0]));
      e=(int) rint(fmin(fmax(l,-900.0),900.0));
    }
    int *out=reinterpret_cast<int *>(part+2*REAL_NORM_BLOCKS);
    out[0]=e;
    out[1]=s0[0] == 0.0 || s1[0] == 0.0;   // a zero input: the product is zero
  }
}

__device__ __forceinline__ int real_exponent(const double *bal)
{
  return bal ? reinterpret_cast<const int *>(bal+2*REAL_NORM_BLOCKS)[0] : 0;
}

__device__ __forceinline__ bool real_zero(const double *bal)
{
  return bal && reinterpret_cast<const int *>(bal+2*REAL_NORM_BLOCKS)[1] != 0;
}

template<typename T>
__global__ void real_pack_kernel(const T *f0, const T *f1, cplx<T> *z, int64_t n,
                                 const double *bal)
{
  const int e=real_exponent(bal);
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i < n) z[i]=mk(f0[i],(T) ldexp((double) f1[i],e));
}

template<typename T>
__global__ void real_unpack_kernel(const cplx<T> *z, T *out, int64_t n, T s,
                                   const double *bal)
{
  const int e=real_exponent(bal);
  const bool zero=real_zero(bal);
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i < n) out[i]=zero ? (T) 0 : (T) ldexp((double) (s*z[i].y),-e);
}

// Element maps of the standalone FFTs, y[b yd + j ys] = s op(x[b xd + j xs])
// for j < n, b < nb, strides in elements of each side's type.  op 0: complex
// copy; 1: real to complex (imaginary part 0); 2: complex to its real part;
// 3: the n-point spectrum of a Hermitian sequence from its n/2+1 stored modes
// (y_j = conj x_{n-j} for j > n/2); 4: real copy.
template<typename T>
__global__ void fft_map_kernel(int op, const void *x, int64_t xs, int64_t xd,
                               void *y, int64_t ys, int64_t yd, int64_t n,
                               int64_t nb, T s)
{
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= n*nb) return;
  int64_t b=i/n, j=i-b*n;
  if(op == 1) {
    T v=reinterpret_cast<const T *>(x)[b*xd+j*xs];
    reinterpret_cast<cplx<T> *>(y)[b*yd+j*ys]=mk(s*v,(T) 0);
    return;
  }
  if(op == 4) {
    reinterpret_cast<T *>(y)[b*yd+j*ys]=s*reinterpret_cast<const T *>(x)[b*xd+j*xs];
    return;
  }
  const cplx<T> *xc=reinterpret_cast<const cplx<T> *>(x)+b*xd;
  if(op == 2) {
    reinterpret_cast<T *>(y)[b*yd+j*ys]=s*xc[j*xs].x;
    return;
  }
  cplx<T> v;
  if(op == 3 && 2*j > n) v=conjg(xc[(n-j)*xs]);
  else v=xc[j*xs];
  reinterpret_cast<cplx<T> *>(y)[b*yd+j*ys]=s*v;
}

// The twiddle step of a four-step DFT of length n = n1 n2: y[b n + k1 n2 + j]
// *= exp(sign 2 pi i j k1 / n), the angle reduced exactly in integers and
// evaluated in FP64.
template<typename T>
__global__ void fft_twiddle_kernel(cplx<T> *y, int64_t n, int64_t n2, int64_t nb, int sign)
{
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= n*nb) return;
  int64_t r=i % n;
  int64_t k1=r/n2, j=r-k1*n2;
  int64_t e=(j*k1) % n;   // j k1 < n
  double sn, cs;
  sincospi(2.0*(double) e/(double) n,&sn,&cs);
  cplx<T> w=mk((T) cs,(T) (sign > 0 ? sn : -sn));
  y[i]=y[i]*w;
}

// Bluestein's algorithm for a length n without a usable divisor: with
// w_k = exp(sign pi i k^2 / n), X_k = w_k sum_j (x_j w_j) conj(w_{k-j}), a
// circular convolution of length len >= 2n - 1.  op 0: the chirp w (y, n
// words); op 1: the convolution kernel b_j = conj(w_j), b_{len-j} = conj(w_j)
// for 0 < j < n, else 0 (y, len words); op 2: y[b yd + j] = x[b xd + j xs] w_j
// for j < n and 0 up to len; op 3: y[b yd + k ys] = s w_k x[b xd + k] for k < n.
// k^2 is reduced modulo 2n in integers (k < 2^31) before the FP64 angle.
template<typename T>
__global__ void fft_chirp_kernel(int op, const cplx<T> *x, int64_t xs, int64_t xd,
                                 cplx<T> *y, int64_t ys, int64_t yd, int64_t n,
                                 int64_t len, int64_t nb, int sign, T s)
{
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  const int64_t cnt=op == 3 ? n : (op == 0 ? n : len);
  if(i >= cnt*nb) return;
  int64_t b=i/cnt, j=i-b*cnt;
  auto chirp=[&](int64_t k) {
    int64_t e=(k*k) % (2*n);
    double sn, cs;
    sincospi((double) e/(double) n,&sn,&cs);
    return mk((T) cs,(T) (sign > 0 ? sn : -sn));
  };
  if(op == 0) { y[j]=chirp(j); return; }
  if(op == 1) {
    cplx<T> v=mk((T) 0,(T) 0);
    if(j < n) v=conjg(chirp(j));
    else if(len-j < n) v=conjg(chirp(len-j));
    y[j]=v;
    return;
  }
  if(op == 2) {
    cplx<T> v=mk((T) 0,(T) 0);
    if(j < n) v=x[b*xd+j*xs]*chirp(j);
    y[b*yd+j]=v;
    return;
  }
  y[b*yd+j*ys]=s*(x[b*xd+j]*chirp(j));
}

// Origin shifts of real 2D/3D transforms.  op 0: negate the reals of row i
// (2D, i odd) or row (i, j) (3D, i + j odd) of an nx x [ny x] rows x rowlen
// array; op 1: zero the Nyquist modes of a shifted 2D spectrum (nx x nyp:
// row 0 for even nx, column nyp - 1 for even ny); op 2: the same in 3D
// (nx x ny x nzp: plane 0, rows j = 0, column nzp - 1 for even nx, ny, nz).
__global__ void fft_shift_kernel(int op, double *data, int64_t nx, int64_t ny, int64_t nz,
                                 int64_t rowlen)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(op == 0) {
    const int64_t rows=nz ? nx*ny : nx;
    if(i >= rows*rowlen) return;
    const int64_t row=i/rowlen;
    const int64_t par=nz ? row/ny+row % ny : row;
    if(par & 1) data[i]=-data[i];
    return;
  }
  cplx<double> *f=reinterpret_cast<cplx<double> *>(data);
  if(op == 1) {
    const int64_t nyp=ny/2+1;
    if(i >= nx*nyp) return;
    const int64_t x=i/nyp, y=i-x*nyp;
    if((nx % 2 == 0 && x == 0) || (ny % 2 == 0 && y == nyp-1)) f[i]=mk(0.0,0.0);
    return;
  }
  const int64_t nzp=nz/2+1;
  if(i >= nx*ny*nzp) return;
  const int64_t z=i % nzp, xy=i/nzp, y=xy % ny, x=xy/ny;
  if((nx % 2 == 0 && x == 0) || (ny % 2 == 0 && y == 0) || (nz % 2 == 0 && z == nzp-1))
    f[i]=mk(0.0,0.0);
}

// out[c][r] = in[r][c] for a rows x cols matrix of W-double elements: 32 x 32
// element tiles through LDS, rows read and written along their length.
template<int W>
__global__ void transpose_tile_kernel(const double *in, double *out, int64_t rows, int64_t cols)
{
  __shared__ double t[32][32*W+1];
  const int64_t r0=(int64_t) blockIdx.y*32, c0=(int64_t) blockIdx.x*32;
  for(int i=threadIdx.y; i < 32; i += 8) {
    const int64_t r=r0+i;
    for(int k=threadIdx.x; k < 32*W; k += 32)
      if(r < rows && c0+k/W < cols) t[i][k]=in[(r*cols+c0)*W+k];
  }
  __syncthreads();
  for(int i=threadIdx.y; i < 32; i += 8) {
    const int64_t c=c0+i;
    for(int k=threadIdx.x; k < 32*W; k += 32)
      if(c < cols && r0+k/W < rows) out[(c*rows+r0)*W+k]=t[k/W][i*W+k%W];
  }
}

// The same for elements of any w doubles, one double per thread in output
// order.
__global__ void transpose_words_kernel(const double *in, double *out, int64_t rows,
                                       int64_t cols, int64_t w)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= rows*cols*w) return;
  const int64_t e=i % w, rc=i/w;
  const int64_t c=rc/rows, r=rc-c*rows;
  out[i]=in[(r*cols+c)*w+e];
}

template<typename T>
__global__ void scale_kernel(cplx<T> *f, int64_t n, T s)
{
  int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i < n) f[i]=s*f[i];
}

// HermitianSymmetrizeX: F=f+x0*Sx; F[-i*Sx]=conj F[i*Sx] (0<i<Hx);
// Im F[0]=0; zero row 0 when x0 == Hx.
__global__ void symx_kernel(int64_t Hx, int64_t Hy, int64_t x0,
                            cplx<double> *f, int64_t Sx)
{
  int64_t t=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  cplx<double> *F=f+x0*Sx;
  if(t == 0) F[0].y=0.0;
  if(t >= 1 && t < Hx) F[-t*Sx]=conjg(F[t*Sx]);
  if(x0 == Hx && t < Hy) {
    // row 0 is the Nyquist row; it is never a source of the mirror above
    // because the mirror reads rows x0+1..x0+Hx-1 only.
    f[t]=mk(0.0,0.0);
  }
}

// HermitianSymmetrizeXY, stage 1: the x line through the origin.
__global__ void symxy1_kernel(int64_t Hx, cplx<double> *f, int64_t origin,
                              int64_t Sx)
{
  int64_t t=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  cplx<double> *F=f+origin;
  if(t == 0) F[0].y=0.0;
  if(t >= 1 && t < Hx) F[-t*Sx]=conjg(F[t*Sx]);
}

// stage 2: f[origin-i*Sx-j*Sy] = conj f[origin+i*Sx+j*Sy], -Hx<i<Hx, 0<j<Hy.
__global__ void symxy2_kernel(int64_t Hx, int64_t Hy, cplx<double> *f,
                              int64_t origin, int64_t Sx, int64_t Sy)
{
  int64_t t=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  int64_t ni=2*Hx-1, nj=Hy-1;
  if(t >= ni*nj) return;
  int64_t i=t/nj-(Hx-1);
  int64_t j=t % nj+1;
  f[origin-i*Sx-j*Sy]=conjg(f[origin+i*Sx+j*Sy]);
}

// stage 3: Nyquist zeroing.
__global__ void symxy3_kernel(int64_t Hx, int64_t Hy, int64_t Hz, int64_t x0,
                              int64_t y0, cplx<double> *f, int64_t Sx,
                              int64_t Sy)
{
  int64_t t=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(x0 == Hx) {
    int64_t Ly=y0+Hy;
    if(t < Ly*Hz) f[Sy*(t/Hz)+t % Hz]=mk(0.0,0.0);
  }
  if(y0 == Hy) {
    int64_t Lx=x0+Hx;
    if(t < Lx*Hz) f[Sx*(t/Hz)+t % Hz]=mk(0.0,0.0);
  }
}

// ------------------------------------------------------------------------
// Self-conjugate plane of a centred Hermitian array (the kz = 0 plane in 3D,
// the ky = 0 line in 2D, k = 0 in 1D): element (i, j), |i| < Hx, |j| < Hy,
// at f[(x0+i) Sx + (y0+j) Sy] pairs with (-i, -j).  "Upper" elements are
// j > 0, or j = 0 and i > 0 (the ones HermitianSymmetrizeX/XY copy from,
// convolve.h:1177-1256); one thread per upper element or the origin.
//   op 0: f(-i,-j) <- conj f(i,j), f(0,0) <- Re f(0,0)   (symmetrize)
//   op 1: both members <- w + conj(partner), w(0,0) <- 2 Re w(0,0)
//   op 2: upper kept, partner zeroed, origin halved (adjoint of op 0)
// ------------------------------------------------------------------------
template<typename T>
__global__ void hplane_kernel(int op, int64_t Hx, int64_t Hy, int64_t x0,
                              int64_t y0, cplx<T> *f, int64_t Sx, int64_t Sy)
{
  int64_t t=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(t >= (2*Hx-1)*Hy) return;
  const int64_t i=t/Hy-(Hx-1), j=t % Hy;
  if(j == 0 && i < 0) return;
  cplx<T> *u=f+(x0+i)*Sx+(y0+j)*Sy;
  if(i == 0 && j == 0) {
    const cplx<T> v=*u;
    if(op == 0) *u=mk(v.x,(T) 0);
    else if(op == 1) *u=mk(2*v.x,(T) 0);
    else *u=mk(v.x/2,v.y/2);
    return;
  }
  cplx<T> *l=f+(x0-i)*Sx+(y0-j)*Sy;
  if(op == 0) *l=conjg(*u);
  else if(op == 1) {
    const cplx<T> a=*u, b=*l;
    *u=a+conjg(b);
    *l=b+conjg(a);
  } else *l=mk((T) 0,(T) 0);
}

// Reference-layout residue blocks (fftBase::forward/backward of the C++
// face).  Entry e of a block reads or writes C columns at word offset
// tab[2e+1] of the block array (complex words, or reals when `real`), and
// its value lives at word (tab[2e] * C + c) of the residue buffer T, where
// tab[2e] = slot m + s for the residue in slot `slot` at sub-FFT position s.
__global__ void block_gather_kernel(const cplx<double> *T, const uint64_t *tab,
                                    int64_t ne, int64_t C, void *out, int real)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= ne*C) return;
  const int64_t e=i/C, c=i-e*C;
  const cplx<double> v=T[(int64_t) tab[2*e]*C+c];
  const int64_t w=(int64_t) tab[2*e+1]+c;
  if(real) reinterpret_cast<double *>(out)[w]=v.x;
  else reinterpret_cast<cplx<double> *>(out)[w]=v;
}

__global__ void block_scatter_kernel(cplx<double> *T, const uint64_t *tab, int64_t ne,
                                     int64_t C, const void *in, int real)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i >= ne*C) return;
  const int64_t e=i/C, c=i-e*C;
  const int64_t w=(int64_t) tab[2*e+1]+c;
  T[(int64_t) tab[2*e]*C+c]=real ? mk(reinterpret_cast<const double *>(in)[w],0.0) :
    reinterpret_cast<const cplx<double> *>(in)[w];
}

// Hermitian data of H stored modes (C columns) to the full centred sequence
// of 2H - 1 modes k = -(H-1) ... H-1 (op 0: f(-k) = conj f(k)); op 1 the
// way back, h(k) (+)= full(k) for k >= 0 (accumulate when acc).
__global__ void herm_expand_kernel(int op, const cplx<double> *a, cplx<double> *b, int64_t H,
                                   int64_t C, int acc)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(op == 0) {
    if(i >= (2*H-1)*C) return;
    const int64_t j=i/C, c=i-j*C, k=j-(H-1);
    b[i]=k >= 0 ? a[k*C+c] : conjg(a[-k*C+c]);
  } else {
    if(i >= H*C) return;
    const cplx<double> v=a[i+(H-1)*C];
    b[i]=acc ? b[i]+v : v;
  }
}

// Item table {i rs, L, i L} of `rows` rows of length L at row stride rs (the
// rows of a 2D/3D inner stage run as the items of the batched row kernel).
__global__ void row_items_kernel(BatchItem *items, int64_t rows, int64_t rs, int64_t L)
{
  const int64_t i=(int64_t) blockIdx.x*blockDim.x+threadIdx.x;
  if(i < rows) {
    BatchItem it;
    it.off=i*rs;
    it.L=L;
    it.poff=i*L;
    items[i]=it;
  }
}

// ------------------------------------------------------------------------
// Host-side helpers of the C ABI
// ------------------------------------------------------------------------
thread_local char last_error[512];

int fail(int code, const char *msg)
{
  strncpy(last_error,msg,sizeof(last_error)-1);
  last_error[sizeof(last_error)-1]=0;
  return code;
}

int hipcheck(hipError_t e, const char *what)
{
  if(e == hipSuccess) return FFTWPP_AMD_OK;
  char buf[512];
  snprintf(buf,sizeof(buf),"%s: %s",what,hipGetErrorString(e));
  return fail(FFTWPP_AMD_ERR_HIP,buf);
}

} // namespace

struct fftwpp_amd_dim {
  DimArgs a64, a32;
  void *z64, *z32;
  size_t L, m, q, N;
  int kind;
};

static int launch_check(const char *what)
{
  return hipcheck(hipGetLastError(),what);
}

static size_t lds_budget() { return 64*1024; }

// Kernels that take more than the default 64 KB of dynamic LDS must raise
// their limit first (gfx950 has 160 KB per workgroup).
// The limit is raised once per kernel to the full 160 KB, so launches
// recorded into a graph later make no further runtime calls.
static size_t lds_max() { return 160*1024; }

template<typename K>
static int allow_lds(K kernel, size_t bytes)
{
  if(bytes <= 64*1024) return FFTWPP_AMD_OK;
  static std::mutex mu;
  static std::vector<const void *> raised;
  std::lock_guard<std::mutex> lock(mu);
  const void *fn=(const void *) kernel;
  if(std::find(raised.begin(),raised.end(),fn) != raised.end()) return FFTWPP_AMD_OK;
  int rc=hipcheck(hipFuncSetAttribute(fn,hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int) lds_max()),"hipFuncSetAttribute");
  if(!rc) raised.push_back(fn);
  return rc;
}

// Blocks of `kernel` that stay resident on the whole device at once
// (occupancy per CU times the CU count of the current device), at least 1:
// the grid of a persistent kernel.
// Cached per (kernel, device).
template<typename K>
static size_t resident_blocks(K kernel, int threads, size_t lds)
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

// Resident blocks per CU of `kernel` at `threads` threads and `lds` bytes of
// dynamic LDS (0 on failure), cached per (kernel, threads, lds).
template<typename K>
static int blocks_per_cu(K kernel, int threads, size_t lds)
{
  static std::mutex mu;
  static std::vector<std::pair<std::pair<const void *,size_t>,int>> known;
  const std::pair<const void *,size_t> key((const void *) kernel,lds*4096+(size_t) threads);
  std::lock_guard<std::mutex> lock(mu);
  for(const auto &e : known) if(e.first == key) return e.second;
  int per=0;
  if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per,(const void *) kernel,threads,lds) != hipSuccess)
    per=0;
  known.push_back({key,per});
  return per;
}

// The register-resident m=256 kernels apply to non-centred complex
// dimensions with p = 1.
static bool fast256(const fftwpp_amd_dim *d, size_t s1=0, size_t s2=0)
{
  // lane offsets (at most 16 strides plus a tile width) must fit in 31 bits
  const size_t lim=(size_t) 1 << 30;
  return d->m == 256 && d->kind == FFTWPP_AMD_KIND_COMPLEX && d->L <= 256 &&
    d->q <= 8 && 16*s1+64 < lim && 16*s2+64 < lim;
}

template<typename T>
static constexpr int rows256() { return sizeof(T) == 8 ? 8 : 16; }

// Whether the register-FFT column kernels take dimension d: power-of-two m in
// [16, 4096], complex or centred, p = ceil(L/m) <= 2.
static bool cols_reg(const fftwpp_amd_dim *d)
{
  const size_t m=d->m;
  if(m < 16 || m > 4096 || (m & (m-1))) return false;
  if(d->kind == FFTWPP_AMD_KIND_HERMITIAN) return false;
  return d->L <= 2*m && m*d->q < ((size_t) 1 << 30);
}

template<typename T, int M>
static int colreg_launch(const DimArgs &a, bool fwd, const void *in,
                         size_t s_in, size_t bsi, void *out, size_t s_out,
                         size_t bso, size_t C, size_t nbatch, size_t r0,
                         size_t nres, int accumulate, double scale,
                         size_t ics, size_t ocs, hipStream_t st)
{
  typedef ColReg<T,M> CR;
  size_t lds=M*sizeof(cplx<T>)+CR::TC*CR::CS*sizeof(T);
  const dim3 grid((unsigned) ((C+CR::TC-1)/CR::TC),(unsigned) nbatch);
  const bool p2=a.L > M;
  // the N residue twiddles in LDS when they fit without lowering occupancy
  auto zfit=[&](auto kernel) {
    const size_t z=lds+(size_t) a.N*sizeof(cplx<T>);
    return z <= lds_max() && blocks_per_cu(kernel,CR::NT,z) == blocks_per_cu(kernel,CR::NT,lds);
  };
  int rc, zl=0;
  if(fwd) {
    if(p2) {
      if((rc=allow_lds(colfwd_reg_kernel<T,M,2>,lds_max()))) return rc;
      if((zl=zfit(colfwd_reg_kernel<T,M,2>))) lds += (size_t) a.N*sizeof(cplx<T>);
      hipLaunchKernelGGL((colfwd_reg_kernel<T,M,2>),grid,dim3(CR::NT),lds,st,a,
                         (const cplx<T> *) in,(int64_t) s_in,(int64_t) C,(int64_t) bsi,
                         (cplx<T> *) out,(int64_t) s_out,(int64_t) bso,(int64_t) r0,(int64_t) nres,
                         (int64_t) ics,(int64_t) ocs,zl);
    } else {
      if((rc=allow_lds(colfwd_reg_kernel<T,M,1>,lds_max()))) return rc;
      if((zl=zfit(colfwd_reg_kernel<T,M,1>))) lds += (size_t) a.N*sizeof(cplx<T>);
      hipLaunchKernelGGL((colfwd_reg_kernel<T,M,1>),grid,dim3(CR::NT),lds,st,a,
                         (const cplx<T> *) in,(int64_t) s_in,(int64_t) C,(int64_t) bsi,
                         (cplx<T> *) out,(int64_t) s_out,(int64_t) bso,(int64_t) r0,(int64_t) nres,
                         (int64_t) ics,(int64_t) ocs,zl);
    }
    return launch_check("colfwd_reg_kernel");
  }
  if(p2) {
    if((rc=allow_lds(colbwd_reg_kernel<T,M,2>,lds_max()))) return rc;
    if((zl=zfit(colbwd_reg_kernel<T,M,2>))) lds += (size_t) a.N*sizeof(cplx<T>);
    hipLaunchKernelGGL((colbwd_reg_kernel<T,M,2>),grid,dim3(CR::NT),lds,st,a,
                       (const cplx<T> *) in,(int64_t) s_in,(int64_t) bsi,(cplx<T> *) out,
                       (int64_t) s_out,(int64_t) C,(int64_t) bso,(int64_t) r0,(int64_t) nres,
                       accumulate,(T) scale,(int64_t) ics,(int64_t) ocs,zl);
  } else {
    if((rc=allow_lds(colbwd_reg_kernel<T,M,1>,lds_max()))) return rc;
    if((zl=zfit(colbwd_reg_kernel<T,M,1>))) lds += (size_t) a.N*sizeof(cplx<T>);
    hipLaunchKernelGGL((colbwd_reg_kernel<T,M,1>),grid,dim3(CR::NT),lds,st,a,
                       (const cplx<T> *) in,(int64_t) s_in,(int64_t) bsi,(cplx<T> *) out,
                       (int64_t) s_out,(int64_t) C,(int64_t) bso,(int64_t) r0,(int64_t) nres,
                       accumulate,(T) scale,(int64_t) ics,(int64_t) ocs,zl);
  }
  return launch_check("colbwd_reg_kernel");
}

template<typename T>
static int colreg_t(const DimArgs &a, bool fwd, const void *in, size_t s_in,
                    size_t bsi, void *out, size_t s_out, size_t bso, size_t C,
                    size_t nbatch, size_t r0, size_t nres, int accumulate,
                    double scale, size_t ics, size_t ocs, hipStream_t st)
{
#define FFTWPP_AMD_COLREG(MM) case MM: return colreg_launch<T,MM>(a,fwd,in,s_in,bsi,out,s_out,bso,C,nbatch,r0,nres,accumulate,scale,ics,ocs,st);
  switch(a.m) {
  FFTWPP_AMD_COLREG(16) FFTWPP_AMD_COLREG(32) FFTWPP_AMD_COLREG(64)
  FFTWPP_AMD_COLREG(128) FFTWPP_AMD_COLREG(256) FFTWPP_AMD_COLREG(512)
  FFTWPP_AMD_COLREG(1024) FFTWPP_AMD_COLREG(2048) FFTWPP_AMD_COLREG(4096)
  default: return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"register FFT size");
  }
#undef FFTWPP_AMD_COLREG
}

// Real columns (colfwd_real_kernel, colbwd_real_kernel): power-of-two m in
// [16, 4096], p = ceil(L/m) <= 2, non-centred data.
static bool cols_real(const fftwpp_amd_dim *d)
{
  const size_t m=d->m;
  if(m < 16 || m > 4096 || (m & (m-1))) return false;
  if(d->kind != FFTWPP_AMD_KIND_COMPLEX) return false;
  return d->L <= 2*m && m*d->q < ((size_t) 1 << 30);
}

template<typename T, int M>
static int colreal_launch(const DimArgs &a, bool fwd, const void *in,
                          size_t s_in, size_t bsi, void *out, size_t s_out,
                          size_t bso, size_t C, size_t nbatch, double scale,
                          size_t ics, size_t ocs, hipStream_t st)
{
  typedef ColReg<T,M> CR;
  const size_t lds=M*sizeof(cplx<T>)+CR::TC*CR::CS*sizeof(T);
  const dim3 grid((unsigned) ((C+CR::TC-1)/CR::TC),(unsigned) nbatch);
  const bool p2=a.L > M;
  int rc;
  if(fwd) {
    if(p2) {
      if((rc=allow_lds(colfwd_real_kernel<T,M,2>,lds))) return rc;
      hipLaunchKernelGGL((colfwd_real_kernel<T,M,2>),grid,dim3(CR::NT),lds,st,a,
                         (const T *) in,(int64_t) s_in,(int64_t) C,(int64_t) bsi,
                         (cplx<T> *) out,(int64_t) s_out,(int64_t) bso,(int64_t) ics,
                         (int64_t) ocs);
    } else {
      if((rc=allow_lds(colfwd_real_kernel<T,M,1>,lds))) return rc;
      hipLaunchKernelGGL((colfwd_real_kernel<T,M,1>),grid,dim3(CR::NT),lds,st,a,
                         (const T *) in,(int64_t) s_in,(int64_t) C,(int64_t) bsi,
                         (cplx<T> *) out,(int64_t) s_out,(int64_t) bso,(int64_t) ics,
                         (int64_t) ocs);
    }
    return launch_check("colfwd_real_kernel");
  }
  if(p2) {
    if((rc=allow_lds(colbwd_real_kernel<T,M,2>,lds))) return rc;
    hipLaunchKernelGGL((colbwd_real_kernel<T,M,2>),grid,dim3(CR::NT),lds,st,a,
                       (const cplx<T> *) in,(int64_t) s_in,(int64_t) bsi,(T *) out,
                       (int64_t) s_out,(int64_t) C,(int64_t) bso,(T) scale,(int64_t) ics,
                       (int64_t) ocs);
  } else {
    if((rc=allow_lds(colbwd_real_kernel<T,M,1>,lds))) return rc;
    hipLaunchKernelGGL((colbwd_real_kernel<T,M,1>),grid,dim3(CR::NT),lds,st,a,
                       (const cplx<T> *) in,(int64_t) s_in,(int64_t) bsi,(T *) out,
                       (int64_t) s_out,(int64_t) C,(int64_t) bso,(T) scale,(int64_t) ics,
                       (int64_t) ocs);
  }
  return launch_check("colbwd_real_kernel");
}

// Single columns of many batch entries become the columns of one launch, as
// in fwd_many_t.
template<typename T>
static int colreal_t(const fftwpp_amd_dim *d, bool fwd, const void *in, size_t s_in,
                     size_t bsi, void *out, size_t s_out, size_t bso, size_t C,
                     size_t nbatch, double scale, hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  size_t ics=1, ocs=1;
  if(C == 1 && nbatch > 1) {
    C=nbatch; ics=bsi; ocs=bso;
    nbatch=1; bsi=bso=0;
  }
#define FFTWPP_AMD_COLREAL(MM) case MM: return colreal_launch<T,MM>(a,fwd,in,s_in,bsi,out,s_out,bso,C,nbatch,scale,ics,ocs,st);
  switch(a.m) {
  FFTWPP_AMD_COLREAL(16) FFTWPP_AMD_COLREAL(32) FFTWPP_AMD_COLREAL(64)
  FFTWPP_AMD_COLREAL(128) FFTWPP_AMD_COLREAL(256) FFTWPP_AMD_COLREAL(512)
  FFTWPP_AMD_COLREAL(1024) FFTWPP_AMD_COLREAL(2048) FFTWPP_AMD_COLREAL(4096)
  default: return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"real residue transforms: m");
  }
#undef FFTWPP_AMD_COLREAL
}

template<typename T>
static int fwd_many_t(const fftwpp_amd_dim *d, const void *f, size_t S,
                      size_t C, size_t nbatch, size_t bsi, void *F,
                      size_t SF, size_t bso, size_t r0, size_t nres,
                      hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  // single columns of many batch entries: the entries become the columns
  // (column strides ics, ocs: the row strides)
  size_t ics=1, ocs=1;
  if(C == 1 && nbatch > 1) {
    C=nbatch; ics=bsi; ocs=bso;
    nbatch=1; bsi=bso=0;
  }
  // the 256-point column kernel fills its twiddle tables per block: taken
  // for interleaved columns with a full tile of them
  if(ics == 1 && ocs == 1 && C >= (size_t) Col256<T>::TC && fast256(d,S,SF)) {
    constexpr int TC=Col256<T>::TC;
    dim3 grid((unsigned) ((C+TC-1)/TC),(unsigned) nbatch);
    size_t lds=col256_lds<T>(nres);
    int rc=allow_lds(colfwd256_kernel<T>,lds);
    if(rc) return rc;
    hipLaunchKernelGGL((colfwd256_kernel<T>),grid,dim3(16*TC),lds,st,
                       (const cplx<T> *) f,(int64_t) S,(int64_t) C,(int64_t) bsi,
                       (cplx<T> *) F,(int64_t) SF,(int64_t) bso,(int) d->L,
                       (int64_t) r0,(int64_t) nres,(const cplx<T> *) a.zN,(int64_t) d->q);
    return launch_check("colfwd256_kernel");
  }
  // register columns: tiles of several columns below m = 2048, so a single
  // short column goes to the LDS path
  if(cols_reg(d) && (C >= 8 || d->m >= 2048))
    return colreg_t<T>(a,true,f,S,bsi,F,SF,bso,C,nbatch,r0,nres,0,1.0,ics,ocs,st);
  size_t elt=sizeof(cplx<T>);
  size_t per=2*(d->m+1)*elt;
  if(per > lds_max()) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"m too large for the LDS transform");
  size_t TC=1;
  while(TC < 16 && TC < C && 2*TC*per <= lds_budget()) TC *= 2;
  size_t lds=TC*per;
  dim3 grid((unsigned) ((C+TC-1)/TC),(unsigned) nbatch);
  int rc=allow_lds(fwd_many_kernel<T>,lds);
  if(rc) return rc;
  hipLaunchKernelGGL(fwd_many_kernel<T>,grid,dim3(256),lds,st,a,
                     (const cplx<T> *) f,(int64_t) S,(int64_t) C,
                     (int64_t) bsi,(cplx<T> *) F,(int64_t) SF,(int64_t) bso,
                     (int64_t) r0,(int64_t) nres,(int) TC,(int64_t) ics,(int64_t) ocs);
  return launch_check("fwd_many_kernel");
}

template<typename T>
static int bwd_many_t(const fftwpp_amd_dim *d, const void *F, size_t SF,
                      size_t bsi, void *f, size_t S, size_t C,
                      size_t nbatch, size_t bso, size_t r0, size_t nres,
                      int accumulate, double scale, hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  // single columns of many batch entries: the entries become the columns
  size_t ics=1, ocs=1;
  if(C == 1 && nbatch > 1) {
    C=nbatch; ics=bsi; ocs=bso;
    nbatch=1; bsi=bso=0;
  }
  if(ics == 1 && ocs == 1 && C >= (size_t) Col256<T>::TC && fast256(d,S,SF)) {
    constexpr int TC=Col256<T>::TC;
    dim3 grid((unsigned) ((C+TC-1)/TC),(unsigned) nbatch);
    size_t lds=col256_lds<T>(nres);
    int rc=allow_lds(colbwd256_kernel<T>,lds);
    if(rc) return rc;
    hipLaunchKernelGGL((colbwd256_kernel<T>),grid,dim3(16*TC),lds,st,
                       (const cplx<T> *) F,(int64_t) SF,(int64_t) bsi,(cplx<T> *) f,
                       (int64_t) S,(int64_t) C,(int64_t) bso,(int) d->L,(int64_t) r0,
                       (int64_t) nres,accumulate,(T) scale,(const cplx<T> *) a.zN,
                       (int64_t) d->q);
    return launch_check("colbwd256_kernel");
  }
  if(cols_reg(d) && (C >= 8 || d->m >= 2048))
    return colreg_t<T>(a,false,F,SF,bsi,f,S,bso,C,nbatch,r0,nres,accumulate,scale,ics,ocs,st);
  size_t elt=sizeof(cplx<T>);
  size_t per=2*(d->m+1)*elt;
  if(per > lds_max()) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"m too large for the LDS transform");
  size_t TC=1;
  while(TC < 16 && TC < C && 2*TC*per <= lds_budget()) TC *= 2;
  size_t lds=TC*per;
  dim3 grid((unsigned) ((C+TC-1)/TC),(unsigned) nbatch);
  int rc=allow_lds(bwd_many_kernel<T>,lds);
  if(rc) return rc;
  hipLaunchKernelGGL(bwd_many_kernel<T>,grid,dim3(256),lds,st,a,
                     (const cplx<T> *) F,(int64_t) SF,(int64_t) bsi,
                     (cplx<T> *) f,(int64_t) S,(int64_t) C,(int64_t) bso,
                     (int64_t) r0,(int64_t) nres,accumulate,(T) scale,(int) TC,
                     (int64_t) ics,(int64_t) ocs);
  return launch_check("bwd_many_kernel");
}

// LDS bytes of batch_conv_kernel for sub-FFT size m, TR items and (with
// LDS accumulation) rows of lacc words; ld = m+1 (bank padding) when that
// fits, else m.
static size_t batch_lds(size_t m, size_t elt, size_t TR, size_t lacc,
                        size_t *ldout)
{
  size_t ld=m+1;
  const size_t base=batch_items_bytes(TR);
  if(base+(4*TR*ld+TR*lacc)*elt > lds_max()) ld=m;
  if(ldout) *ldout=ld;
  return base+(4*TR*ld+TR*lacc)*elt;
}

struct BatchLaunch {
  const BatchItem *items;
  size_t count;
  const void *f, *g;
  void *out, *part;
  size_t TR, threads, nblocks, lacc, D, ptotal, ld, lds;
  double scale;
};

template<typename T, int MULT>
static int batch_rows_launch(const DimArgs &a, const BatchLaunch &b,
                             hipStream_t st)
{
  int rc;
  if(b.lacc) {
    if((rc=allow_lds(batch_conv_kernel<T,MULT,true>,b.lds))) return rc;
    hipLaunchKernelGGL((batch_conv_kernel<T,MULT,true>),dim3((unsigned) b.nblocks),
                       dim3((unsigned) b.threads),b.lds,st,a,b.items,(int64_t) b.count,
                       (const cplx<T> *) b.f,(const cplx<T> *) b.g,(cplx<T> *) b.out,
                       (T) b.scale,(int) b.TR,(int) b.ld,(int) b.lacc,(int) b.D,
                       (cplx<T> *) b.part,(int64_t) b.ptotal);
  } else {
    if((rc=allow_lds(batch_conv_kernel<T,MULT,false>,b.lds))) return rc;
    hipLaunchKernelGGL((batch_conv_kernel<T,MULT,false>),dim3((unsigned) b.nblocks),
                       dim3((unsigned) b.threads),b.lds,st,a,b.items,(int64_t) b.count,
                       (const cplx<T> *) b.f,(const cplx<T> *) b.g,(cplx<T> *) b.out,
                       (T) b.scale,(int) b.TR,(int) b.ld,0,(int) b.D,
                       (cplx<T> *) b.part,(int64_t) b.ptotal);
  }
  rc=launch_check("batch_conv_kernel");
  if(rc) return rc;
  const size_t G=(a.q+b.D-1)/b.D;
  if(G > 1) {
    const size_t nb=b.count < 4096 ? b.count : 4096;
    hipLaunchKernelGGL(batch_reduce_kernel<T>,dim3((unsigned) nb),dim3(256),0,st,
                       b.items,(int64_t) b.count,(const cplx<T> *) b.part,
                       (int64_t) b.ptotal,(int) G,(cplx<T> *) b.out,(T) b.scale);
    rc=launch_check("batch_reduce_kernel");
  }
  return rc;
}

// Threads per sequence of the register-FFT batch kernel for sub-FFT size m
// (a power of two in [16, 4096]) and element size elt, 0 when the LDS kernel
// runs.
static size_t batch_tps(size_t m, size_t elt)
{
  if(m < 16 || m > 4096 || (m & (m-1))) return 0;
  return m/(elt == sizeof(cplx<double>) ? BatchV<double>::V : BatchV<float>::V);
}

// Threads per block and LDS bytes of the register-FFT kernel for q residues.
static void batch_reg_shape(size_t m, size_t q, size_t elt, size_t threads,
                            size_t *nt, size_t *lds)
{
  const size_t tps=batch_tps(m,elt);
  size_t n=tps > 64 ? tps : std::max(threads,tps)/tps*tps;
  if(n > 512) n=512/tps*tps;
  const size_t pad=m+m/16;
  *nt=n;
  *lds=m*elt+(n/tps)*2*pad*(elt/2)+q*(m/tps)*elt;
}

template<typename T, int M, bool ZM>
static int batch_reg_launch_z(const DimArgs &a, const BatchLaunch &b, int mult,
                              hipStream_t st)
{
  size_t nt, lds;
  batch_reg_shape(M,(size_t) a.q,sizeof(cplx<T>),b.threads,&nt,&lds);
  if(lds > lds_max()) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"batch tile does not fit in LDS");
  int rc=allow_lds(batch_reg_kernel<T,M,ZM>,lds);
  if(rc) return rc;
  hipLaunchKernelGGL((batch_reg_kernel<T,M,ZM>),dim3((unsigned) b.nblocks),dim3((unsigned) nt),
                     lds,st,a,b.items,(int64_t) b.count,(const cplx<T> *) b.f,
                     (const cplx<T> *) b.g,(cplx<T> *) b.out,(T) b.scale,mult,(int) b.D,
                     (cplx<T> *) b.part,(int64_t) b.ptotal);
  return launch_check("batch_reg_kernel");
}

template<typename T, int M>
static int batch_reg_launch(const DimArgs &a, const BatchLaunch &b, int mult,
                            hipStream_t st)
{
  return b.D == 0 ? batch_reg_launch_z<T,M,true>(a,b,mult,st)
                  : batch_reg_launch_z<T,M,false>(a,b,mult,st);
}

template<typename T>
static int batch_reg_t(const DimArgs &a, const BatchLaunch &b, int mult,
                       hipStream_t st)
{
  int rc;
  switch(a.m) {
  case 16: rc=batch_reg_launch<T,16>(a,b,mult,st); break;
  case 32: rc=batch_reg_launch<T,32>(a,b,mult,st); break;
  case 64: rc=batch_reg_launch<T,64>(a,b,mult,st); break;
  case 128: rc=batch_reg_launch<T,128>(a,b,mult,st); break;
  case 256: rc=batch_reg_launch<T,256>(a,b,mult,st); break;
  case 512: rc=batch_reg_launch<T,512>(a,b,mult,st); break;
  case 1024: rc=batch_reg_launch<T,1024>(a,b,mult,st); break;
  case 2048: rc=batch_reg_launch<T,2048>(a,b,mult,st); break;
  case 4096: rc=batch_reg_launch<T,4096>(a,b,mult,st); break;
  default: return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"register FFT size");
  }
  if(rc) return rc;
  if(b.D == 0) {
    const size_t nb=b.count < 4096 ? b.count : 4096;
    const dim3 grid((unsigned) nb), blk(256);
    const BatchItem *it=b.items;
    const int64_t n=(int64_t) b.count;
    const cplx<T> *Z=(const cplx<T> *) b.part, *zN=(const cplx<T> *) a.zN;
    cplx<T> *out=(cplx<T> *) b.out;
    const int q=(int) a.q, m=(int) a.m;
    if(q <= 4)
      hipLaunchKernelGGL((batch_unfold_kernel<T,4>),grid,blk,0,st,it,n,Z,q,m,zN,out,(T) b.scale);
    else if(q <= 8)
      hipLaunchKernelGGL((batch_unfold_kernel<T,8>),grid,blk,0,st,it,n,Z,q,m,zN,out,(T) b.scale);
    else if(q <= 16)
      hipLaunchKernelGGL((batch_unfold_kernel<T,16>),grid,blk,0,st,it,n,Z,q,m,zN,out,(T) b.scale);
    else
      hipLaunchKernelGGL((batch_unfold_kernel<T,0>),grid,blk,0,st,it,n,Z,q,m,zN,out,(T) b.scale);
    return launch_check("batch_unfold_kernel");
  }
  const size_t G=(a.q+b.D-1)/b.D;
  if(G > 1) {
    const size_t nb=b.count < 4096 ? b.count : 4096;
    hipLaunchKernelGGL(batch_reduce_kernel<T>,dim3((unsigned) nb),dim3(256),0,st,
                       b.items,(int64_t) b.count,(const cplx<T> *) b.part,
                       (int64_t) b.ptotal,(int) G,(cplx<T> *) b.out,(T) b.scale);
    rc=launch_check("batch_reduce_kernel");
  }
  return rc;
}

template<typename T>
static int batch_rows_t(const fftwpp_amd_dim *d, int mult, BatchLaunch b,
                        hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  if(batch_tps(d->m,sizeof(cplx<T>))) return batch_reg_t<T>(a,b,mult,st);
  b.lds=batch_lds(d->m,sizeof(cplx<T>),b.TR,b.lacc,&b.ld);
  if(b.lds > lds_max())
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"batch tile does not fit in LDS");
  if(mult == FFTWPP_AMD_MULT_CORRELATION) return batch_rows_launch<T,3>(a,b,st);
  if(mult == FFTWPP_AMD_MULT_NONE) return batch_rows_launch<T,0>(a,b,st);
  return batch_rows_launch<T,1>(a,b,st);
}

// Whether the register-FFT row kernel takes rows of dimension d: power-of-two
// m in [16, 4096] covering the row in one block.
static bool rows_reg(const fftwpp_amd_dim *d)
{
  const size_t m=d->m;
  if(m < 16 || m > 4096 || (m & (m-1))) return false;
  if(m*d->q >= ((size_t) 1 << 30)) return false;   // 32-bit twiddle exponents
  if(d->kind == FFTWPP_AMD_KIND_HERMITIAN) return (d->L+1)/2 <= m;
  return d->L <= m;
}

template<typename T>
static int conv_rows_t(const fftwpp_amd_dim *d, int mult, const void *f,
                       const void *g, size_t rows, size_t rs, void *out,
                       size_t os, double scale, hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  // the register row kernels (rowsfused.h), the m = 256 ones with the
  // multiplier fixed at compile time
  {
    const bool hm=d->kind == FFTWPP_AMD_KIND_HERMITIAN;
    int rc;
    if(rows256_take(a,hm,rs,os) && mult == FFTWPP_AMD_MULT_CORRELATION)
      rc=rows_fused<T>(a,hm,f,g,rows,rs,out,os,scale,LibMult<3>(),st);
    else if(rows256_take(a,hm,rs,os) && mult == FFTWPP_AMD_MULT_NONE)
      rc=rows_fused<T>(a,hm,f,g,rows,rs,out,os,scale,LibMult<0>(),st);
    else if(rows256_take(a,hm,rs,os))
      rc=rows_fused<T>(a,hm,f,g,rows,rs,out,os,scale,LibMult<1>(),st);
    else
      rc=rows_fused<T>(a,hm,f,g,rows,rs,out,os,scale,RunMult{mult},st);
    if(rc < 0) return hipcheck((hipError_t) -rc,"fused row kernel");
    if(rc == 0) return FFTWPP_AMD_OK;
  }
  size_t elt=sizeof(cplx<T>);
  bool herm=d->kind == FFTWPP_AMD_KIND_HERMITIAN;
  size_t lout=herm ? (d->L+1)/2 : d->L;
  // pad the leading dimension against bank conflicts when it still fits
  size_t ld=d->m+1;
  if((4*ld+lout)*elt > lds_max()) ld=d->m;
  size_t per=(4*ld+lout)*elt;
  if(per > lds_max()) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"m too large for the fused row kernel");
  size_t TR=1;
  while(TR < 16 && TR < rows && 2*TR*per <= lds_budget()) TR *= 2;
  size_t lds=TR*per;
  dim3 grid((unsigned) ((rows+TR-1)/TR));
  int rc=FFTWPP_AMD_OK;
  if(herm) rc=allow_lds(hconv_rows_kernel<T>,lds);
  else if(mult == FFTWPP_AMD_MULT_CORRELATION) rc=allow_lds(conv_rows_kernel<T,3>,lds);
  else if(mult == FFTWPP_AMD_MULT_NONE) rc=allow_lds(conv_rows_kernel<T,0>,lds);
  else rc=allow_lds(conv_rows_kernel<T,1>,lds);
  if(rc) return rc;
  if(herm)
    hipLaunchKernelGGL(hconv_rows_kernel<T>,grid,dim3(256),lds,st,a,
                       (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,
                       (int64_t) rs,(cplx<T> *) out,(int64_t) os,(T) scale,(int) TR,(int) ld);
  else if(mult == FFTWPP_AMD_MULT_CORRELATION)
    hipLaunchKernelGGL((conv_rows_kernel<T,3>),grid,dim3(256),lds,st,a,
                       (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,
                       (int64_t) rs,(cplx<T> *) out,(int64_t) os,(T) scale,(int) TR,(int) ld);
  else if(mult == FFTWPP_AMD_MULT_NONE)
    hipLaunchKernelGGL((conv_rows_kernel<T,0>),grid,dim3(256),lds,st,a,
                       (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,
                       (int64_t) rs,(cplx<T> *) out,(int64_t) os,(T) scale,(int) TR,(int) ld);
  else
    hipLaunchKernelGGL((conv_rows_kernel<T,1>),grid,dim3(256),lds,st,a,
                       (const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,
                       (int64_t) rs,(cplx<T> *) out,(int64_t) os,(T) scale,(int) TR,(int) ld);
  return launch_check("conv_rows_kernel");
}

template<typename T, int M>
static int hrows_long_launch(const fftwpp_amd_dim *d, const void *f, const void *g,
                             size_t rows, size_t rs, void *out, size_t os,
                             void *part, double scale, hipStream_t st)
{
  const DimArgs &a=sizeof(T) == 8 ? d->a64 : d->a32;
  constexpr int TPS=M/RowV::V;
  const size_t nt=TPS > 64 ? TPS : 256;
  const size_t lds=M*sizeof(cplx<T>)+(nt/TPS)*2*RPlan<M,RowV::V>::PAD*sizeof(T);
  int rc=allow_lds(hrows_long_kernel<T,M>,lds);
  if(rc) return rc;
  const size_t passes=(d->q+1)/2;
  const size_t units=rows*passes;
  const size_t nb=std::min((units+nt/TPS-1)/(nt/TPS),(size_t) 65536);
  hipLaunchKernelGGL((hrows_long_kernel<T,M>),dim3((unsigned) nb),dim3((unsigned) nt),lds,st,
                     a,(const cplx<T> *) f,(const cplx<T> *) g,(int64_t) rows,(int64_t) rs,
                     (cplx<T> *) part);
  if((rc=launch_check("hrows_long_kernel"))) return rc;
  hipLaunchKernelGGL(hrows_reduce_kernel<T>,dim3((unsigned) std::min(rows,(size_t) 65536)),
                     dim3(256),0,st,(const cplx<T> *) part,(int64_t) rows,(int) a.H,
                     (int) passes,(cplx<T> *) out,(int64_t) os,(T) scale);
  return launch_check("hrows_reduce_kernel");
}

template<typename T>
static int hrows_long_t(const fftwpp_amd_dim *d, const void *f, const void *g,
                        size_t rows, size_t rs, void *out, size_t os, void *part,
                        double scale, hipStream_t st)
{
  switch(d->m) {
  case 16: return hrows_long_launch<T,16>(d,f,g,rows,rs,out,os,part,scale,st);
  case 32: return hrows_long_launch<T,32>(d,f,g,rows,rs,out,os,part,scale,st);
  case 64: return hrows_long_launch<T,64>(d,f,g,rows,rs,out,os,part,scale,st);
  case 128: return hrows_long_launch<T,128>(d,f,g,rows,rs,out,os,part,scale,st);
  case 256: return hrows_long_launch<T,256>(d,f,g,rows,rs,out,os,part,scale,st);
  case 512: return hrows_long_launch<T,512>(d,f,g,rows,rs,out,os,part,scale,st);
  case 1024: return hrows_long_launch<T,1024>(d,f,g,rows,rs,out,os,part,scale,st);
  case 2048: return hrows_long_launch<T,2048>(d,f,g,rows,rs,out,os,part,scale,st);
  case 4096: return hrows_long_launch<T,4096>(d,f,g,rows,rs,out,os,part,scale,st);
  default: return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"long Hermitian rows need a power-of-two m in [16, 4096]");
  }
}

// ------------------------------------------------------------------------
// The y-z plane stage of 3D convolutions (plane256.h): one cooperative
// launch of 8 teams of W = CUs/8 blocks.
// ------------------------------------------------------------------------
static bool plane_dim(const fftwpp_amd_dim *d)
{
  return d && d->m == 256 && d->q == 2 && d->kind == FFTWPP_AMD_KIND_COMPLEX && d->L == 256;
}

static size_t plane_ctl_bytes() { return 8192; }
static_assert(sizeof(PlaneCtl) <= 8192,"plane control block");

template<typename T>
static size_t plane_work_t() { return plane_ctl_bytes()+8*PlaneSub<T>::n*PLANE_E*sizeof(cplx<T>); }

template<typename T, int MULT>
static int plane_launch(const cplx<T> *Xf, const cplx<T> *Xg, cplx<T> *Xh,
                        int64_t planes, const cplx<T> *zN, void *work, hipStream_t st)
{
  int dev=0, cus=0, per=0;
  int rc=hipcheck(hipGetDevice(&dev),"hipGetDevice");
  if(!rc) rc=hipcheck(hipDeviceGetAttribute(&cus,hipDeviceAttributeMultiprocessorCount,dev),
                      "device attribute");
  if(rc) return rc;
  const size_t lds=plane_lds<T>();
  if((rc=allow_lds(plane256_kernel<T,MULT>,lds))) return rc;
  if((rc=hipcheck(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per,(const void *) plane256_kernel<T,MULT>,PLANE_THREADS,lds),
                  "occupancy"))) return rc;
  if(per < PlaneSub<T>::n || cus != 8*PLANE_W)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"plane kernel: needs 8 XCDs of 32 CUs, two resident blocks per CU");
  PlaneCtl *ctl=(PlaneCtl *) work;
  cplx<T> *E=(cplx<T> *) ((char *) work+plane_ctl_bytes());
  if((rc=hipcheck(hipMemsetAsync(ctl,0,sizeof(PlaneCtl),st),"clear plane control"))) return rc;
  void *args[]={(void *) &Xf,(void *) &Xg,(void *) &Xh,(void *) &planes,(void *) &zN,
                (void *) &E,(void *) &ctl};
  return hipcheck(hipLaunchCooperativeKernel((const void *) plane256_kernel<T,MULT>,
                                             dim3((unsigned) (PlaneSub<T>::n*cus)),dim3(PLANE_THREADS),args,
                                             (unsigned) lds,st),
                  "plane256_kernel cooperative launch");
}

template<typename T>
static int plane_conv_t(const fftwpp_amd_dim *dz, int mult, const void *Xf,
                        const void *Xg, void *Xh, size_t planes, void *work,
                        hipStream_t st)
{
  const cplx<T> *zN=(const cplx<T> *) (sizeof(T) == 8 ? dz->a64.zN : dz->a32.zN);
  const cplx<T> *f=(const cplx<T> *) Xf, *g=(const cplx<T> *) Xg;
  cplx<T> *h=(cplx<T> *) Xh;
  if(mult == FFTWPP_AMD_MULT_CORRELATION)
    return plane_launch<T,3>(f,g,h,(int64_t) planes,zN,work,st);
  if(mult == FFTWPP_AMD_MULT_NONE)
    return plane_launch<T,0>(f,g,h,(int64_t) planes,zN,work,st);
  return plane_launch<T,1>(f,g,h,(int64_t) planes,zN,work,st);
}

extern "C" {

int fftwpp_amd_internal_block_map(int scatter, const void *T, const void *tab,
                                   size_t ne, size_t C, void *arr, int real,
                                   void *stream)
{
  if(!T || !tab || !arr) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(ne == 0 || C == 0) return FFTWPP_AMD_OK;
  const unsigned nb=(unsigned) ((ne*C+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(scatter)
    hipLaunchKernelGGL(block_scatter_kernel,dim3(nb),dim3(256),0,st,(cplx<double> *) T,
                       (const uint64_t *) tab,(int64_t) ne,(int64_t) C,(const void *) arr,real);
  else
    hipLaunchKernelGGL(block_gather_kernel,dim3(nb),dim3(256),0,st,(const cplx<double> *) T,
                       (const uint64_t *) tab,(int64_t) ne,(int64_t) C,arr,real);
  return launch_check("block map kernel");
}

int fftwpp_amd_internal_herm_expand(int op, const void *a, void *b, size_t H, size_t C,
                                    int acc, void *stream)
{
  if(!a || !b || H == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad argument");
  const size_t n=(op == 0 ? 2*H-1 : H)*C;
  hipLaunchKernelGGL(herm_expand_kernel,dim3((unsigned) ((n+255)/256)),dim3(256),0,
                     (hipStream_t) stream,op,(const cplx<double> *) a,(cplx<double> *) b,
                     (int64_t) H,(int64_t) C,acc);
  return launch_check("herm_expand_kernel");
}

int fftwpp_amd_internal_row_items(void *items, size_t rows, size_t rs, size_t L,
                                  void *stream)
{
  if(!items) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(rows == 0) return FFTWPP_AMD_OK;
  hipLaunchKernelGGL(row_items_kernel,dim3((unsigned) ((rows+255)/256)),dim3(256),0,
                     (hipStream_t) stream,(BatchItem *) items,(int64_t) rows,(int64_t) rs,
                     (int64_t) L);
  return launch_check("row_items_kernel");
}

const char *fftwpp_amd_last_error(void) { return last_error; }

// Used by the host C++ layer to record its own failures in the same slot.
void fftwpp_amd_internal_set_error(const char *msg) { fail(0,msg); }

// Whether fftwpp_amd_conv_rows can take one row of dimension d in dtype: the
// m = 256 register kernels always can; the generic kernel needs four
// sequences of m words plus the L (or H) outputs in LDS.
int fftwpp_amd_internal_rows_fit(const fftwpp_amd_dim *d, int dtype)
{
  if(!d) return 0;
  if(fast256(d) || rows_reg(d)) return 1;
  const size_t elt=dtype == FFTWPP_AMD_F32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const size_t lout=d->kind == FFTWPP_AMD_KIND_HERMITIAN ? (d->L+1)/2 : d->L;
  return (4*d->m+lout)*elt <= lds_max();
}

// Which engine the dispatchers above pick for dimension d (strides below
// 2^30): stage 0 = column transforms (forward_many / backward_many), 1 =
// fused rows (conv_rows).  Returns 0 for the generic LDS kernels, 1 for the
// m = 256 register kernels, 2 for the register-FFT kernels of rowsreg.h.
// Column transforms take the register kernels when a tile has at least 8
// columns (or m >= 2048); fewer columns run on the LDS kernels.
int fftwpp_amd_internal_engine(const fftwpp_amd_dim *d, int stage)
{
  if(!d) return -1;
  if(fast256(d)) return 1;
  if(stage == 0) return cols_reg(d) ? 2 : 0;
  return rows_reg(d) ? 2 : 0;
}

size_t fftwpp_amd_plane_work(int dtype)
{
  return dtype == FFTWPP_AMD_F32 ? plane_work_t<float>() : plane_work_t<double>();
}

// Whether the plane kernel runs on the current device for dimensions dy, dz:
// the geometry rule, and once per device a launch with no planes that forms
// the teams and checks that every XCD holds blocks of the grid (this
// synchronises the device the first time).
int fftwpp_amd_plane_supported(const fftwpp_amd_dim *dy, const fftwpp_amd_dim *dz)
{
  if(!plane_dim(dy) || !plane_dim(dz) || dy->N != dz->N) return 0;
  const char *env=getenv("FFTWPP_AMD_PLANE");
  if(env && env[0] == '0') return 0;
  static std::mutex mu;
  static std::vector<std::pair<int,int>> known;   // (device, usable)
  int dev=0;
  if(hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  for(const auto &e : known) if(e.first == dev) return e.second;
  int ok=0;
  void *work=nullptr;
  // both precisions must keep their sub-teams resident (checked here so that
  // a plan never meets the restriction at execution time)
  int per64=0, per32=0;
  if(allow_lds(plane256_kernel<double,1>,plane_lds<double>()) ||
     allow_lds(plane256_kernel<float,1>,plane_lds<float>()) ||
     hipOccupancyMaxActiveBlocksPerMultiprocessor(&per64,(const void *) plane256_kernel<double,1>,
                                                  PLANE_THREADS,plane_lds<double>()) != hipSuccess ||
     hipOccupancyMaxActiveBlocksPerMultiprocessor(&per32,(const void *) plane256_kernel<float,1>,
                                                  PLANE_THREADS,plane_lds<float>()) != hipSuccess ||
     per64 < PlaneSub<double>::n || per32 < PlaneSub<float>::n) {
    (void) hipGetLastError();
    known.push_back({dev,0});
    return 0;
  }
  if(hipMalloc(&work,plane_ctl_bytes()) == hipSuccess) {
    int rc=plane_conv_t<double>(dz,FFTWPP_AMD_MULT_BINARY,nullptr,nullptr,nullptr,0,work,0);
    PlaneCtl c;
    memset(&c,0,sizeof(c));
    if(!rc) rc=hipcheck(hipDeviceSynchronize(),"plane probe sync");
    if(!rc) rc=hipcheck(hipMemcpy(&c,work,sizeof(c),hipMemcpyDeviceToHost),"plane probe copy");
    if(!rc) {
      ok=c.abort_ == 0;
      for(int i=0; i < 8; ++i) if(c.tk[i] != (unsigned) (PlaneSub<double>::n*PLANE_W)) ok=0;
      if(!ok) {
        char buf[256];
        snprintf(buf,sizeof(buf),"plane probe: abort %u, blocks per XCD %u %u %u %u %u %u %u %u",
                 c.abort_,c.tk[0],c.tk[1],c.tk[2],c.tk[3],c.tk[4],c.tk[5],c.tk[6],c.tk[7]);
        fail(FFTWPP_AMD_ERR_UNSUPPORTED,buf);
      }
    }
    (void) hipFree(work);
  }
  (void) hipGetLastError();
  known.push_back({dev,ok});
  return ok;
}

int fftwpp_amd_plane_conv(const fftwpp_amd_dim *dy, const fftwpp_amd_dim *dz,
                          int dtype, int mult, const void *Xf, const void *Xg,
                          void *Xh, size_t planes, size_t Ly, size_t lz,
                          void *work, void *stream)
{
  if(!Xf || !Xg || !Xh || !work) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(!plane_dim(dy) || !plane_dim(dz) || dy->N != dz->N)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"plane kernel: y and z need m = 256, q = 2, L = 256");
  if(mult != FFTWPP_AMD_MULT_BINARY && mult != FFTWPP_AMD_MULT_CORRELATION &&
     mult != FFTWPP_AMD_MULT_NONE)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"plane kernel: multNone, multBinary or multcorrelation");
  if(Ly != dy->L || lz != dz->L || planes == 0)
    return fail(FFTWPP_AMD_ERR_ARG,"plane kernel: shape does not match the dimensions");
  if(planes*Ly*lz > ((size_t) 1 << 40)) return fail(FFTWPP_AMD_ERR_ARG,"plane kernel: too many planes");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    plane_conv_t<float>(dz,mult,Xf,Xg,Xh,planes,work,st) :
    plane_conv_t<double>(dz,mult,Xf,Xg,Xh,planes,work,st);
}

// Abort flag of the last plane launch on `work` (0: completed; 1: a team
// wait timed out; 2: an XCD held no block).  Synchronises the stream.
int fftwpp_amd_plane_status(void *work, void *stream)
{
  if(!work) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  PlaneCtl c;
  int rc=hipcheck(hipStreamSynchronize((hipStream_t) stream),"plane status sync");
  if(!rc) rc=hipcheck(hipMemcpy(&c,work,sizeof(c),hipMemcpyDeviceToHost),"plane status copy");
  return rc ? rc : (int) c.abort_;
}

int fftwpp_amd_dim_create(fftwpp_amd_dim **out, size_t L, size_t m,
                          size_t q, int kind)
{
  if(!out) return fail(FFTWPP_AMD_ERR_ARG,"null output handle");
  *out=NULL;
  if(L == 0 || m == 0 || q == 0)
    return fail(FFTWPP_AMD_ERR_ARG,"L, m and q must be positive");
  if(kind < FFTWPP_AMD_KIND_COMPLEX || kind > FFTWPP_AMD_KIND_HERMITIAN)
    return fail(FFTWPP_AMD_ERR_ARG,"unknown transform kind");
  size_t N=m*q;
  if(m > (1u << 20) || N > ((size_t) 1 << 31))
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"padded length too large");
  fftwpp_amd_dim *d=new fftwpp_amd_dim;
  memset(d,0,sizeof(*d));
  d->L=L; d->m=m; d->q=q; d->N=N; d->kind=kind;
  DimArgs a;
  memset(&a,0,sizeof(a));
  a.L=(int64_t) L; a.m=(int64_t) m; a.q=(int64_t) q; a.N=(int64_t) N;
  a.k0=kind == FFTWPP_AMD_KIND_CENTERED ? -(int64_t) (L/2) : 0;
  a.H=kind == FFTWPP_AMD_KIND_HERMITIAN ? (int64_t) ((L+1)/2) : (int64_t) L;
  // radix plan: 8s, then 4, 2, 3, 5, then remaining primes
  size_t r=m;
  int ns=0;
  while(r % 8 == 0 && ns < MAXSTAGES) { a.radix[ns++]=8; r /= 8; }
  while(r % 4 == 0 && ns < MAXSTAGES) { a.radix[ns++]=4; r /= 4; }
  while(r % 2 == 0 && ns < MAXSTAGES) { a.radix[ns++]=2; r /= 2; }
  for(size_t p=3; r > 1 && ns < MAXSTAGES; p += 2)
    while(r % p == 0 && ns < MAXSTAGES) { a.radix[ns++]=(int) p; r /= p; }
  if(r != 1) { delete d; return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"too many FFT stages"); }
  a.nstages=ns;
  int rc=hipcheck(hipMalloc(&d->z64,sizeof(cplx<double>)*N),"hipMalloc twiddles");
  if(rc) { delete d; return rc; }
  rc=hipcheck(hipMalloc(&d->z32,sizeof(cplx<float>)*N),"hipMalloc twiddles");
  if(rc) { (void) hipFree(d->z64); delete d; return rc; }
  unsigned nb=(unsigned) ((N+255)/256);
  twiddle_kernel<<<nb,256>>>((int64_t) N,(cplx<double> *) d->z64,
                             (cplx<float> *) d->z32);
  rc=launch_check("twiddle_kernel");
  if(!rc) rc=hipcheck(hipDeviceSynchronize(),"twiddle sync");
  if(rc) { (void) hipFree(d->z64); (void) hipFree(d->z32); delete d; return rc; }
  d->a64=a; d->a64.zN=d->z64;
  d->a32=a; d->a32.zN=d->z32;
  *out=d;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_dim_destroy(fftwpp_amd_dim *d)
{
  if(!d) return FFTWPP_AMD_OK;
  (void) hipFree(d->z64);
  (void) hipFree(d->z32);
  delete d;
  return FFTWPP_AMD_OK;
}

size_t fftwpp_amd_max_m(int dtype)
{
  // one sequence in the fused row kernel needs 4 buffers of m plus L
  return dtype == FFTWPP_AMD_F32 ? 4096 : 2048;
}


int fftwpp_amd_forward_many(const fftwpp_amd_dim *d, int dtype,
                            const void *f, size_t S, size_t C,
                            size_t nbatch, size_t bstride_in, void *F,
                            size_t SF, size_t bstride_out, size_t r0,
                            size_t nres, void *stream)
{
  if(!d || !f || !F) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(d->kind == FFTWPP_AMD_KIND_HERMITIAN)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"Hermitian dimensions are innermost only");
  if(r0+nres > d->q || C == 0 || nbatch == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad residue range or shape");
  if(S < C || SF < C) return fail(FFTWPP_AMD_ERR_ARG,"stride smaller than count");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    fwd_many_t<float>(d,f,S,C,nbatch,bstride_in,F,SF,bstride_out,r0,nres,st) :
    fwd_many_t<double>(d,f,S,C,nbatch,bstride_in,F,SF,bstride_out,r0,nres,st);
}


int fftwpp_amd_backward_many(const fftwpp_amd_dim *d, int dtype,
                             const void *F, size_t SF, size_t bstride_in,
                             void *f, size_t S, size_t C, size_t nbatch,
                             size_t bstride_out, size_t r0, size_t nres,
                             int accumulate, double scale, void *stream)
{
  if(!d || !f || !F) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(d->kind == FFTWPP_AMD_KIND_HERMITIAN)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"Hermitian dimensions are innermost only");
  if(r0+nres > d->q || nres == 0 || C == 0 || nbatch == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad residue range or shape");
  if(S < C || SF < C) return fail(FFTWPP_AMD_ERR_ARG,"stride smaller than count");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    bwd_many_t<float>(d,F,SF,bstride_in,f,S,C,nbatch,bstride_out,r0,nres,accumulate,scale,st) :
    bwd_many_t<double>(d,F,SF,bstride_in,f,S,C,nbatch,bstride_out,r0,nres,accumulate,scale,st);
}


const void *fftwpp_amd_dim_kernel_args(const fftwpp_amd_dim *d, int dtype)
{
  if(!d) return NULL;
  return dtype == FFTWPP_AMD_F32 ? (const void *) &d->a32 : (const void *) &d->a64;
}

size_t fftwpp_amd_real_planes(const fftwpp_amd_dim *d)
{
  if(!d || !cols_real(d)) return 0;
  return (size_t) RealKept::planes((int64_t) d->m,(int64_t) d->q);
}

int fftwpp_amd_forward_real_many(const fftwpp_amd_dim *d, int dtype,
                                 const void *f, size_t S, size_t C,
                                 size_t nbatch, size_t bstride_in, void *F,
                                 size_t SF, size_t bstride_out, void *stream)
{
  if(!d || !f || !F) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(!cols_real(d))
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"real residue transforms take power-of-two m in [16, 4096] and L <= 2m");
  if(C == 0 || nbatch == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad shape");
  if(S < C || SF < C) return fail(FFTWPP_AMD_ERR_ARG,"stride smaller than count");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    colreal_t<float>(d,true,f,S,bstride_in,F,SF,bstride_out,C,nbatch,1.0,st) :
    colreal_t<double>(d,true,f,S,bstride_in,F,SF,bstride_out,C,nbatch,1.0,st);
}

int fftwpp_amd_backward_real_many(const fftwpp_amd_dim *d, int dtype,
                                  const void *F, size_t SF, size_t bstride_in,
                                  void *f, size_t S, size_t C, size_t nbatch,
                                  size_t bstride_out, double scale, void *stream)
{
  if(!d || !f || !F) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(!cols_real(d))
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"real residue transforms take power-of-two m in [16, 4096] and L <= 2m");
  if(C == 0 || nbatch == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad shape");
  if(S < C || SF < C) return fail(FFTWPP_AMD_ERR_ARG,"stride smaller than count");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    colreal_t<float>(d,false,F,SF,bstride_in,f,S,bstride_out,C,nbatch,scale,st) :
    colreal_t<double>(d,false,F,SF,bstride_in,f,S,bstride_out,C,nbatch,scale,st);
}

int fftwpp_amd_conv_rows(const fftwpp_amd_dim *d, int dtype, int mult,
                         const void *f, const void *g, size_t rows,
                         size_t rowstride, void *out, size_t outstride,
                         double scale, void *stream)
{
  if(!d || !f || !g || !out) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(rows == 0) return FFTWPP_AMD_OK;
  bool herm=d->kind == FFTWPP_AMD_KIND_HERMITIAN;
  if(herm && mult != FFTWPP_AMD_MULT_REAL_BINARY)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"Hermitian rows take realMultBinary");
  if(!herm && mult != FFTWPP_AMD_MULT_BINARY && mult != FFTWPP_AMD_MULT_CORRELATION &&
     mult != FFTWPP_AMD_MULT_NONE)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"complex rows take multNone, multBinary or multcorrelation");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    conv_rows_t<float>(d,mult,f,g,rows,rowstride,out,outstride,scale,st) :
    conv_rows_t<double>(d,mult,f,g,rows,rowstride,out,outstride,scale,st);
}

size_t fftwpp_amd_hconv_rows_long_part(const fftwpp_amd_dim *d, size_t rows)
{
  if(!d || d->kind != FFTWPP_AMD_KIND_HERMITIAN) return 0;
  const size_t m=d->m;
  if(m < 16 || m > 4096 || (m & (m-1)) || m*d->q >= ((size_t) 1 << 30)) return 0;
  return (d->q+1)/2*rows*((d->L+1)/2);
}

int fftwpp_amd_hconv_rows_long(const fftwpp_amd_dim *d, int dtype, const void *f,
                               const void *g, size_t rows, size_t rowstride,
                               void *out, size_t outstride, void *part,
                               double scale, void *stream)
{
  if(!d || !f || !g || !out || !part) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(rows == 0) return FFTWPP_AMD_OK;
  if(!fftwpp_amd_hconv_rows_long_part(d,1))
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"long Hermitian rows need a Hermitian dimension with power-of-two m in [16, 4096]");
  hipStream_t st=(hipStream_t) stream;
  return dtype == FFTWPP_AMD_F32 ?
    hrows_long_t<float>(d,f,g,rows,rowstride,out,outstride,part,scale,st) :
    hrows_long_t<double>(d,f,g,rows,rowstride,out,outstride,part,scale,st);
}

size_t fftwpp_amd_batch_rows_lds(const fftwpp_amd_dim *d, int dtype,
                                 size_t TR, size_t lacc)
{
  if(!d || TR == 0) return 0;
  const size_t elt=dtype == FFTWPP_AMD_F32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  if(batch_tps(d->m,elt)) {
    if(TR > 1 || lacc) return 0;
    size_t nt, lds;
    batch_reg_shape(d->m,d->q,elt,64,&nt,&lds);
    return lds > lds_max() ? 0 : lds;
  }
  const size_t b=batch_lds(d->m,elt,TR,lacc,NULL);
  return b > lds_max() ? 0 : b;
}

size_t fftwpp_amd_batch_rows_tps(const fftwpp_amd_dim *d, int dtype)
{
  return d ? batch_tps(d->m,dtype == FFTWPP_AMD_F32 ? sizeof(cplx<float>) : sizeof(cplx<double>)) : 0;
}

int fftwpp_amd_batch_rows(const fftwpp_amd_dim *d, int dtype, int mult,
                          const void *items, size_t count, const void *f,
                          const void *g, void *out, size_t TR, size_t threads,
                          size_t nblocks, size_t lacc, size_t D, void *part,
                          size_t ptotal, double scale, void *stream)
{
  if(!d || !items || !f || !g || !out) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(D == 0 && !batch_tps(d->m,dtype == FFTWPP_AMD_F32 ? sizeof(cplx<float>) : sizeof(cplx<double>)))
    return fail(FFTWPP_AMD_ERR_ARG,"the sequence mode (D = 0) needs a register-FFT sub-FFT size");
  if(D < d->q && !part) return fail(FFTWPP_AMD_ERR_ARG,"residue groups need a partial-sum buffer");
  if(count == 0) return FFTWPP_AMD_OK;
  if(d->kind != FFTWPP_AMD_KIND_COMPLEX)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"batched rows are non-centred complex");
  if(mult != FFTWPP_AMD_MULT_BINARY && mult != FFTWPP_AMD_MULT_CORRELATION &&
     mult != FFTWPP_AMD_MULT_NONE)
    return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"batched rows take multNone, multBinary or multcorrelation");
  if(TR == 0 || nblocks == 0 || nblocks > (1u << 30) || threads < 64 ||
     threads > 1024 || threads % 64 != 0)
    return fail(FFTWPP_AMD_ERR_ARG,"bad tile, block or thread count");
  hipStream_t st=(hipStream_t) stream;
  BatchLaunch b;
  b.items=(const BatchItem *) items; b.count=count; b.f=f; b.g=g; b.out=out;
  b.part=part; b.TR=TR; b.threads=threads; b.nblocks=nblocks; b.lacc=lacc;
  b.D=D < d->q ? D : d->q; b.ptotal=ptotal; b.ld=0; b.lds=0; b.scale=scale;
  if(D == 0 && d->q == 1) b.D=1;   // one residue: the sequence is the output
  return dtype == FFTWPP_AMD_F32 ? batch_rows_t<float>(d,mult,b,st) :
    batch_rows_t<double>(d,mult,b,st);
}

// Scratch of fftwpp_amd_batch_rows_noalloc: a file-scope device array, one
// copy per GPU.  Host code never passes the array's name to a kernel (on the
// host it only stands for the symbol); the device address comes from
// hipGetSymbolAddress, looked up once per device on the entry point's first
// call there.
static __device__ __align__(256) unsigned char scratch[FFTWPP_AMD_NOALLOC_SCRATCH_BYTES];

static int noalloc_scratch(void **ptr)
{
  static std::mutex mu;
  static std::vector<std::pair<int,void *>> known;   // (device, address)
  int dev=0;
  int rc=hipcheck(hipGetDevice(&dev),"hipGetDevice");
  if(rc) return rc;
  std::lock_guard<std::mutex> lock(mu);
  for(const auto &e : known)
    if(e.first == dev) { *ptr=e.second; return FFTWPP_AMD_OK; }
  void *p=nullptr;
  rc=hipcheck(hipGetSymbolAddress(&p,HIP_SYMBOL(scratch)),"hipGetSymbolAddress(scratch)");
  if(rc) return rc;
  known.push_back({dev,p});
  *ptr=p;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_batch_rows_noalloc(const fftwpp_amd_dim *d, int dtype, int mult,
                                  const void *items, size_t count, const void *f,
                                  const void *g, void *out, size_t TR,
                                  size_t threads, size_t nblocks, size_t lacc,
                                  size_t D, size_t ptotal, double scale,
                                  void *stream)
{
  if(!d) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  const size_t elt=dtype == FFTWPP_AMD_F32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  size_t words=0;
  if(D == 0) words=count*d->q*d->m;
  else if(D < d->q) words=((d->q+D-1)/D)*ptotal;
  if(words > FFTWPP_AMD_NOALLOC_SCRATCH_BYTES/elt)
    return fail(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_batch_rows_noalloc: partial results exceed the "
                "16 MiB scratch array (use fftwpp_amd_batch_rows with a caller buffer)");
  void *part=nullptr;
  if(words) {
    int rc=noalloc_scratch(&part);
    if(rc) return rc;
  }
  return fftwpp_amd_batch_rows(d,dtype,mult,items,count,f,g,out,TR,threads,nblocks,lacc,D,
                               part,ptotal,scale,stream);
}

size_t fftwpp_amd_real_balance_bytes(void)
{
  return (2*REAL_NORM_BLOCKS+1)*sizeof(double);   // partials, then e and the zero flag
}

int fftwpp_amd_real_balance(int dtype, const void *f0, const void *f1, size_t n,
                            void *bal, void *stream)
{
  if(!f0 || !f1 || !bal) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(real_norms_kernel<float>,dim3(REAL_NORM_BLOCKS),dim3(256),0,st,
                       (const float *) f0,(const float *) f1,(int64_t) n,(double *) bal);
  else
    hipLaunchKernelGGL(real_norms_kernel<double>,dim3(REAL_NORM_BLOCKS),dim3(256),0,st,
                       (const double *) f0,(const double *) f1,(int64_t) n,(double *) bal);
  int rc=launch_check("real_norms_kernel");
  if(rc) return rc;
  hipLaunchKernelGGL(real_balance_kernel,dim3(1),dim3(256),0,st,(double *) bal);
  return launch_check("real_balance_kernel");
}

int fftwpp_amd_real_pack(int dtype, const void *f0, const void *f1, void *z,
                         size_t n, const void *bal, void *stream)
{
  if(n == 0) return FFTWPP_AMD_OK;
  if(!f0 || !f1 || !z) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  const unsigned nb=(unsigned) ((n+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(real_pack_kernel<float>,dim3(nb),dim3(256),0,st,
                       (const float *) f0,(const float *) f1,(cplx<float> *) z,(int64_t) n,
                       (const double *) bal);
  else
    hipLaunchKernelGGL(real_pack_kernel<double>,dim3(nb),dim3(256),0,st,
                       (const double *) f0,(const double *) f1,(cplx<double> *) z,(int64_t) n,
                       (const double *) bal);
  return launch_check("real_pack_kernel");
}

int fftwpp_amd_real_unpack(int dtype, const void *z, void *out, size_t n,
                           double scale, const void *bal, void *stream)
{
  if(n == 0) return FFTWPP_AMD_OK;
  if(!z || !out) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  const unsigned nb=(unsigned) ((n+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(real_unpack_kernel<float>,dim3(nb),dim3(256),0,st,
                       (const cplx<float> *) z,(float *) out,(int64_t) n,(float) scale,
                       (const double *) bal);
  else
    hipLaunchKernelGGL(real_unpack_kernel<double>,dim3(nb),dim3(256),0,st,
                       (const cplx<double> *) z,(double *) out,(int64_t) n,scale,
                       (const double *) bal);
  return launch_check("real_unpack_kernel");
}

// The element maps of fft_map_kernel (standalone FFTs; internal to the
// library, not declared in the public header).
int fftwpp_amd_internal_fft_map(int dtype, int op, const void *x, size_t xs,
                                size_t xd, void *y, size_t ys, size_t yd,
                                size_t n, size_t nb, double scale, void *stream)
{
  if(n == 0 || nb == 0) return FFTWPP_AMD_OK;
  if(!x || !y || op < 0 || op > 4) return fail(FFTWPP_AMD_ERR_ARG,"bad element map");
  const size_t total=n*nb;
  if((total+255)/256 > 0xffffffffu) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"element map too large");
  const unsigned blocks=(unsigned) ((total+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(fft_map_kernel<float>,dim3(blocks),dim3(256),0,st,op,x,(int64_t) xs,
                       (int64_t) xd,y,(int64_t) ys,(int64_t) yd,(int64_t) n,(int64_t) nb,
                       (float) scale);
  else
    hipLaunchKernelGGL(fft_map_kernel<double>,dim3(blocks),dim3(256),0,st,op,x,(int64_t) xs,
                       (int64_t) xd,y,(int64_t) ys,(int64_t) yd,(int64_t) n,(int64_t) nb,
                       scale);
  return launch_check("fft_map_kernel");
}

// The twiddle step of fft_twiddle_kernel over nb contiguous transforms
// (internal to the library).
int fftwpp_amd_internal_fft_twiddle(int dtype, void *y, size_t n, size_t n2, size_t nb,
                                    int sign, void *stream)
{
  if(n == 0 || nb == 0) return FFTWPP_AMD_OK;
  if(!y || n2 == 0 || n % n2) return fail(FFTWPP_AMD_ERR_ARG,"bad twiddle step");
  const size_t total=n*nb;
  if((total+255)/256 > 0xffffffffu) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"twiddle step too large");
  const unsigned blocks=(unsigned) ((total+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(fft_twiddle_kernel<float>,dim3(blocks),dim3(256),0,st,(cplx<float> *) y,
                       (int64_t) n,(int64_t) n2,(int64_t) nb,sign);
  else
    hipLaunchKernelGGL(fft_twiddle_kernel<double>,dim3(blocks),dim3(256),0,st,(cplx<double> *) y,
                       (int64_t) n,(int64_t) n2,(int64_t) nb,sign);
  return launch_check("fft_twiddle_kernel");
}

// The steps of fft_chirp_kernel (internal to the library).
int fftwpp_amd_internal_fft_chirp(int dtype, int op, const void *x, size_t xs, size_t xd,
                                  void *y, size_t ys, size_t yd, size_t n, size_t len,
                                  size_t nb, int sign, double scale, void *stream)
{
  if(n == 0 || nb == 0) return FFTWPP_AMD_OK;
  if(!y || op < 0 || op > 3 || (op >= 2 && !x) || len < n || n >= ((size_t) 1 << 31))
    return fail(FFTWPP_AMD_ERR_ARG,"bad chirp step");
  const size_t cnt=(op == 0 || op == 3) ? n : len;
  const size_t total=cnt*nb;
  if((total+255)/256 > 0xffffffffu) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"chirp step too large");
  const unsigned blocks=(unsigned) ((total+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(fft_chirp_kernel<float>,dim3(blocks),dim3(256),0,st,op,
                       (const cplx<float> *) x,(int64_t) xs,(int64_t) xd,(cplx<float> *) y,
                       (int64_t) ys,(int64_t) yd,(int64_t) n,(int64_t) len,(int64_t) nb,sign,
                       (float) scale);
  else
    hipLaunchKernelGGL(fft_chirp_kernel<double>,dim3(blocks),dim3(256),0,st,op,
                       (const cplx<double> *) x,(int64_t) xs,(int64_t) xd,(cplx<double> *) y,
                       (int64_t) ys,(int64_t) yd,(int64_t) n,(int64_t) len,(int64_t) nb,sign,
                       scale);
  return launch_check("fft_chirp_kernel");
}

int fftwpp_amd_fft_shift(int op, double *data, size_t nx, size_t ny, size_t nz,
                         size_t rowlen, void *stream)
{
  if(!data || op < 0 || op > 2 || nx == 0 || ny == 0 || (op == 2 && nz == 0))
    return fail(FFTWPP_AMD_ERR_ARG,"fft_shift: bad arguments");
  if(op == 0 && (nx % 2 || (nz && ny % 2)))
    return fail(FFTWPP_AMD_ERR_ARG,"shift of the Fourier origin needs even nx (and even ny in 3D)");
  size_t total=op == 0 ? (nz ? nx*ny : nx)*rowlen : op == 1 ? nx*(ny/2+1) : nx*ny*(nz/2+1);
  if(total == 0) return FFTWPP_AMD_OK;
  if((total+255)/256 > 0xffffffffu) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"fft_shift too large");
  hipLaunchKernelGGL(fft_shift_kernel,dim3((unsigned) ((total+255)/256)),dim3(256),0,
                     (hipStream_t) stream,op,data,(int64_t) nx,(int64_t) ny,(int64_t) nz,
                     (int64_t) rowlen);
  return launch_check("fft_shift_kernel");
}

int fftwpp_amd_transpose(size_t rows, size_t cols, size_t words, const void *in,
                         void *out, void *stream)
{
  if(rows == 0 || cols == 0 || words == 0) return FFTWPP_AMD_OK;
  if(!in || !out || in == out) return fail(FFTWPP_AMD_ERR_ARG,"transpose: distinct device arrays required");
  hipStream_t st=(hipStream_t) stream;
  if((words == 1 || words == 2) && (rows+31)/32 <= 65535 && (cols+31)/32 <= 0x7fffffff) {
    dim3 grid((unsigned) ((cols+31)/32),(unsigned) ((rows+31)/32));
    if(words == 1)
      hipLaunchKernelGGL(transpose_tile_kernel<1>,grid,dim3(32,8),0,st,(const double *) in,
                         (double *) out,(int64_t) rows,(int64_t) cols);
    else
      hipLaunchKernelGGL(transpose_tile_kernel<2>,grid,dim3(32,8),0,st,(const double *) in,
                         (double *) out,(int64_t) rows,(int64_t) cols);
    return launch_check("transpose_tile_kernel");
  }
  const size_t total=rows*cols*words;
  if((total+255)/256 > 0xffffffffu) return fail(FFTWPP_AMD_ERR_UNSUPPORTED,"transpose too large");
  hipLaunchKernelGGL(transpose_words_kernel,dim3((unsigned) ((total+255)/256)),dim3(256),0,st,
                     (const double *) in,(double *) out,(int64_t) rows,(int64_t) cols,
                     (int64_t) words);
  return launch_check("transpose_words_kernel");
}

int fftwpp_amd_multiply(int dtype, int mult, void *F0, const void *F1,
                        size_t n, void *stream)
{
  if(mult == FFTWPP_AMD_MULT_NONE || n == 0) return FFTWPP_AMD_OK;
  if(!F0 || !F1) return fail(FFTWPP_AMD_ERR_ARG,"null argument");
  unsigned nb=(unsigned) ((n+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(multiply_kernel<float>,dim3(nb),dim3(256),0,st,mult,
                       (cplx<float> *) F0,(const cplx<float> *) F1,(int64_t) n);
  else
    hipLaunchKernelGGL(multiply_kernel<double>,dim3(nb),dim3(256),0,st,mult,
                       (cplx<double> *) F0,(const cplx<double> *) F1,(int64_t) n);
  return launch_check("multiply_kernel");
}

int fftwpp_amd_scale(int dtype, void *f, size_t n, double s, void *stream)
{
  if(n == 0) return FFTWPP_AMD_OK;
  unsigned nb=(unsigned) ((n+255)/256);
  hipStream_t st=(hipStream_t) stream;
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(scale_kernel<float>,dim3(nb),dim3(256),0,st,
                       (cplx<float> *) f,(int64_t) n,(float) s);
  else
    hipLaunchKernelGGL(scale_kernel<double>,dim3(nb),dim3(256),0,st,
                       (cplx<double> *) f,(int64_t) n,s);
  return launch_check("scale_kernel");
}

int fftwpp_amd_symmetrize_x(size_t Hx, size_t Hy, size_t x0, void *f,
                            size_t Sx, void *stream)
{
  if(!f || Hx == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad argument");
  if(x0+1 < Hx) return fail(FFTWPP_AMD_ERR_ARG,"x0 must be at least Hx-1");
  size_t n=Hx > Hy ? Hx : Hy;
  unsigned nb=(unsigned) ((n+255)/256);
  hipLaunchKernelGGL(symx_kernel,dim3(nb),dim3(256),0,(hipStream_t) stream,
                     (int64_t) Hx,(int64_t) Hy,(int64_t) x0,(cplx<double> *) f,
                     (int64_t) Sx);
  return launch_check("symx_kernel");
}

int fftwpp_amd_symmetrize_xy(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                             size_t y0, void *f, size_t Sx, size_t Sy,
                             void *stream)
{
  if(!f || Hx == 0 || Hy == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad argument");
  if(x0+1 < Hx || y0+1 < Hy) return fail(FFTWPP_AMD_ERR_ARG,"origin too small");
  hipStream_t st=(hipStream_t) stream;
  int64_t origin=(int64_t) (x0*Sx+y0*Sy);
  hipLaunchKernelGGL(symxy1_kernel,dim3((unsigned) ((Hx+255)/256)),dim3(256),0,st,
                     (int64_t) Hx,(cplx<double> *) f,origin,(int64_t) Sx);
  int rc=launch_check("symxy1_kernel");
  if(rc) return rc;
  size_t n2=(2*Hx-1)*(Hy-1);
  if(n2 > 0) {
    hipLaunchKernelGGL(symxy2_kernel,dim3((unsigned) ((n2+255)/256)),dim3(256),0,st,
                       (int64_t) Hx,(int64_t) Hy,(cplx<double> *) f,origin,
                       (int64_t) Sx,(int64_t) Sy);
    rc=launch_check("symxy2_kernel");
    if(rc) return rc;
  }
  size_t n3=((y0+Hy) > (x0+Hx) ? (y0+Hy) : (x0+Hx))*Hz;
  if(x0 == Hx || y0 == Hy) {
    hipLaunchKernelGGL(symxy3_kernel,dim3((unsigned) ((n3+255)/256)),dim3(256),0,st,
                       (int64_t) Hx,(int64_t) Hy,(int64_t) Hz,(int64_t) x0,
                       (int64_t) y0,(cplx<double> *) f,(int64_t) Sx,(int64_t) Sy);
    rc=launch_check("symxy3_kernel");
  }
  return rc;
}

int fftwpp_amd_hermitian_plane(int dtype, int op, size_t Hx, size_t Hy,
                               size_t x0, size_t y0, void *f, size_t Sx,
                               size_t Sy, void *stream)
{
  if(!f || Hx == 0 || Hy == 0) return fail(FFTWPP_AMD_ERR_ARG,"bad argument");
  if(op < 0 || op > 2) return fail(FFTWPP_AMD_ERR_ARG,"unknown plane operation");
  if(x0+1 < Hx || y0+1 < Hy) return fail(FFTWPP_AMD_ERR_ARG,"origin too small");
  hipStream_t st=(hipStream_t) stream;
  const size_t n=(2*Hx-1)*Hy;
  const dim3 grid((unsigned) ((n+255)/256));
  if(dtype == FFTWPP_AMD_F32)
    hipLaunchKernelGGL(hplane_kernel<float>,grid,dim3(256),0,st,op,(int64_t) Hx,
                       (int64_t) Hy,(int64_t) x0,(int64_t) y0,(cplx<float> *) f,
                       (int64_t) Sx,(int64_t) Sy);
  else
    hipLaunchKernelGGL(hplane_kernel<double>,grid,dim3(256),0,st,op,(int64_t) Hx,
                       (int64_t) Hy,(int64_t) x0,(int64_t) y0,(cplx<double> *) f,
                       (int64_t) Sx,(int64_t) Sy);
  return launch_check("hplane_kernel");
}

// ---- memory and stream plumbing for host code that includes no HIP header
int fftwpp_amd_malloc(void **p, size_t bytes)
{
  if(!p) return fail(FFTWPP_AMD_ERR_ARG,"null pointer");
  *p=NULL;
  if(bytes == 0) return FFTWPP_AMD_OK;
  return hipcheck(hipMalloc(p,bytes),"hipMalloc");
}

int fftwpp_amd_free(void *p)
{
  if(!p) return FFTWPP_AMD_OK;
  return hipcheck(hipFree(p),"hipFree");
}

int fftwpp_amd_memcpy(void *dst, const void *src, size_t bytes, void *stream)
{
  if(bytes == 0) return FFTWPP_AMD_OK;
  return hipcheck(hipMemcpyAsync(dst,src,bytes,hipMemcpyDefault,(hipStream_t) stream),"hipMemcpyAsync");
}

int fftwpp_amd_memset(void *dst, size_t bytes, void *stream)
{
  if(bytes == 0) return FFTWPP_AMD_OK;
  return hipcheck(hipMemsetAsync(dst,0,bytes,(hipStream_t) stream),"hipMemsetAsync");
}

int fftwpp_amd_synchronize(void *stream)
{
  return hipcheck(hipStreamSynchronize((hipStream_t) stream),"hipStreamSynchronize");
}

int fftwpp_amd_is_device_pointer(const void *p)
{
  if(!p) return 0;
  hipPointerAttribute_t attr;
  hipError_t e=hipPointerGetAttributes(&attr,p);
  if(e != hipSuccess) {
    (void) hipGetLastError();
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int fftwpp_amd_event_create(void **event)
{
  if(!event) return fail(FFTWPP_AMD_ERR_ARG,"null event");
  return hipcheck(hipEventCreate((hipEvent_t *) event),"hipEventCreate");
}

int fftwpp_amd_event_record(void *event, void *stream)
{
  return hipcheck(hipEventRecord((hipEvent_t) event,(hipStream_t) stream),"hipEventRecord");
}

int fftwpp_amd_event_elapsed(float *ms, void *start, void *stop)
{
  if(!ms) return fail(FFTWPP_AMD_ERR_ARG,"null output");
  int rc=hipcheck(hipEventSynchronize((hipEvent_t) stop),"hipEventSynchronize");
  if(rc) return rc;
  return hipcheck(hipEventElapsedTime(ms,(hipEvent_t) start,(hipEvent_t) stop),"hipEventElapsedTime");
}

int fftwpp_amd_event_destroy(void *event)
{
  if(!event) return FFTWPP_AMD_OK;
  return hipcheck(hipEventDestroy((hipEvent_t) event),"hipEventDestroy");
}

int fftwpp_amd_device_count(void)
{
  int n=0;
  if(hipGetDeviceCount(&n) != hipSuccess) {
    (void) hipGetLastError();
    return 0;
  }
  return n;
}

} // extern "C"
