/*
 * oracle.c -- CPU oracle for the hybrid implicitly dealiased convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for the baseline, as the timed CPU port).  The product path
 * (fftwpp_amd/) never links, loads or calls it.
 *
 * Contents
 *   1. Direct O(N^2) convolutions, restated from the reference test oracle
 *      tests/direct.h:11-156 and tests/direct.cc:5-98, plus the Hermitian
 *      symmetrisation helpers of convolve.h:1168-1267.
 *   2. A scalar mixed-radix Stockham FFT (own code, no FFTW).
 *   3. A CPU port of the implicitly padded residue transforms and of the
 *      Convolution/Convolution2/Convolution3 residue loops
 *      (convolve.cc:836-1466 forward family, :1468-1965 backward family,
 *      :7513-7575 Convolution::convolveRaw, convolve.h:1475-1535 and
 *      :1812-1879 for the 2D/3D drivers), OpenMP-parallel over columns and
 *      rows.  It is the "port" CPU baseline and a second oracle at sizes the
 *      direct sums cannot reach.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   direct_conv        numpy.convolve; closed form of tests/hybridconv.cc:50-69
 *   direct_convC       numpy.convolve on the centred index range
 *   direct_convh       full-spectrum expansion + numpy.convolve
 *   direct_conv2/3     scipy.signal.convolve
 *   direct_convh2/3    full-spectrum expansion + scipy.signal.convolve
 *   symmetrize_x/xy    Hermitian-symmetry invariant of the result + idempotence
 *   fft                numpy.fft at every radix used (2,3,4,5,7 and a generic prime)
 *   hybrid_conv1/2/3   direct_conv* at small sizes, numpy explicit padding at
 *                      larger sizes, independence from (m,q)
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------------ */
/* 1. Direct convolutions                                                    */
/* ------------------------------------------------------------------------ */

/* h[i] = sum_{j<=i} f[j] g[i-j], i < L.  (tests/direct.h:21-28) */
void oracle_direct_conv(cplx *h, const cplx *f, const cplx *g, size_t L)
{
  for(size_t i=0; i < L; ++i) {
    cplx sum=0.0;
    for(size_t j=0; j <= i; ++j) sum += f[j]*g[i-j];
    h[i]=sum;
  }
}

/* Centred 1D convolution: index j <-> wavenumber j-floor(L/2).
   (tests/direct.h:29-42) */
void oracle_direct_convC(cplx *h, const cplx *f, const cplx *g, size_t L)
{
  for(size_t i=0; i < L/2; ++i) {
    cplx sum=0.0;
    for(size_t j=i+1; j < L; ++j) sum += f[j]*g[L+i-j];
    h[i+(L+1)/2]=sum;
  }
  for(size_t i=L/2; i < L; ++i) {
    cplx sum=0.0;
    for(size_t j=0; j <= i; ++j) sum += f[j]*g[i-j];
    h[i-L/2]=sum;
  }
}

/* 1D Hermitian convolution of the H non-negative modes.
   (tests/direct.cc:5-15) */
void oracle_direct_convh(cplx *h, const cplx *f, const cplx *g, size_t H)
{
  for(size_t i=0; i < H; ++i) {
    cplx sum=0.0;
    for(size_t j=0; j <= i; ++j) sum += f[j]*g[i-j];
    for(size_t j=i+1; j < H; ++j) sum += f[j]*conj(g[j-i]);
    for(size_t j=1; j < H-i; ++j) sum += conj(f[j])*g[i+j];
    h[i]=sum;
  }
}

/* 2D complex convolution, input row stride Sx, output row stride Ly.
   (tests/direct.h:72-85) */
void oracle_direct_conv2(cplx *h, const cplx *f, const cplx *g,
                         size_t Lx, size_t Ly, size_t Sx)
{
  if(Sx == 0) Sx=Ly;
  for(size_t i=0; i < Lx; ++i)
    for(size_t j=0; j < Ly; ++j) {
      cplx sum=0.0;
      for(size_t k=0; k <= i; ++k)
        for(size_t p=0; p <= j; ++p)
          sum += f[Sx*k+p]*g[Sx*(i-k)+j-p];
      h[i*Ly+j]=sum;
    }
}

/* 3D complex convolution.  (tests/direct.h:119-136) */
void oracle_direct_conv3(cplx *h, const cplx *f, const cplx *g,
                         size_t Lx, size_t Ly, size_t Lz,
                         size_t Sx, size_t Sy)
{
  if(Sy == 0) Sy=Lz;
  if(Sx == 0) Sx=Ly*Sy;
  for(size_t i=0; i < Lx; ++i)
    for(size_t j=0; j < Ly; ++j)
      for(size_t k=0; k < Lz; ++k) {
        cplx sum=0.0;
        for(size_t r=0; r <= i; ++r)
          for(size_t p=0; p <= j; ++p)
            for(size_t q=0; q <= k; ++q)
              sum += f[r*Sx+p*Sy+q]*g[(i-r)*Sx+(j-p)*Sy+(k-q)];
        h[i*Ly*Lz+j*Lz+k]=sum;
      }
}

/* Enforce f(-kx,0)=conj f(kx,0), Im f(0,0)=0 and zero the Nyquist row when
   x0 == Hx.  (convolve.h:1177-1199) */
void oracle_symmetrize_x(size_t Hx, size_t Hy, size_t x0, cplx *f, size_t Sx)
{
  cplx *F=f+x0*Sx;
  for(size_t i=Sx; i < Hx*Sx; i += Sx)
    *(F-i)=conj(F[i]);
  F[0]=creal(F[0]);
  if(x0 == Hx)
    for(size_t j=0; j < Hy; ++j) f[j]=0.0;
}

/* 3D analogue on the kz=0 plane.  (convolve.h:1203-1256) */
void oracle_symmetrize_xy(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                          size_t y0, cplx *f, size_t Sx, size_t Sy)
{
  size_t origin=x0*Sx+y0*Sy;
  cplx *F=f+origin;
  size_t stop=Hx*Sx;
  for(size_t i=Sx; i < stop; i += Sx)
    *(F-i)=conj(F[i]);
  F[0]=creal(F[0]);
  for(ptrdiff_t i=(-(ptrdiff_t) Hx+1)*(ptrdiff_t) Sx; i < (ptrdiff_t) stop;
      i += (ptrdiff_t) Sx) {
    size_t mi=origin-i;
    size_t pi=origin+i;
    for(size_t j=Sy; j < Sy*Hy; j += Sy)
      f[mi-j]=conj(f[pi+j]);
  }
  if(x0 == Hx) {
    size_t Ly=y0+Hy;
    for(size_t j=0; j < Ly; ++j)
      for(size_t k=0; k < Hz; ++k) f[Sy*j+k]=0.0;
  }
  if(y0 == Hy) {
    size_t Lx=x0+Hx;
    for(size_t i=0; i < Lx; ++i)
      for(size_t k=0; k < Hz; ++k) f[Sx*i+k]=0.0;
  }
}

/* 2D centred Hermitian convolution: x centred (xorigin=Hx-xcompact), y holds
   the Hy non-negative modes.  Output row stride is Hy.
   (tests/direct.cc:17-55) */
void oracle_direct_convh2(cplx *h, cplx *f, cplx *g, size_t Hx, size_t Hy,
                          int xcompact, size_t Sx, int symmetrize)
{
  if(Sx == 0) Sx=Hy;
  size_t xorigin=Hx-(xcompact != 0);
  if(symmetrize) {
    oracle_symmetrize_x(Hx,Hy,xorigin,f,Hy);
    oracle_symmetrize_x(Hx,Hy,xorigin,g,Hy);
  }
  int xstart=-(int) xorigin, ystart=1-(int) Hy;
  int xstop=(int) Hx, ystop=(int) Hy;
  int pstart=xstart+!xcompact;
  for(int kx=xstart; kx < xstop; ++kx)
    for(int ky=0; ky < ystop; ++ky) {
      cplx sum=0.0;
      for(int px=pstart; px < xstop; ++px) {
        int qx=kx-px;
        if(qx < pstart || qx >= xstop) continue;
        for(int py=ystart; py < ystop; ++py) {
          int qy=ky-py;
          if(qy < ystart || qy >= ystop) continue;
          cplx a=py >= 0 ? f[(xorigin+px)*Sx+py] : conj(f[(xorigin-px)*Sx-py]);
          cplx b=qy >= 0 ? g[(xorigin+qx)*Sx+qy] : conj(g[(xorigin-qx)*Sx-qy]);
          sum += a*b;
        }
      }
      h[(xorigin+kx)*Hy+ky]=sum;
    }
}

/* 3D centred Hermitian convolution: x and y centred, z holds Hz
   non-negative modes.  Output strides (2Hy-ycompact)*Hz and Hz.
   (tests/direct.cc:57-98) */
void oracle_direct_convh3(cplx *h, cplx *f, cplx *g, size_t Hx, size_t Hy,
                          size_t Hz, int xcompact, int ycompact,
                          size_t Sx, size_t Sy, int symmetrize)
{
  if(Sy == 0) Sy=Hz;
  if(Sx == 0) Sx=(2*Hy-(ycompact != 0))*Sy;
  size_t xorigin=Hx-(xcompact != 0), yorigin=Hy-(ycompact != 0);
  size_t ny=2*Hy-(ycompact != 0);
  if(symmetrize) {
    size_t Ly=yorigin+Hy;
    oracle_symmetrize_xy(Hx,Hy,Hz,xorigin,yorigin,f,Ly*Hz,Hz);
    oracle_symmetrize_xy(Hx,Hy,Hz,xorigin,yorigin,g,Ly*Hz,Hz);
  }
  int xstart=-(int) xorigin, ystart=-(int) yorigin, zstart=1-(int) Hz;
  int xstop=(int) Hx, ystop=(int) Hy, zstop=(int) Hz;
  int pxs=xstart+!xcompact, pys=ystart+!ycompact;
  for(int kx=xstart; kx < xstop; ++kx)
    for(int ky=ystart; ky < ystop; ++ky)
      for(int kz=0; kz < zstop; ++kz) {
        cplx sum=0.0;
        for(int px=pxs; px < xstop; ++px) {
          int qx=kx-px;
          if(qx < pxs || qx >= xstop) continue;
          for(int py=pys; py < ystop; ++py) {
            int qy=ky-py;
            if(qy < pys || qy >= ystop) continue;
            for(int pz=zstart; pz < zstop; ++pz) {
              int qz=kz-pz;
              if(qz < zstart || qz >= zstop) continue;
              cplx a=pz >= 0 ? f[Sx*(xorigin+px)+Sy*(yorigin+py)+pz] :
                conj(f[Sx*(xorigin-px)+Sy*(yorigin-py)-pz]);
              cplx b=qz >= 0 ? g[Sx*(xorigin+qx)+Sy*(yorigin+qy)+qz] :
                conj(g[Sx*(xorigin-qx)+Sy*(yorigin-qy)-qz]);
              sum += a*b;
            }
          }
        }
        h[((xorigin+kx)*ny+yorigin+ky)*Hz+kz]=sum;
      }
}

/* ------------------------------------------------------------------------ */
/* 2. Mixed-radix Stockham FFT (sign +1 forward as in convolve.cc:576-577)   */
/* ------------------------------------------------------------------------ */

typedef struct {
  size_t n;
  int nstages;
  int radix[64];
  cplx *tw;  /* tw[j]=exp(2 pi i j/n), j<n */
} offt;

static cplx expi_frac(uint64_t num, uint64_t den)
{
  /* exp(2 pi i num/den) with the fraction reduced before the sincos */
  num %= den;
  double a=2.0*M_PI*((double) num/(double) den);
  return cos(a)+I*sin(a);
}

static void offt_init(offt *p, size_t n)
{
  p->n=n;
  p->nstages=0;
  size_t r=n;
  while(r % 4 == 0) { p->radix[p->nstages++]=4; r /= 4; }
  while(r % 2 == 0) { p->radix[p->nstages++]=2; r /= 2; }
  for(size_t f=3; r > 1; f += 2)
    while(r % f == 0) { p->radix[p->nstages++]=(int) f; r /= f; }
  p->tw=(cplx *) malloc(sizeof(cplx)*(n ? n : 1));
  for(size_t j=0; j < n; ++j) p->tw[j]=expi_frac(j,n);
}

static void offt_free(offt *p) { free(p->tw); }

/* In-place transform of x (length n) using work (length n).
   sign=+1: X_k = sum_j x_j e^{+2 pi i jk/n}; sign=-1: conjugate kernel. */
static void offt_exec(const offt *p, cplx *x, cplx *work, int sign)
{
  size_t n=p->n;
  if(n <= 1) return;
  cplx *a=x, *b=work;
  size_t Ns=1;
  for(int st=0; st < p->nstages; ++st) {
    int R=p->radix[st];
    size_t nR=n/R;
    size_t twstep=n/(Ns*R);
    for(size_t j=0; j < nR; ++j) {
      size_t k=j % Ns;
      cplx v[64], y[64];
      cplx *vv=R <= 64 ? v : (cplx *) malloc(sizeof(cplx)*R);
      cplx *yy=R <= 64 ? y : (cplx *) malloc(sizeof(cplx)*R);
      for(int t=0; t < R; ++t) {
        cplx w=p->tw[(t*k*twstep) % n];
        if(sign < 0) w=conj(w);
        vv[t]=a[j+t*nR]*w;
      }
      if(R == 2) {
        yy[0]=vv[0]+vv[1]; yy[1]=vv[0]-vv[1];
      } else if(R == 4) {
        cplx s02=vv[0]+vv[2], d02=vv[0]-vv[2];
        cplx s13=vv[1]+vv[3], d13=vv[1]-vv[3];
        cplx jd13=(sign > 0 ? I : -I)*d13;
        yy[0]=s02+s13; yy[2]=s02-s13;
        yy[1]=d02+jd13; yy[3]=d02-jd13;
      } else {
        size_t base=n/R;
        for(int u=0; u < R; ++u) {
          cplx s=0.0;
          for(int t=0; t < R; ++t) {
            cplx w=p->tw[(size_t) ((u*t) % R)*base];
            if(sign < 0) w=conj(w);
            s += vv[t]*w;
          }
          yy[u]=s;
        }
      }
      size_t obase=(j/Ns)*Ns*R+k;
      for(int u=0; u < R; ++u) b[obase+u*Ns]=yy[u];
      if(R > 64) { free(vv); free(yy); }
    }
    cplx *tmp=a; a=b; b=tmp;
    Ns *= R;
  }
  if(a != x) memcpy(x,a,sizeof(cplx)*n);
}

/* Exported for the FFT pinning test. */
void oracle_fft(cplx *x, size_t n, int sign)
{
  offt p;
  offt_init(&p,n);
  cplx *w=(cplx *) malloc(sizeof(cplx)*(n ? n : 1));
  offt_exec(&p,x,w,sign);
  free(w);
  offt_free(&p);
}

/* ------------------------------------------------------------------------ */
/* 3. Implicitly padded residue transforms and convolution drivers           */
/* ------------------------------------------------------------------------ */

/* One padded dimension.  Data index j in [0,L) holds wavenumber k=j+k0
   (k0=0 for fftPad, -floor(L/2) for fftPadCentered).  The padded transform
   evaluates F_l = sum_k f_k zeta_N^{k l}, zeta_N=e^{2 pi i/N}, N=m*q, and
   residue r (0<=r<q) is the set l=q*s+r, s<m.  Writing k=m*t+sigma:
     F_{qs+r} = DFT_m[ zeta_N^{sigma r} sum_t zeta_q^{t r} f_{mt+sigma} ](s),
   which covers the p=1 pre-twiddle (convolve.cc:905-921), the p=2 fold
   (convolve.cc:1088-1108) and the p>2 block DFT (convolve.cc:1281-1326). */
typedef struct {
  size_t L, m, q, N;
  ptrdiff_t k0;
  offt plan;
  cplx *zN;  /* zN[j]=zeta_N^j, j<N */
} odim;

static void odim_init(odim *d, size_t L, size_t m, size_t q, int centered)
{
  d->L=L; d->m=m; d->q=q; d->N=m*q;
  d->k0=centered ? -(ptrdiff_t) (L/2) : 0;
  offt_init(&d->plan,m);
  d->zN=(cplx *) malloc(sizeof(cplx)*d->N);
  for(size_t j=0; j < d->N; ++j) d->zN[j]=expi_frac(j,d->N);
}

static void odim_free(odim *d) { offt_free(&d->plan); free(d->zN); }

static inline size_t modN(ptrdiff_t a, size_t N)
{
  ptrdiff_t r=a % (ptrdiff_t) N;
  return (size_t) (r < 0 ? r+(ptrdiff_t) N : r);
}

/* Forward residue r of one column: in[j*stride], j<L -> W[s], s<m. */
static void odim_forward(const odim *d, const cplx *in, size_t stride,
                         size_t r, cplx *W, cplx *work)
{
  size_t m=d->m, N=d->N;
  for(size_t s=0; s < m; ++s) W[s]=0.0;
  for(size_t j=0; j < d->L; ++j) {
    ptrdiff_t k=(ptrdiff_t) j+d->k0;
    /* zeta_N^{k r} = zeta_N^{sigma r} zeta_q^{t r} */
    size_t sigma=modN(k,m);
    W[sigma] += d->zN[modN(k*(ptrdiff_t) r,N)]*in[j*stride];
  }
  offt_exec(&d->plan,W,work,1);
}

/* Backward residue r: out[j*stride] (+)= sum_s F[s] zeta_N^{-k(qs+r)}. */
static void odim_backward(const odim *d, cplx *F, cplx *out, size_t stride,
                          size_t r, int accumulate, cplx *work)
{
  size_t m=d->m, N=d->N;
  offt_exec(&d->plan,F,work,-1);
  for(size_t j=0; j < d->L; ++j) {
    ptrdiff_t k=(ptrdiff_t) j+d->k0;
    size_t sigma=modN(k,m);
    cplx v=conj(d->zN[modN(k*(ptrdiff_t) r,N)])*F[sigma];
    if(accumulate) out[j*stride] += v;
    else out[j*stride]=v;
  }
}

/* Hermitian dimension (fftPadHermitian, convolve.cc:4309-5447): data holds
   H=ceil(L/2) non-negative modes of a Hermitian spectrum; the forward
   transform produces the real samples u_l = sum_{|k|<H} f_k zeta_N^{k l}. */
static void odimh_forward(const odim *d, const cplx *in, size_t stride,
                          size_t r, double *u, cplx *W, cplx *work)
{
  size_t m=d->m, N=d->N, H=(d->L+1)/2;
  for(size_t s=0; s < m; ++s) W[s]=0.0;
  for(size_t j=0; j < H; ++j) {
    cplx v=in[j*stride];
    ptrdiff_t k=(ptrdiff_t) j;
    W[modN(k,m)] += d->zN[modN(k*(ptrdiff_t) r,N)]*v;
    if(j > 0)
      W[modN(-k,m)] += d->zN[modN(-k*(ptrdiff_t) r,N)]*conj(v);
  }
  offt_exec(&d->plan,W,work,1);
  for(size_t s=0; s < m; ++s) u[s]=creal(W[s]);
}

static void odimh_backward(const odim *d, const double *u, cplx *out,
                           size_t stride, size_t r, int accumulate,
                           cplx *W, cplx *work)
{
  size_t m=d->m, N=d->N, H=(d->L+1)/2;
  for(size_t s=0; s < m; ++s) W[s]=u[s];
  offt_exec(&d->plan,W,work,-1);
  for(size_t j=0; j < H; ++j) {
    ptrdiff_t k=(ptrdiff_t) j;
    cplx v=conj(d->zN[modN(k*(ptrdiff_t) r,N)])*W[modN(k,m)];
    if(accumulate) out[j*stride] += v;
    else out[j*stride]=v;
  }
}

/* Fused 1D convolution of one row (Convolution::convolveRaw,
   convolve.cc:7513-7575): for every residue, forward both inputs, multiply
   (multBinary, convolve.cc:33-56), backward and accumulate.  Unnormalised. */
static void conv_row(const odim *d, cplx *f, const cplx *g, size_t stride,
                     cplx *Wf, cplx *Wg, cplx *acc, cplx *work)
{
  for(size_t r=0; r < d->q; ++r) {
    odim_forward(d,f,stride,r,Wf,work);
    odim_forward(d,g,stride,r,Wg,work);
    for(size_t s=0; s < d->m; ++s) Wf[s] *= Wg[s];
    odim_backward(d,Wf,acc,1,r,r > 0,work);
  }
  for(size_t j=0; j < d->L; ++j) f[j*stride]=acc[j];
}

/* Hermitian row: realMultBinary on the real samples (convolve.cc:60-83). */
static void hconv_row(const odim *d, cplx *f, const cplx *g, size_t stride,
                      double *u, double *v, cplx *W, cplx *acc, cplx *work)
{
  size_t H=(d->L+1)/2;
  for(size_t r=0; r < d->q; ++r) {
    odimh_forward(d,f,stride,r,u,W,work);
    odimh_forward(d,g,stride,r,v,W,work);
    for(size_t s=0; s < d->m; ++s) u[s] *= v[s];
    odimh_backward(d,u,acc,1,r,r > 0,W,work);
  }
  for(size_t j=0; j < H; ++j) f[j*stride]=acc[j];
}

static int nthreads_req=0;
void oracle_set_threads(int n) { nthreads_req=n; }
int oracle_get_threads(void)
{
#ifdef _OPENMP
  return nthreads_req > 0 ? nthreads_req : omp_get_max_threads();
#else
  return 1;
#endif
}

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(dynamic,4) num_threads(oracle_get_threads())")
#else
#define OMP_FOR
#endif

/* 1D: f <- (f*g) normalised by 1/N (Convolution::convolve). */
void oracle_hybrid_conv1(cplx *f, cplx *g, size_t L, size_t m, size_t q,
                         int centered)
{
  odim d;
  odim_init(&d,L,m,q,centered);
  cplx *buf=(cplx *) malloc(sizeof(cplx)*(3*m+L));
  conv_row(&d,f,g,1,buf,buf+m,buf+2*m+m,buf+2*m);
  double scale=1.0/(double) d.N;
  for(size_t j=0; j < L; ++j) f[j] *= scale;
  free(buf);
  odim_free(&d);
}

void oracle_hybrid_hconv1(cplx *f, cplx *g, size_t L, size_t m, size_t q)
{
  odim d;
  odim_init(&d,L,m,q,1);
  size_t H=(L+1)/2;
  double *u=(double *) malloc(sizeof(double)*2*m);
  cplx *buf=(cplx *) malloc(sizeof(cplx)*(2*m+H));
  hconv_row(&d,f,g,1,u,u+m,buf,buf+2*m,buf+m);
  double scale=1.0/(double) d.N;
  for(size_t j=0; j < H; ++j) f[j] *= scale;
  free(u); free(buf);
  odim_free(&d);
}

/* Outer-dimension forward/backward over C columns of stride 1 and element
   stride S (the "Many" routines).  F has layout F[s*C+c], s<m. */
static void many_forward(const odim *d, const cplx *f, size_t C, size_t S,
                         size_t r, cplx *F)
{
  size_t m=d->m;
  OMP_FOR
  for(size_t c=0; c < C; ++c) {
    cplx *W=(cplx *) malloc(sizeof(cplx)*2*m);
    odim_forward(d,f+c,S,r,W,W+m);
    for(size_t s=0; s < m; ++s) F[s*C+c]=W[s];
    free(W);
  }
}

static void many_backward(const odim *d, const cplx *F, size_t C, cplx *out,
                          size_t S, size_t r, int accumulate)
{
  size_t m=d->m;
  OMP_FOR
  for(size_t c=0; c < C; ++c) {
    cplx *W=(cplx *) malloc(sizeof(cplx)*2*m);
    for(size_t s=0; s < m; ++s) W[s]=F[s*C+c];
    odim_backward(d,W,out+c,S,r,accumulate,W+m);
    free(W);
  }
}

/* 2D plane convolution of Fx rows (Ly x Sy-stride rows of length Lz... ) */
static void plane_conv2(const odim *dy, const odim *dz, int hermitian,
                        cplx *f, cplx *g, size_t Sy, cplx *Fy, cplx *Gy,
                        cplx *V)
{
  /* f,g: Ly rows of stride Sy, length lz (complex) */
  size_t my=dy->m;
  size_t lz=hermitian ? (dz->L+1)/2 : dz->L;
  size_t mz=dz->m;
  cplx *W=(cplx *) malloc(sizeof(cplx)*(4*(my > mz ? my : mz)+lz));
  double *u=(double *) malloc(sizeof(double)*2*mz);
  for(size_t ry=0; ry < dy->q; ++ry) {
    for(size_t c=0; c < lz; ++c) {
      odim_forward(dy,f+c,Sy,ry,W,W+my);
      for(size_t s=0; s < my; ++s) Fy[s*lz+c]=W[s];
      odim_forward(dy,g+c,Sy,ry,W,W+my);
      for(size_t s=0; s < my; ++s) Gy[s*lz+c]=W[s];
    }
    for(size_t s=0; s < my; ++s) {
      if(hermitian)
        hconv_row(dz,Fy+s*lz,Gy+s*lz,1,u,u+mz,W,W+2*mz,W+mz);
      else
        conv_row(dz,Fy+s*lz,Gy+s*lz,1,W,W+mz,W+3*mz,W+2*mz);
    }
    for(size_t c=0; c < lz; ++c) {
      for(size_t s=0; s < my; ++s) W[s]=Fy[s*lz+c];
      odim_backward(dy,W,V+c,lz,ry,ry > 0,W+my);
    }
  }
  for(size_t j=0; j < dy->L; ++j)
    for(size_t c=0; c < lz; ++c) f[j*Sy+c]=V[j*lz+c];
  free(W); free(u);
}

/* 2D: x outer (fftPad or fftPadCentered), y inner (fftPad or Hermitian).
   kind 0 = complex (conv2d), kind 1 = centred Hermitian (hconv2d).
   f,g contiguous Lx x ly, ly = Ly or ceil(Ly/2).  Result normalised in f. */
void oracle_hybrid_conv2(cplx *f, cplx *g, size_t Lx, size_t Ly, size_t mx,
                         size_t qx, size_t my, size_t qy, int kind)
{
  int herm=kind == 1;
  odim dx, dy;
  odim_init(&dx,Lx,mx,qx,herm);
  odim_init(&dy,Ly,my,qy,herm);
  size_t ly=herm ? (Ly+1)/2 : Ly;
  size_t C=ly;
  cplx *Fx=(cplx *) malloc(sizeof(cplx)*mx*C);
  cplx *Gx=(cplx *) malloc(sizeof(cplx)*mx*C);
  cplx *V=(cplx *) malloc(sizeof(cplx)*Lx*C);
  for(size_t rx=0; rx < qx; ++rx) {
    many_forward(&dx,f,C,C,rx,Fx);
    many_forward(&dx,g,C,C,rx,Gx);
    OMP_FOR
    for(size_t i=0; i < mx; ++i) {
      cplx *row=Fx+i*C, *grow=Gx+i*C;
      if(herm) {
        double *u=(double *) malloc(sizeof(double)*2*my);
        cplx *W=(cplx *) malloc(sizeof(cplx)*(2*my+ly));
        hconv_row(&dy,row,grow,1,u,u+my,W,W+2*my,W+my);
        free(u); free(W);
      } else {
        cplx *W=(cplx *) malloc(sizeof(cplx)*(3*my+Ly));
        conv_row(&dy,row,grow,1,W,W+my,W+3*my,W+2*my);
        free(W);
      }
    }
    many_backward(&dx,Fx,C,V,C,rx,rx > 0);
  }
  double scale=1.0/((double) dx.N*(double) dy.N);
  for(size_t i=0; i < Lx*C; ++i) f[i]=V[i]*scale;
  free(Fx); free(Gx); free(V);
  odim_free(&dx); odim_free(&dy);
}

/* 3D: x and y outer, z inner.  kind 0 = complex (conv3d), 1 = centred
   Hermitian (hconv3d, z Hermitian).  Layout Lx x Ly x lz contiguous. */
void oracle_hybrid_conv3(cplx *f, cplx *g, size_t Lx, size_t Ly, size_t Lz,
                         size_t mx, size_t qx, size_t my, size_t qy,
                         size_t mz, size_t qz, int kind)
{
  int herm=kind == 1;
  odim dx, dy, dz;
  odim_init(&dx,Lx,mx,qx,herm);
  odim_init(&dy,Ly,my,qy,herm);
  odim_init(&dz,Lz,mz,qz,herm);
  size_t lz=herm ? (Lz+1)/2 : Lz;
  size_t C=Ly*lz;
  cplx *Fx=(cplx *) malloc(sizeof(cplx)*mx*C);
  cplx *Gx=(cplx *) malloc(sizeof(cplx)*mx*C);
  cplx *V=(cplx *) malloc(sizeof(cplx)*Lx*C);
  for(size_t rx=0; rx < qx; ++rx) {
    many_forward(&dx,f,C,C,rx,Fx);
    many_forward(&dx,g,C,C,rx,Gx);
    OMP_FOR
    for(size_t i=0; i < mx; ++i) {
      cplx *Fy=(cplx *) malloc(sizeof(cplx)*my*lz);
      cplx *Gy=(cplx *) malloc(sizeof(cplx)*my*lz);
      cplx *Vp=(cplx *) malloc(sizeof(cplx)*Ly*lz);
      plane_conv2(&dy,&dz,herm,Fx+i*C,Gx+i*C,lz,Fy,Gy,Vp);
      free(Fy); free(Gy); free(Vp);
    }
    many_backward(&dx,Fx,C,V,C,rx,rx > 0);
  }
  double scale=1.0/((double) dx.N*(double) dy.N*(double) dz.N);
  for(size_t i=0; i < Lx*C; ++i) f[i]=V[i]*scale;
  free(Fx); free(Gx); free(V);
  odim_free(&dx); odim_free(&dy); odim_free(&dz);
}

/* Forward residue r of a single padded transform (for the transform
   identity check of tests/hybrid.cc:60-110).  out has m entries. */
void oracle_pad_forward(const cplx *f, size_t L, size_t m, size_t q,
                        int centered, size_t r, cplx *out)
{
  odim d;
  odim_init(&d,L,m,q,centered);
  cplx *work=(cplx *) malloc(sizeof(cplx)*m);
  odim_forward(&d,f,1,r,out,work);
  free(work);
  odim_free(&d);
}
