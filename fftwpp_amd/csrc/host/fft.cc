// fft.cc -- standalone FFTs (the transform classes of fftw++.h:556-1444) on
// the residue transforms of the kernel layer.
//
// A DFT of length n = m q with sign +1 is every residue of the padded
// transform of a dimension with L = N = n: residue r, position s holds
// sum_k f_k zeta_n^{k(q s + r)}, so one forward residue pass per r, written
// with output stride q and offset r, leaves the transform in natural order.
// Sign -1 is the backward transform of the same dimension read the same
// way: residue r takes the inputs at indices q s + r, and the passes
// accumulate into the output.  m is the largest divisor of n that the
// sub-FFT layer takes (at most 4096), so powers of two up to 4096 are one
// register pass and up to 8192 two.  Longer contiguous 1D transforms
// (q > 2) run in four steps instead: DFTs of length n1 = q down the columns
// of the n1 x n2 matrix (n2 = m), the twiddles exp(sign 2 pi i j k1 / n),
// and DFTs of length n2 along the rows written transposed, three passes.
// A 1D length with no divisor in [n/4096, 4096] (a prime above 4096, say)
// runs through Bluestein's algorithm: chirp multiply, a circular
// convolution of a power-of-two length >= 2n - 1 by two such transforms,
// chirp multiply.
#include "fftwpp_amd.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

extern "C" void fftwpp_amd_internal_set_error(const char *msg);
extern "C" int fftwpp_amd_internal_fft_map(int dtype, int op, const void *x, size_t xs,
                                           size_t xd, void *y, size_t ys, size_t yd,
                                           size_t n, size_t nb, double scale, void *stream);
extern "C" int fftwpp_amd_internal_fft_twiddle(int dtype, void *y, size_t n, size_t n2,
                                               size_t nb, int sign, void *stream);
extern "C" int fftwpp_amd_internal_fft_chirp(int dtype, int op, const void *x, size_t xs,
                                             size_t xd, void *y, size_t ys, size_t yd,
                                             size_t n, size_t len, size_t nb, int sign,
                                             double scale, void *stream);

struct fftwpp_amd_fft {
  int kind, dims, sign, dtype;
  size_t n[3], q[3];
  size_t howmany, istride, ostride, idist, odist;
  fftwpp_amd_dim *d[3];
  // four-step split of a long 1D transform: lengths n1 x n2, one dimension each
  size_t n1, n2;
  fftwpp_amd_dim *f1, *f2;
  // Bluestein: transforms of the power of two blen >= 2n - 1 (sign -1, +1),
  // the chirp w (n words), the transformed kernel / blen (blen words) and
  // blen words per transform of work
  size_t blen;
  fftwpp_amd_fft *bf, *bi;
  void *chirp, *bhat, *bwork;
  // real 2D/3D transforms: r2c / c2r of the last axis over all leading
  // indices (contiguous real rows, and rows padded to 2 (n/2+1) reals as an
  // in-place transform stores them), complex transforms of the leading axes
  fftwpp_amd_fft *rows, *rowsPad;
  size_t lead, half;
  void *tmp[3];
  size_t tmpBytes;
};

namespace {

const size_t maxSub=4096;     // largest sub-FFT size m
const size_t maxPasses=4096;  // largest residue count q = n/m
const size_t maxBatch=65535;  // grid rows per launch (single columns of many
                              // entries run as columns and need no split)

int error(int rc, const char *msg)
{
  fftwpp_amd_internal_set_error(msg);
  return rc;
}

size_t subsize(size_t n)
{
  if(n <= maxSub) return n;
  for(size_t m=maxSub; m > 1; --m)
    if(n % m == 0) return m;
  return 1;
}

size_t elt(int dtype) { return dtype == FFTWPP_AMD_F32 ? 8 : 16; }

// One axis: element (b, j, c) of x at b*xb + j*xS + c (j < n, c < C,
// b < nb), of y at b*yb + j*yS + c.
int axis(const fftwpp_amd_fft *h, int a, int sign, const char *x, size_t xS,
         size_t C, size_t nb, size_t xb, char *y, size_t yS, size_t yb,
         void *stream)
{
  const size_t q=h->q[a], E=elt(h->dtype);
  const size_t step=C == 1 ? nb : maxBatch;
  for(size_t b0=0; b0 < nb; b0 += step) {
    const size_t cnt=std::min(step,nb-b0);
    const char *xx=x+b0*xb*E;
    char *yy=y+b0*yb*E;
    for(size_t r=0; r < q; ++r) {
      int rc=sign > 0 ?
        fftwpp_amd_forward_many(h->d[a],h->dtype,xx,xS,C,cnt,xb,yy+r*yS*E,q*yS,yb,r,1,stream) :
        fftwpp_amd_backward_many(h->d[a],h->dtype,xx+r*xS*E,q*xS,xb,yy,yS,C,cnt,yb,r,1,
                                 r > 0,1.0,stream);
      if(rc) return rc;
    }
  }
  return FFTWPP_AMD_OK;
}

// One DFT of length n (1 or 2 residue passes) of a q = 1 dimension: column
// form as in axis().
int single(const fftwpp_amd_fft *h, fftwpp_amd_dim *d, int sign, const char *x, size_t xS,
           size_t C, size_t nb, size_t xb, char *y, size_t yS, size_t yb, void *stream)
{
  const size_t E=elt(h->dtype);
  const size_t step=C == 1 ? nb : maxBatch;
  for(size_t b0=0; b0 < nb; b0 += step) {
    const size_t cnt=std::min(step,nb-b0);
    int rc=sign > 0 ?
      fftwpp_amd_forward_many(d,h->dtype,x+b0*xb*E,xS,C,cnt,xb,y+b0*yb*E,yS,yb,0,1,stream) :
      fftwpp_amd_backward_many(d,h->dtype,x+b0*xb*E,xS,xb,y+b0*yb*E,yS,C,cnt,yb,0,1,0,1.0,
                               stream);
    if(rc) return rc;
  }
  return FFTWPP_AMD_OK;
}

// H contiguous transforms of length n = n1 n2 from x to y (x may equal y)
// through the temporary t: element j1 n2 + j2 is row j1, column j2.
int fourstep(const fftwpp_amd_fft *h, const char *x, char *y, char *t, size_t H, void *stream)
{
  const size_t n=h->n[0], n1=h->n1, n2=h->n2, E=elt(h->dtype);
  // columns: t[k1 n2 + j2] = sum_j1 x[j1 n2 + j2] w_n1^{j1 k1}
  int rc=single(h,h->f1,h->sign,x,n2,n2,H,n,t,n2,n,stream);
  if(rc) return rc;
  if((rc=fftwpp_amd_internal_fft_twiddle(h->dtype,t,n,n2,H,h->sign,stream))) return rc;
  // rows k1, written transposed: y[k2 n1 + k1] = sum_j2 t[k1 n2 + j2] w_n2^{j2 k2}
  for(size_t b=0; b < H; ++b)
    if((rc=single(h,h->f2,h->sign,t+b*n*E,1,1,n1,n2,y+b*n*E,n1,1,stream))) return rc;
  return FFTWPP_AMD_OK;
}

// Bluestein's algorithm for H transforms of length n: element j of
// transform t at x[t xd + j xs], output at y[t yd + k ys].
int bluestein(const fftwpp_amd_fft *h, const void *x, size_t xs, size_t xd, void *y,
              size_t ys, size_t yd, size_t H, double scale, void *stream)
{
  const size_t n=h->n[0], len=h->blen, E=elt(h->dtype);
  int rc=fftwpp_amd_internal_fft_chirp(h->dtype,2,x,xs,xd,h->bwork,1,len,n,len,H,h->sign,1.0,
                                       stream);
  if(!rc) rc=fftwpp_amd_fft_execute(h->bf,h->bwork,h->bwork,1.0,stream);
  for(size_t t=0; !rc && t < H; ++t)
    rc=fftwpp_amd_multiply(h->dtype,FFTWPP_AMD_MULT_BINARY,(char *) h->bwork+t*len*E,h->bhat,
                           len,stream);
  if(!rc) rc=fftwpp_amd_fft_execute(h->bi,h->bwork,h->bwork,1.0,stream);
  if(!rc) rc=fftwpp_amd_internal_fft_chirp(h->dtype,3,h->bwork,1,len,y,ys,yd,n,len,H,h->sign,
                                           scale,stream);
  return rc;
}

int bluesteinSetup(fftwpp_amd_fft *h)
{
  const size_t n=h->n[0], H=h->howmany, E=elt(h->dtype);
  size_t len=1;
  while(len < 2*n-1) len *= 2;
  h->blen=len;
  int rc=fftwpp_amd_fft_create(&h->bf,FFTWPP_AMD_FFT_C2C,1,&len,-1,h->dtype,H,1,1,len,len);
  if(!rc) rc=fftwpp_amd_fft_create(&h->bi,FFTWPP_AMD_FFT_C2C,1,&len,1,h->dtype,H,1,1,len,len);
  if(!rc) rc=fftwpp_amd_malloc(&h->chirp,n*E);
  if(!rc) rc=fftwpp_amd_malloc(&h->bhat,len*E);
  if(!rc) rc=fftwpp_amd_malloc(&h->bwork,len*H*E);
  if(rc) return rc;
  // the kernel's transform, scaled by 1/len for the inverse
  fftwpp_amd_fft *one=NULL;
  rc=fftwpp_amd_fft_create(&one,FFTWPP_AMD_FFT_C2C,1,&len,-1,h->dtype,1,1,1,len,len);
  if(!rc) rc=fftwpp_amd_internal_fft_chirp(h->dtype,0,NULL,0,0,h->chirp,1,n,n,len,1,h->sign,1.0,
                                           NULL);
  if(!rc) rc=fftwpp_amd_internal_fft_chirp(h->dtype,1,NULL,0,0,h->bhat,1,len,n,len,1,h->sign,
                                           1.0,NULL);
  if(!rc) rc=fftwpp_amd_fft_execute(one,h->bhat,h->bhat,1.0/(double) len,NULL);
  if(!rc) rc=fftwpp_amd_synchronize(NULL);
  fftwpp_amd_fft_destroy(one);
  return rc;
}

int temporaries(fftwpp_amd_fft *h, size_t words)
{
  const size_t bytes=words*elt(h->dtype);
  if(bytes <= h->tmpBytes) return FFTWPP_AMD_OK;
  for(void *&p : h->tmp) {
    if(p) fftwpp_amd_free(p);
    p=NULL;
  }
  h->tmpBytes=0;
  for(void *&p : h->tmp) {
    int rc=fftwpp_amd_malloc(&p,bytes);
    if(rc) return rc;
  }
  h->tmpBytes=bytes;
  return FFTWPP_AMD_OK;
}

size_t total(const fftwpp_amd_fft *h)
{
  size_t t=1;
  for(int a=0; a < h->dims; ++a) t *= h->n[a];
  return t;
}

} // namespace

extern "C" {

int fftwpp_amd_fft_create(fftwpp_amd_fft **out, int kind, int dims, const size_t *n,
                          int sign, int dtype, size_t howmany, size_t istride,
                          size_t ostride, size_t idist, size_t odist)
{
  if(!out || !n) return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_create: null argument");
  *out=NULL;
  if(kind < FFTWPP_AMD_FFT_C2C || kind > FFTWPP_AMD_FFT_C2R || dims < 1 || dims > 3 ||
     (dtype != FFTWPP_AMD_F64 && dtype != FFTWPP_AMD_F32) || (sign != 1 && sign != -1))
    return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_create: bad kind, dims, sign or dtype");
  if(kind == FFTWPP_AMD_FFT_R2C) sign=-1;
  if(kind == FFTWPP_AMD_FFT_C2R) sign=1;
  if(dims > 1 && howmany > 1)
    return error(FFTWPP_AMD_ERR_UNSUPPORTED,"fftwpp_amd_fft_create: multidimensional transforms are single transforms");
  if(howmany == 0) howmany=1;
  for(int a=0; a < dims; ++a)
    if(n[a] == 0) return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_create: zero length");
  const size_t hw=kind == FFTWPP_AMD_FFT_C2C ? n[0] : n[0]/2+1;
  const size_t inlen=kind == FFTWPP_AMD_FFT_C2R ? hw : n[0];
  const size_t outlen=kind == FFTWPP_AMD_FFT_R2C ? hw : n[0];
  if(dims == 1) {
    if(istride == 0) istride=1;
    if(ostride == 0) ostride=1;
    if(idist == 0) idist=inlen;
    if(odist == 0) odist=outlen;
  } else {
    istride=ostride=1;
    idist=odist=0;
  }
  fftwpp_amd_fft *h=new(std::nothrow) fftwpp_amd_fft();
  if(!h) return error(FFTWPP_AMD_ERR_HIP,"fftwpp_amd_fft_create: out of memory");
  h->kind=kind; h->dims=dims; h->sign=sign; h->dtype=dtype;
  h->howmany=howmany; h->istride=istride; h->ostride=ostride; h->idist=idist; h->odist=odist;
  for(int a=0; a < dims; ++a) h->n[a]=n[a];
  int axes=dims;
  if(kind != FFTWPP_AMD_FFT_C2C && dims > 1) {
    // the last axis as 1D real transforms of all leading indices
    const size_t nl=n[dims-1], hl=nl/2+1;
    h->lead=1;
    for(int a=0; a+1 < dims; ++a) h->lead *= n[a];
    h->half=hl;
    const bool r2c=kind == FFTWPP_AMD_FFT_R2C;
    int rc=fftwpp_amd_fft_create(&h->rows,kind,1,&nl,sign,dtype,h->lead,1,1,r2c ? nl : hl,
                                 r2c ? hl : nl);
    if(!rc) rc=fftwpp_amd_fft_create(&h->rowsPad,kind,1,&nl,sign,dtype,h->lead,1,1,
                                     r2c ? 2*hl : hl,r2c ? hl : 2*hl);
    if(rc) { fftwpp_amd_fft_destroy(h); return rc; }
    axes=dims-1;
  }
  for(int a=0; a < axes; ++a) {
    const size_t m=subsize(n[a]);
    h->q[a]=n[a]/m;
    if(h->q[a] > maxPasses) {
      // Bluestein's power of two must itself split into four steps (<= 2^24)
      if(dims > 1 || n[a] > ((size_t) 1 << 23)) {
        fftwpp_amd_fft_destroy(h);
        return error(FFTWPP_AMD_ERR_UNSUPPORTED,"fftwpp_amd_fft_create: no divisor in [n/4096, 4096] (a multidimensional axis, or a 1D length above 2^23)");
      }
      int rc=bluesteinSetup(h);
      if(rc) { fftwpp_amd_fft_destroy(h); return rc; }
      continue;
    }
    int rc=fftwpp_amd_dim_create(&h->d[a],n[a],m,h->q[a],FFTWPP_AMD_KIND_COMPLEX);
    if(rc) { fftwpp_amd_fft_destroy(h); return rc; }
    if(dims == 1 && h->q[0] > 2) {
      h->n1=h->q[0];
      h->n2=m;
      rc=fftwpp_amd_dim_create(&h->f1,h->n1,h->n1,1,FFTWPP_AMD_KIND_COMPLEX);
      if(!rc) rc=fftwpp_amd_dim_create(&h->f2,h->n2,h->n2,1,FFTWPP_AMD_KIND_COMPLEX);
      if(rc) { fftwpp_amd_fft_destroy(h); return rc; }
    }
  }
  *out=h;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_fft_destroy(fftwpp_amd_fft *h)
{
  if(!h) return FFTWPP_AMD_OK;
  for(int a=0; a < 3; ++a)
    if(h->d[a]) fftwpp_amd_dim_destroy(h->d[a]);
  if(h->f1) fftwpp_amd_dim_destroy(h->f1);
  if(h->f2) fftwpp_amd_dim_destroy(h->f2);
  fftwpp_amd_fft_destroy(h->bf);
  fftwpp_amd_fft_destroy(h->bi);
  fftwpp_amd_fft_destroy(h->rows);
  fftwpp_amd_fft_destroy(h->rowsPad);
  for(void *p : {h->chirp,h->bhat,h->bwork})
    if(p) fftwpp_amd_free(p);
  for(void *p : h->tmp)
    if(p) fftwpp_amd_free(p);
  delete h;
  return FFTWPP_AMD_OK;
}

size_t fftwpp_amd_fft_normalization(const fftwpp_amd_fft *h)
{
  return h ? total(h) : 0;
}

int fftwpp_amd_fft_normalize(const fftwpp_amd_fft *h, void *out, void *stream)
{
  if(!h || !out) return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_normalize: null argument");
  if(!fftwpp_amd_is_device_pointer(out))
    return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_normalize: out must be a device pointer");
  const double s=1.0/(double) total(h);
  const int op=h->kind == FFTWPP_AMD_FFT_C2R ? 4 : 0;
  if(h->dims == 1) {
    const size_t len=h->kind == FFTWPP_AMD_FFT_R2C ? h->n[0]/2+1 : h->n[0];
    return fftwpp_amd_internal_fft_map(h->dtype,op,out,h->ostride,h->odist,out,h->ostride,
                                       h->odist,len,h->howmany,s,stream);
  }
  // multidimensional: contiguous complex words, or contiguous real rows
  const size_t words=h->kind == FFTWPP_AMD_FFT_R2C ? h->lead*h->half : total(h);
  return fftwpp_amd_internal_fft_map(h->dtype,op,out,1,words,out,1,words,words,1,s,stream);
}

int fftwpp_amd_fft_execute(fftwpp_amd_fft *h, const void *in, void *out, double scale,
                           void *stream)
{
  if(!h || !in || !out) return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_execute: null argument");
  if(!fftwpp_amd_is_device_pointer(in) || !fftwpp_amd_is_device_pointer(out))
    return error(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_fft_execute: in and out must be device pointers");
  const int dt=h->dtype;
  const size_t n=h->n[0], H=h->howmany;
  const char *x=(const char *) in;
  char *y=(char *) out;
  int rc;
  if(h->rows) {
    // real 2D/3D: rows of the last axis, then the leading axes on lead x half
    // complex words (in place: real rows padded to 2 half words)
    fftwpp_amd_fft *rows=x == y ? h->rowsPad : h->rows;
    const size_t words=h->lead*h->half;
    if((rc=temporaries(h,words))) return rc;
    const int la=h->dims-1;
    auto leading=[&](const char *src, char *last) -> const char * {
      const char *cur=src;
      for(int a=0; a < la; ++a) {
        size_t outer=1, inner=h->half;
        for(int e=0; e < a; ++e) outer *= h->n[e];
        for(int e=a+1; e < la; ++e) inner *= h->n[e];
        char *dst=a == la-1 && last ? last : (char *) h->tmp[(a+1) % 2];
        if(cur == dst) dst=(char *) h->tmp[a % 2];
        if((rc=axis(h,a,h->sign,cur,inner,inner,outer,h->n[a]*inner,dst,inner,
                    h->n[a]*inner,stream))) return NULL;
        cur=dst;
      }
      return cur;
    };
    if(h->kind == FFTWPP_AMD_FFT_R2C) {
      char *t0=(char *) h->tmp[0];
      if((rc=fftwpp_amd_fft_execute(rows,x,t0,1.0,stream))) return rc;
      if(!leading(t0,y)) return rc;
      if(scale != 1.0)
        rc=fftwpp_amd_internal_fft_map(dt,0,y,1,words,y,1,words,words,1,scale,stream);
      return rc;
    }
    const char *cur=leading(x,NULL);
    if(!cur) return rc;
    return fftwpp_amd_fft_execute(rows,cur,y,scale,stream);
  }
  if(h->kind != FFTWPP_AMD_FFT_C2C) {
    // through complex sequences of length n in two temporaries
    if((rc=temporaries(h,n*H))) return rc;
    char *t0=(char *) h->tmp[0], *t1=(char *) h->tmp[1];
    const bool r2c=h->kind == FFTWPP_AMD_FFT_R2C;
    if((rc=fftwpp_amd_internal_fft_map(dt,r2c ? 1 : 3,x,h->istride,h->idist,t0,1,n,n,H,1.0,stream)))
      return rc;
    rc=h->blen ? bluestein(h,t0,1,n,t1,1,n,H,1.0,stream) :
      h->f1 ? fourstep(h,t0,t1,(char *) h->tmp[2],H,stream) :
      axis(h,0,h->sign,t0,1,1,H,n,t1,1,n,stream);
    if(rc) return rc;
    return fftwpp_amd_internal_fft_map(dt,r2c ? 0 : 2,t1,1,n,y,h->ostride,h->odist,
                                       r2c ? n/2+1 : n,H,scale,stream);
  }
  if(h->dims == 1 && h->blen)
    return bluestein(h,x,h->istride,h->idist,y,h->ostride,h->odist,H,scale,stream);
  if(h->dims == 1) {
    const bool alias=x == y;
    if(h->f1 && h->istride == 1 && h->ostride == 1 && h->idist == n && h->odist == n) {
      if((rc=temporaries(h,n*H))) return rc;
      rc=fourstep(h,x,y,(char *) h->tmp[0],H,stream);
    } else if(alias) {
      if((rc=temporaries(h,n*H))) return rc;
      if((rc=fftwpp_amd_internal_fft_map(dt,0,x,h->istride,h->idist,h->tmp[0],1,n,n,H,1.0,stream)))
        return rc;
      rc=axis(h,0,h->sign,(const char *) h->tmp[0],1,1,H,n,y,h->ostride,h->odist,stream);
    } else if(h->idist == 1 && h->odist == 1 && h->istride >= H && h->ostride >= H)
      rc=axis(h,0,h->sign,x,h->istride,H,1,0,y,h->ostride,0,stream);   // interleaved columns
    else
      rc=axis(h,0,h->sign,x,h->istride,1,H,h->idist,y,h->ostride,h->odist,stream);
    if(rc) return rc;
    if(scale != 1.0)
      rc=fftwpp_amd_internal_fft_map(dt,0,y,h->ostride,h->odist,y,h->ostride,h->odist,n,H,scale,stream);
    return rc;
  }
  const size_t all=total(h);
  if((rc=temporaries(h,all))) return rc;
  const char *cur=x;
  for(int a=0; a < h->dims; ++a) {
    size_t outer=1, inner=1;
    for(int e=0; e < a; ++e) outer *= h->n[e];
    for(int e=a+1; e < h->dims; ++e) inner *= h->n[e];
    char *dst=a == h->dims-1 ? y : (char *) h->tmp[a % 2];
    if((rc=axis(h,a,h->sign,cur,inner,inner,outer,h->n[a]*inner,dst,inner,h->n[a]*inner,stream)))
      return rc;
    cur=dst;
  }
  if(scale != 1.0)
    rc=fftwpp_amd_internal_fft_map(dt,0,y,1,all,y,1,all,all,1,scale,stream);
  return rc;
}

} // extern "C"
