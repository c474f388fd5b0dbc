/* fftw++.h -- the transform classes of FFTW++'s fftw++.h (fftw++.h:556-1585:
 * fft1d, mfft1d, fft2d, fft3d, rcfft1d, crfft1d, mrcfft1d, mcrfft1d, rcfft2d,
 * crfft2d, rcfft3d, crfft3d, Transpose) on the
 * MI355X, over the standalone FFT entry points of fftwpp_amd.h.
 *
 * Same constructors and fft / fftNormalized methods; the transforms run on
 * the device.  Arrays may be device pointers (the work is
 * enqueued on the default stream and waited for) or host pointers (staged
 * through device buffers).  A NULL output means in place.  Differences:
 * the `threads` arguments and the in/out arrays passed to constructors are
 * accepted and unused (they chose FFTW plans); fft0 on host arrays leaves
 * the host input unshifted (the reference negates it in place).
 */
#ifndef FFTWPP_AMD_FFTWPP_H
#define FFTWPP_AMD_FFTWPP_H

#include <algorithm>

#include "convolve.h"

namespace fftwpp {

// Shared by the transform classes (the role of class fftw, fftw++.h:119-434).
class fftwPlan {
protected:
  fftwpp_amd_fft *h=NULL;
  size_t inWords=0, outWords=0;   // array extents in elements
  size_t inSize=0, outSize=0;     // element bytes (sizeof(Complex) or sizeof(double))
  // Reals that Normalize scales for a complex-to-real plan whose output rows
  // are padded to 2 (n/2+1) reals (in place), as the reference records at
  // construction (the `doubles` of crfft1d/2d/3d, fftw++.h:957-975,
  // 1318-1331, 1522-1535); 0: the plan's contiguous output.
  size_t paddedReals=0;

  // The reference's in-place rule (Inplace, fftw++.h:88-96): no output
  // array, or the output is the input.
  static bool inplaceArgs(const void *in, const void *out) { return !out || in == out; }
  void padOutput(size_t rows, size_t last) { paddedReals=2*rows*(last/2+1); }

  void create(int kind, int dims, const size_t *n, int sign, size_t howmany,
              size_t istride, size_t ostride, size_t idist, size_t odist) {
    int rc=fftwpp_amd_fft_create(&h,kind,dims,n,sign,FFTWPP_AMD_F64,howmany,istride,
                                 ostride,idist,odist);
    if(rc) throw Error(rc,fftwpp_amd_last_error());
    inSize=kind == FFTWPP_AMD_FFT_R2C ? sizeof(double) : sizeof(Complex);
    outSize=kind == FFTWPP_AMD_FFT_C2R ? sizeof(double) : sizeof(Complex);
    const size_t len=n[0], half=n[0]/2+1;
    size_t inl=kind == FFTWPP_AMD_FFT_C2R ? half : len;
    size_t outl=kind == FFTWPP_AMD_FFT_R2C ? half : len;
    if(dims > 1) {
      // real transforms: lead x (n/2+1) complex words against lead x n reals
      const size_t all=fftwpp_amd_fft_normalization(h), last=n[dims-1];
      const size_t cw=all/last*(last/2+1);
      inWords=kind == FFTWPP_AMD_FFT_C2R ? cw : all;
      outWords=kind == FFTWPP_AMD_FFT_R2C ? cw : all;
      return;
    }
    if(istride == 0) istride=1;
    if(ostride == 0) ostride=1;
    if(idist == 0) idist=inl;
    if(odist == 0) odist=outl;
    if(howmany == 0) howmany=1;
    inWords=(howmany-1)*idist+(inl-1)*istride+1;
    outWords=(howmany-1)*odist+(outl-1)*ostride+1;
  }

  // Origin shift of the real 2D/3D classes (fftw++.h:144-222): side 1
  // negates the input reals before an r2c transform, side 2 the output
  // reals after a c2r one; sdims, sn: the real array's shape.
  int shiftSide=0, sdims=0;
  size_t sn[3]={0,0,0};

  void origin(int side, int dims, const size_t *n) {
    shiftSide=side;
    sdims=dims;
    for(int d=0; d < dims; ++d) sn[d]=n[d];
  }

  int shiftReals(void *data, bool inplace) {
    const size_t last=sn[sdims-1];
    const size_t rowlen=inplace ? 2*(last/2+1) : last;
    return fftwpp_amd_fft_shift(0,(double *) data,sn[0],sdims == 2 ? last : sn[1],
                                sdims == 3 ? last : 0,rowlen,NULL);
  }

  void denyquist(Complex *f) {
    const size_t last=sn[sdims-1];
    const size_t bytes=(sdims == 2 ? sn[0] : sn[0]*sn[1])*(last/2+1)*sizeof(Complex);
    const bool dev=fftwpp_amd_is_device_pointer(f);
    void *d=f;
    int rc=dev ? FFTWPP_AMD_OK : fftwpp_amd_malloc(&d,bytes);
    if(!rc && !dev) rc=fftwpp_amd_memcpy(d,f,bytes,NULL);
    if(!rc) rc=fftwpp_amd_fft_shift(sdims == 2 ? 1 : 2,(double *) d,sn[0],
                                    sdims == 2 ? last : sn[1],sdims == 3 ? last : 0,0,NULL);
    if(!rc && !dev) rc=fftwpp_amd_memcpy(f,d,bytes,NULL);
    if(!rc) rc=fftwpp_amd_synchronize(NULL);
    if(!dev && d) fftwpp_amd_free(d);
    if(rc) throw Error(rc,fftwpp_amd_last_error());
  }

  void normalizeOut(void *out) {
    const bool dev=fftwpp_amd_is_device_pointer(out);
    const size_t bytes=paddedReals ? paddedReals*sizeof(double) : outWords*outSize;
    void *d=out;
    int rc=dev ? FFTWPP_AMD_OK : fftwpp_amd_malloc(&d,bytes);
    if(!rc && !dev) rc=fftwpp_amd_memcpy(d,out,bytes,NULL);
    // padded rows: every real of the padded array (pairs of reals scaled as
    // complex words), as the reference's Normalize(double *) does
    if(!rc) rc=paddedReals ?
      fftwpp_amd_scale(FFTWPP_AMD_F64,d,paddedReals/2,
                       1.0/(double) fftwpp_amd_fft_normalization(h),NULL) :
      fftwpp_amd_fft_normalize(h,d,NULL);
    if(!rc && !dev) rc=fftwpp_amd_memcpy(out,d,bytes,NULL);
    if(!rc) rc=fftwpp_amd_synchronize(NULL);
    if(!dev && d) fftwpp_amd_free(d);
    if(rc) throw Error(rc,fftwpp_amd_last_error());
  }

  void run(const void *in, void *out, double scale, bool shift=false) {
    if(!out) out=(void *) in;
    const bool dev=fftwpp_amd_is_device_pointer(in) && fftwpp_amd_is_device_pointer(out);
    shift=shift && shiftSide;
    int rc=FFTWPP_AMD_OK;
    if(dev) {
      // on device arrays the input shift is made in place, as the reference does
      if(shift && shiftSide == 1) rc=shiftReals((void *) in,in == out);
      if(!rc) rc=fftwpp_amd_fft_execute(h,in,out,scale,NULL);
      if(!rc && shift && shiftSide == 2) rc=shiftReals(out,in == out);
      if(!rc) rc=fftwpp_amd_synchronize(NULL);
    } else {
      // host arrays: one device buffer per distinct array
      const bool inplace=in == out;
      // in place, real rows are padded to the complex row length
      size_t ib=inWords*inSize, ob=outWords*outSize;
      if(inplace) ib=ob=std::max(ib,ob);
      void *din=NULL, *dout=NULL;
      rc=fftwpp_amd_malloc(&din,inplace ? std::max(ib,ob) : ib);
      if(!rc && !inplace) rc=fftwpp_amd_malloc(&dout,ob);
      if(inplace) dout=din;
      if(!rc) rc=fftwpp_amd_memcpy(din,in,ib,NULL);
      // host arrays: the shift applies to the device copy (the host input is
      // left as it was)
      if(!rc && shift && shiftSide == 1) rc=shiftReals(din,inplace);
      if(!rc) rc=fftwpp_amd_fft_execute(h,din,dout,scale,NULL);
      if(!rc && shift && shiftSide == 2) rc=shiftReals(dout,inplace);
      if(!rc) rc=fftwpp_amd_memcpy(out,dout,ob,NULL);
      if(!rc) rc=fftwpp_amd_synchronize(NULL);
      if(din) fftwpp_amd_free(din);
      if(dout && !inplace) fftwpp_amd_free(dout);
    }
    if(rc) throw Error(rc,fftwpp_amd_last_error());
  }

public:
  fftwPlan() {}
  fftwPlan(const fftwPlan &)=delete;
  fftwPlan &operator=(const fftwPlan &)=delete;
  virtual ~fftwPlan() { fftwpp_amd_fft_destroy(h); }

  // out /= normalization() over the output layout (fftw++.h:357-434).
  void Normalize(Complex *out) { normalizeOut(out); }
  void Normalize(double *out) { normalizeOut(out); }

  // The product of the transform lengths (what fftNormalized divides by).
  size_t normalization() const { return fftwpp_amd_fft_normalization(h); }
  double Normalization() const { return 1.0/(double) normalization(); }

  void fft(Complex *in, Complex *out=NULL) { run(in,out,1.0); }
  void fft(double *in, Complex *out=NULL) { run(in,out,1.0); }
  void fft(Complex *in, double *out) { run(in,out,1.0); }
  void fftNormalized(Complex *in, Complex *out=NULL) { run(in,out,Normalization()); }
  void fftNormalized(double *in, Complex *out=NULL) { run(in,out,Normalization()); }
  void fftNormalized(Complex *in, double *out) { run(in,out,Normalization()); }
  // With the Fourier origin shifted to (nx/2, 0) in 2D or (nx/2, ny/2, 0) in
  // 3D for the real 2D/3D classes; the plain transforms for the others.
  void fft0(Complex *in, Complex *out=NULL) { run(in,out,1.0,true); }
  void fft0(double *in, Complex *out=NULL) { run(in,out,1.0,true); }
  void fft0(Complex *in, double *out) { run(in,out,1.0,true); }
  void fft0Normalized(Complex *in, Complex *out=NULL) { run(in,out,Normalization(),true); }
  void fft0Normalized(Complex *in, double *out) { run(in,out,Normalization(),true); }
  void fft0Normalized(double *in, Complex *out) { run(in,out,Normalization(),true); }
};

// fftw++.h:436-497: out[c][r] = in[r][c] for a rows x cols matrix of
// elements of `length` T's (T a multiple of a double in size); in place when
// out is NULL or in (through a device buffer).
class Transpose {
  size_t rows, cols, words;
  bool inplace;
public:
  template<class T>
  Transpose(size_t rows, size_t cols, size_t length, T *in, T *out=NULL,
            size_t =fftw::maxthreads) : rows(rows), cols(cols) {
    static_assert(sizeof(T) % sizeof(double) == 0,"Transpose takes multiples of doubles");
    words=length*(sizeof(T)/sizeof(double));
    if(!out) out=in;
    inplace=out == in;
  }

  template<class T>
  void transpose(T *in, T *out=NULL) {
    if(rows == 0 || cols == 0 || words == 0) return;
    if(!out) out=in;
    if(inplace != (out == in))
      throw Error(FFTWPP_AMD_ERR_ARG,"Transpose: in-place mode differs from construction");
    const size_t bytes=rows*cols*words*sizeof(double);
    const bool dev=fftwpp_amd_is_device_pointer(in) && fftwpp_amd_is_device_pointer(out);
    int rc=FFTWPP_AMD_OK;
    if(dev && !inplace) {
      rc=fftwpp_amd_transpose(rows,cols,words,in,out,NULL);
    } else {
      // a device copy of the input (host arrays, or an in-place transpose)
      void *din=NULL, *dout=NULL;
      rc=fftwpp_amd_malloc(&din,bytes);
      if(!rc) rc=fftwpp_amd_memcpy(din,in,bytes,NULL);
      if(!rc && !dev) rc=fftwpp_amd_malloc(&dout,bytes);
      if(!rc) rc=fftwpp_amd_transpose(rows,cols,words,din,dev ? (void *) out : dout,NULL);
      if(!rc && !dev) rc=fftwpp_amd_memcpy(out,dout,bytes,NULL);
      if(!rc) rc=fftwpp_amd_synchronize(NULL);
      if(din) fftwpp_amd_free(din);
      if(dout) fftwpp_amd_free(dout);
    }
    if(!rc) rc=fftwpp_amd_synchronize(NULL);
    if(rc) throw Error(rc,fftwpp_amd_last_error());
  }
};

// fftw++.h:556-574: one complex transform of length nx.
class fft1d : public fftwPlan {
public:
  fft1d(size_t nx, int sign, Complex * =NULL, Complex * =NULL,
        size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_C2C,1,&nx,sign,1,1,1,0,0);
  }
};

// fftw++.h:827-896: M complex transforms, element j of transform t at
// j*stride + t*dist (defaults: contiguous transforms, dist = nx).
class mfft1d : public fftwPlan {
public:
  mfft1d(size_t nx, int sign, size_t M=1, Complex * =NULL, Complex * =NULL,
         size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_C2C,1,&nx,sign,M,1,1,nx,nx);
  }
  mfft1d(size_t nx, int sign, size_t M, size_t stride, size_t dist=0,
         Complex * =NULL, Complex * =NULL, size_t =fftw::maxthreads) {
    if(dist == 0) dist=nx;
    create(FFTWPP_AMD_FFT_C2C,1,&nx,sign,M,stride,stride,dist,dist);
  }
};

// fftw++.h:1199-1247, 1390-1444: complex 2D and 3D transforms of
// contiguous row-major arrays.
class fft2d : public fftwPlan {
public:
  fft2d(size_t nx, size_t ny, int sign, Complex * =NULL, Complex * =NULL,
        size_t =fftw::maxthreads) {
    const size_t n[2]={nx,ny};
    create(FFTWPP_AMD_FFT_C2C,2,n,sign,1,1,1,0,0);
  }
};

class fft3d : public fftwPlan {
public:
  fft3d(size_t nx, size_t ny, size_t nz, int sign, Complex * =NULL,
        Complex * =NULL, size_t =fftw::maxthreads) {
    const size_t n[3]={nx,ny,nz};
    create(FFTWPP_AMD_FFT_C2C,3,n,sign,1,1,1,0,0);
  }
};

// fftw++.h:917-975: real-to-complex (sign -1, nx/2+1 outputs) and
// complex-to-real (sign +1) transforms.
class rcfft1d : public fftwPlan {
public:
  rcfft1d(size_t nx, Complex * =NULL, size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_R2C,1,&nx,-1,1,1,1,0,0);
  }
  rcfft1d(size_t nx, double *, Complex * =NULL, size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_R2C,1,&nx,-1,1,1,1,0,0);
  }
};

class crfft1d : public fftwPlan {
public:
  crfft1d(size_t nx, double * =NULL, size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_C2R,1,&nx,1,1,1,1,0,0);
    padOutput(1,nx);
  }
  crfft1d(size_t nx, Complex *in, double *out=NULL, size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_C2R,1,&nx,1,1,1,1,0,0);
    if(inplaceArgs(in,out)) padOutput(1,nx);
  }
};

// fftw++.h:1249-1388, 1436-1585: real 2D and 3D transforms of row-major
// arrays, nx x ... x (nz/2+1) complex words against nx x ... x nz reals (in
// place: real rows padded to 2 (nz/2+1) doubles, as FFTW stores them).
class rcfft2d : public fftwPlan {
public:
  rcfft2d(size_t nx, size_t ny, Complex * =NULL, size_t =fftw::maxthreads) {
    const size_t n[2]={nx,ny};
    create(FFTWPP_AMD_FFT_R2C,2,n,-1,1,1,1,0,0);
    origin(1,2,n);
  }
  rcfft2d(size_t nx, size_t ny, double *, Complex * =NULL, size_t =fftw::maxthreads) {
    const size_t n[2]={nx,ny};
    create(FFTWPP_AMD_FFT_R2C,2,n,-1,1,1,1,0,0);
    origin(1,2,n);
  }
  // fftw++.h:1342-1360 (2D), 1466-1495 (3D): zero the Nyquist modes of a
  // shifted spectrum
  void deNyquist(Complex *f) { denyquist(f); }
};

class crfft2d : public fftwPlan {
public:
  crfft2d(size_t nx, size_t ny, double * =NULL, size_t =fftw::maxthreads) {
    const size_t n[2]={nx,ny};
    create(FFTWPP_AMD_FFT_C2R,2,n,1,1,1,1,0,0);
    origin(2,2,n);
    padOutput(nx,ny);
  }
  crfft2d(size_t nx, size_t ny, Complex *in, double *out=NULL,
          size_t =fftw::maxthreads) {
    const size_t n[2]={nx,ny};
    create(FFTWPP_AMD_FFT_C2R,2,n,1,1,1,1,0,0);
    origin(2,2,n);
    if(inplaceArgs(in,out)) padOutput(nx,ny);
  }
  // fftw++.h:1342-1360 (2D), 1466-1495 (3D): zero the Nyquist modes of a
  // shifted spectrum
  void deNyquist(Complex *f) { denyquist(f); }
};

class rcfft3d : public fftwPlan {
public:
  rcfft3d(size_t nx, size_t ny, size_t nz, Complex * =NULL, size_t =fftw::maxthreads) {
    const size_t n[3]={nx,ny,nz};
    create(FFTWPP_AMD_FFT_R2C,3,n,-1,1,1,1,0,0);
    origin(1,3,n);
  }
  rcfft3d(size_t nx, size_t ny, size_t nz, double *, Complex * =NULL,
          size_t =fftw::maxthreads) {
    const size_t n[3]={nx,ny,nz};
    create(FFTWPP_AMD_FFT_R2C,3,n,-1,1,1,1,0,0);
    origin(1,3,n);
  }
  // fftw++.h:1342-1360 (2D), 1466-1495 (3D): zero the Nyquist modes of a
  // shifted spectrum
  void deNyquist(Complex *f) { denyquist(f); }
};

class crfft3d : public fftwPlan {
public:
  crfft3d(size_t nx, size_t ny, size_t nz, double * =NULL, size_t =fftw::maxthreads) {
    const size_t n[3]={nx,ny,nz};
    create(FFTWPP_AMD_FFT_C2R,3,n,1,1,1,1,0,0);
    origin(2,3,n);
    padOutput(nx*ny,nz);
  }
  crfft3d(size_t nx, size_t ny, size_t nz, Complex *in, double *out=NULL,
          size_t =fftw::maxthreads) {
    const size_t n[3]={nx,ny,nz};
    create(FFTWPP_AMD_FFT_C2R,3,n,1,1,1,1,0,0);
    origin(2,3,n);
    if(inplaceArgs(in,out)) padOutput(nx*ny,nz);
  }
  // fftw++.h:1342-1360 (2D), 1466-1495 (3D): zero the Nyquist modes of a
  // shifted spectrum
  void deNyquist(Complex *f) { denyquist(f); }
};

// fftw++.h:1023-1072, 1122-1171: M real-to-complex / complex-to-real
// transforms with input and output strides and distances.
class mrcfft1d : public fftwPlan {
public:
  mrcfft1d(size_t nx, size_t M, size_t istride, size_t ostride, size_t idist,
           size_t odist, double * =NULL, Complex * =NULL,
           size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_R2C,1,&nx,-1,M,istride,ostride,idist,odist);
  }
};

class mcrfft1d : public fftwPlan {
public:
  mcrfft1d(size_t nx, size_t M, size_t istride, size_t ostride, size_t idist,
           size_t odist, Complex * =NULL, double * =NULL,
           size_t =fftw::maxthreads) {
    create(FFTWPP_AMD_FFT_C2R,1,&nx,1,M,istride,ostride,idist,odist);
  }
};

} // namespace fftwpp

#endif
