/* convolve.h -- drop-in C++ API of the MI355X hybrid implicitly dealiased
 * convolution library.  Mirrors the reference's convolve.h: Application,
 * fftBase, fftPad, fftPadCentered, fftPadHermitian, Convolution,
 * Convolution2, Convolution3, the multiplier signature and the Hermitian
 * symmetrisers, with the same constructors, parameter algebra and validity
 * rules.  The arithmetic runs on the GPU through the C ABI of fftwpp_amd.h;
 * this header includes no HIP header.
 *
 * Data pointers handed to convolve()/convolveRaw() may be device pointers
 * (used in place) or host pointers (staged through device buffers owned by
 * the Convolution object).  forward()/backward() take device pointers.
 *
 * Differences from the reference, by design:
 *   - errors throw fftwpp::Error (carrying a status code) instead of exit();
 *   - the parameter optimiser picks (m, D, I) with the GPU heuristic/autotune
 *     cache instead of timing CPU convolutions (convolve.cc:177-470);
 *   - the four library multipliers run fused on the device; a user-written
 *     multiplier is a DEVICE multiplier: it is called once per convolution
 *     with device arrays F[a] holding the whole padded transforms (not once
 *     per residue block), reads their layout from indices->layout and
 *     enqueues its work on indices->stream (A inputs, B outputs, complex
 *     and centred dimensions of contiguous data);
 *   - caller-provided F/W/V work arrays are accepted and not used (the device
 *     keeps its own work buffers).
 */
#ifndef FFTWPP_AMD_CONVOLVE_H
#define FFTWPP_AMD_CONVOLVE_H

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "Complex.h"
#include "fftwpp_amd.h"

namespace fftwpp {

class Error : public std::runtime_error {
public:
  int code;
  Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
};

extern const double twopi;
extern bool showOptTimes;
extern bool showRoutines;

const Complex I(0.0,1.0);

namespace fftw {
extern size_t maxthreads;  // kept for API compatibility (fftw++.cc:17)
}

namespace utils {
inline size_t ceilquotient(size_t a, size_t b) { return (a+b-1)/b; }
size_t ceilpow2(size_t n);
// 32-byte alignment as set by parallel.cc:18 (2*sizeof(Complex)).
inline size_t align(size_t n) { return ceilquotient(n*sizeof(Complex),32)*32/sizeof(Complex); }
Complex *ComplexAlign(size_t size);
Complex **ComplexAlign(size_t n, size_t size);
double *doubleAlign(size_t size);
void deleteAlign(void *p);
}

size_t nextfftsize(size_t m);

class fftBase;

class Indices {
public:
  fftBase *fft;
  size_t *index;
  size_t size, maxsize;
  size_t r;
  size_t offset;
  // Set for a user multiplier, which runs on the device: the layout of the
  // padded arrays it receives and the stream (hipStream_t) for its work.
  const fftwpp_amd_layout *layout;
  void *stream;

  Indices() : fft(NULL), index(NULL), size(0), maxsize(0), r(0), offset(0),
              layout(NULL), stream(NULL) {}
  void copy(Indices *indices, size_t size0);
  ~Indices();
};

typedef void multiplier(Complex **F, size_t n, Indices *indices,
                        size_t threads);

// The library multipliers.  Called directly they act on DEVICE arrays of n
// complex words (realMultBinary: n doubles, as in convolve.cc:60-83).  Passed
// to an Application they select the fused device multiplier.
multiplier multNone, multBinary, realMultBinary, multcorrelation;

class ThreadBase {
protected:
  size_t threads;
public:
  ThreadBase() : threads(fftw::maxthreads) {}
  ThreadBase(size_t threads) : threads(threads) {}
  size_t Threads() { return threads; }
  void Threads(size_t n) { threads=n; }
};

class Application : public ThreadBase {
public:
  size_t A;
  size_t B;
  multiplier *mult;
  bool verbose;
  size_t m;
  size_t D;
  ptrdiff_t I;
  size_t maxthreads;

  void check();

  Application(size_t A, size_t B, multiplier *mult,
              size_t threads=fftw::maxthreads, bool verbose=false,
              size_t m=0, size_t D=0, ptrdiff_t I=-1) :
    ThreadBase(threads), A(A), B(B), mult(mult), verbose(verbose), m(m), D(D),
    I(I), maxthreads(threads) { check(); }

  Application(size_t A, size_t B, multiplier *mult, Application &parent,
              size_t m=0, size_t D=0, ptrdiff_t I=-1) :
    ThreadBase(1), A(A), B(B), mult(mult), verbose(parent.verbose), m(m),
    D(D), I(I), maxthreads(parent.maxthreads) { check(); }
};

// Device staging buffers shared by the convolution drivers.
class DeviceWork {
public:
  std::vector<void *> bufs;
  std::vector<size_t> sizes;
  std::vector<std::vector<size_t>> keys;  // what a slot holds (empty: nothing known), e.g. an item table
  void *get(size_t slot, size_t bytes);
  ~DeviceWork();

};

static const struct centeredConstructor {} noOptimize={};

class fftBase : public ThreadBase {
public:
  size_t L;   // number of unpadded Complex data values
  size_t M;   // minimum (after construction: actual) padded length
  size_t C;   // number of FFTs computed in parallel
  size_t S;   // stride between successive elements
  size_t m;
  size_t p;
  size_t q;
  size_t n;   // number of residues
  size_t R;   // number of residue blocks
  size_t dr;  // r increment
  size_t D;   // number of residues stored in F at a time
  size_t D0;  // remainder
  size_t Cm, Sm;
  size_t l;   // block size of a single FFT
  size_t b;   // total block size, including stride
  bool inplace;
  Application app;
  bool centered;
  bool overwrite;

  // device geometry (twiddles, radix plan) of this padded dimension
  fftwpp_amd_dim *dev;

  static void parameters(size_t L, size_t M, size_t m, bool centered,
                         size_t &p, size_t &n, size_t &q);
  void checkParameters();
  void common();
  void invalid();

  fftBase(size_t L, size_t M, Application &app, size_t C=1, size_t S=0,
          bool centered=false);
  fftBase(size_t L, size_t M, Application &app, size_t C, size_t S,
          size_t m, size_t D, bool inplace, bool centered=false);
  virtual ~fftBase();

  // Return transformed index for residue r at position i (convolve.h:297-326)
  virtual size_t index(size_t r, size_t i);

  size_t normalization() { return M; }
  size_t paddedSize() { return m*q; }

  virtual bool conjugates() { return D > 1 && (p <= 2 || (centered && p % 2 == 0)); }
  virtual size_t residueBlocks() { return conjugates() ? utils::ceilquotient(n,2) : n; }
  size_t Dr() { return conjugates() ? D/2 : D; }
  virtual size_t increment(size_t r) {
    return r > 0 ? dr : (conjugates() ? utils::ceilquotient(D0,2) : D0);
  }
  size_t nloops() {
    size_t count=0;
    for(size_t r=0; r < R; r += increment(r)) ++count;
    return count;
  }
  bool loop2() { return nloops() == 2 && app.A > app.B && !overwrite; }
  virtual size_t inputLength() { return L; }
  virtual size_t wordSize() { return 2; }
  virtual size_t doubles() { return wordSize()*S*inputLength(); }
  virtual size_t outputSize() { return b*D; }
  virtual size_t blocksize(size_t) { return l; }
  virtual size_t noutputs(size_t r) { return blocksize(r)*(r == 0 ? D0 : D); }
  virtual size_t span(size_t r) { return S*noutputs(r); }
  size_t workSizeV() { return nloops() == 1 || loop2() ? 0 : utils::ceilquotient(doubles(),2); }
  virtual size_t workSizeW() { return inplace ? 0 : outputSize(); }
  size_t repad() { return !inplace && L < m; }

  // Device transform kind of this dimension.
  virtual int kind() const { return FFTWPP_AMD_KIND_COMPLEX; }
  // Real input data (fftPadReal): the convolution drivers take double arrays.
  virtual bool realInput() const { return false; }

  // Residue block r of the padded transform of DEVICE data f (stride S,
  // count C; Hermitian: C columns of ceil(L/2) modes) into F, laid out as
  // the reference's routines lay it out (convolve.cc:771-5447, read by
  // tests/hybrid.cc:60-110 and tests/hybridh.cc:60-100):
  //   complex and centred: F[S*k+c] holds the padded transform at
  //     index(r,k), k < noutputs(r), conjugate pairs of D > 1 blocks included;
  //   Hermitian: copy d < (r == 0 ? D0 : D) of the block holds the real
  //     padded transform at index(r, k + noutputs(0) d) in real word C*k+c
  //     of (double *) (F + b*d);
  //   overwrite (in-place) plans: forward(f,F,0) computes every block at
  //     once, block 0 over f, the middle blocks over f + b, the last in F.
  // Runs on the default stream.
  void forward(Complex *f, Complex *F, size_t r=0, Complex *W=NULL);
  // Adjoint of forward over the same layout: the r=0 block writes f, later
  // blocks accumulate (overwrite plans: every block at r = 0).
  void backward(Complex *F, Complex *f, size_t r=0, Complex *W=NULL);
  void pad(Complex *) {}

protected:
  void initDevice();
  // The reference-layout face: entries of block r (padded position, array,
  // word offset), their residues and device tables, built once per block.
  struct FaceBlock {
    bool ready=false;
    std::vector<size_t> res;        // residues of the block, by slot
    void *tab[3]={NULL,NULL,NULL};  // per array: {slot m + s, word} pairs
    size_t ne[3]={0,0,0};
  };
  std::vector<FaceBlock> faceBlocks;
  fftwpp_amd_dim *hdev=NULL;        // centred twin of a Hermitian dimension
  DeviceWork faceWork;
  FaceBlock &faceBlock(size_t r);
};

class fftPad : public fftBase {
public:
  static bool valid(size_t m, size_t p, size_t q, size_t n, size_t D, size_t S);

  // Normal entry point: C FFTs of length L, stride S >= C, padded to >= M.
  fftPad(size_t L, size_t M, Application &app, size_t C=1, size_t S=0,
         bool Explicit=false);
  // Padded to N=m*q >= M >= L with the given m, D and in-place flag.
  fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
         size_t m, size_t D, bool inplace);
  // Centered entry points: defer parameter selection to the subclass.
  fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
         centeredConstructor);
  fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
         size_t m, size_t D, bool inplace, centeredConstructor);

  void init();
};

class fftPadCentered : public fftPad {
public:
  static bool validCentered(size_t m, size_t p, size_t q, size_t n, size_t D,
                            size_t S) {
    return (q == 1 || p % 2 == 0) && fftPad::valid(m,p,q,n,D,S);
  }

  fftPadCentered(size_t L, size_t M, Application &app, size_t C=1,
                 size_t S=0, bool Explicit=false);
  fftPadCentered(size_t L, size_t M, Application &app, size_t C, size_t S,
                 size_t m, size_t D, bool inplace);

  bool conjugates() { return D > 1 && (p == 1 || p % 2 == 0); }
  int kind() const { return FFTWPP_AMD_KIND_CENTERED; }
};

// Real data (fftPadReal, convolve.h:805-980 of the reference): the outer (or
// only) dimension of a real convolution, with the reference's constructors.
// The dimension is planned as a complex fftPad.  When its geometry has a
// power-of-two m in [16, 4096] and L <= 2m, the drivers run the real residue
// transforms (fftwpp_amd_forward_real_many / backward_real_many): N/2 + 1
// kept complex planes, over which the inner dimensions convolve as complex
// data.  Other geometries take the binary convolution of real f and g as the
// imaginary part of the complex self-convolution of z = f + i g, halved
// ((f + i g) * (f + i g) = f*f - g*g + 2 i f*g), with the drivers packing and
// unpacking the real arrays.
class fftPadReal : public fftPad {
public:
  fftPadReal(size_t L, size_t M, Application &app, size_t C=1, size_t S=0,
             bool Explicit=false) : fftPad(L,M,app,C,S,Explicit) {}
  fftPadReal(size_t L, size_t M, Application &app, size_t C, size_t S,
             size_t m, size_t D, bool inplace) : fftPad(L,M,app,C,S,m,D,inplace) {}
  bool realInput() const { return true; }
  size_t wordSize() { return 1; }
};

class fftPadHermitian : public fftBase {
public:
  size_t e;
  size_t B;  // work block size

  static bool validHermitian(size_t m, size_t p, size_t q, size_t n,
                             size_t D, size_t C) {
    return (D == 1 && q == 1) || (D == 2 && p % 2 == 0 && (p == 2 || C == 1));
  }

  fftPadHermitian(size_t L, size_t M, Application &app, size_t C=1,
                  bool Explicit=false);
  fftPadHermitian(size_t L, size_t M, Application &app, size_t C,
                  size_t m, size_t D, bool inplace);

  void init();

  size_t inputLength() { return utils::ceilquotient(L,2); }
  size_t blocksize(size_t) { return m*(q == 1 ? 1 : p/2); }
  size_t noutputs(size_t) { return blocksize(0); }
  size_t span(size_t) { return 2*b; }
  size_t workSizeW() { return inplace ? 0 : B*D; }
  int kind() const { return FFTWPP_AMD_KIND_HERMITIAN; }
};

// Optional per-launch timing of a driver: an event after every launch on the
// launch stream.  stage(i) is the duration of launch i in milliseconds.
class Timeline {
public:
  std::vector<void *> events;
  std::vector<std::string> names;
  size_t used=0;
  void *stream=NULL;
  void begin(void *s);
  void mark(const char *name);
  double stage(size_t i);
  size_t stages() const { return names.size(); }
  ~Timeline();
private:
  void record();
};


// A convolution with a user multiplier on contiguous complex device data of
// nd dimensions (dims outermost first, complex or centred): every input
// f[a], a < A, is expanded to all residues of every dimension, `mult` is
// called once on the padded arrays (layout as fftwpp_amd_layout describes),
// and the B products are transformed back into f[0..B-1], multiplied by
// `scale`.  f holds max(A,B) arrays.
typedef std::function<void(void **F, size_t n, const fftwpp_amd_layout &layout)>
  DeviceMultiplier;
void generalConvolution(fftBase *const *dims, size_t nd, size_t A, size_t B,
                        DeviceWork &work, int dtype, void *const *f,
                        double scale, void *stream, const DeviceMultiplier &mult);

// A caller-supplied inner row stage (fftwpp_amd_conv_execute_fused): when a
// convolution's rowHook is set, its fused innermost-row stage is handed to
// hook(args, user) instead of the library's row kernels.
struct RowHook {
  fftwpp_amd_rows_hook hook;
  void *user;
};

class Convolution : public ThreadBase {
public:
  fftBase *fft;
  size_t L;
  size_t A;
  size_t B;
  multiplier *mult;
  double scale;
  Indices indices;

  // F, W, V: optional work arrays of the reference API (accepted, unused).
  Convolution(fftBase *fft, Complex **F=NULL, Complex *W=NULL, Complex *V=NULL);
  virtual ~Convolution();

  double normalization() { return fft->normalization(); }

  void convolveRaw(Complex **f) { run(f,0,false); }
  void convolveRaw(Complex **f, Indices *) { run(f,0,false); }
  void convolveRaw(Complex **f, size_t offset) { run(f,offset,false); }
  void convolveRaw(Complex **f, size_t offset, Indices *) { run(f,offset,false); }
  void convolveRaw(double **f) { run((Complex **) f,0,false); }
  void convolve(Complex **f) { run(f,0,true); }
  void convolve(Complex **f, size_t offset) { run(f,offset,true); }
  void convolve(double **f) { run((Complex **) f,0,true); }

  // Device-pointer entry (stream = hipStream_t as void*).
  void convolveDevice(Complex *f0, const Complex *f1, bool normalize,
                      void *stream=NULL);
  // out <- f0 (*) f1 for dtype FFTWPP_AMD_F64/F32 on complex128/complex64
  // device arrays (out may equal f0); mult >= 0 overrides the Application's
  // multiplier (the adjoint uses multcorrelation).
  void execDevice(void *out, const void *f0, const void *f1, bool normalize,
                  int dtype, void *stream, int mult=-1);
  size_t arraySize() const;  // complex words per input array

  Timeline *timeline=NULL;  // when set, every launch is timed
  const RowHook *rowHook=NULL;  // when set, the row stage runs through it

protected:
  int multCode;
  DeviceWork work;
  void run(Complex **f, size_t offset, bool normalize);
  // The device convolution of complex arrays, or with real = true of real
  // arrays through the real residue transforms of the outer dimension.
  void execComplex(void *out, const void *f0, const void *f1, bool normalize,
                   int dtype, void *stream, int code, bool real=false);
};

void HermitianSymmetrize(Complex *f);
void HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0, Complex *f,
                          size_t Sx, size_t threads=fftw::maxthreads);
void HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0, Complex *f);
void HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                           size_t y0, Complex *f, size_t Sx, size_t Sy,
                           size_t threads=fftw::maxthreads);
void HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                           size_t y0, Complex *f);

class Convolution2 : public ThreadBase {
public:
  fftBase *fftx, *ffty;
  size_t Lx, Ly;
  size_t Sx;
  size_t A;
  size_t B;
  multiplier *mult;
  double scale;
  Indices indices;

  Convolution2(fftBase *fftx, fftBase *ffty, Complex **F=NULL,
               Complex *W=NULL, Complex *V=NULL);
  virtual ~Convolution2();

  double normalization() { return (double) fftx->normalization()*(double) ffty->normalization(); }

  void convolveRaw(Complex **f, size_t offset=0, Indices * =NULL) { run(f,offset,false); }
  void convolveRaw(double **f, size_t offset=0, Indices * =NULL) { run((Complex **) f,offset,false); }
  void convolve(Complex **f, size_t offset=0) { run(f,offset,true); }
  void convolve(double **f, size_t offset=0) { run((Complex **) f,offset,true); }

  void convolveDevice(Complex *f0, const Complex *f1, bool normalize,
                      void *stream=NULL);
  // out <- f0 (*) f1 for dtype FFTWPP_AMD_F64/F32 on complex128/complex64
  // device arrays (out may equal f0); mult >= 0 overrides the Application's
  // multiplier (the adjoint uses multcorrelation).
  void execDevice(void *out, const void *f0, const void *f1, bool normalize,
                  int dtype, void *stream, int mult=-1);
  size_t arraySize() const;

  Timeline *timeline=NULL;  // when set, every launch is timed
  const RowHook *rowHook=NULL;  // when set, the row stage runs through it

  // Device bytes the plan's work buffers may hold (0: no limit); over it the
  // x residues run in blocks of Dx per pass (the reference's residue loop,
  // convolve.h:1475-1535) with a copy of the input an in-place call
  // overwrites.
  size_t memoryLimit=0;
  // Work bytes with Dx x residues per pass (0: all of them).
  size_t workBytes(int dtype, size_t Dx=0) const;
  // x residues per pass (0: none fits the limit).
  size_t residueBlock(int dtype, bool real) const;

protected:
  int multCode;
  size_t ly;  // stored length of a y row (Ly or ceil(Ly/2))
  DeviceWork work;
  void run(Complex **f, size_t offset, bool normalize);
  // The device convolution of complex arrays, or with real = true of real
  // arrays through the real residue transforms of the outer dimension.
  void execComplex(void *out, const void *f0, const void *f1, bool normalize,
                   int dtype, void *stream, int code, bool real=false);
};

class Convolution3 : public ThreadBase {
public:
  fftBase *fftx, *ffty, *fftz;
  size_t Lx, Ly, Lz;
  size_t Sx, Sy;
  size_t A;
  size_t B;
  multiplier *mult;
  double scale;
  Indices indices;

  Convolution3(fftBase *fftx, fftBase *ffty, fftBase *fftz,
               Complex **F=NULL, Complex *W=NULL, Complex *V=NULL,
               bool mpi=false);
  virtual ~Convolution3();

  double normalization() {
    return (double) fftx->normalization()*(double) ffty->normalization()*
      (double) fftz->normalization();
  }
  bool contiguous() const { return Sy == Lz; }

  void convolveRaw(Complex **f, size_t offset=0, Indices * =NULL) { run(f,offset,false); }
  void convolveRaw(double **f, size_t offset=0, Indices * =NULL) { run((Complex **) f,offset,false); }
  void convolve(Complex **f, size_t offset=0) { run(f,offset,true); }
  void convolve(double **f, size_t offset=0) { run((Complex **) f,offset,true); }

  void convolveDevice(Complex *f0, const Complex *f1, bool normalize,
                      void *stream=NULL);
  // out <- f0 (*) f1 for dtype FFTWPP_AMD_F64/F32 on complex128/complex64
  // device arrays (out may equal f0); mult >= 0 overrides the Application's
  // multiplier (the adjoint uses multcorrelation).
  void execDevice(void *out, const void *f0, const void *f1, bool normalize,
                  int dtype, void *stream, int mult=-1);
  size_t arraySize() const;

  Timeline *timeline=NULL;  // when set, every launch is timed
  const RowHook *rowHook=NULL;  // when set, the row stage runs through it

  // Device bytes the plan's work buffers may hold (0: no limit).  The
  // separate y, z and y launches need 2 + 2 q_y x-transformed arrays; the
  // plane kernel (y and z with m = 256, q = 2, L = 256) needs two plus
  // fftwpp_amd_plane_work(), and is taken when the limit excludes the first;
  // otherwise the x residues run in blocks of Dx per pass (the reference's
  // residue loop, convolve.h:1812-1879), which needs 2 + 2 q_y arrays of Dx
  // residues and a copy of the input an in-place call overwrites.
  size_t memoryLimit=0;
  // Work bytes of a complex execution in dtype: plane engine, or Dx x
  // residues per pass (0: all).
  size_t workBytes(int dtype, bool plane, size_t Dx=0) const;
  // Whether an execution in dtype takes the plane kernel.
  bool usesPlanes(int dtype) const;
  // x residues per pass of the separate stages (0: none fits the limit).
  size_t residueBlock(int dtype, bool real) const;

protected:
  int multCode;
  size_t lz;  // stored length of a z row (Lz or ceil(Lz/2))
  bool planes=false;  // the y-z stage can run as one plane-kernel launch
  DeviceWork work;
  void run(Complex **f, size_t offset, bool normalize);
  // The device convolution of complex arrays, or with real = true of real
  // arrays through the real residue transforms of the outer dimension.
  void execComplex(void *out, const void *f0, const void *f1, bool normalize,
                   int dtype, void *stream, int code, bool real=false);
};

} // namespace fftwpp

#endif
