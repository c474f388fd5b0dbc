// convolve.cc -- host side of the drop-in C++ API (include/convolve.h).
// Pure host C++ (compiled with g++): every device action goes through the
// C ABI of include/fftwpp_amd.h.
#include "convolve.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

// kernel-layer queries (kernels.hip): does a fused row kernel take a row;
// which engine takes the rows (stage 1) of a dimension (2: register FFT).
extern "C" int fftwpp_amd_internal_rows_fit(const fftwpp_amd_dim *d, int dtype);
extern "C" int fftwpp_amd_internal_engine(const fftwpp_amd_dim *d, int stage);
// the item table {i rs, L, i L} of `rows` rows for the batched row kernel
extern "C" int fftwpp_amd_internal_row_items(void *items, size_t rows, size_t rs,
                                             size_t L, void *stream);
// reference-layout blocks: T <-> block array through a table; Hermitian
// half <-> full centred spectrum
extern "C" int fftwpp_amd_internal_block_map(int scatter, const void *T, const void *tab,
                                             size_t ne, size_t C, void *arr, int real,
                                             void *stream);
extern "C" int fftwpp_amd_internal_herm_expand(int op, const void *a, void *b, size_t H,
                                               size_t C, int acc, void *stream);

namespace fftwpp {

const double twopi=2.0*3.14159265358979323846264338327950288;
bool showOptTimes=false;
bool showRoutines=false;

namespace fftw {
size_t maxthreads=1;
}

static void check(int rc, const char *what)
{
  if(rc != FFTWPP_AMD_OK) {
    std::ostringstream s;
    s << what << ": " << fftwpp_amd_last_error();
    throw Error(rc,s.str());
  }
}

namespace utils {

size_t ceilpow2(size_t n)
{
  --n;
  n |= n >> 1; n |= n >> 2; n |= n >> 4; n |= n >> 8; n |= n >> 16; n |= n >> 32;
  return ++n;
}

Complex *ComplexAlign(size_t size)
{
  if(size == 0) return NULL;
  void *p=NULL;
  if(posix_memalign(&p,32,size*sizeof(Complex)) != 0) throw Error(FFTWPP_AMD_ERR_ARG,"out of host memory");
  Complex *v=(Complex *) p;
  for(size_t i=0; i < size; ++i) v[i]=Complex(0.0,0.0);
  return v;
}

Complex **ComplexAlign(size_t n, size_t size)
{
  if(n == 0 || size == 0) return NULL;
  Complex **v=new Complex*[n];
  size_t Size=2*ceilquotient(size,2);
  Complex *B=ComplexAlign((n-1)*Size+size);
  for(size_t i=0; i < n; ++i) v[i]=B+i*Size;
  return v;
}

double *doubleAlign(size_t size)
{
  if(size == 0) return NULL;
  void *p=NULL;
  if(posix_memalign(&p,32,size*sizeof(double)) != 0) throw Error(FFTWPP_AMD_ERR_ARG,"out of host memory");
  return (double *) p;
}

void deleteAlign(void *p) { free(p); }

} // namespace utils

// Smallest N >= m of the form 2^a 3^b 5^c 7^d (convolve.cc:114-124).
size_t nextfftsize(size_t m)
{
  size_t N=utils::ceilpow2(m);
  if(m == N) return m;
  for(size_t ni=1; ni < N; ni *= 7)
    for(size_t nj=ni; nj < N; nj *= 5)
      for(size_t nk=nj; nk < N; nk *= 3)
        N=std::min(N,nk*utils::ceilpow2(utils::ceilquotient(m,nk)));
  return N;
}

void Indices::copy(Indices *indices, size_t size0)
{
  size=indices ? indices->size : size0;
  if(size > maxsize) {
    if(maxsize > 0) delete [] index;
    index=new size_t[size];
    maxsize=size;
  }
  if(indices)
    for(size_t d=1; d < size; ++d) index[d]=indices->index[d];
}

Indices::~Indices()
{
  if(maxsize > 0) delete [] index;
}

// ---- multipliers on device arrays (convolve.cc:26-110)
void multNone(Complex **, size_t, Indices *, size_t) {}

void multBinary(Complex **F, size_t n, Indices *, size_t)
{
  check(fftwpp_amd_multiply(FFTWPP_AMD_F64,FFTWPP_AMD_MULT_BINARY,F[0],F[1],n,NULL),"multBinary");
}

void realMultBinary(Complex **F, size_t n, Indices *, size_t)
{
  // n counts doubles, as in the reference
  check(fftwpp_amd_multiply(FFTWPP_AMD_F64,FFTWPP_AMD_MULT_REAL_BINARY,F[0],F[1],n,NULL),"realMultBinary");
}

void multcorrelation(Complex **F, size_t n, Indices *, size_t)
{
  check(fftwpp_amd_multiply(FFTWPP_AMD_F64,FFTWPP_AMD_MULT_CORRELATION,F[0],F[1],n,NULL),"multcorrelation");
}

static int multiplierCode(multiplier *mult)
{
  if(mult == multNone) return FFTWPP_AMD_MULT_NONE;
  if(mult == multBinary) return FFTWPP_AMD_MULT_BINARY;
  if(mult == realMultBinary) return FFTWPP_AMD_MULT_REAL_BINARY;
  if(mult == multcorrelation) return FFTWPP_AMD_MULT_CORRELATION;
  return -1;
}

void Application::check()
{
  if(m == 1) {
    std::cerr << std::endl << "WARNING: m=1 changed to m=0 to force optimization." << std::endl;
    m=0;
  }
}

// ------------------------------------------------------------------------
// fftBase: parameter algebra (convolve.cc:403-410, 480-509; convolve.h:297-451)
// ------------------------------------------------------------------------
void fftBase::parameters(size_t L, size_t M, size_t m, bool centered,
                         size_t &p, size_t &n, size_t &q)
{
  p=utils::ceilquotient(L,m);
  size_t P=(centered && p == 2*(p/2)) || p == 2 ? (p/2) : p;
  n=utils::ceilquotient(M,P*m);
  q=P*n;
}

void fftBase::checkParameters()
{
  if(L > M) {
    std::ostringstream s;
    s << "L=" << L << " is greater than M=" << M << ".";
    throw Error(FFTWPP_AMD_ERR_ARG,s.str());
  }
  if(S < C) throw Error(FFTWPP_AMD_ERR_ARG,"stride S cannot be less than count C");
}

void fftBase::common()
{
  parameters(L,M,m,centered,p,n,q);
  if(q*m < M) {
    std::ostringstream s;
    s << "Invalid parameters: q=" << q << " m=" << m << " M=" << M;
    throw Error(FFTWPP_AMD_ERR_ARG,s.str());
  }
  Cm=C*m;
  Sm=S*m;
  M=m*q;
  overwrite=false;
}

void fftBase::invalid()
{
  std::ostringstream s;
  s << "Invalid parameters: m=" << m << " p=" << p << " q=" << q << " n=" << n
    << "  D=" << D << " S=" << S;
  throw Error(FFTWPP_AMD_ERR_ARG,s.str());
}

fftBase::fftBase(size_t L, size_t M, Application &app, size_t C, size_t S,
                 bool centered) :
  ThreadBase(app.Threads()), L(L), M(M), C(C), S(S == 0 ? C : S), m(0), p(0),
  q(0), n(0), R(0), dr(0), D(0), D0(0), Cm(0), Sm(0), l(0), b(0),
  inplace(false), app(app), centered(centered), overwrite(false), dev(NULL)
{
  checkParameters();
}

fftBase::fftBase(size_t L, size_t M, Application &app, size_t C, size_t S,
                 size_t m, size_t D, bool inplace, bool centered) :
  ThreadBase(app.Threads()), L(L), M(M), C(C), S(S == 0 ? C : S), m(m), p(0),
  q(0), n(0), R(0), dr(0), D(D), D0(0), Cm(0), Sm(0), l(0), b(0),
  inplace(inplace), app(app), centered(centered), overwrite(false), dev(NULL)
{
  checkParameters();
  this->app.D=D;
}

fftBase::~fftBase()
{
  fftwpp_amd_dim_destroy(dev);
  fftwpp_amd_dim_destroy(hdev);
  for(FaceBlock &fb : faceBlocks)
    for(void *t : fb.tab) fftwpp_amd_free(t);
}

size_t fftBase::index(size_t r, size_t i)
{
  if(q == 1) return i;
  size_t P=utils::ceilquotient(p,2);
  if(overwrite && centered && r == 2) {
    if(i < m) i=(i-1) % m;
    if(p > 2) i=(i+m) % (P*m);
  }
  size_t s=i % m;
  size_t u;
  if(D > 1 && ((centered && p % 2 == 0) || p <= 2)) {
    u=(i/m) % P;
    size_t offset=r == 0 && i >= P*m && D0 % 2 == 1 ? 1 : 0;
    size_t incr=(i+P*m*offset)/(2*P*m);
    r += incr;
    if(i/(P*m)-2*incr+offset == 1) {
      if((!centered && p == 2) || (r > 0 && u == 0))
        s=s > 0 ? s-1 : m-1;
      if(r == 0)
        r=n/2;
      else {
        r=n-r;
        u=u > 0 ? u-1 : P-1;
      }
    }
  } else {
    u=(i/m) % p;
    r += i/(p*m);
  }
  return q*s+n*u+r;
}

void fftBase::initDevice()
{
  fftwpp_amd_dim_destroy(dev);
  dev=NULL;
  check(fftwpp_amd_dim_create(&dev,L,m,q,kind()),"device twiddles");
}

// Reference-layout block r: entries (padded position, array, word offset)
// in the order the reference's routines write them (see forward()).
// Arrays: 0 the block array F, 1 the input f (overwrite: block 0), 2 f + b
// (overwrite: the middle blocks).
fftBase::FaceBlock &fftBase::faceBlock(size_t r)
{
  if(r >= faceBlocks.size()) faceBlocks.resize(r+1);
  FaceBlock &fb=faceBlocks[r];
  if(fb.ready) return fb;
  struct Entry { size_t pos; int arr; size_t word; };
  std::vector<Entry> e;
  if(kind() == FFTWPP_AMD_KIND_HERMITIAN) {
    const size_t stride=noutputs(0), D1=r == 0 ? D0 : D;
    for(size_t d=0; d < D1; ++d)
      for(size_t k=0; k < stride; ++k)
        e.push_back({index(r,k+stride*d),0,2*b*d+C*k});
  } else if(overwrite) {
    for(size_t rr=0; rr < R; rr += increment(rr))
      for(size_t k=0; k < noutputs(rr); ++k)
        e.push_back({index(rr,k),rr == n-1 ? 0 : (rr == 0 ? 1 : 2),S*k});
  } else {
    for(size_t k=0; k < noutputs(r); ++k)
      e.push_back({index(r,k),0,S*k});
  }
  std::vector<uint64_t> h[3];
  for(const Entry &x : e) {
    if(x.pos >= m*q) throw Error(FFTWPP_AMD_ERR_ARG,"padded index beyond m q");
    const size_t rr=x.pos % q, s=x.pos/q;
    size_t slot=0;
    while(slot < fb.res.size() && fb.res[slot] != rr) ++slot;
    if(slot == fb.res.size()) fb.res.push_back(rr);
    h[x.arr].push_back(slot*m+s);
    h[x.arr].push_back(x.word);
  }
  for(int a=0; a < 3; ++a) {
    fb.ne[a]=h[a].size()/2;
    if(fb.ne[a] == 0) continue;
    check(fftwpp_amd_malloc(&fb.tab[a],h[a].size()*sizeof(uint64_t)),"block table");
    check(fftwpp_amd_memcpy(fb.tab[a],h[a].data(),h[a].size()*sizeof(uint64_t),NULL),"block table");
  }
  check(fftwpp_amd_synchronize(NULL),"block table");
  fb.ready=true;
  return fb;
}

void fftBase::forward(Complex *f, Complex *F, size_t r, Complex *)
{
  if(!fftwpp_amd_is_device_pointer(f) || !fftwpp_amd_is_device_pointer(F))
    throw Error(FFTWPP_AMD_ERR_ARG,"forward() takes device pointers");
  if(overwrite && r != 0)
    throw Error(FFTWPP_AMD_ERR_ARG,"an in-place (overwrite) plan computes every block at r = 0");
  if(r >= R) throw Error(FFTWPP_AMD_ERR_ARG,"residue block beyond R");
  FaceBlock &fb=faceBlock(r);
  const bool herm=kind() == FFTWPP_AMD_KIND_HERMITIAN;
  const size_t nres=fb.res.size();
  Complex *T=(Complex *) faceWork.get(0,nres*m*C*sizeof(Complex));
  const fftwpp_amd_dim *dim=dev;
  const void *src=f;
  size_t Sin=S;
  if(herm) {
    // the full centred spectrum of 2H - 1 modes on the centred twin
    const size_t H=inputLength();
    if(!hdev) check(fftwpp_amd_dim_create(&hdev,2*H-1,m,q,FFTWPP_AMD_KIND_CENTERED),"Hermitian twin");
    void *full=faceWork.get(1,(2*H-1)*C*sizeof(Complex));
    check(fftwpp_amd_internal_herm_expand(0,f,full,H,C,0,NULL),"Hermitian expand");
    dim=hdev;
    src=full;
    Sin=C;
  }
  for(size_t slot=0; slot < nres; ++slot)
    check(fftwpp_amd_forward_many(dim,FFTWPP_AMD_F64,src,Sin,C,1,0,T+slot*m*C,C,0,
                                  fb.res[slot],1,NULL),"forward");
  Complex *base[3]={F,f,f+b};
  for(int a=0; a < 3; ++a)
    if(fb.ne[a])
      check(fftwpp_amd_internal_block_map(0,T,fb.tab[a],fb.ne[a],C,base[a],herm,NULL),
            "forward layout");
}

void fftBase::backward(Complex *F, Complex *f, size_t r, Complex *)
{
  if(!fftwpp_amd_is_device_pointer(f) || !fftwpp_amd_is_device_pointer(F))
    throw Error(FFTWPP_AMD_ERR_ARG,"backward() takes device pointers");
  if(overwrite && r != 0)
    throw Error(FFTWPP_AMD_ERR_ARG,"an in-place (overwrite) plan computes every block at r = 0");
  if(r >= R) throw Error(FFTWPP_AMD_ERR_ARG,"residue block beyond R");
  FaceBlock &fb=faceBlock(r);
  const bool herm=kind() == FFTWPP_AMD_KIND_HERMITIAN;
  const size_t nres=fb.res.size();
  const size_t tb=nres*m*C*sizeof(Complex);
  Complex *T=(Complex *) faceWork.get(0,tb);
  check(fftwpp_amd_memset(T,tb,NULL),"backward clear");
  Complex *base[3]={F,f,f+b};
  for(int a=0; a < 3; ++a)
    if(fb.ne[a])
      check(fftwpp_amd_internal_block_map(1,T,fb.tab[a],fb.ne[a],C,base[a],herm,NULL),
            "backward layout");
  if(herm) {
    const size_t H=inputLength();
    if(!hdev) check(fftwpp_amd_dim_create(&hdev,2*H-1,m,q,FFTWPP_AMD_KIND_CENTERED),"Hermitian twin");
    void *full=faceWork.get(2,(2*H-1)*C*sizeof(Complex));
    for(size_t slot=0; slot < nres; ++slot)
      check(fftwpp_amd_backward_many(hdev,FFTWPP_AMD_F64,T+slot*m*C,C,0,full,C,C,1,0,
                                     fb.res[slot],1,slot > 0,1.0,NULL),"backward");
    check(fftwpp_amd_internal_herm_expand(1,full,f,H,C,r > 0,NULL),"Hermitian fold");
    return;
  }
  for(size_t slot=0; slot < nres; ++slot)
    check(fftwpp_amd_backward_many(dev,FFTWPP_AMD_F64,T+slot*m*C,C,0,f,S,C,1,0,
                                   fb.res[slot],1,r > 0 || slot > 0,1.0,NULL),"backward");
}

// ------------------------------------------------------------------------
// GPU parameter choice (replaces the CPU timing optimiser, convolve.cc:177-470)
// ------------------------------------------------------------------------
struct Choice { size_t m, D; bool inplace; };

static size_t maxm() { return fftwpp_amd_max_m(FFTWPP_AMD_F64); }

// kind: 0 complex, 1 centred, 2 Hermitian
static Choice choose(size_t L, size_t M, Application &app, size_t C, size_t S,
                     bool Explicit, int kind)
{
  Choice c;
  c.inplace=app.I >= 0 ? app.I != 0 : false;
  bool centered=kind != 0;
  if(Explicit) {
    c.m=(app.m >= 1 && app.m >= M) ? app.m : nextfftsize(M);
    c.D=1;
    return c;
  }
  if(app.m >= 1) {
    c.m=app.m;
    size_t p, n, q;
    fftBase::parameters(L,M,c.m,centered,p,n,q);
    if(app.D > 0) c.D=app.D;
    else c.D=(kind == 2 && q > 1) ? 2 : 1;
    return c;
  }
  size_t best_m=0, best_N=0;
  size_t lim=maxm();
  // candidate sub-FFT sizes: every split of L into p blocks of smooth m
  for(size_t pp=1; pp <= 64; ++pp) {
    size_t m=nextfftsize(utils::ceilquotient(L,pp));
    if(m > lim) continue;
    size_t p, n, q;
    fftBase::parameters(L,M,m,centered,p,n,q);
    bool ok;
    if(kind == 0) ok=fftPad::valid(m,p,q,n,1,S);
    else if(kind == 1) ok=fftPadCentered::validCentered(m,p,q,n,1,S);
    else ok=fftPadHermitian::validHermitian(m,p,q,n,q > 1 ? 2 : 1,C);
    if(!ok) continue;
    size_t N=m*q;
    // prefer the smallest padded length, then the larger sub-FFT
    if(best_m == 0 || N < best_N || (N == best_N && m > best_m)) {
      best_m=m;
      best_N=N;
    }
  }
  // The register-resident kernels take power-of-two m up to 4096 (rows when
  // one block of m covers the row: L <= m, or H <= m for Hermitian data);
  // they run several times faster than the LDS engine, so when the smallest
  // padded length needs another m, or a power of two that splits the data
  // into p > 1 blocks (a length of 67 padded to 136 with m = 4, p = 17), a
  // covering power of two is preferred while its padded length stays within
  // twice the smallest one.
  const size_t cover=kind == 2 ? (L+1)/2 : L;
  if(best_m != 0 && ((best_m & (best_m-1)) != 0 || best_m < cover)) {
    size_t m=utils::ceilpow2(std::max<size_t>(cover,16));
    if(m <= 4096) {
      size_t p, n, q;
      fftBase::parameters(L,M,m,centered,p,n,q);
      bool ok;
      if(kind == 0) ok=fftPad::valid(m,p,q,n,1,S);
      else if(kind == 1) ok=fftPadCentered::validCentered(m,p,q,n,1,S);
      else ok=fftPadHermitian::validHermitian(m,p,q,n,q > 1 ? 2 : 1,C);
      if(ok && m*q <= 2*best_N) {
        best_m=m;
        best_N=m*q;
      }
    }
  }
  // Complex rows (C = 1) longer than 2048 take m = 4096: the fused register
  // row kernel covers L <= 4096, and longer rows run on the batched register
  // kernel with one block per residue (the LDS kernels for other m run a
  // long row in one block, or unfused).
  if(C == 1 && kind == 0 && L > 2048) {
    size_t p, n, q;
    fftBase::parameters(L,M,4096,centered,p,n,q);
    if(fftPad::valid(4096,p,q,n,1,S)) {
      best_m=4096;
      best_N=4096*q;
    }
  }
  // Complex and centred rows (C = 1) run fused with their residues in
  // sequence.  At m = 1024 the register row kernel is faster at m = 2048 when
  // that lowers q, although the padded length may grow: measured 10% for a
  // 1024-point row with M = 2048 (q 2 -> 1).  (Hermitian rows of length 2047
  // gain too at q 3 -> 2, but that needs p = 1 with conjugate pairs, which
  // the reference's validity rules exclude, convolve.h:721-723.)
  if(C == 1 && kind != 2 && best_m == 1024 && L <= 2048) {
    size_t p, n, q, p2, n2, q2;
    fftBase::parameters(L,M,best_m,centered,p,n,q);
    fftBase::parameters(L,M,2048,centered,p2,n2,q2);
    bool ok=kind == 0 ? fftPad::valid(2048,p2,q2,n2,1,S) :
      fftPadCentered::validCentered(2048,p2,q2,n2,1,S);
    if(ok && q2 < q) {
      best_m=2048;
      best_N=2048*q2;
    }
  }
  // Hermitian rows (C = 1) longer than 2 x 4096 run on the long Hermitian row
  // kernel (residue pairs in parallel), which takes power-of-two m and any p:
  // m = 4096.  (The reference's D = 2 code paths need even p,
  // convolve.h:721-723; the pairs (r, r + 1) of this kernel do not.)
  if(C == 1 && kind == 2 && (L+1)/2 > 4096) {
    size_t p, n, q;
    fftBase::parameters(L,M,4096,centered,p,n,q);
    best_m=4096;
    best_N=4096*q;
  }
  if(best_m == 0) {
    // explicit padding as the last resort (convolve.cc:431-434)
    c.m=nextfftsize(M);
    c.D=1;
    return c;
  }
  c.m=best_m;
  size_t p, n, q;
  fftBase::parameters(L,M,c.m,centered,p,n,q);
  c.D=(kind == 2 && q > 1) ? 2 : 1;
  return c;
}

static void report(fftBase *fft, const char *name)
{
  if(!fft->app.verbose) return;
  size_t mpL=fft->m*fft->p-fft->L;
  std::cout << std::endl << "Optimal padding: ";
  if(fft->p == fft->q) std::cout << "Explicit" << std::endl;
  else if(mpL > 0) std::cout << "Hybrid" << std::endl;
  else std::cout << "Implicit" << std::endl;
  std::cout << "m=" << fft->m << std::endl << "p=" << fft->p << std::endl
            << "q=" << fft->q << std::endl << "C=" << fft->C << std::endl
            << "S=" << fft->S << std::endl << "D=" << fft->D << std::endl
            << "I=" << fft->inplace << std::endl << "Padding: " << mpL
            << std::endl;
  if(showRoutines)
    std::cout << "Forwards Routine: " << name << "::device" << std::endl
              << "Backwards Routine: " << name << "::device" << std::endl;
}

// ------------------------------------------------------------------------
// fftPad / fftPadCentered / fftPadHermitian
// ------------------------------------------------------------------------
bool fftPad::valid(size_t, size_t, size_t q, size_t n, size_t D, size_t S)
{
  if(q == 1) return D == 1;
  return D == 1 || (S == 1 && ((D < n && D % 2 == 0) || D == n));
}

fftPad::fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
               bool Explicit) : fftBase(L,M,app,C,S)
{
  Choice c=choose(L,M,app,C,this->S,Explicit,0);
  m=c.m; D=c.D; inplace=c.inplace;
  if(Explicit) this->M=m;
  parameters(L,this->M,m,centered,p,n,q);
  init();
}

fftPad::fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
               size_t m, size_t D, bool inplace) :
  fftBase(L,M,app,C,S,m,D,inplace)
{
  parameters(L,M,m,centered,p,n,q);
  if(q > 1 && !valid(m,p,q,n,D,this->S)) invalid();
  init();
}

fftPad::fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
               centeredConstructor) : fftBase(L,M,app,C,S,true) {}

fftPad::fftPad(size_t L, size_t M, Application &app, size_t C, size_t S,
               size_t m, size_t D, bool inplace, centeredConstructor) :
  fftBase(L,M,app,C,S,m,D,inplace,true) {}

// Bookkeeping of fftPad::init (convolve.cc:511-719).
void fftPad::init()
{
  common();
  if(q == 1) {
    dr=D=D0=R=1;
    l=M;
    b=S*l;
  } else {
    size_t P;
    if(p == 2) P=1;
    else if(centered) P=p/2;
    else P=p;
    l=m*P;
    b=S*l;
    overwrite=inplace && L == p*m && n == (centered ? 3 : p+1) && D == 1 &&
      app.A >= app.B;
    if(!centered && p > 1) overwrite=false;
    dr=Dr();
    R=residueBlocks();
    D0=n % D;
    if(D0 == 0) D0=D;
  }
  initDevice();
  report(this,centered ? "fftPadCentered" : "fftPad");
}

fftPadCentered::fftPadCentered(size_t L, size_t M, Application &app,
                               size_t C, size_t S, bool Explicit) :
  fftPad(L,M,app,C,S,noOptimize)
{
  Choice c=choose(L,M,app,C,this->S,Explicit,1);
  m=c.m; D=c.D; inplace=c.inplace;
  if(Explicit) this->M=m;
  parameters(L,this->M,m,centered,p,n,q);
  init();
}

fftPadCentered::fftPadCentered(size_t L, size_t M, Application &app,
                               size_t C, size_t S, size_t m, size_t D,
                               bool inplace) :
  fftPad(L,M,app,C,S,m,D,inplace,noOptimize)
{
  parameters(L,M,m,centered,p,n,q);
  if(q > 1 && !validCentered(m,p,q,n,D,this->S)) invalid();
  init();
}

fftPadHermitian::fftPadHermitian(size_t L, size_t M, Application &app,
                                 size_t C, bool Explicit) :
  fftBase(L,M,app,C,C,true), e(0), B(0)
{
  Choice c=choose(L,M,app,C,C,Explicit,2);
  m=c.m; D=c.D; inplace=c.inplace;
  if(Explicit) this->M=m;
  parameters(L,this->M,m,centered,p,n,q);
  init();
}

fftPadHermitian::fftPadHermitian(size_t L, size_t M, Application &app,
                                 size_t C, size_t m, size_t D, bool inplace) :
  fftBase(L,M,app,C,C,m,D,inplace,true), e(0), B(0)
{
  parameters(L,M,m,centered,p,n,q);
  if(q > 1 && !validHermitian(m,p,q,n,D,C)) invalid();
  init();
}

// Bookkeeping of fftPadHermitian::init (convolve.cc:4309-4421).
void fftPadHermitian::init()
{
  common();
  S=C;
  e=m/2+1;
  if(q == 1) {
    B=b=C*e;
    dr=D0=R=1;
  } else {
    dr=Dr();
    size_t p2=p/2;
    b=utils::align(utils::ceilquotient(p2*Cm,2));
    B=utils::align(p2*C*e);
    if(inplace) b=B;
    R=residueBlocks();
    D0=n % D;
    if(D0 == 0) D0=D;
  }
  l=blocksize(0);
  initDevice();
  report(this,"fftPadHermitian");
}

// ------------------------------------------------------------------------
// Timeline
// ------------------------------------------------------------------------
void Timeline::record()
{
  if(used == events.size()) {
    void *e=NULL;
    check(fftwpp_amd_event_create(&e),"event");
    events.push_back(e);
  }
  check(fftwpp_amd_event_record(events[used++],stream),"event record");
}

void Timeline::begin(void *s)
{
  stream=s;
  used=0;
  names.clear();
  record();
}

void Timeline::mark(const char *name)
{
  names.push_back(name);
  record();
}

double Timeline::stage(size_t i)
{
  float ms=0.0f;
  check(fftwpp_amd_event_elapsed(&ms,events[i],events[i+1]),"event elapsed");
  return ms;
}

Timeline::~Timeline()
{
  for(void *e : events) fftwpp_amd_event_destroy(e);
}

// ------------------------------------------------------------------------
// Device work buffers
// ------------------------------------------------------------------------
void *DeviceWork::get(size_t slot, size_t bytes)
{
  if(slot >= bufs.size()) {
    bufs.resize(slot+1,NULL);
    sizes.resize(slot+1,0);
  }
  if(slot >= keys.size()) keys.resize(slot+1);
  if(sizes[slot] < bytes) {
    fftwpp_amd_free(bufs[slot]);
    bufs[slot]=NULL;
    sizes[slot]=0;
    keys[slot].clear();
    check(fftwpp_amd_malloc(&bufs[slot],bytes),"device work buffer");
    sizes[slot]=bytes;
  }
  return bufs[slot];
}

DeviceWork::~DeviceWork()
{
  for(void *p : bufs) fftwpp_amd_free(p);
}

static size_t eltsize(int dtype) { return dtype == FFTWPP_AMD_F32 ? 8 : 16; }

// Stage host arrays through device buffers; run `exec` on device pointers.
// Elements are esize bytes: complex words, or real words for real data
// (the arrays of fftPadReal convolutions are double arrays).
template<class Exec>
static void staged(DeviceWork &work, Complex **f, size_t offset, size_t n,
                   size_t nin, size_t esize, Exec exec)
{
  char *a=(char *) f[0]+offset*esize, *b=(char *) f[1]+offset*esize;
  bool dev0=fftwpp_amd_is_device_pointer(a);
  bool dev1=fftwpp_amd_is_device_pointer(b);
  if(dev0 && dev1) {
    exec(a,b);
    check(fftwpp_amd_synchronize(NULL),"synchronize");
    return;
  }
  size_t bytes=n*esize;
  void *d0=work.get(100,bytes);
  void *d1=work.get(101,bytes);
  check(fftwpp_amd_memcpy(d0,a,bytes,NULL),"copy in");
  if(nin > 1) check(fftwpp_amd_memcpy(d1,b,bytes,NULL),"copy in");
  exec(d0,d1);
  check(fftwpp_amd_memcpy(a,d0,bytes,NULL),"copy out");
  check(fftwpp_amd_synchronize(NULL),"synchronize");
}

// A binary convolution of real arrays f0, f1 (n real words each) into out:
// z = f0 + i 2^e f1 in work slot 6 with the norm-balancing exponent e (work
// slot 7), the complex self-convolution `self` of z in place, out =
// 2^-e Im z / 2.
template<class Self>
static void realConvolution(DeviceWork &work, size_t n, int dtype, int code,
                            void *out, const void *f0, const void *f1,
                            void *stream, Self self)
{
  if(code != FFTWPP_AMD_MULT_BINARY)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"real convolutions take multBinary");
  void *z=work.get(6,n*(dtype == FFTWPP_AMD_F32 ? 8 : 16));
  void *bal=work.get(7,fftwpp_amd_real_balance_bytes());
  check(fftwpp_amd_real_balance(dtype,f0,f1,n,bal,stream),"real balance");
  check(fftwpp_amd_real_pack(dtype,f0,f1,z,n,bal,stream),"real pack");
  self(z);
  check(fftwpp_amd_real_unpack(dtype,z,out,n,0.5,bal,stream),"real unpack");
}

// Whether the real residue transforms take the outer dimension of a real
// convolution (else the f + i g packing of realConvolution runs).
static bool realPlanes(fftBase *fft)
{
  return fft->realInput() && fftwpp_amd_real_planes(fft->dev) > 0;
}

static void requireBinary(size_t A, size_t B, int code)
{
  if(code < 0)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a user multiplier runs through convolve() or fftwpp_amd_conv_execute_general");
  if(A != 2 || B != 1)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"only binary convolutions (A=2, B=1) run on the device");
}

// ------------------------------------------------------------------------
// User multipliers (the multiplier/Indices contract, convolve.h:48-82 and
// Convolution::operate, convolve.h:1120-1135): every dimension expanded to
// all its residues, the multiplier called once on the padded arrays.
// ------------------------------------------------------------------------
// forward_many/backward_many take one grid row per batch entry: batches of
// several columns are split to stay within the grid's y extent (single
// columns of many entries run as the columns of one launch).
static const size_t maxBatch=65535;

void generalConvolution(fftBase *const *dims, size_t nd, size_t A, size_t B,
                        DeviceWork &work, int dtype, void *const *f,
                        double scale, void *stream, const DeviceMultiplier &mult)
{
  if(nd < 1 || nd > 3) throw Error(FFTWPP_AMD_ERR_ARG,"one to three dimensions");
  if(A < 1 || B < 1 || A > 8 || B > 8) throw Error(FFTWPP_AMD_ERR_ARG,"1 <= A, B <= 8 arrays");
  size_t Ld[3], Nd[3];
  size_t total=1;
  for(size_t d=0; d < nd; ++d) {
    if(dims[d]->kind() == FFTWPP_AMD_KIND_HERMITIAN || dims[d]->realInput())
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"user multipliers take complex or centred dimensions");
    Ld[d]=dims[d]->L;
    Nd[d]=dims[d]->m*dims[d]->q;
    total *= Nd[d];
  }
  // Stage d turns outer(d) x L_d x inner(d) words into outer(d) x N_d x
  // inner(d): outer(d) = N_0...N_{d-1}, inner(d) = L_{d+1}...L_{nd-1}.
  auto outer=[&](size_t d) { size_t w=1; for(size_t e=0; e < d; ++e) w *= Nd[e]; return w; };
  auto inner=[&](size_t d) { size_t w=1; for(size_t e=d+1; e < nd; ++e) w *= Ld[e]; return w; };
  const size_t E=eltsize(dtype);
  const size_t na=std::max(A,B);
  std::vector<void *> F(na);
  for(size_t a=0; a < na; ++a) F[a]=work.get(20+a,total*E);
  // P[d]: the output of forward stage d < nd-1 (and input of backward d+1)
  char *P[2]={NULL,NULL};
  for(size_t d=0; d+1 < nd; ++d)
    P[d]=(char *) work.get(18+d,outer(d+1)*inner(d)*E);

  for(size_t a=0; a < A; ++a)
    for(size_t d=0; d < nd; ++d) {
      const size_t out=outer(d), in=inner(d);
      const char *src=d == 0 ? (const char *) f[a] : P[d-1];
      char *dst=d+1 == nd ? (char *) F[a] : P[d];
      const size_t step=in == 1 ? out : maxBatch;
      for(size_t b0=0; b0 < out; b0 += step)
        check(fftwpp_amd_forward_many(dims[d]->dev,dtype,src+b0*Ld[d]*in*E,in,in,
                                      std::min(step,out-b0),Ld[d]*in,
                                      dst+b0*Nd[d]*in*E,in,Nd[d]*in,0,dims[d]->q,
                                      stream),"user multiplier: forward");
    }

  fftwpp_amd_layout layout;
  memset(&layout,0,sizeof(layout));
  layout.dims=(int) nd;
  layout.dtype=dtype;
  layout.A=A;
  layout.B=B;
  for(size_t d=0; d < nd; ++d) {
    layout.N[d]=Nd[d];
    layout.m[d]=dims[d]->m;
    layout.q[d]=dims[d]->q;
    size_t st=1;
    for(size_t e=d+1; e < nd; ++e) st *= Nd[e];
    layout.stride[d]=st;
  }
  mult(F.data(),total,layout);

  for(size_t b=0; b < B; ++b)
    for(size_t d=nd; d-- > 0;) {
      const size_t out=outer(d), in=inner(d);
      const char *src=d+1 == nd ? (const char *) F[b] : P[d];
      char *dst=d == 0 ? (char *) f[b] : P[d-1];
      const size_t step=in == 1 ? out : maxBatch;
      for(size_t b0=0; b0 < out; b0 += step)
        check(fftwpp_amd_backward_many(dims[d]->dev,dtype,src+b0*Nd[d]*in*E,in,Nd[d]*in,
                                       dst+b0*Ld[d]*in*E,in,in,std::min(step,out-b0),
                                       Ld[d]*in,0,dims[d]->q,0,d == 0 ? scale : 1.0,
                                       stream),"user multiplier: backward");
    }
}

// The C++ face of a user multiplier: f[0..max(A,B)-1] host or device arrays
// of `words` complex words from `offset`; host arrays are staged through
// device buffers.  The multiplier sees the padded device arrays and finds
// their layout and stream in `indices`.
// generalConvolution on f[0..max(A,B)-1], host or device arrays of `words`
// complex words from `offset`; host arrays are staged through device buffers.
static void stagedGeneral(fftBase *const *dims, size_t nd, size_t A, size_t B,
                          DeviceWork &work, Complex **f, size_t offset, size_t words,
                          double scale, const DeviceMultiplier &mult)
{
  const size_t na=std::max(A,B);
  const size_t bytes=words*sizeof(Complex);
  std::vector<void *> d(na);
  std::vector<bool> host(na);
  for(size_t a=0; a < na; ++a) {
    Complex *p=f[a]+offset;
    host[a]=!fftwpp_amd_is_device_pointer(p);
    d[a]=host[a] ? work.get(40+a,bytes) : (void *) p;
    if(host[a] && a < A)
      check(fftwpp_amd_memcpy(d[a],p,bytes,NULL),"copy in");
  }
  generalConvolution(dims,nd,A,B,work,FFTWPP_AMD_F64,d.data(),scale,NULL,mult);
  for(size_t b=0; b < B; ++b)
    if(host[b])
      check(fftwpp_amd_memcpy(f[b]+offset,d[b],bytes,NULL),"copy out");
  check(fftwpp_amd_synchronize(NULL),"synchronize");
}

static void userConvolution(fftBase *const *dims, size_t nd, size_t A, size_t B,
                            multiplier *mult, Indices &indices, size_t threads,
                            DeviceWork &work, Complex **f, size_t offset,
                            size_t words, double scale)
{
  stagedGeneral(dims,nd,A,B,work,f,offset,words,scale,
                [&](void **F, size_t n, const fftwpp_amd_layout &layout) {
                  indices.layout=&layout;
                  indices.stream=NULL;
                  indices.r=0;
                  indices.offset=0;
                  mult((Complex **) F,n,&indices,threads);
                  indices.layout=NULL;
                });
}

// A library multiplier with A != 2 or B != 1 (the reference runs any A and B
// through its residue loop, convolve.cc:7513-7575): all A inputs are
// transformed, multBinary / multcorrelation act on F[0] and F[1] as in
// convolve.cc:33-110 (multNone on none), and the first B arrays are
// transformed back.  Complex and centred dimensions, as user multipliers.
static void libraryGeneral(fftBase *const *dims, size_t nd, size_t A, size_t B, int code,
                           DeviceWork &work, Complex **f, size_t offset, size_t words,
                           double scale)
{
  if(code == FFTWPP_AMD_MULT_REAL_BINARY)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"realMultBinary with A != 2 or B != 1 (Hermitian data)");
  if(code != FFTWPP_AMD_MULT_NONE && A < 2)
    throw Error(FFTWPP_AMD_ERR_ARG,"a binary multiplier needs A >= 2 inputs");
  stagedGeneral(dims,nd,A,B,work,f,offset,words,scale,
                [&](void **F, size_t n, const fftwpp_amd_layout &) {
                  if(code != FFTWPP_AMD_MULT_NONE)
                    check(fftwpp_amd_multiply(FFTWPP_AMD_F64,code,F[0],F[1],n,NULL),"multiply");
                });
}

// ------------------------------------------------------------------------
// Convolution (1D)
// ------------------------------------------------------------------------
Convolution::Convolution(fftBase *fft, Complex **, Complex *, Complex *) :
  ThreadBase(fft->Threads()), fft(fft), L(fft->L), A(fft->app.A),
  B(fft->app.B), mult(fft->app.mult)
{
  scale=1.0/normalization();
  multCode=multiplierCode(mult);
  indices.copy(NULL,0);
  indices.fft=fft;
}

Convolution::~Convolution() {}

size_t Convolution::arraySize() const
{
  return fft->S*(fft->inputLength()-1)+fft->C;
}

void Convolution::convolveDevice(Complex *f0, const Complex *f1,
                                 bool normalize, void *stream)
{
  execDevice(f0,f0,f1,normalize,FFTWPP_AMD_F64,stream);
}

// Whether the fused row stage takes the rows of fft: a fused row kernel, the
// long Hermitian row path or the batched register kernel (Hermitian or
// complex rows longer than one block of a power-of-two m <= 4096).
static bool rowsSupported(fftBase *fft, int dtype)
{
  return fftwpp_amd_internal_rows_fit(fft->dev,dtype) ||
    fftwpp_amd_hconv_rows_long_part(fft->dev,1) != 0 ||
    (fft->kind() == FFTWPP_AMD_KIND_COMPLEX && fftwpp_amd_batch_rows_tps(fft->dev,dtype) != 0);
}

// The row stage handed to a caller's hook (fftwpp_amd_conv_execute_fused):
// `outer` carries the outer layout of the rows, the rest is filled here.
static void hookRows(const RowHook *hook, fftBase *fft, int dtype, const void *f,
                     const void *g, size_t rows, size_t rs, void *out, size_t os,
                     double s, void *stream, fftwpp_amd_rows_args outer, const char *what)
{
  outer.dim=fft->dev;
  outer.dtype=dtype;
  outer.hermitian=fft->kind() == FFTWPP_AMD_KIND_HERMITIAN;
  outer.f=f;
  outer.g=g;
  outer.out=out;
  outer.rows=rows;
  outer.rowstride=rs;
  outer.outstride=os;
  outer.scale=s;
  outer.stream=stream;
  const int rc=hook->hook(&outer,hook->user);
  if(rc)
    throw Error(rc < 0 ? rc : FFTWPP_AMD_ERR_UNSUPPORTED,
                std::string(what)+": the fused row stage refused the geometry or failed to launch");
}

// Outer layout of the rows of a 2D (one outer dimension) or 3D (two) row stage.
static fftwpp_amd_rows_args outerLayout(fftBase *x, size_t nx, bool realx, fftBase *y=NULL)
{
  fftwpp_amd_rows_args a;
  memset(&a,0,sizeof(a));
  a.nouter=y ? 2 : 1;
  a.om[0]=x->m;
  a.oq[0]=x->q;
  a.ocount[0]=nx;
  a.oreal[0]=realx;
  a.ostride[0]=y ? y->q*y->m : 1;
  if(y) {
    a.om[1]=y->m;
    a.oq[1]=y->q;
    a.ocount[1]=y->q*y->m;
    a.ostride[1]=1;
  }
  return a;
}

// The inner row stage without the fused kernels: rows no row engine takes
// (complex or centred rows beyond the register, LDS and batched engines) and
// inner transforms of C > 1 interleaved columns (element stride C).  Every
// residue of f and g forward into full padded rows, the multiplier on the
// device arrays, every residue backward into f -- the residue loop of
// Convolution::convolveRaw (convolve.cc:7513-7575) with all residues at
// once.  Row i starts at i rs; work slots `slot` and slot + 1.
static void unfusedRows(fftBase *fft, DeviceWork &work, size_t slot, int dtype, int code,
                        void *f, void *g, size_t rows, size_t C, size_t rs,
                        void *stream, const char *what)
{
  if(fft->kind() == FFTWPP_AMD_KIND_HERMITIAN)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,std::string(what)+": Hermitian rows beyond the row kernels");
  if(code == FFTWPP_AMD_MULT_REAL_BINARY)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,std::string(what)+": realMultBinary takes Hermitian rows");
  const size_t N=fft->m*fft->q, E=eltsize(dtype), q=fft->q;
  char *Y[2]={(char *) work.get(slot,rows*N*C*E),(char *) work.get(slot+1,rows*N*C*E)};
  char *X[2]={(char *) f,(char *) g};
  // single columns: the rows become the columns of one launch; else batches
  // of at most maxBatch rows (the grid's y extent)
  const size_t step=C == 1 ? rows : maxBatch;
  for(int a=0; a < 2; ++a)
    for(size_t b0=0; b0 < rows; b0 += step)
      check(fftwpp_amd_forward_many(fft->dev,dtype,X[a]+b0*rs*E,C,C,std::min(step,rows-b0),rs,
                                    Y[a]+b0*N*C*E,C,N*C,0,q,stream),what);
  check(fftwpp_amd_multiply(dtype,code,Y[0],Y[1],rows*N*C,stream),what);
  for(size_t b0=0; b0 < rows; b0 += step)
    check(fftwpp_amd_backward_many(fft->dev,dtype,Y[0]+b0*N*C*E,C,N*C,X[0]+b0*rs*E,C,C,
                                   std::min(step,rows-b0),rs,0,q,0,1.0,stream),what);
}

// The fused row stage of dimension fft: the register row kernels when they
// cover the row, else the long Hermitian path for Hermitian rows of a
// power-of-two m (residue passes in parallel, partial sums in work slot
// `slot`), else the LDS row kernel.
static void fusedRows(fftBase *fft, DeviceWork &work, size_t slot, int dtype,
                      int code, const void *f, const void *g, size_t rows,
                      size_t rs, void *out, size_t os, double s, void *stream,
                      const char *what)
{
  const bool reg=fftwpp_amd_internal_engine(fft->dev,1) > 0;
  const bool longh=fftwpp_amd_hconv_rows_long_part(fft->dev,1) != 0;
  const size_t tps=fft->kind() == FFTWPP_AMD_KIND_COMPLEX && rs == os &&
    code != FFTWPP_AMD_MULT_REAL_BINARY ? fftwpp_amd_batch_rows_tps(fft->dev,dtype) : 0;
  if(reg || (!longh && !tps && fftwpp_amd_internal_rows_fit(fft->dev,dtype))) {
    check(fftwpp_amd_conv_rows(fft->dev,dtype,code,f,g,rows,rs,out,os,s,stream),what);
    return;
  }
  if(tps) {
    // complex rows longer than one block of a power-of-two m: the batched
    // register kernel with the rows as items (row i at i rs).  The residues
    // always go to partial sums that its second launch adds into out (at
    // least two groups when q > 1), since out may alias f or g and a unit
    // that wrote its outputs would re-read them in its next residue's fold:
    // one residue per unit when rows are few, two groups otherwise.
    const size_t L=fft->L, q=fft->q;
    const size_t D=rows >= 512 ? (q+1)/2 : 1;
    void *items=work.get(slot+1,rows*3*sizeof(int64_t));
    const std::vector<size_t> key={rows,rs,L};
    if(work.keys[slot+1] != key) {
      // item i = {i rs, L, i L}, written on the device (no host staging)
      check(fftwpp_amd_internal_row_items(items,rows,rs,L,stream),"item table");
      work.keys[slot+1]=key;
    }
    const size_t G=(q+D-1)/D;
    void *part=G > 1 ? work.get(slot,G*rows*L*(dtype == FFTWPP_AMD_F32 ? 8 : 16)) : NULL;
    const size_t units=rows*G;
    check(fftwpp_amd_batch_rows(fft->dev,dtype,code,items,rows,f,g,out,1,
                                tps < 64 ? 64 : tps,std::min<size_t>(units,2048),0,D,
                                part,rows*L,s,stream),what);
    return;
  }
  const size_t words=fftwpp_amd_hconv_rows_long_part(fft->dev,rows);
  if(words == 0)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"inner sub-FFT size beyond the fused row kernels");
  if(code != FFTWPP_AMD_MULT_REAL_BINARY)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"Hermitian rows take realMultBinary");
  void *part=work.get(slot,words*(dtype == FFTWPP_AMD_F32 ? 8 : 16));
  check(fftwpp_amd_hconv_rows_long(fft->dev,dtype,f,g,rows,rs,out,os,part,s,stream),what);
}

void Convolution::execDevice(void *out, const void *f0, const void *f1, bool normalize,
                             int dtype, void *stream, int mult)
{
  int code=mult >= 0 ? mult : multCode;
  requireBinary(A,B,code);
  if(rowHook && fft->realInput() && !realPlanes(fft))
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a fused multiplier needs the real residue transforms (power-of-two m in [16, 4096], L <= 2m)");
  if(realPlanes(fft)) {
    if(code != FFTWPP_AMD_MULT_BINARY && code != FFTWPP_AMD_MULT_CORRELATION)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"real convolutions take multBinary or multcorrelation");
    execComplex(out,f0,f1,normalize,dtype,stream,code,true);
  } else if(fft->realInput())
    realConvolution(work,arraySize(),dtype,code,out,f0,f1,stream,[&](void *z) {
      execComplex(z,z,z,normalize,dtype,stream,code);
    });
  else
    execComplex(out,f0,f1,normalize,dtype,stream,code);
}

void Convolution::execComplex(void *out, const void *f0, const void *f1, bool normalize,
                              int dtype, void *stream, int code, bool real)
{
  if(fft->C != 1)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"1D Convolution with C > 1");
  double s=normalize ? scale : 1.0;
  size_t n=fft->inputLength();
  if(timeline) timeline->begin(stream);
  if(rowHook) {
    // the whole 1D array is the one row of the stage
    if(real)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a 1D real convolution has no row stage for a fused multiplier");
    fftwpp_amd_rows_args outer;
    memset(&outer,0,sizeof(outer));
    hookRows(rowHook,fft,dtype,f0,f1,1,n,out,n,s,stream,outer,"rows");
    if(timeline) timeline->mark("rows");
    return;
  }
  if(real) {
    // kept planes of f and g, their products, the real backward transform
    // (fftPadReal::forward/backward around Convolution::operate)
    const size_t K=fftwpp_amd_real_planes(fft->dev), E=eltsize(dtype);
    void *T0=work.get(0,K*E);
    void *T1=work.get(1,K*E);
    check(fftwpp_amd_forward_real_many(fft->dev,dtype,f0,1,1,1,0,T0,1,0,stream),"real forward");
    check(fftwpp_amd_forward_real_many(fft->dev,dtype,f1,1,1,1,0,T1,1,0,stream),"real forward");
    if(timeline) timeline->mark("real forward");
    check(fftwpp_amd_multiply(dtype,code,T0,T1,K,stream),"multiply");
    check(fftwpp_amd_backward_real_many(fft->dev,dtype,T0,1,0,out,1,1,1,0,s,stream),"real backward");
    if(timeline) timeline->mark("real backward");
    return;
  }
  if(fft->kind() == FFTWPP_AMD_KIND_HERMITIAN) {
    fusedRows(fft,work,4,dtype,code,f0,f1,1,n,out,n,s,stream,"Hermitian rows");
    if(timeline) timeline->mark("rows");
    return;
  }
  if(fftwpp_amd_internal_rows_fit(fft->dev,dtype)) {
    check(fftwpp_amd_conv_rows(fft->dev,dtype,code,f0,f1,1,n,out,n,s,
                               stream),"conv_rows");
    if(timeline) timeline->mark("rows");
    return;
  }
  size_t E=eltsize(dtype);
  if(fft->kind() == FFTWPP_AMD_KIND_COMPLEX && fftwpp_amd_batch_rows_tps(fft->dev,dtype) &&
     code != FFTWPP_AMD_MULT_REAL_BINARY) {
    // a row longer than a power-of-two m <= 4096: the batched register kernel
    fusedRows(fft,work,2,dtype,code,f0,f1,1,n,out,n,s,stream,"batched row");
    if(timeline) timeline->mark("batched row");
    return;
  }
  // unfused: all residues forward, multiply, backward
  size_t words=fft->q*fft->m;
  void *T0=work.get(0,words*E);
  void *T1=work.get(1,words*E);
  check(fftwpp_amd_forward_many(fft->dev,dtype,f0,1,1,1,0,T0,1,0,0,fft->q,stream),"forward");
  check(fftwpp_amd_forward_many(fft->dev,dtype,f1,1,1,1,0,T1,1,0,0,fft->q,stream),"forward");
  check(fftwpp_amd_multiply(dtype,code,T0,T1,code == FFTWPP_AMD_MULT_REAL_BINARY ? 2*words : words,stream),"multiply");
  check(fftwpp_amd_backward_many(fft->dev,dtype,T0,1,0,out,1,1,1,0,0,fft->q,0,s,stream),"backward");
  if(timeline) timeline->mark("unfused");
}

void Convolution::run(Complex **f, size_t offset, bool normalize)
{
  if(multCode < 0 || A != 2 || B != 1) {
    fftBase *dims[1]={fft};
    if(fft->C != 1)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"1D Convolution with C > 1");
    if(multCode < 0)
      userConvolution(dims,1,A,B,mult,indices,threads,work,f,offset,arraySize(),
                      normalize ? scale : 1.0);
    else
      libraryGeneral(dims,1,A,B,multCode,work,f,offset,arraySize(),normalize ? scale : 1.0);
    return;
  }
  staged(work,f,offset,arraySize(),2,fft->realInput() ? sizeof(double) : sizeof(Complex),
         [&](void *a, void *b) { execDevice(a,a,b,normalize,FFTWPP_AMD_F64,NULL); });
}

// ------------------------------------------------------------------------
// Hermitian symmetrisers (convolve.h:1168-1267) on device or host data
// ------------------------------------------------------------------------
template<class Exec>
static void onDevice(Complex *f, size_t extent, Exec exec)
{
  if(fftwpp_amd_is_device_pointer(f)) {
    exec(f);
    check(fftwpp_amd_synchronize(NULL),"synchronize");
    return;
  }
  void *d=NULL;
  size_t bytes=extent*sizeof(Complex);
  check(fftwpp_amd_malloc(&d,bytes),"symmetrize buffer");
  int rc=fftwpp_amd_memcpy(d,f,bytes,NULL);
  if(!rc) { exec((Complex *) d); rc=fftwpp_amd_memcpy(f,d,bytes,NULL); }
  if(!rc) rc=fftwpp_amd_synchronize(NULL);
  fftwpp_amd_free(d);
  check(rc,"symmetrize");
}

void HermitianSymmetrize(Complex *f)
{
  onDevice(f,1,[&](Complex *d) {
    check(fftwpp_amd_symmetrize_x(1,1,0,d,1,NULL),"HermitianSymmetrize");
  });
}

void HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0, Complex *f,
                          size_t Sx, size_t)
{
  size_t extent=(x0+Hx-1)*Sx+Hy;
  onDevice(f,extent,[&](Complex *d) {
    check(fftwpp_amd_symmetrize_x(Hx,Hy,x0,d,Sx,NULL),"HermitianSymmetrizeX");
  });
}

void HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0, Complex *f)
{
  HermitianSymmetrizeX(Hx,Hy,x0,f,Hy);
}

void HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                           size_t y0, Complex *f, size_t Sx, size_t Sy,
                           size_t)
{
  size_t extent=(x0+Hx-1)*Sx+(y0+Hy-1)*Sy+Hz;
  onDevice(f,extent,[&](Complex *d) {
    check(fftwpp_amd_symmetrize_xy(Hx,Hy,Hz,x0,y0,d,Sx,Sy,NULL),"HermitianSymmetrizeXY");
  });
}

void HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                           size_t y0, Complex *f)
{
  size_t Ly=y0+Hy;
  HermitianSymmetrizeXY(Hx,Hy,Hz,x0,y0,f,Ly*Hz,Hz);
}

// ------------------------------------------------------------------------
// Convolution2 (convolve.h:1269-1548): x outer transform (multNone), inner y
// rows fused.  All x residues are expanded at once (nloops == 1), so the
// backward transform writes f[0] directly.
// ------------------------------------------------------------------------
Convolution2::Convolution2(fftBase *fftx, fftBase *ffty, Complex **,
                           Complex *, Complex *) :
  ThreadBase(fftx->Threads()), fftx(fftx), ffty(ffty), A(fftx->app.A),
  B(fftx->app.B), mult(ffty->app.mult)
{
  Lx=fftx->L;
  Ly=fftx->C;
  Sx=fftx->S;
  ly=ffty->inputLength();
  scale=1.0/normalization();
  multCode=multiplierCode(mult);
  indices.copy(NULL,1);
  indices.fft=ffty;
  // C > 1: the y rows hold C interleaved sequences (element stride C)
  if(ffty->C > 1 && ffty->S != ffty->C)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"inner transform with C > 1 takes interleaved columns (S = C)");
  if(ly*ffty->C != Ly)
    throw Error(FFTWPP_AMD_ERR_ARG,"fftx->C must equal the stored length of the y transform times its C");
}

Convolution2::~Convolution2() {}

size_t Convolution2::arraySize() const { return Sx*(Lx-1)+Ly; }

void Convolution2::convolveDevice(Complex *f0, const Complex *f1,
                                  bool normalize, void *stream)
{
  execDevice(f0,f0,f1,normalize,FFTWPP_AMD_F64,stream);
}

void Convolution2::execDevice(void *out, const void *f0, const void *f1, bool normalize,
                              int dtype, void *stream, int mult)
{
  int code=mult >= 0 ? mult : multCode;
  requireBinary(A,B,code);
  if(rowHook && fftx->realInput() && !realPlanes(fftx))
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a fused multiplier needs the real residue transforms (power-of-two m in [16, 4096], L <= 2m)");
  if(realPlanes(fftx)) {
    if(code != FFTWPP_AMD_MULT_BINARY && code != FFTWPP_AMD_MULT_CORRELATION)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"real convolutions take multBinary or multcorrelation");
    execComplex(out,f0,f1,normalize,dtype,stream,code,true);
  } else if(fftx->realInput())
    realConvolution(work,arraySize(),dtype,code,out,f0,f1,stream,[&](void *z) {
      execComplex(z,z,z,normalize,dtype,stream,code);
    });
  else
    execComplex(out,f0,f1,normalize,dtype,stream,code);
}

void Convolution2::execComplex(void *out, const void *f0, const void *f1, bool normalize,
                               int dtype, void *stream, int code, bool real)
{
  const size_t Cy=ffty->C, lyc=ly*Cy;   // words of a stored y row
  const bool fused=Cy == 1 && rowsSupported(ffty,dtype);
  if(rowHook && !fused)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a fused multiplier needs y rows of the fused row kernels");
  double s=normalize ? scale : 1.0;
  size_t E=eltsize(dtype);
  size_t qx=fftx->q, mx=fftx->m;
  // x residues per pass (all unless the memory limit asks for blocks)
  const size_t Dx=residueBlock(dtype,real);
  if(Dx == 0)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"2D convolution: work buffers exceed the memory limit");
  // x-transformed rows per pass: Dx residues, or the kept planes of real data
  const size_t maxrows=real ? fftwpp_amd_real_planes(fftx->dev) : Dx*mx;
  void *F[2]={work.get(0,maxrows*lyc*E),work.get(1,maxrows*lyc*E)};
  const void *in[2]={f0,f1};
  if(Dx < qx && (out == f0 || out == f1)) {
    // residue blocks write partial results into out: keep the input it aliases
    void *keep=work.get(10,arraySize()*E);
    check(fftwpp_amd_memcpy(keep,out,arraySize()*E,stream),"input copy");
    in[out == f0 ? 0 : 1]=keep;
  }
  if(timeline) timeline->begin(stream);
  for(size_t r0=0; r0 < qx; r0 += Dx) {
    const size_t nres=std::min(Dx,qx-r0);
    const size_t rows=real ? maxrows : nres*mx;
    const bool first=r0 == 0, last=r0+nres == qx;
    for(int a=0; a < 2; ++a) {
      if(real)
        check(fftwpp_amd_forward_real_many(fftx->dev,dtype,in[a],Sx,lyc,1,0,F[a],lyc,0,stream),"x forward");
      else
        check(fftwpp_amd_forward_many(fftx->dev,dtype,in[a],Sx,lyc,1,0,F[a],lyc,0,r0,nres,stream),"x forward");
      if(timeline) timeline->mark(real ? (a == 0 ? "x real forward f" : "x real forward g") : (a == 0 ? "x forward f" : "x forward g"));
    }
    if(rowHook)
      hookRows(rowHook,ffty,dtype,F[0],F[1],rows,ly,F[0],ly,1.0,stream,outerLayout(fftx,rows,real),"y rows");
    else if(fused)
      fusedRows(ffty,work,2,dtype,code,F[0],F[1],rows,ly,F[0],ly,1.0,stream,"y rows");
    else
      unfusedRows(ffty,work,8,dtype,code,F[0],F[1],rows,Cy,lyc,stream,"y rows");
    if(timeline) timeline->mark(fused ? "y rows" : "y rows (unfused)");
    // this block's x backward added to the earlier blocks' sum; the last block
    // applies the normalisation to the whole sum
    if(real)
      check(fftwpp_amd_backward_real_many(fftx->dev,dtype,F[0],lyc,0,out,Sx,lyc,1,0,s,stream),"x backward");
    else
      check(fftwpp_amd_backward_many(fftx->dev,dtype,F[0],lyc,0,out,Sx,lyc,1,0,r0,nres,!first,
                                     last ? s : 1.0,stream),"x backward");
    if(timeline) timeline->mark(real ? "x real backward" : "x backward");
  }
}

// Work bytes of an execution with Dx x residues per pass (0: all): the two
// x-transformed arrays, the unfused row stage's two expanded arrays when it
// runs, and with residue blocks a copy of the input an in-place call
// overwrites.
size_t Convolution2::workBytes(int dtype, size_t Dx) const
{
  const size_t E=eltsize(dtype);
  const bool real=realPlanes(fftx);
  if(Dx == 0 || real) Dx=fftx->q;
  const size_t rows=real ? fftwpp_amd_real_planes(fftx->dev) : Dx*fftx->m;
  const size_t Cy=ffty->C;
  size_t w=2*rows*ly*Cy*E;
  if(Cy != 1 || !rowsSupported(ffty,dtype)) w += 2*rows*ffty->m*ffty->q*Cy*E;
  if(Dx < fftx->q) w += arraySize()*E;
  return w;
}

size_t Convolution2::residueBlock(int dtype, bool real) const
{
  const size_t qx=fftx->q;
  if(!memoryLimit) return qx;
  if(real || rowHook) return workBytes(dtype,qx) <= memoryLimit ? qx : 0;
  for(size_t D=qx; D > 0; --D)
    if(workBytes(dtype,D) <= memoryLimit) return D;
  return 0;
}

void Convolution2::run(Complex **f, size_t offset, bool normalize)
{
  if(multCode < 0 || A != 2 || B != 1) {
    fftBase *dims[2]={fftx,ffty};
    if(Sx != Ly || ffty->C != 1)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"user multipliers, and A != 2 or B != 1, take contiguous arrays of one column");
    if(multCode < 0)
      userConvolution(dims,2,A,B,mult,indices,threads,work,f,offset,arraySize(),
                      normalize ? scale : 1.0);
    else
      libraryGeneral(dims,2,A,B,multCode,work,f,offset,arraySize(),normalize ? scale : 1.0);
    return;
  }
  staged(work,f,offset,arraySize(),2,fftx->realInput() ? sizeof(double) : sizeof(Complex),
         [&](void *a, void *b) { execDevice(a,a,b,normalize,FFTWPP_AMD_F64,NULL); });
}

// ------------------------------------------------------------------------
// Convolution3 (convolve.h:1550-1893): x and y outer transforms, z rows fused.
// ------------------------------------------------------------------------
Convolution3::Convolution3(fftBase *fftx, fftBase *ffty, fftBase *fftz,
                           Complex **, Complex *, Complex *, bool) :
  ThreadBase(fftx->Threads()), fftx(fftx), ffty(ffty), fftz(fftz),
  A(fftx->app.A), B(fftx->app.B), mult(fftz->app.mult)
{
  Lx=fftx->L;
  Ly=ffty->L;
  Lz=ffty->C;
  Sx=fftx->S;
  Sy=ffty->S;
  lz=fftz->inputLength();
  scale=1.0/normalization();
  multCode=multiplierCode(mult);
  indices.copy(NULL,2);
  indices.fft=fftz;
  // C > 1: the z rows hold C interleaved sequences (element stride C)
  if(fftz->C > 1 && fftz->S != fftz->C)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"inner transform with C > 1 takes interleaved columns (S = C)");
  if(lz*fftz->C != Lz)
    throw Error(FFTWPP_AMD_ERR_ARG,"ffty->C must equal the stored length of the z transform times its C");
  if(Sx < Ly*Sy) throw Error(FFTWPP_AMD_ERR_ARG,"Sx cannot be less than Ly*Sy");
  if(fftx->C != (contiguous() ? Ly*Sy : Lz)) throw Error(FFTWPP_AMD_ERR_ARG,"fftx->C is invalid");
  // y and z both m = 256, q = 2, complex, rows of lz words back to back: the
  // y-z stage of every x-plane runs as one plane-kernel launch
  planes=contiguous() && Sx == Ly*Sy && !fftx->realInput() && fftz->C == 1 &&
    fftwpp_amd_plane_supported(ffty->dev,fftz->dev);
}

Convolution3::~Convolution3() {}

size_t Convolution3::workBytes(int dtype, bool plane, size_t Dx) const
{
  const size_t E=eltsize(dtype);
  const bool real=realPlanes(fftx);
  if(Dx == 0 || real) Dx=fftx->q;
  const size_t planesX=real ? fftwpp_amd_real_planes(fftx->dev) : Dx*fftx->m;
  const size_t X=planesX*Ly*Lz*E;                     // one x-transformed array
  if(plane) return 2*X+fftwpp_amd_plane_work(dtype);
  // residue blocks: the input that out aliases is kept (in-place calls)
  const size_t keep=Dx < fftx->q ? arraySize()*E : 0;
  // the unfused z stage expands every z row to its padded length
  const size_t zrows=planesX*ffty->q*ffty->m;
  const size_t unf=fftz->C != 1 || !rowsSupported(fftz,dtype) ? 2*zrows*fftz->m*fftz->q*fftz->C*E : 0;
  return 2*X+2*ffty->q*X+keep+unf;
}

// x residues per pass: all of them, or under a memory limit the largest
// block whose work buffers fit (0: none fits).  Real outer dimensions keep
// all their planes in one pass.
size_t Convolution3::residueBlock(int dtype, bool real) const
{
  const size_t qx=fftx->q;
  if(!memoryLimit || real || rowHook) return workBytes(dtype,false,qx) <= memoryLimit || !memoryLimit ? qx : 0;
  for(size_t D=qx; D > 0; --D)
    if(workBytes(dtype,false,D) <= memoryLimit) return D;
  return 0;
}

// The plane kernel runs when it applies and either FFTWPP_AMD_PLANE=1 asks for
// it or the memory limit excludes the separate y, z and y launches (which
// are faster on one MI355X: their stages stream through HBM at near copy
// speed, while the plane kernel's team barriers cost more than the HBM
// traffic it saves).
bool Convolution3::usesPlanes(int dtype) const
{
  if(!planes) return false;
  const char *env=getenv("FFTWPP_AMD_PLANE");
  if(env && env[0] == '1') return true;
  return memoryLimit != 0 && workBytes(dtype,false,fftx->q) > memoryLimit &&
    workBytes(dtype,true) <= memoryLimit;
}

size_t Convolution3::arraySize() const { return Sx*(Lx-1)+Sy*(Ly-1)+Lz; }

void Convolution3::convolveDevice(Complex *f0, const Complex *f1,
                                  bool normalize, void *stream)
{
  execDevice(f0,f0,f1,normalize,FFTWPP_AMD_F64,stream);
}

void Convolution3::execDevice(void *out, const void *f0, const void *f1, bool normalize,
                              int dtype, void *stream, int mult)
{
  int code=mult >= 0 ? mult : multCode;
  requireBinary(A,B,code);
  if(rowHook && fftx->realInput() && !realPlanes(fftx))
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a fused multiplier needs the real residue transforms (power-of-two m in [16, 4096], L <= 2m)");
  if(realPlanes(fftx)) {
    if(code != FFTWPP_AMD_MULT_BINARY && code != FFTWPP_AMD_MULT_CORRELATION)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"real convolutions take multBinary or multcorrelation");
    execComplex(out,f0,f1,normalize,dtype,stream,code,true);
  } else if(fftx->realInput())
    realConvolution(work,arraySize(),dtype,code,out,f0,f1,stream,[&](void *z) {
      execComplex(z,z,z,normalize,dtype,stream,code);
    });
  else
    execComplex(out,f0,f1,normalize,dtype,stream,code);
}

void Convolution3::execComplex(void *out, const void *f0, const void *f1, bool normalize,
                               int dtype, void *stream, int code, bool real)
{
  // z rows of Lz = lz C words (C interleaved sequences when C > 1)
  const size_t Cz=fftz->C;
  const bool fused=Cz == 1 && rowsSupported(fftz,dtype);
  if(rowHook && !fused)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"a fused multiplier needs z rows of the fused row kernels");
  double s=normalize ? scale : 1.0;
  size_t E=eltsize(dtype);
  size_t qx=fftx->q, mx=fftx->m, qy=ffty->q, my=ffty->m;
  size_t plane=Ly*Lz;
  size_t yplane=qy*my*Lz;        // y-transformed plane
  const void *in[2]={f0,f1};
  const bool viaPlanes=!real && !rowHook && code != FFTWPP_AMD_MULT_REAL_BINARY && usesPlanes(dtype);
  // x residues per pass of the separate y, z, y stages (all of them unless
  // the memory limit asks for residue blocks)
  const size_t Dx=viaPlanes ? qx : residueBlock(dtype,real);
  if(Dx == 0)
    throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"3D convolution: work buffers exceed the memory limit");
  if(viaPlanes) {
    // x forward (f, g) -> y-z plane convolutions (one launch) -> x backward
    const size_t nplanes=qx*mx;
    void *X[2]={work.get(0,nplanes*plane*E),work.get(1,nplanes*plane*E)};
    void *P=work.get(5,fftwpp_amd_plane_work(dtype));
    if(timeline) timeline->begin(stream);
    for(int a=0; a < 2; ++a) {
      if(contiguous())
        check(fftwpp_amd_forward_many(fftx->dev,dtype,in[a],Sx,plane,1,0,X[a],plane,0,0,qx,stream),"x forward");
      else
        check(fftwpp_amd_forward_many(fftx->dev,dtype,in[a],Sx,Lz,Ly,Sy,X[a],plane,Lz,0,qx,stream),"x forward");
      if(timeline) timeline->mark(a == 0 ? "x forward f" : "x forward g");
    }
    check(fftwpp_amd_plane_conv(ffty->dev,fftz->dev,dtype,code,X[0],X[1],X[0],nplanes,Ly,lz,P,stream),
          "y-z planes");
    if(timeline) timeline->mark("y-z planes");
    check(fftwpp_amd_backward_many(fftx->dev,dtype,X[0],plane,0,out,Sx,plane,1,0,0,qx,0,s,stream),
          "x backward");
    if(timeline) timeline->mark("x backward");
    return;
  }
  // x-transformed planes of Ly x Lz per pass: Dx residues, or the kept planes
  // of real data
  const size_t maxplanes=real ? fftwpp_amd_real_planes(fftx->dev) : Dx*mx;
  void *X[2]={work.get(0,maxplanes*plane*E),work.get(1,maxplanes*plane*E)};
  void *Y[2]={work.get(2,maxplanes*yplane*E),work.get(3,maxplanes*yplane*E)};
  if(Dx < qx && (out == f0 || out == f1)) {
    // residue blocks write partial results into out: keep the input it aliases
    void *keep=work.get(9,arraySize()*E);
    check(fftwpp_amd_memcpy(keep,out,arraySize()*E,stream),"input copy");
    in[out == f0 ? 0 : 1]=keep;
  }
  if(timeline) timeline->begin(stream);
  for(size_t r0=0; r0 < qx; r0 += Dx) {
    const size_t nres=std::min(Dx,qx-r0);
    const size_t nplanes=real ? maxplanes : nres*mx;
    const bool first=r0 == 0, last=r0+nres == qx;
    // x forward: columns (y,z); non-contiguous y strides become batches
    for(int a=0; a < 2; ++a) {
      if(real && contiguous())
        check(fftwpp_amd_forward_real_many(fftx->dev,dtype,in[a],Sx,plane,1,0,X[a],plane,0,stream),"x forward");
      else if(real)
        check(fftwpp_amd_forward_real_many(fftx->dev,dtype,in[a],Sx,Lz,Ly,Sy,X[a],plane,Lz,stream),"x forward");
      else if(contiguous())
        check(fftwpp_amd_forward_many(fftx->dev,dtype,in[a],Sx,plane,1,0,X[a],plane,0,r0,nres,stream),"x forward");
      else
        check(fftwpp_amd_forward_many(fftx->dev,dtype,in[a],Sx,Lz,Ly,Sy,X[a],plane,Lz,r0,nres,stream),"x forward");
      if(timeline) timeline->mark(real ? (a == 0 ? "x real forward f" : "x real forward g") : (a == 0 ? "x forward f" : "x forward g"));
      check(fftwpp_amd_forward_many(ffty->dev,dtype,X[a],Lz,Lz,nplanes,plane,Y[a],Lz,yplane,0,qy,stream),"y forward");
      if(timeline) timeline->mark(a == 0 ? "y forward f" : "y forward g");
    }
    if(rowHook)
      hookRows(rowHook,fftz,dtype,Y[0],Y[1],nplanes*qy*my,lz,Y[0],lz,1.0,stream,
               outerLayout(fftx,nplanes,real,ffty),"z rows");
    else if(fused)
      fusedRows(fftz,work,4,dtype,code,Y[0],Y[1],nplanes*qy*my,lz,Y[0],lz,1.0,stream,"z rows");
    else
      unfusedRows(fftz,work,10,dtype,code,Y[0],Y[1],nplanes*qy*my,Cz,Lz,stream,"z rows");
    if(timeline) timeline->mark(fused ? "z rows" : "z rows (unfused)");
    check(fftwpp_amd_backward_many(ffty->dev,dtype,Y[0],Lz,yplane,X[0],Lz,Lz,nplanes,plane,0,qy,0,1.0,stream),"y backward");
    if(timeline) timeline->mark("y backward");
    // x backward of this block, added to the earlier blocks' sum; the last
    // block applies the normalisation to the whole sum
    const double sb=last ? s : 1.0;
    if(real && contiguous())
      check(fftwpp_amd_backward_real_many(fftx->dev,dtype,X[0],plane,0,out,Sx,plane,1,0,s,stream),"x backward");
    else if(real)
      check(fftwpp_amd_backward_real_many(fftx->dev,dtype,X[0],plane,Lz,out,Sx,Lz,Ly,Sy,s,stream),"x backward");
    else if(contiguous())
      check(fftwpp_amd_backward_many(fftx->dev,dtype,X[0],plane,0,out,Sx,plane,1,0,r0,nres,!first,sb,stream),"x backward");
    else
      check(fftwpp_amd_backward_many(fftx->dev,dtype,X[0],plane,Lz,out,Sx,Lz,Ly,Sy,r0,nres,!first,sb,stream),"x backward");
    if(timeline) timeline->mark(real ? "x real backward" : "x backward");
  }
}

void Convolution3::run(Complex **f, size_t offset, bool normalize)
{
  if(multCode < 0 || A != 2 || B != 1) {
    fftBase *dims[3]={fftx,ffty,fftz};
    if(Sy != Lz || Sx != Ly*Lz || fftz->C != 1)
      throw Error(FFTWPP_AMD_ERR_UNSUPPORTED,"user multipliers, and A != 2 or B != 1, take contiguous arrays of one column");
    if(multCode < 0)
      userConvolution(dims,3,A,B,mult,indices,threads,work,f,offset,arraySize(),
                      normalize ? scale : 1.0);
    else
      libraryGeneral(dims,3,A,B,multCode,work,f,offset,arraySize(),normalize ? scale : 1.0);
    return;
  }
  staged(work,f,offset,arraySize(),2,fftx->realInput() ? sizeof(double) : sizeof(Complex),
         [&](void *a, void *b) { execDevice(a,a,b,normalize,FFTWPP_AMD_F64,NULL); });
}

} // namespace fftwpp
