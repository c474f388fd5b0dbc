// batch_plan.hip -- plans for batches of independent 1D complex dealiased
// convolutions of uneven lengths (no reference counterpart: FFTW++ convolves
// one set of arrays per Convolution object and call).
//
// Items are packed back to back: item i occupies [off_i, off_i + L_i) of f,
// g and out.  Each item gets a sub-FFT size m; with the reference's parameter
// rule (fftBase::parameters, convolve.cc:403-410, non-centred data) that
// fixes p = ceil(L/m), P = (p == 2 ? 1 : p), n = ceil(M/(P m)), q = P n and
// the padded length N = m q >= M.  Items with equal (m, q) form a class that
// one persistent launch of fftwpp_amd_batch_rows convolves.  The classes are
// forked over side streams (default priority) and joined back; with
// FFTWPP_AMD_BATCH_GRAPH the first execute() on a set of arrays runs eagerly
// and records the same fork/join into a hipGraph that later calls replay
// with one launch.
//
// Geometry and tile shapes: with FFTWPP_AMD_BATCH_TUNE the plan times, for
// every power-of-two length bucket (2^(k-1), 2^k], the candidate sub-FFT
// sizes (explicit N = 2^ceil(log2 M) with q = 1, and m = 2^k, 2^(k-1),
// 2^(k-2) within the LDS limit) and for each resulting class the tile shapes
// (items per block TR, threads per block), keeping the fastest; this replaces
// the reference's CPU timing optimizer (convolve.cc:177-470).  Without it a
// fixed rule is used: m = min(2^ceil(log2 L), max_m).
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fftwpp_amd.h"
#include "batch.h"

extern "C" void fftwpp_amd_internal_set_error(const char *msg);

using fftwpp_amd::BatchItem;

namespace {

int bfail(int code, const char *msg)
{
  fftwpp_amd_internal_set_error(msg);
  return code;
}

int bhip(hipError_t e, const char *what)
{
  if(e == hipSuccess) return FFTWPP_AMD_OK;
  char buf[256];
  snprintf(buf,sizeof(buf),"%s: %s",what,hipGetErrorString(e));
  return bfail(FFTWPP_AMD_ERR_HIP,buf);
}

size_t ceilquot(size_t a, size_t b) { return (a+b-1)/b; }

size_t ceilpow2(size_t n)
{
  size_t p=1;
  while(p < n) p <<= 1;
  return p;
}

// q for an item of length L, minimum padded length M and sub-FFT size m
// (fftBase::parameters, convolve.cc:403-410; P = p/2 only for p == 2 in the
// non-centred case).
size_t residues(size_t L, size_t M, size_t m)
{
  size_t p=ceilquot(L,m);
  size_t P=p == 2 ? 1 : p;
  size_t n=ceilquot(M,P*m);
  return P*n;
}

struct Tile {
  size_t TR=1, threads=256, nblocks=1, lacc=0, D=1;
};

struct Cls {
  size_t m=0, q=0;
  fftwpp_amd_dim *dim=nullptr;
  std::vector<size_t> ids;      // items, longest first
  BatchItem *items=nullptr;     // device copy
  size_t Lmax=0, ptotal=0;      // longest item, sum of lengths
  void *part=nullptr;           // partial sums when the tile splits residues
  Tile tile;
  double work=0;                // N log2 N summed over items
  double ms=0;                  // tuned launch time
};

struct Graph {
  const void *out, *f, *g;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

const int NSIDE=4;

// Largest sub-FFT size: the register-resident FFT of the batch kernel.
const size_t BATCH_MAXM=4096;

} // namespace

// Recorded graphs kept per batch handle, one per (out, f, g) triple; the
// least recently used one is replaced beyond this.
constexpr size_t MAXGRAPHS=16;

struct fftwpp_amd_batch {
  int dtype=0, mult=0, flags=0;
  size_t n=0, total=0;
  std::vector<size_t> L, M, off, mof;
  std::vector<Cls> cls;
  hipStream_t side[NSIDE]={};
  hipStream_t cap=nullptr;
  hipEvent_t fork=nullptr, join[NSIDE]={};
  std::vector<Graph> graphs;
  int cus=256;
  // stream schedule of the classes: 0 = dealt round robin over the side
  // streams; 1 = all on one stream in order; 2 = the classes with a
  // persistent grid (fewer blocks than work units) in order on one stream,
  // the others dealt over the remaining side streams
  int sched=0;
};

namespace {

void free_class(Cls &c)
{
  if(c.items) (void) hipFree(c.items);
  if(c.part) (void) hipFree(c.part);
  if(c.dim) fftwpp_amd_dim_destroy(c.dim);
  c.items=nullptr;
  c.part=nullptr;
  c.dim=nullptr;
}

// Work units per item: residue groups of D residues, or q single residues in
// the sequence mode (D = 0).
size_t groups(const Cls &c, const Tile &t) { return t.D ? ceilquot(c.q,t.D) : c.q; }

// Words of the class's partial buffer for tile t: G slices of the outputs, or
// q m per item in the sequence mode; 0 when the tile writes out directly.
size_t part_words(const Cls &c, const Tile &t)
{
  if(t.D == 0) return c.ids.size()*c.q*c.m;
  size_t G=groups(c,t);
  return G > 1 ? G*c.ptotal : 0;
}

// Group items by (m, q) for the sub-FFT sizes `mof`; creates twiddle tables
// and device item lists.
int build_classes(const fftwpp_amd_batch &h, const std::vector<size_t> &ids,
                  const std::vector<size_t> &mof, std::vector<Cls> &out)
{
  std::map<std::pair<size_t,size_t>,size_t> index;
  for(size_t i : ids) {
    size_t m=mof[i], q=residues(h.L[i],h.M[i],m);
    auto key=std::make_pair(m,q);
    auto it=index.find(key);
    if(it == index.end()) {
      index[key]=out.size();
      Cls c;
      c.m=m;
      c.q=q;
      out.push_back(c);
      it=index.find(key);
    }
    Cls &c=out[it->second];
    c.ids.push_back(i);
    c.Lmax=std::max(c.Lmax,h.L[i]);
    double N=(double) (m*q);
    c.work += N*std::log2(N);
  }
  for(Cls &c : out) {
    std::sort(c.ids.begin(),c.ids.end(),[&](size_t a, size_t b) {
      return h.L[a] != h.L[b] ? h.L[a] > h.L[b] : a < b;
    });
    int rc=fftwpp_amd_dim_create(&c.dim,c.Lmax,c.m,c.q,FFTWPP_AMD_KIND_COMPLEX);
    if(rc) return rc;
    std::vector<BatchItem> host(c.ids.size());
    c.ptotal=0;
    for(size_t k=0; k < c.ids.size(); ++k) {
      host[k].off=(int64_t) h.off[c.ids[k]];
      host[k].L=(int64_t) h.L[c.ids[k]];
      host[k].poff=(int64_t) c.ptotal;
      c.ptotal += h.L[c.ids[k]];
    }
    rc=bhip(hipMalloc(&c.items,host.size()*sizeof(BatchItem)),"hipMalloc batch items");
    if(rc) return rc;
    rc=bhip(hipMemcpy(c.items,host.data(),host.size()*sizeof(BatchItem),
                      hipMemcpyHostToDevice),"copy batch items");
    if(rc) return rc;
  }
  return FFTWPP_AMD_OK;
}

// Tile of TR items and D residues per work unit, with LDS accumulation
// when it fits.
Tile make_tile(const fftwpp_amd_batch &h, const Cls &c, size_t TR,
               size_t threads, size_t D)
{
  Tile t;
  t.TR=TR;
  t.threads=threads;
  t.D=std::min(D,c.q);   // D = 0: the sequence mode
  t.lacc=fftwpp_amd_batch_rows_lds(c.dim,h.dtype,TR,c.Lmax) ? c.Lmax : 0;
  size_t units=ceilquot(c.ids.size(),TR)*groups(c,t);
  // the register path runs threads/tps work units per block at once
  size_t tps=fftwpp_amd_batch_rows_tps(c.dim,h.dtype);
  size_t slots=tps ? std::max<size_t>(1,std::max(threads,tps)/tps) : 1;
  if(tps > 64) slots=1;
  t.nblocks=std::min(ceilquot(units,slots),(size_t) (8*h.cus));
  // several residue groups per item on one block each: a multiple of 8
  // blocks keeps an item's groups on one XCD (the kernel's unit mapping)
  if(tps && slots == 1 && groups(c,t) > 1) t.nblocks=ceilquot(t.nblocks,8)*8;
  return t;
}

// Residue group sizes tried: q, q/2, q/4, q/8 (rounded up), at most 8 groups,
// and for the register kernel with q > 1 the sequence mode (D = 0).
std::vector<size_t> group_sizes(size_t q, bool reg)
{
  std::vector<size_t> v;
  for(size_t G=1; G <= 8 && G <= q; G *= 2) {
    size_t D=ceilquot(q,G);
    if(std::find(v.begin(),v.end(),D) == v.end()) v.push_back(D);
  }
  if(reg && q > 1) v.push_back(0);
  return v;
}

bool tile_fits(const fftwpp_amd_batch &h, const Cls &c, size_t TR)
{
  return fftwpp_amd_batch_rows_lds(c.dim,h.dtype,TR,0) != 0;
}

// Without tuning: small tiles of short items, and enough residue groups
// for about two work units per compute unit.
Tile default_tile(const fftwpp_amd_batch &h, const Cls &c)
{
  if(size_t tps=fftwpp_amd_batch_rows_tps(c.dim,h.dtype)) {
    size_t threads=std::max<size_t>(256,tps);
    size_t slots=tps > 64 ? 1 : threads/tps;
    size_t G=std::min<size_t>({8,c.q,std::max<size_t>(1,ceilquot(2*(size_t) h.cus*slots,
                                                                  c.ids.size()))});
    return make_tile(h,c,1,threads,ceilquot(c.q,G));
  }
  size_t TR=std::max<size_t>(1,std::min<size_t>(32,2048/c.m));
  while(TR > 1 && !tile_fits(h,c,TR)) TR /= 2;
  size_t threads=std::min<size_t>(256,std::max<size_t>(64,TR*c.m/4));
  threads=(threads+63)/64*64;
  size_t tiles=ceilquot(c.ids.size(),TR);
  size_t G=std::min<size_t>({8,c.q,std::max<size_t>(1,ceilquot(2*(size_t) h.cus,tiles))});
  return make_tile(h,c,TR,threads,ceilquot(c.q,G));
}

int launch_class(const fftwpp_amd_batch &h, const Cls &c, const Tile &t,
                 void *part, void *out, const void *f, const void *g,
                 hipStream_t s)
{
  return fftwpp_amd_batch_rows(c.dim,h.dtype,h.mult,c.items,c.ids.size(),f,g,out,
                               t.TR,t.threads,t.nblocks,t.lacc,t.D,part,c.ptotal,
                               1.0/(double) (c.m*c.q),s);
}

// Partial-sum (or sequence) buffer for the class's chosen tile (none when
// the tile writes out directly).
int alloc_part(const fftwpp_amd_batch &h, Cls &c)
{
  if(c.part) { (void) hipFree(c.part); c.part=nullptr; }
  size_t n=part_words(c,c.tile);
  if(n == 0) return FFTWPP_AMD_OK;
  size_t elt=h.dtype == FFTWPP_AMD_F32 ? 8 : 16;
  return bhip(hipMalloc(&c.part,n*elt),"hipMalloc partial sums");
}

// Average time of one launch (ms), or a negative value on failure.
// A buffer larger than the Infinity Cache, overwritten before every timed
// launch so that launches are timed with cold caches, as they run in the
// batch (set while a plan is tuned).
thread_local void *scrub_buf=nullptr;
// Candidate set of tune_class: false = residue groups only (no sequence mode,
// no persistent grids).
thread_local bool tune_wide=true;
thread_local size_t scrub_bytes=0;
const size_t SCRUB_BYTES=(size_t) 320 << 20;

// Per-class timings with (true) or without (false) the cache scrub.
thread_local bool tune_cold=true;

int scrub(hipStream_t s, int k)
{
  if(!scrub_buf || !tune_cold) return FFTWPP_AMD_OK;
  return bhip(hipMemsetAsync(scrub_buf,k & 0xff,scrub_bytes,s),"scrub caches");
}

double time_launch(const fftwpp_amd_batch &h, const Cls &c, const Tile &t,
                   void *part, void *out, const void *f, const void *g,
                   hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
  if(launch_class(h,c,t,part,out,f,g,s)) return -1;
  const int reps=3;
  double total=0;
  for(int k=0; k < reps; ++k) {
    if(scrub(s,k)) return -1;
    if(hipEventRecord(e0,s) != hipSuccess) return -1;
    if(launch_class(h,c,t,part,out,f,g,s)) return -1;
    if(hipEventRecord(e1,s) != hipSuccess) return -1;
    if(hipEventSynchronize(e1) != hipSuccess) return -1;
    float ms=0;
    if(hipEventElapsedTime(&ms,e0,e1) != hipSuccess) return -1;
    total += ms;
  }
  return total/reps;
}

// Best tile of a class by timing the candidates (items per block, threads,
// residues per work unit); part holds up to 8 partial-sum slices.  Returns
// the best time.
double tune_class(const fftwpp_amd_batch &h, Cls &c, void *part, void *out,
                  const void *f, const void *g, hipStream_t s, hipEvent_t e0,
                  hipEvent_t e1)
{
  double best=-1;
  Tile bt=default_tile(h,c);
  const bool reg=fftwpp_amd_batch_rows_tps(c.dim,h.dtype) != 0;
  for(size_t D : group_sizes(c.q,reg && tune_wide)) {
    // the timing arrays hold 8 slices of the batch for partial results
    if(D == 0 && c.ids.size()*c.q*c.m > 8*h.total) continue;
    for(size_t TR=1; TR <= 64; TR *= 2) {
      if(TR > 1 && TR*c.m > 16384) break;
      if(TR > 1 && TR > c.ids.size()) break;
      if(!tile_fits(h,c,TR)) break;
      for(size_t threads=64; threads <= 512; threads *= 2) {
        if(threads > 64 && threads > TR*c.m/2) break;
        Tile t=make_tile(h,c,TR,threads,D);
        double ms=time_launch(h,c,t,part,out,f,g,s,e0,e1);
        if(ms >= 0 && (best < 0 || ms < best)) { best=ms; bt=t; }
        // persistent grids of 1, 2 or 4 blocks per CU for one-sequence blocks
        // with several units per item: fewer items in flight per XCD keep an
        // item's inputs in its L2 for the re-reads of its other residues
        const size_t tps=fftwpp_amd_batch_rows_tps(c.dim,h.dtype);
        if(!tune_wide || !reg || tps <= 64 || groups(c,t) < 2) continue;
        for(size_t per=1; per <= 4; per *= 2) {
          Tile tc=t;
          tc.nblocks=per*(size_t) h.cus;
          if(tc.nblocks >= t.nblocks) break;
          const double mc=time_launch(h,c,tc,part,out,f,g,s,e0,e1);
          if(mc >= 0 && (best < 0 || mc < best)) { best=mc; bt=tc; }
        }
      }
    }
  }
  c.tile=bt;
  c.ms=best;
  return best;
}

// Run every class: fork from s over the side streams, join back into s.
int run_classes(fftwpp_amd_batch &h, void *out, const void *f, const void *g,
                hipStream_t s)
{
  int rc=bhip(hipEventRecord(h.fork,s),"record fork");
  if(rc) return rc;
  for(int k=0; k < NSIDE; ++k)
    if((rc=bhip(hipStreamWaitEvent(h.side[k],h.fork,0),"wait fork"))) return rc;
  size_t other=0;
  for(size_t i=0; i < h.cls.size(); ++i) {
    const Cls &c=h.cls[i];
    int k;
    if(h.sched == 1) k=0;
    else if(h.sched == 2) {
      const size_t tps=fftwpp_amd_batch_rows_tps(c.dim,h.dtype);
      const size_t units=ceilquot(c.ids.size(),c.tile.TR)*groups(c,c.tile);
      const bool persistent=tps > 64 && c.tile.nblocks < units;
      k=persistent ? 0 : 1+(int) (other++ % (NSIDE-1));
    } else k=(int) (i % NSIDE);
    if((rc=launch_class(h,c,c.tile,c.part,out,f,g,h.side[k]))) return rc;
  }
  for(int k=0; k < NSIDE; ++k) {
    if((rc=bhip(hipEventRecord(h.join[k],h.side[k]),"record join"))) return rc;
    if((rc=bhip(hipStreamWaitEvent(s,h.join[k],0),"wait join"))) return rc;
  }
  return FFTWPP_AMD_OK;
}

// Stream schedule with the least time for the whole batch (its classes'
// tiles already chosen), timed on the tuning arrays.
int tune_schedule(fftwpp_amd_batch &h, void *out, const void *f, const void *g,
                  hipEvent_t e0, hipEvent_t e1, double *tbest)
{
  double best=-1;
  int bs=0;
  for(int sc=0; sc < 3; ++sc) {
    h.sched=sc;
    int rc=run_classes(h,out,f,g,h.cap);
    double total=0;
    for(int k=0; k < 3 && !rc; ++k) {
      rc=scrub(h.cap,k);
      if(!rc) rc=bhip(hipEventRecord(e0,h.cap),"record");
      if(!rc) rc=run_classes(h,out,f,g,h.cap);
      if(!rc) rc=bhip(hipEventRecord(e1,h.cap),"record");
      if(!rc) rc=bhip(hipEventSynchronize(e1),"sync");
      float ms=0;
      if(!rc) rc=bhip(hipEventElapsedTime(&ms,e0,e1),"elapsed");
      total += ms;
    }
    if(rc) return rc;
    if(best < 0 || total < best) { best=total; bs=sc; }
  }
  h.sched=bs;
  *tbest=best;
  return FFTWPP_AMD_OK;
}

void destroy(fftwpp_amd_batch *h)
{
  for(Graph &g : h->graphs) {
    (void) hipGraphExecDestroy(g.exec);
    (void) hipGraphDestroy(g.graph);
  }
  for(Cls &c : h->cls) free_class(c);
  for(int k=0; k < NSIDE; ++k) {
    if(h->side[k]) (void) hipStreamDestroy(h->side[k]);
    if(h->join[k]) (void) hipEventDestroy(h->join[k]);
  }
  if(h->cap) (void) hipStreamDestroy(h->cap);
  if(h->fork) (void) hipEventDestroy(h->fork);
  delete h;
}

// Candidate sub-FFT sizes for the bucket whose longest item is Lmax.
std::vector<size_t> candidates(size_t Lmax, size_t Mmax, size_t maxm)
{
  std::vector<size_t> c;
  size_t e=ceilpow2(Mmax);
  if(e <= maxm) c.push_back(e);
  size_t top=std::min(ceilpow2(Lmax),maxm);
  for(int s=0; s < 3; ++s) {
    size_t m=top >> s;
    if(m < 16) break;
    if(std::find(c.begin(),c.end(),m) == c.end()) c.push_back(m);
  }
  return c;
}

// Scratch arrays for timing candidates: f, g, out and 8 partial-sum slices
// of the whole batch, zero-filled.
struct TuneArrays {
  void *buf=nullptr;
  void *f=nullptr, *g=nullptr, *out=nullptr, *part=nullptr;
  hipEvent_t e0=nullptr, e1=nullptr;
  int init(const fftwpp_amd_batch &h) {
    const size_t elt=h.dtype == FFTWPP_AMD_F32 ? 8 : 16;
    const size_t bytes=11*h.total*elt;
    int rc=bhip(hipMalloc(&buf,bytes),"hipMalloc tuning arrays");
    if(rc) return rc;
    char *b=(char *) buf;
    f=b; g=b+h.total*elt; out=b+2*h.total*elt; part=b+3*h.total*elt;
    rc=bhip(hipMemsetAsync(buf,0,bytes,h.cap),"clear tuning arrays");
    if(!rc) rc=bhip(hipEventCreate(&e0),"event");
    if(!rc) rc=bhip(hipEventCreate(&e1),"event");
    // the cache scrub buffer is optional: without it launches are timed warm
    if(!rc && hipMalloc(&scrub_buf,SCRUB_BYTES) == hipSuccess) scrub_bytes=SCRUB_BYTES;
    else { scrub_buf=nullptr; (void) hipGetLastError(); }
    return rc;
  }
  double time(const fftwpp_amd_batch &h, Cls &c) {
    return tune_class(h,c,part,out,f,g,h.cap,e0,e1);
  }
  void release(const fftwpp_amd_batch &h) {
    if(e0) (void) hipEventDestroy(e0);
    if(e1) (void) hipEventDestroy(e1);
    (void) hipStreamSynchronize(h.cap);
    if(buf) (void) hipFree(buf);
    if(scrub_buf) (void) hipFree(scrub_buf);
    scrub_buf=nullptr;
    scrub_bytes=0;
    buf=nullptr; e0=e1=nullptr;
  }
};

// Sub-FFT size per power-of-two length bucket: the candidate whose classes
// (each with its best tile) take the least total time.
int tune_geometry(fftwpp_amd_batch &h, TuneArrays &ta)
{
  const size_t maxm=BATCH_MAXM;
  std::map<size_t,std::vector<size_t>> buckets;
  for(size_t i=0; i < h.n; ++i) buckets[ceilpow2(h.L[i])].push_back(i);
  int rc=FFTWPP_AMD_OK;
  for(auto &kv : buckets) {
    const std::vector<size_t> &ids=kv.second;
    size_t Lmax=0, Mmax=0;
    for(size_t i : ids) { Lmax=std::max(Lmax,h.L[i]); Mmax=std::max(Mmax,h.M[i]); }
    double best=-1;
    size_t bestm=0;
    for(size_t m : candidates(Lmax,Mmax,maxm)) {
      std::vector<size_t> mof(h.n,0);
      for(size_t i : ids) mof[i]=m;
      std::vector<Cls> cls;
      rc=build_classes(h,ids,mof,cls);
      double total=0;
      for(Cls &c : cls) {
        if(rc) break;
        double t=ta.time(h,c);
        if(t < 0) { total=-1; break; }
        total += t;
      }
      for(Cls &c : cls) free_class(c);
      if(rc) return rc;
      if(total >= 0 && (best < 0 || total < best)) { best=total; bestm=m; }
    }
    if(bestm == 0) return bfail(FFTWPP_AMD_ERR_HIP,"batch tuning found no working launch");
    for(size_t i : ids) h.mof[i]=bestm;
  }
  return rc;
}

// ------------------------------------------------------------------------
// Persisted tuning results (the role of FFTW wisdom, fftw++.cc:31-72, for the
// batch autotuner): a JSON file of entries keyed by a hash of (precision,
// multiplier, device architecture and CU count, the item lengths L and
// padded lengths M), each holding the chosen sub-FFT size per length bucket
// and the tile of every class.  A tuned plan whose key is present is built
// from the entry without timing; otherwise the tuner runs and appends its
// result.  The file is FFTWPP_AMD_TUNE_CACHE when set (empty: no cache),
// else $XDG_CACHE_HOME/fftwpp_amd/batch_tune.json or
// ~/.cache/fftwpp_amd/batch_tune.json.  One entry per line:
//   {"key":"<16 hex>","buckets":[[b,m],...],"tiles":[[m,q,TR,threads,D,nblocks,lacc],...],
//    "sched":s}
// ------------------------------------------------------------------------
std::string cache_path()
{
  if(const char *e=getenv("FFTWPP_AMD_TUNE_CACHE")) return e;
  std::string base;
  if(const char *x=getenv("XDG_CACHE_HOME")) base=x;
  else if(const char *hm=getenv("HOME")) base=std::string(hm)+"/.cache";
  else return "";
  return base+"/fftwpp_amd/batch_tune.json";
}

uint64_t fnv1a(uint64_t hsh, const void *data, size_t n)
{
  const unsigned char *p=(const unsigned char *) data;
  for(size_t i=0; i < n; ++i) { hsh ^= p[i]; hsh *= 1099511628211ull; }
  return hsh;
}

std::string tune_key(const fftwpp_amd_batch &h)
{
  uint64_t k=1469598103934665603ull;
  int dev=0;
  hipDeviceProp_t prop;
  memset(&prop,0,sizeof(prop));
  if(hipGetDevice(&dev) == hipSuccess) (void) hipGetDeviceProperties(&prop,dev);
  const int vals[3]={h.dtype,h.mult,h.cus};
  k=fnv1a(k,vals,sizeof(vals));
  k=fnv1a(k,prop.gcnArchName,strnlen(prop.gcnArchName,sizeof(prop.gcnArchName)));
  k=fnv1a(k,&h.n,sizeof(h.n));
  k=fnv1a(k,h.L.data(),h.n*sizeof(size_t));
  k=fnv1a(k,h.M.data(),h.n*sizeof(size_t));
  char buf[24];
  snprintf(buf,sizeof(buf),"%016llx",(unsigned long long) k);
  return buf;
}

struct TuneEntry {
  std::map<size_t,size_t> buckets;        // ceilpow2(L) -> m
  std::vector<std::vector<size_t>> tiles; // m, q, TR, threads, D, nblocks, lacc
  int sched=0;                            // stream schedule
};

std::mutex cache_mu;

// Integers of the bracketed list that follows `name` in line (nested lists
// flattened), or false when absent.
bool json_ints(const std::string &line, const char *name, std::vector<std::vector<size_t>> &out)
{
  size_t at=line.find(std::string("\"")+name+"\"");
  if(at == std::string::npos) return false;
  at=line.find('[',at);
  if(at == std::string::npos) return false;
  int depth=0;
  std::vector<size_t> cur;
  std::string num;
  for(size_t i=at; i < line.size(); ++i) {
    const char ch=line[i];
    if(ch >= '0' && ch <= '9') { num += ch; continue; }
    if(!num.empty()) { cur.push_back(strtoull(num.c_str(),nullptr,10)); num.clear(); }
    if(ch == '[') { ++depth; if(depth == 2) cur.clear(); }
    else if(ch == ']') {
      if(depth == 2) out.push_back(cur);
      if(--depth == 0) return true;
    }
  }
  return false;
}

bool cache_lookup(const std::string &key, TuneEntry &e)
{
  const std::string path=cache_path();
  if(path.empty()) return false;
  std::lock_guard<std::mutex> lock(cache_mu);
  std::ifstream in(path);
  std::string line, want="\"key\":\""+key+"\"";
  while(std::getline(in,line)) {
    if(line.find(want) == std::string::npos) continue;
    std::vector<std::vector<size_t>> b, t;
    if(!json_ints(line,"buckets",b) || !json_ints(line,"tiles",t)) continue;
    e=TuneEntry();
    for(auto &v : b) if(v.size() == 2) e.buckets[v[0]]=v[1];
    for(auto &v : t) if(v.size() == 7) e.tiles.push_back(v);
    const size_t at=line.find("\"sched\":");
    if(at != std::string::npos) e.sched=atoi(line.c_str()+at+8);
    return true;
  }
  return false;
}

void cache_store(const std::string &key, const fftwpp_amd_batch &h)
{
  const std::string path=cache_path();
  if(path.empty()) return;
  std::ostringstream os;
  os << "{\"key\":\"" << key << "\",\"buckets\":[";
  std::map<size_t,size_t> b;
  for(size_t i=0; i < h.n; ++i) b[ceilpow2(h.L[i])]=h.mof[i];
  bool first=true;
  for(auto &kv : b) {
    os << (first ? "" : ",") << "[" << kv.first << "," << kv.second << "]";
    first=false;
  }
  os << "],\"tiles\":[";
  first=true;
  for(const Cls &c : h.cls) {
    const Tile &t=c.tile;
    os << (first ? "" : ",") << "[" << c.m << "," << c.q << "," << t.TR << "," << t.threads
       << "," << t.D << "," << t.nblocks << "," << t.lacc << "]";
    first=false;
  }
  os << "],\"sched\":" << h.sched << "}\n";
  std::lock_guard<std::mutex> lock(cache_mu);
  const size_t slash=path.rfind('/');
  if(slash != std::string::npos) {
    // create the directory chain (errors surface as a failed open below)
    std::string dir=path.substr(0,slash);
    for(size_t i=1; i <= dir.size(); ++i)
      if(i == dir.size() || dir[i] == '/') (void) mkdir(dir.substr(0,i).c_str(),0755);
  }
  std::ofstream out(path,std::ios::app);
  if(out) out << os.str();
}


// ------------------------------------------------------------------------
// Timed sub-FFT sizes of single convolutions (fftwpp_amd_conv_create_tuned).
// The reference's optimizer times candidate parameters (m, D, I) on the CPU
// when a plan is made (fftBase::OptBase, convolve.h:164-236, convolve.cc:239-470);
// on the GPU the kernels fix D and
// the in-place choice, so the sizes m of the dimensions are what is timed.
// ------------------------------------------------------------------------

// Smallest 2^a 3^b 5^c 7^d >= n (the sizes the LDS engine transforms).
size_t next_smooth(size_t n)
{
  for(size_t v=std::max<size_t>(n,1);; ++v) {
    size_t r=v;
    for(size_t f : {2,3,5,7}) while(r % f == 0) r /= f;
    if(r == 1) return v;
  }
}

// Code of the 1D plan that has dimension d's own rule (fftwpp_amd_conv_create
// kinds: 0 complex, 1 centred Hermitian, 2 centred complex, 3 real).
int dim_kind(int hermitian, int dim, int d)
{
  if(hermitian == 1) return d == dim-1 ? 1 : 2;
  if(hermitian == 3) return d == 0 ? 3 : 0;
  return hermitian;
}

// Element bytes and stored words of one input of a plan.
size_t conv_words(int dim, int hermitian, const size_t *L)
{
  size_t w=1;
  for(int d=0; d < dim; ++d)
    w *= hermitian == 1 && d == dim-1 ? (L[d]+1)/2 : L[d];
  return w;
}

size_t conv_elt(int hermitian, int dtype)
{
  const size_t c=dtype == FFTWPP_AMD_F64 ? 16 : 8;
  return hermitian == 3 ? c/2 : c;
}

std::string conv_key(int dim, int hermitian, const size_t *L, const size_t *M, int dtype)
{
  uint64_t k=1469598103934665603ull;
  int dev=0;
  hipDeviceProp_t prop;
  memset(&prop,0,sizeof(prop));
  if(hipGetDevice(&dev) == hipSuccess) (void) hipGetDeviceProperties(&prop,dev);
  const int vals[4]={dim,hermitian,dtype,prop.multiProcessorCount};
  k=fnv1a(k,vals,sizeof(vals));
  k=fnv1a(k,prop.gcnArchName,strnlen(prop.gcnArchName,sizeof(prop.gcnArchName)));
  size_t lm[6]={0,0,0,0,0,0};
  for(int d=0; d < dim; ++d) { lm[d]=L[d]; lm[3+d]=M ? M[d] : 0; }
  k=fnv1a(k,lm,sizeof(lm));
  char buf[24];
  snprintf(buf,sizeof(buf),"c%016llx",(unsigned long long) k);
  return buf;
}

// The cached sizes of a convolution key ("m":[...] of its line), if any.
bool conv_cache_lookup(const std::string &key, std::vector<size_t> &m)
{
  const std::string path=cache_path();
  if(path.empty()) return false;
  std::lock_guard<std::mutex> lock(cache_mu);
  std::ifstream in(path);
  std::string line, want="\"key\":\""+key+"\"";
  while(std::getline(in,line)) {
    if(line.find(want) == std::string::npos) continue;
    const size_t at=line.find("\"m\":[");
    if(at == std::string::npos) continue;
    m.clear();
    const char *c=line.c_str()+at+5;
    while(*c && *c != ']') {
      char *end;
      const unsigned long long v=strtoull(c,&end,10);
      if(end == c) break;
      m.push_back((size_t) v);
      c=*end == ',' ? end+1 : end;
    }
    return !m.empty();
  }
  return false;
}

void conv_cache_store(const std::string &key, const size_t *m, int dim, double ms)
{
  const std::string path=cache_path();
  if(path.empty()) return;
  std::ostringstream os;
  os << "{\"key\":\"" << key << "\",\"m\":[";
  for(int d=0; d < dim; ++d) os << (d ? "," : "") << m[d];
  os << "],\"ms\":" << ms << "}\n";
  std::lock_guard<std::mutex> lock(cache_mu);
  const size_t slash=path.rfind('/');
  if(slash != std::string::npos) {
    std::string dir=path.substr(0,slash);
    for(size_t i=1; i <= dir.size(); ++i)
      if(i == dir.size() || dir[i] == '/') (void) mkdir(dir.substr(0,i).c_str(),0755);
  }
  std::ofstream out(path,std::ios::app);
  if(out) out << os.str();
}

// The sub-FFT sizes a plan runs with, into chosen (when not null).
int conv_chosen(const fftwpp_amd_conv *h, int dim, size_t *chosen)
{
  if(!chosen) return FFTWPP_AMD_OK;
  size_t N[3], m[3], q[3];
  const int rc=fftwpp_amd_conv_geometry(h,N,m,q);
  if(rc) return rc;
  for(int d=0; d < dim; ++d) chosen[d]=m[d];
  return FFTWPP_AMD_OK;
}

// Candidate sizes of dimension d: the smooth size of every split of L into
// p <= 8 blocks and the powers of two 16..4096 that cover L in at most 8
// blocks, kept when a 1D plan of the dimension's rule accepts them, ordered
// by padded length (then larger m), at most `keep` of them within 1.5 times
// the shortest padded length.  The library's own choice is always kept.
std::vector<size_t> dim_candidates(int hermitian, int dim, int d, size_t L, size_t M,
                                   int dtype, size_t mine, size_t keep)
{
  std::vector<size_t> raw;
  for(size_t p=1; p <= 8; ++p) raw.push_back(next_smooth(ceilquot(L,p)));
  for(size_t v=16; v <= 4096; v *= 2)
    if(8*v >= L) raw.push_back(v);
  raw.push_back(mine);
  std::sort(raw.begin(),raw.end());
  raw.erase(std::unique(raw.begin(),raw.end()),raw.end());
  const size_t lim=std::max<size_t>(fftwpp_amd_max_m(dtype),4096);
  std::vector<std::pair<size_t,size_t>> ok;           // (N, m)
  for(size_t m : raw) {
    if(m < 2 || m > lim) continue;              // m = 1 means "choose" to Application
    fftwpp_amd_conv *h=nullptr;
    const size_t Ld=L, Md=M, md=m;
    if(fftwpp_amd_conv_create(&h,1,dim_kind(hermitian,dim,d),&Ld,&Md,&md,dtype) != FFTWPP_AMD_OK)
      continue;
    size_t N[3], mm[3], q[3];
    if(fftwpp_amd_conv_geometry(h,N,mm,q) == FFTWPP_AMD_OK && mm[0] == m) ok.push_back({N[0],m});
    fftwpp_amd_conv_destroy(h);
  }
  std::sort(ok.begin(),ok.end(),[](const std::pair<size_t,size_t> &a,
                                   const std::pair<size_t,size_t> &b) {
    return a.first < b.first || (a.first == b.first && a.second > b.second);
  });
  std::vector<size_t> out{mine};
  for(const auto &c : ok) {
    if(out.size() >= keep) break;
    if(c.second == mine || 2*c.first > 3*ok[0].first) continue;
    out.push_back(c.second);
  }
  return out;
}

} // namespace

extern "C" {

int fftwpp_amd_batch_create(fftwpp_amd_batch **out, size_t n, const size_t *L,
                            const size_t *M, int dtype, int mult, int flags)
{
  if(!out) return bfail(FFTWPP_AMD_ERR_ARG,"null output handle");
  *out=nullptr;
  if(n == 0 || !L) return bfail(FFTWPP_AMD_ERR_ARG,"empty batch");
  if(dtype != FFTWPP_AMD_F64 && dtype != FFTWPP_AMD_F32)
    return bfail(FFTWPP_AMD_ERR_ARG,"unknown dtype");
  if(mult != FFTWPP_AMD_MULT_BINARY && mult != FFTWPP_AMD_MULT_CORRELATION &&
     mult != FFTWPP_AMD_MULT_NONE)
    return bfail(FFTWPP_AMD_ERR_UNSUPPORTED,"batched convolutions take multNone, multBinary or multcorrelation");
  int ndev=0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bfail(FFTWPP_AMD_ERR_NODEVICE,"no HIP device");
  fftwpp_amd_batch *h=new fftwpp_amd_batch;
  h->dtype=dtype;
  h->mult=mult;
  h->flags=flags;
  h->n=n;
  h->L.assign(L,L+n);
  h->M.resize(n);
  h->off.resize(n+1);
  h->off[0]=0;
  for(size_t i=0; i < n; ++i) {
    if(L[i] == 0 || L[i] > ((size_t) 1 << 30)) {
      destroy(h);
      return bfail(FFTWPP_AMD_ERR_ARG,"item length out of range");
    }
    // default M = 2L - 1 (wrappers/HybridConvolution.h:19)
    h->M[i]=M ? M[i] : 2*L[i]-1;
    if(h->M[i] < L[i] || h->M[i] > ((size_t) 1 << 31)) {
      destroy(h);
      return bfail(FFTWPP_AMD_ERR_ARG,"padded length must satisfy L <= M <= 2^31");
    }
    h->off[i+1]=h->off[i]+L[i];
  }
  h->total=h->off[n];
  int dev=0;
  int rc=bhip(hipGetDevice(&dev),"hipGetDevice");
  if(!rc) rc=bhip(hipDeviceGetAttribute(&h->cus,hipDeviceAttributeMultiprocessorCount,dev),
                  "device attribute");
  if(!rc) rc=bhip(hipStreamCreateWithFlags(&h->cap,hipStreamNonBlocking),"stream");
  for(int k=0; k < NSIDE && !rc; ++k) {
    rc=bhip(hipStreamCreateWithFlags(&h->side[k],hipStreamNonBlocking),"stream");
    if(!rc) rc=bhip(hipEventCreateWithFlags(&h->join[k],hipEventDisableTiming),"event");
  }
  if(!rc) rc=bhip(hipEventCreateWithFlags(&h->fork,hipEventDisableTiming),"event");
  if(rc) { destroy(h); return rc; }
  const size_t maxm=BATCH_MAXM;
  h->mof.resize(n);
  for(size_t i=0; i < n; ++i) h->mof[i]=std::min(ceilpow2(h->L[i]),maxm);
  const bool tuned=flags & FFTWPP_AMD_BATCH_TUNE;
  TuneArrays ta;
  bool timing=false;                      // ta holds the timing arrays
  // a stored tuning result for this batch, if any
  std::string key;
  TuneEntry entry;
  bool cached=false;
  if(tuned) {
    key=tune_key(*h);
    cached=cache_lookup(key,entry);
    for(size_t i=0; i < n && cached; ++i) {
      auto it=entry.buckets.find(ceilpow2(h->L[i]));
      if(it == entry.buckets.end() || it->second == 0 || it->second > maxm) cached=false;
      else h->mof[i]=it->second;
    }
    if(!cached) {
      for(size_t i=0; i < n; ++i) h->mof[i]=std::min(ceilpow2(h->L[i]),maxm);
      rc=ta.init(*h);
      timing=true;
      if(!rc) rc=tune_geometry(*h,ta);
      if(rc) { ta.release(*h); destroy(h); return rc; }
    }
  }
  std::vector<size_t> all(n);
  for(size_t i=0; i < n; ++i) all[i]=i;
  rc=build_classes(*h,all,h->mof,h->cls);
  std::sort(h->cls.begin(),h->cls.end(),[](const Cls &a, const Cls &b) { return a.work > b.work; });
  bool complete=true;                     // every class's tile came from the cache
  for(Cls &c : h->cls) {
    if(rc) break;
    bool have=false;
    if(tuned && cached) {
      for(const auto &v : entry.tiles) {
        if(v[0] != c.m || v[1] != c.q) continue;
        Tile t=make_tile(*h,c,std::max<size_t>(1,v[2]),v[3],v[4]);
        if(v[5]) t.nblocks=v[5];
        c.tile=t;
        c.ms=-1;
        have=true;
        break;
      }
    }
    if(!tuned) c.tile=default_tile(*h,c);
    else if(!have) {
      complete=false;
      if(!timing) { rc=ta.init(*h); timing=true; }
      if(!rc && ta.time(*h,c) < 0) rc=bfail(FFTWPP_AMD_ERR_HIP,"batch tuning found no working launch");
    }
    if(!rc) rc=alloc_part(*h,c);
  }
  if(tuned && cached && complete) h->sched=entry.sched;
  else if(tuned && !rc) {
    // The classes run concurrently in the batch, which per-class timings do
    // not see (nor how warm the caches are then): the tiles are chosen four
    // ways -- the wide candidate set or residue groups only, each timed with
    // cold or warm caches -- and the whole batch is timed with each choice;
    // the fastest is kept.
    if(!timing) { rc=ta.init(*h); timing=true; }
    double best=-1;
    std::vector<Tile> bt;
    std::vector<double> bms;
    int bsched=0;
    for(int way=0; way < 4 && !rc; ++way) {
      tune_wide=way < 2;
      tune_cold=way % 2 == 0;
      if(way > 0) {                          // way 0: the tiles tuned above
        for(Cls &c : h->cls) {
          if(rc) break;
          if(ta.time(*h,c) < 0) rc=bfail(FFTWPP_AMD_ERR_HIP,"batch tuning found no working launch");
          if(!rc) rc=alloc_part(*h,c);
        }
      }
      tune_cold=true;
      double t=-1;
      if(!rc) rc=tune_schedule(*h,ta.out,ta.f,ta.g,ta.e0,ta.e1,&t);
      if(!rc && (best < 0 || t < best)) {
        best=t;
        bsched=h->sched;
        bt.clear();
        bms.clear();
        for(const Cls &c : h->cls) { bt.push_back(c.tile); bms.push_back(c.ms); }
      }
    }
    tune_wide=true;
    tune_cold=true;
    for(size_t i=0; i < h->cls.size() && !rc; ++i) {
      h->cls[i].tile=bt[i];
      h->cls[i].ms=bms[i];
      rc=alloc_part(*h,h->cls[i]);
    }
    h->sched=bsched;
  }
  if(timing) ta.release(*h);
  if(rc) { destroy(h); return rc; }
  if(tuned && !(cached && complete)) cache_store(key,*h);
  *out=h;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_batch_destroy(fftwpp_amd_batch *h)
{
  if(!h) return FFTWPP_AMD_OK;
  (void) hipDeviceSynchronize();
  destroy(h);
  return FFTWPP_AMD_OK;
}

size_t fftwpp_amd_batch_size(const fftwpp_amd_batch *h)
{
  return h ? h->total : 0;
}

int fftwpp_amd_batch_offsets(const fftwpp_amd_batch *h, size_t *off)
{
  if(!h || !off) return bfail(FFTWPP_AMD_ERR_ARG,"null argument");
  memcpy(off,h->off.data(),(h->n+1)*sizeof(size_t));
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_batch_geometry(const fftwpp_amd_batch *h, size_t i, size_t *m,
                              size_t *q, size_t *N)
{
  if(!h || i >= h->n) return bfail(FFTWPP_AMD_ERR_ARG,"bad batch item");
  size_t mm=h->mof[i], qq=residues(h->L[i],h->M[i],mm);
  if(m) *m=mm;
  if(q) *q=qq;
  if(N) *N=mm*qq;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_batch_describe(const fftwpp_amd_batch *h, char *buf, size_t len)
{
  if(!h || !buf || len == 0) return bfail(FFTWPP_AMD_ERR_ARG,"null argument");
  std::string s;
  char line[160];
  for(const Cls &c : h->cls) {
    snprintf(line,sizeof(line),
             "m=%zu q=%zu items=%zu Lmax=%zu TR=%zu threads=%zu D=%zu blocks=%zu lacc=%zu ms=%.4f\n",
             c.m,c.q,c.ids.size(),c.Lmax,c.tile.TR,c.tile.threads,c.tile.D,c.tile.nblocks,
             c.tile.lacc,c.ms);
    s += line;
  }
  snprintf(line,sizeof(line),"schedule=%d\n",h->sched);
  s += line;
  size_t k=std::min(len-1,s.size());
  memcpy(buf,s.data(),k);
  buf[k]=0;
  return FFTWPP_AMD_OK;
}

int fftwpp_amd_batch_execute(fftwpp_amd_batch *h, void *out, const void *f,
                             const void *g, void *stream)
{
  if(!h || !out || !f || !g) return bfail(FFTWPP_AMD_ERR_ARG,"null argument");
  if(!fftwpp_amd_is_device_pointer(out) || !fftwpp_amd_is_device_pointer(f) ||
     !fftwpp_amd_is_device_pointer(g))
    return bfail(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_batch_execute: arguments must be device pointers");
  if(out == f || out == g)
    return bfail(FFTWPP_AMD_ERR_ARG,"fftwpp_amd_batch_execute: out must not be f or g");
  hipStream_t s=(hipStream_t) stream;
  if(!(h->flags & FFTWPP_AMD_BATCH_GRAPH)) return run_classes(*h,out,f,g,s);
  for(size_t i=0; i < h->graphs.size(); ++i) {
    Graph &gr=h->graphs[i];
    if(gr.out == out && gr.f == f && gr.g == g) {
      // most recently used last
      std::rotate(h->graphs.begin()+i,h->graphs.begin()+i+1,h->graphs.end());
      return bhip(hipGraphLaunch(h->graphs.back().exec,s),"hipGraphLaunch");
    }
  }
  // first call on these arrays: run eagerly, then record the graph
  int rc=run_classes(*h,out,f,g,s);
  if(rc) return rc;
  Graph gr;
  gr.out=out; gr.f=f; gr.g=g;
  rc=bhip(hipStreamBeginCapture(h->cap,hipStreamCaptureModeThreadLocal),"begin capture");
  if(rc) return rc;
  int rrc=run_classes(*h,out,f,g,h->cap);
  rc=bhip(hipStreamEndCapture(h->cap,&gr.graph),"end capture");
  if(rrc) {
    if(!rc) (void) hipGraphDestroy(gr.graph);
    return rrc;
  }
  if(rc) return rc;
  rc=bhip(hipGraphInstantiate(&gr.exec,gr.graph,nullptr,nullptr,0),"hipGraphInstantiate");
  if(rc) { (void) hipGraphDestroy(gr.graph); return rc; }
  if(h->graphs.size() >= MAXGRAPHS) {
    // the least recently used graph goes, after any launch of it, on
    // whichever stream, has completed
    (void) hipDeviceSynchronize();
    Graph &old=h->graphs.front();
    (void) hipGraphExecDestroy(old.exec);
    (void) hipGraphDestroy(old.graph);
    h->graphs.erase(h->graphs.begin());
  }
  h->graphs.push_back(gr);
  return FFTWPP_AMD_OK;
}


/* Plan creation with timed sub-FFT sizes: see fftwpp_amd.h. */
int fftwpp_amd_conv_create_tuned(fftwpp_amd_conv **out, int dim, int hermitian,
                                 const size_t *L, const size_t *M, int dtype,
                                 int flags, size_t *chosen)
{
  if(!out || !L) return bfail(FFTWPP_AMD_ERR_ARG,"null argument");
  *out=nullptr;
  if(dim < 1 || dim > 3) return bfail(FFTWPP_AMD_ERR_ARG,"dim must be 1, 2 or 3");
  const std::string key=conv_key(dim,hermitian,L,M,dtype);
  std::vector<size_t> cached;
  if(!(flags & FFTWPP_AMD_TUNE_RETIME) && conv_cache_lookup(key,cached) &&
     (int) cached.size() == dim &&
     fftwpp_amd_conv_create(out,dim,hermitian,L,M,cached.data(),dtype) == FFTWPP_AMD_OK)
    return conv_chosen(*out,dim,chosen);
  // the library's own choice: the first candidate and the plan to beat
  fftwpp_amd_conv *base=nullptr;
  int rc=fftwpp_amd_conv_create(&base,dim,hermitian,L,M,nullptr,dtype);
  if(rc) return rc;
  size_t N0[3], m0[3], q0[3];
  rc=fftwpp_amd_conv_geometry(base,N0,m0,q0);
  fftwpp_amd_conv_destroy(base);
  if(rc) return rc;
  const size_t keep=dim == 1 ? 6 : (dim == 2 ? 4 : 3);
  std::vector<std::vector<size_t>> cand(dim);
  for(int d=0; d < dim; ++d) {
    cand[d]=dim_candidates(hermitian,dim,d,L[d],M ? M[d] : 0,dtype,m0[d],keep);
    if(getenv("FFTWPP_AMD_TUNE_VERBOSE")) {
      fprintf(stderr,"fftwpp_amd tune: dimension %d candidates",d);
      for(size_t m : cand[d]) fprintf(stderr," %zu",m);
      fprintf(stderr,"\n");
    }
  }
  // device inputs (finite bit pattern) shared by every candidate
  const size_t bytes=conv_words(dim,hermitian,L)*conv_elt(hermitian,dtype);
  void *a=nullptr, *b=nullptr, *o=nullptr;
  hipStream_t s=nullptr;
  hipEvent_t e0=nullptr, e1=nullptr;
  rc=bhip(hipMalloc(&a,bytes),"tune inputs");
  if(!rc) rc=bhip(hipMalloc(&b,bytes),"tune inputs");
  if(!rc) rc=bhip(hipMalloc(&o,bytes),"tune output");
  if(!rc) rc=bhip(hipMemset(a,0x3e,bytes),"tune inputs");
  if(!rc) rc=bhip(hipMemset(b,0x3d,bytes),"tune inputs");
  if(!rc) rc=bhip(hipStreamCreateWithFlags(&s,hipStreamNonBlocking),"tune stream");
  if(!rc) rc=bhip(hipEventCreate(&e0),"tune event");
  if(!rc) rc=bhip(hipEventCreate(&e1),"tune event");
  // average ms of one execution of the plan with sizes m (nullptr: the
  // library's choice), or -1 when it is refused or slower than 3x `best`:
  // one untimed run (allocates the work buffers), one timed, then
  // repetitions filling about 20 ms, at least three
  const bool verbose=getenv("FFTWPP_AMD_TUNE_VERBOSE") != nullptr;
  auto show=[&](const size_t *m, const char *what, double ms) {
    if(!verbose) return;
    fprintf(stderr,"fftwpp_amd tune: m =");
    for(int d=0; d < dim; ++d) fprintf(stderr," %zu",m ? m[d] : m0[d]);
    if(ms >= 0) fprintf(stderr,"%s: %.4f ms\n",what,ms);
    else fprintf(stderr,"%s: %s\n",what,fftwpp_amd_last_error());
  };
  auto time_plan=[&](const size_t *m, double best) -> double {
    fftwpp_amd_conv *h=nullptr;
    if(fftwpp_amd_conv_create(&h,dim,hermitian,L,M,m,dtype) != FFTWPP_AMD_OK) {
      (void) hipGetLastError();
      show(m," refused",-1);
      return -1;
    }
    float ms=0;
    double avg=-1;
    bool good=fftwpp_amd_conv_execute_out(h,o,a,b,s) == FFTWPP_AMD_OK &&
      hipEventRecord(e0,s) == hipSuccess &&
      fftwpp_amd_conv_execute_out(h,o,a,b,s) == FFTWPP_AMD_OK &&
      hipEventRecord(e1,s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
      hipEventElapsedTime(&ms,e0,e1) == hipSuccess;
    if(good && best >= 0 && ms >= 3*best) show(m," skipped, warm-up run",ms);
    if(good && (best < 0 || ms < 3*best)) {
      const int reps=std::max(3,std::min(50,(int) (20.0/std::max(ms,0.001f))));
      good=hipEventRecord(e0,s) == hipSuccess;
      for(int i=0; good && i < reps; ++i)
        good=fftwpp_amd_conv_execute_out(h,o,a,b,s) == FFTWPP_AMD_OK;
      good=good && hipEventRecord(e1,s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms,e0,e1) == hipSuccess;
      if(good) avg=ms/reps;
    }
    (void) hipStreamSynchronize(s);
    (void) hipGetLastError();                 // a refused candidate leaves no error behind
    fftwpp_amd_conv_destroy(h);
    if(avg >= 0) show(m,m ? "" : " (library choice)",avg);
    return avg;
  };
  double best=rc ? -1 : time_plan(nullptr,-1);
  bool chose=false;                           // a candidate beat the library's choice
  size_t bestm[3]={m0[0],m0[1],m0[2]};
  size_t idx[3]={0,0,0};
  for(bool more=!rc; more; ) {
    size_t m[3]={0,0,0};
    bool same=true;
    for(int d=0; d < dim; ++d) {
      m[d]=cand[d][idx[d]];
      same=same && m[d] == m0[d];
    }
    if(!same) {
      const double t=time_plan(m,best);
      if(t >= 0 && (best < 0 || t < best)) {
        best=t;
        chose=true;
        for(int d=0; d < dim; ++d) bestm[d]=m[d];
      }
    }
    more=false;
    for(int d=dim-1; d >= 0; --d) {
      if(++idx[d] < cand[d].size()) { more=true; break; }
      idx[d]=0;
    }
  }
  if(e0) (void) hipEventDestroy(e0);
  if(e1) (void) hipEventDestroy(e1);
  if(s) (void) hipStreamDestroy(s);
  if(a) (void) hipFree(a);
  if(b) (void) hipFree(b);
  if(o) (void) hipFree(o);
  if(rc) return rc;
  if(best < 0) return bfail(FFTWPP_AMD_ERR_HIP,"convolution tuning found no working plan");
  rc=fftwpp_amd_conv_create(out,dim,hermitian,L,M,chose ? bestm : nullptr,dtype);
  if(rc) return rc;
  // zeros record that the library's own choice was fastest
  const size_t none[3]={0,0,0};
  conv_cache_store(key,chose ? bestm : none,dim,best);
  return conv_chosen(*out,dim,chosen);
}

} // extern "C"
