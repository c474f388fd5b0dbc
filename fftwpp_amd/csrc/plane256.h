// plane256.h -- the inner two dimensions of a 3D convolution as one
// persistent kernel whose padded intermediates stay in the L2 cache.
//
// Shape: after the outer (x) forward residue transforms, each of the P planes
// X_a[p] (a = f, g; 256 x 256 complex, y rows of 256 contiguous words) is
// convolved as a 2D problem: y forward residues (m = 256, q = 2, p = 1,
// convolve.cc:836-1030), z forward residues, the multiplier
// (convolve.cc:26-110), z backward, y backward accumulated over residues
// (convolve.cc:1468-1613) -- the x-plane sub-convolution of
// Convolution3::convolveRaw (convolve.h:1812-1879, the inner Convolution2
// over the y-z plane).  The plane result overwrites X_h[p].
//
// Work decomposition: each of the 8 XCDs runs PlaneSub<T>::n sub-teams of
// PLANE_W = 32 blocks (one block of every sub-team per CU); sub-team s of XCD
// x works through the planes x + 8 s, x + 8 s + 8 n, ..., with one plane's
// y-transformed data E (both residues of f and g: 2 x 2 x 256 x 256 complex,
// 4 MB in FP64, 2 MB in FP32) exchanged through the XCD's own L2 instead of
// HBM:
//   A (columns): block k owns z columns [8k, 8k+8); group a (f or g) loads
//     its column of X_a once, transforms residue 0, stores E[0][a][ky][z],
//     announces it, transforms residue 1, stores E[1][a], announces it;
//   B (rows): rows (ry, ky) with ky in [8k, 8k+8), residue ry once its
//     columns are announced by every block: both z residues, multiplier,
//     inverse, residue sum (the rowconv256p layout), written over the f row;
//     then the block announces its rows;
//   C (columns): once every block's rows are in, group ry reads its columns
//     of E, inverts residue ry, and the two residues are summed in LDS and
//     written to X_h[p].
// C of plane p and A of the next plane touch only the block's own columns,
// so they need no team synchronisation; the next plane's columns are loaded
// while the last barrier is pending.  HBM sees X_f, X_g read once and X_h
// written once per plane (6 of the 39 units of E L^3 bytes that the
// separate y, z and y kernels move).
//
// Team synchronisation (no reference counterpart): a block's XCD comes from
// the XCC_ID hardware register, so a team is exactly the blocks that share an
// L2, whatever the dispatcher does; blocks take ranks by atomic tickets per
// XCD, meet once in a device-wide barrier (agent-scope fences) and check the
// team sizes.  Inside a sub-team, barriers are split: arriving means every
// thread's stores have reached L2 (s_waitcnt) and the block increments a
// monotone counter of the event; waiting means thread 0 polls the counter
// for (plane index + 1) * PLANE_W arrivals, and work that does not depend on
// other blocks runs in between.  Reads of E after a wait use device-scope
// (sc1) buffer loads, which bypass the CU's L1 and hit the shared L2, so no
// L2 writeback or invalidation is needed.  Every wait is bounded: a timeout
// raises the abort flag, which every waiting block also polls, so the grid
// always drains; the host reports it.  The kernel is launched cooperatively
// (all blocks resident).
#pragma once
#include "kernels256.h"

namespace fftwpp_amd {

// Control block of one launch (zeroed before it).
struct PlaneCtl {
  unsigned tk[8];        // blocks per XCD (tickets)
  unsigned arrive;       // device-wide start barrier
  unsigned abort_;       // 1: a wait timed out; 2: an XCD holds another block count
  unsigned pad[6];
  unsigned ctr[72][16];  // counters (XCD, sub-team, event), one 64-byte line each
};

// PlaneSub<T>::n sub-teams of PLANE_W blocks of PLANE_THREADS threads per
// XCD: each CU holds one block of each, working on different planes, so one
// sub-team's waits and memory latency overlap the others' arithmetic (two in
// FP64, whose 245 VGPRs allow two waves per SIMD; three in FP32).
constexpr int PLANE_THREADS=256;
constexpr int PLANE_W=32;
template<typename T> struct PlaneSub { static constexpr int n=sizeof(T) == 8 ? 2 : 3; };
constexpr int PLANE_SUBMAX=3;
constexpr size_t PLANE_E=2*2*256*256;   // exchange words (complex) per sub-team

__device__ __forceinline__ unsigned xcc_id()
{
  // s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4)
  return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4-1) << 11));
}

__device__ __forceinline__ unsigned ld_dev(unsigned *p)
{
  return __hip_atomic_load(p,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
}

// Wait until *p >= target; false on timeout (which raises *ab) or when
// another block raised *ab.
__device__ inline bool plane_wait(unsigned *p, unsigned target, unsigned *ab)
{
  for(unsigned it=0; it < (1u << 22); ++it) {
    if(ld_dev(p) >= target) return true;
    if(ld_dev(ab)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(ab,1u,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// Split-phase team barrier.  plane_arrive: every store of the block has
// reached L2, then the block counts itself on ctr.  plane_await: thread 0
// waits for `target` arrivals, the block follows.  Between the two the block
// does work that does not depend on the other blocks.
__device__ __forceinline__ void plane_arrive(unsigned *ctr)
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE,"workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if(threadIdx.x == 0)
    __hip_atomic_fetch_add(ctr,1u,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool plane_await(unsigned *ctr, unsigned target,
                                            unsigned *ab, unsigned *s_ok)
{
  if(threadIdx.x == 0) *s_ok=plane_wait(ctr,target,ab) ? 1u : 0u;
  __syncthreads();
  return *s_ok != 0;
}

// Device-scope (L1-bypassing) load of one complex word at byte offset off
// of the buffer r.
template<typename T>
__device__ __forceinline__ cplx<T> ld_l2(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff=0);
template<>
__device__ __forceinline__ cplx<double> ld_l2<double>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                      unsigned soff)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v=__builtin_bit_cast(d2,__builtin_amdgcn_raw_buffer_load_b128(r,off,soff,16));
  return mk(v.x,v.y);
}
template<>
__device__ __forceinline__ cplx<float> ld_l2<float>(__amdgpu_buffer_rsrc_t r, unsigned off,
                                                    unsigned soff)
{
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v=__builtin_bit_cast(f2,__builtin_amdgcn_raw_buffer_load_b64(r,off,soff,16));
  return mk(v.x,v.y);
}

// Column areas of the plane kernel: 8 columns per group of 128 threads.
// FP64: areas 276 words apart, index XORed by 2 for columns 4..7 (the layout
// of the m = 256 column kernels); FP32: areas 292 words apart (292 = 4 mod 32
// puts the 32 lanes (c, j < 4) of a write pass, word 292 c + 17 j + u, and of
// a read pass, 292 c + 17 t + j, on 32 distinct banks).
template<typename T> struct PlaneCol {
  static constexpr int TC=8;
  static constexpr int CS=sizeof(T) == 8 ? 276 : 292;
  __device__ static int swz(int c) { return sizeof(T) == 8 ? ((c >> 1) & 2) : 0; }
};

// Exchange-area scalar words: two column groups (phases A and C), 16 row
// groups of 16 x 17 words (phase B), or the 2048 complex sums of phase C.
template<typename T>
__host__ __device__ constexpr int plane_xwords()
{
  constexpr int cols=2*8*PlaneCol<T>::CS, rows=16*16*17, sums=2*2048*2;
  return cols > rows ? (cols > sums ? cols : sums) : (rows > sums ? rows : sums);
}

// Dynamic LDS: twiddles (768 complex), the exchange area and two flags (the
// kernel has no static LDS).
template<typename T>
inline size_t plane_lds()
{
  return (size_t) 768*sizeof(cplx<T>)+(size_t) plane_xwords<T>()*sizeof(T)+16;
}

// Planes of 256 x 256 words (Ly = lz = 256).  Sub-team s of XCD x takes
// the planes x + 8 (2 i + s); its block k owns z columns [8k, 8k+8) in
// phases A and C and rows [16k, 16k+16) of its E in phase B.
template<typename T, int MULT>
__global__ void __launch_bounds__(PLANE_THREADS,PlaneSub<T>::n)
plane256_kernel(const cplx<T> *Xf, const cplx<T> *Xg, cplx<T> *Xh,
                int64_t planes, const cplx<T> *zN, cplx<T> *E, PlaneCtl *ctl)
{
  constexpr unsigned EW=sizeof(cplx<T>);
  constexpr size_t PW=65536;               // words per plane
  extern __shared__ char smem_raw[];
  cplx<T> *rtab=reinterpret_cast<cplx<T> *>(smem_raw);   // 2 x 256
  cplx<T> *tw16=rtab+512;                                 // 256
  T *xb=reinterpret_cast<T *>(tw16+256);
  unsigned &s_rank=reinterpret_cast<unsigned *>(xb+plane_xwords<T>())[0];
  unsigned &s_ok=reinterpret_cast<unsigned *>(xb+plane_xwords<T>())[1];
  const int tid=threadIdx.x;
  fill_rtab(rtab,zN,0,2,tid,PLANE_THREADS);
  fill_tw16(tw16,zN,2,tid,PLANE_THREADS);

  // team formation: ranks by tickets per XCD, one device-wide barrier, and
  // every XCD holding exactly SUB * PLANE_W blocks
  constexpr int SUB=PlaneSub<T>::n;
  const unsigned x=xcc_id() & 7;
  if(tid == 0) {
    s_rank=__hip_atomic_fetch_add(&ctl->tk[x],1u,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    __hip_atomic_fetch_add(&ctl->arrive,1u,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
    bool ok=plane_wait(&ctl->arrive,gridDim.x,&ctl->abort_);
    __threadfence();
    for(int i=0; i < 8 && ok; ++i)
      if(ld_dev(&ctl->tk[i]) != (unsigned) (SUB*PLANE_W)) {
        ok=false;
        __hip_atomic_store(&ctl->abort_,2u,__ATOMIC_RELAXED,__HIP_MEMORY_SCOPE_AGENT);
      }
    s_ok=ok ? 1u : 0u;
  }
  __syncthreads();
  if(!s_ok || planes == 0) return;
  const int sub=(int) (s_rank % SUB), k=(int) (s_rank/SUB);
  const int team=SUB*(int) x+sub;
  // events of the sub-team: residue 0 / residue 1 columns written (A),
  // every row written (B)
  unsigned *cA0=&ctl->ctr[3*team][0];
  unsigned *cA1=&ctl->ctr[3*team+1][0];
  unsigned *cB=&ctl->ctr[3*team+2][0];
  const __amdgpu_buffer_rsrc_t re=buf_rsrc(E+(size_t) team*PLANE_E,PLANE_E);

  // column phases: group g = array a in A, residue ry in C; lane (c, j)
  const int g=tid >> 7, t=tid & 127;
  const int c=t & 7, j=t >> 3;
  T *crow=xb+(g*8+c)*PlaneCol<T>::CS;
  const int h=PlaneCol<T>::swz(c);
  const int z=8*k+c;                       // the lane's z column
  const unsigned lcol=(unsigned) (z+256*j)*EW;   // + 4096 kk words: y = j + 16 kk
  // row phase: wave w holds rows ky = 8k + 2w, 8k + 2w + 1 of residue ry
  // (pass ry)
  const int w=tid >> 6, lane=tid & 63;
  const int jr=lane & 15, grp=lane >> 4, rz=grp & 1;
  T *xrow=xb+(w*4+grp)*16*17;
  const int ky=8*k+2*w+(grp >> 1);
  unsigned target=0;

  // column z of X_a of the next plane, loaded ahead (while the previous
  // plane's last barrier is pending)
  cplx<T> x0[16];
  const int64_t p0=x+8*sub, pstep=8*SUB;
  if(p0 >= planes) return;
  {
    const __amdgpu_buffer_rsrc_t rx=buf_rsrc((g == 0 ? Xf : Xg)+p0*(int64_t) PW,PW);
#pragma unroll
    for(int kk=0; kk < 16; ++kk) x0[kk]=ld_buf<T>(rx,lcol,4096u*EW*kk);
  }
  for(int64_t p=p0; p < planes; p += pstep) {
    target += PLANE_W;
    // A: y forward residues 0 and 1 of column z of X_a (read once), to
    // E[ry][a][ky][z]; each residue announced as soon as it is stored
    {
      const cplx<T> *rt=rtab+256;
      cplx<T> v[16];
#pragma unroll
      for(int kk=0; kk < 16; ++kk) v[kk]=x0[kk];
      fft256s<T,1,0>(v,crow,j,h,tw16);
      const unsigned eo=(unsigned) (g*65536+z+256*j)*EW;
#pragma unroll
      for(int u=0; u < 16; ++u) st_buf<T>(re,eo,4096u*EW*u,v[u]);
      plane_arrive(cA0);
#pragma unroll
      for(int kk=0; kk < 16; ++kk) v[kk]=x0[kk]*(rt+16*kk)[j];
      fft256s<T,1,0>(v,crow,j,h,tw16);
#pragma unroll
      for(int u=0; u < 16; ++u) st_buf<T>(re,eo,(2u*65536u+4096u*u)*EW,v[u]);
      plane_arrive(cA1);
    }
    // B: rows (ry, ky) of E: both z residues of f and g, multiplier,
    // inverse, residue sum over the group pair, written over the f row
    {
      const cplx<T> *rt=rtab+256*rz;
#pragma unroll 1
      for(int ry=0; ry < 2; ++ry) {
        if(!plane_await(ry == 0 ? cA0 : cA1,target,&ctl->abort_,&s_ok)) return;
        const unsigned frow=(unsigned) ((ry*2*256+ky)*256+jr)*EW;
        // the g row is loaded after the f row's transform (16 loads in
        // flight, not 32: FP64 then fits 256 VGPRs)
        cplx<T> av[16], bv[16];
#pragma unroll
        for(int kk=0; kk < 16; ++kk) av[kk]=ld_l2<T>(re,frow,16u*EW*kk)*(rt+16*kk)[jr];
        fft256s<T,1,1,17>(av,xrow,jr,0,tw16);
#pragma unroll
        for(int kk=0; kk < 16; ++kk) bv[kk]=ld_l2<T>(re,frow,(65536u+16u*kk)*EW)*(rt+16*kk)[jr];
        fft256s<T,1,1,17>(bv,xrow,jr,0,tw16);
#pragma unroll
        for(int kk=0; kk < 16; ++kk) {
          if(MULT == 1) av[kk]=av[kk]*bv[kk];
          else if(MULT == 3) av[kk]=mulc(av[kk],bv[kk]);
        }
        fft256s<T,-1,1,17>(av,xrow,jr,0,tw16);
#pragma unroll
        for(int kk=0; kk < 16; ++kk) {
          const cplx<T> y=mulc(av[kk],(rt+16*kk)[jr]);
          av[kk]=mk(pair16_sum(y.x),pair16_sum(y.y));
        }
#pragma unroll
        for(int kk=0; kk < 16; ++kk)
          if((kk >> 3) == rz) st_buf<T>(re,frow,16u*EW*kk,av[kk]);
      }
      plane_arrive(cB);
    }
    {
      // (the last plane loads itself again: x0 is then dead through phase B)
      const int64_t pn=p+pstep < planes ? p+pstep : p;
      const __amdgpu_buffer_rsrc_t rx=buf_rsrc((g == 0 ? Xf : Xg)+pn*(int64_t) PW,PW);
#pragma unroll
      for(int kk=0; kk < 16; ++kk) x0[kk]=ld_buf<T>(rx,lcol,4096u*EW*kk);
    }
    // C: y backward of column z (group g = residue ry), the two residues'
    // terms put in LDS and summed by all threads into X_h.  Every block
    // has written every row once cB counts all of them.
    if(!plane_await(cB,target,&ctl->abort_,&s_ok)) return;
    {
      const int ry=g;
      const cplx<T> *rt=rtab+256*ry;
      cplx<T> *sum=reinterpret_cast<cplx<T> *>(xb);   // 2 x 16 x 128 words after the FFTs
      cplx<T> v[16];
      const unsigned ho=(unsigned) (ry*2*65536+z+256*j)*EW;
#pragma unroll
      for(int u=0; u < 16; ++u) v[u]=ld_l2<T>(re,ho,4096u*EW*u);
      fft256s<T,-1,0>(v,crow,j,h,tw16);
      if(ry) {
#pragma unroll
        for(int kk=0; kk < 16; ++kk) v[kk]=mulc(v[kk],(rt+16*kk)[j]);
      }
#pragma unroll
      for(int kk=0; kk < 16; ++kk) sum[(ry*16+kk)*128+t]=v[kk];
      __syncthreads();
      const __amdgpu_buffer_rsrc_t rh=buf_rsrc(Xh+p*(int64_t) PW,PW);
#pragma unroll
      for(int s8=0; s8 < 8; ++s8) {
        // word i = kk 128 + t' of the block's 16 x 128 outputs: y = j' + 16 kk
        const int i=tid+PLANE_THREADS*s8, kk=i >> 7, tt=i & 127;
        const unsigned off=(unsigned) (8*k+(tt & 7)+256*((tt >> 3)+16*kk))*EW;
        st_buf<T>(rh,off,0,sum[i]+sum[2048+i]);
      }
      __syncthreads();
    }
  }
}

} // namespace fftwpp_amd
