/* fftwpp_amd.h -- C ABI of the MI355X (gfx950) hybrid dealiased convolution
 * library libfftwpp_amd.so.
 *
 * Three layers are declared here:
 *
 *   1. The kernel layer: padded residue transforms of one dimension, the
 *      fused row convolutions, the device multipliers and the Hermitian
 *      symmetrisers.  All pointers are device pointers; `stream` is a
 *      hipStream_t passed as void* (NULL = the default stream).  The C++
 *      drop-in classes (convolve.h) are written against this layer only.
 *   2. The convolution layer: whole 1D/2D/3D convolutions (complex,
 *      centred, centred Hermitian or real, FP64 or FP32) on device pointers,
 *      the adjoint (correlation) entry points, user device multipliers and
 *      the batched entry points.  These extend the reference C wrapper
 *      (cfftw++.h) with device pointers, a precision choice, status codes,
 *      batching, general multipliers and adjoints.
 *   3. Standalone FFTs and transposes (the transform classes of fftw++.h).
 *
 * Every function returns 0 (FFTWPP_AMD_OK) on success or a negative status
 * code; fftwpp_amd_last_error() describes the last failure of the calling
 * thread.  Nothing here calls exit(), unlike the reference
 * (convolve.cc:480-501, fftw++.h:251-254).
 */
#ifndef FFTWPP_AMD_H
#define FFTWPP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FFTWPP_AMD_OK = 0,
  FFTWPP_AMD_ERR_ARG = -1,          /* invalid argument / parameters     */
  FFTWPP_AMD_ERR_HIP = -2,          /* HIP runtime failure (OOM, launch) */
  FFTWPP_AMD_ERR_UNSUPPORTED = -3,  /* valid but not implemented here    */
  FFTWPP_AMD_ERR_NODEVICE = -4      /* no GPU visible                    */
};

enum { FFTWPP_AMD_F64 = 0, FFTWPP_AMD_F32 = 1 };

/* fftPad, fftPadCentered, fftPadHermitian (convolve.h:471-803) */
enum {
  FFTWPP_AMD_KIND_COMPLEX = 0,
  FFTWPP_AMD_KIND_CENTERED = 1,
  FFTWPP_AMD_KIND_HERMITIAN = 2
};

/* multNone, multBinary, realMultBinary, multcorrelation (convolve.h:81) */
enum {
  FFTWPP_AMD_MULT_NONE = 0,
  FFTWPP_AMD_MULT_BINARY = 1,
  FFTWPP_AMD_MULT_REAL_BINARY = 2,
  FFTWPP_AMD_MULT_CORRELATION = 3
};

const char *fftwpp_amd_last_error(void);

/* ---------------------------------------------------------------------- */
/* 1. Kernel layer                                                         */
/* ---------------------------------------------------------------------- */

typedef struct fftwpp_amd_dim fftwpp_amd_dim;

/* One padded dimension: L data values padded to N = m*q (residues q), with
 * FP64-generated twiddles zeta_N^j on the device.  Replaces the per-fftBase
 * tables initZetaqm/Zetaqp/ZetaqmS (convolve.cc:126-140,613-636) and the
 * FFTW plans of fftPad::init (convolve.cc:511-719). */
int fftwpp_amd_dim_create(fftwpp_amd_dim **d, size_t L, size_t m, size_t q,
                          int kind);
int fftwpp_amd_dim_destroy(fftwpp_amd_dim *d);

/* Largest m the fused LDS row kernel accepts for a precision. */
size_t fftwpp_amd_max_m(int dtype);

/* Forward residues [r0, r0+nres) of nbatch x C columns.
 * Input element (b, j, c) at f[b*bstride_in + j*S + c], j < L.
 * Output F[b*bstride_out + ((r-r0)*m + s)*SF + c] = sum_k f_k zeta_N^{k(q s+r)}.
 * Replaces fftPad::forward*Many / fftPadCentered::forward*Many
 * (convolve.cc:771-1466, 2199-4307). */
int fftwpp_amd_forward_many(const fftwpp_amd_dim *d, int dtype,
                            const void *f, size_t S, size_t C,
                            size_t nbatch, size_t bstride_in, void *F,
                            size_t SF, size_t bstride_out, size_t r0,
                            size_t nres, void *stream);

/* Backward residues [r0, r0+nres): f (+)= sum_r inverse(F_r) conj-twiddled;
 * `accumulate` adds to the existing output; `scale` multiplies the final
 * value.  Replaces fftPad::backward*Many (convolve.cc:1468-1965). */
int fftwpp_amd_backward_many(const fftwpp_amd_dim *d, int dtype,
                             const void *F, size_t SF, size_t bstride_in,
                             void *f, size_t S, size_t C, size_t nbatch,
                             size_t bstride_out, size_t r0, size_t nres,
                             int accumulate, double scale, void *stream);

/* Real residue transforms (fftPadReal::forward / backward and their Many,
 * Inner and Explicit variants, convolve.cc:5449-7483, declared at
 * convolve.h:805-980) for a non-centred dimension with power-of-two m in
 * [16, 4096] and L <= 2m.  Real data carry conjugate transforms at padded
 * indices l and N - l, so only fftwpp_amd_real_planes(d) = N/2 + 1 complex
 * planes are kept, residue-major: residue 0 at s = 0..m/2, residues
 * 0 < r < q/2 at s = 0..m-1, and (q even) residue q/2 at s = 0..m/2-1; kept
 * plane base(r) + s holds sum_j f_j exp(2 pi i j (q s + r)/N).
 * fftwpp_amd_real_planes returns 0 for geometries these kernels do not take.
 *
 * Forward: f real (element j of column c of batch b at f[b bstride_in + c +
 * j S], real words), F complex (plane k at F[b bstride_out + c + k SF]).
 * Backward: f_j = scale * sum over kept l of c_l Re(F_l exp(-2 pi i j l/N)),
 * c_l = 1 for the self-conjugate planes l = 0 and l = N/2, else 2 (a kept
 * plane stands for itself and its conjugate partner), written for j < L. */
size_t fftwpp_amd_real_planes(const fftwpp_amd_dim *d);
int fftwpp_amd_forward_real_many(const fftwpp_amd_dim *d, int dtype,
                                 const void *f, size_t S, size_t C,
                                 size_t nbatch, size_t bstride_in, void *F,
                                 size_t SF, size_t bstride_out, void *stream);
int fftwpp_amd_backward_real_many(const fftwpp_amd_dim *d, int dtype,
                                  const void *F, size_t SF, size_t bstride_in,
                                  void *f, size_t S, size_t C, size_t nbatch,
                                  size_t bstride_out, double scale, void *stream);

/* Fused innermost-dimension convolution of `rows` contiguous rows: every
 * residue forward for both inputs, the multiplier, backward, accumulate.
 * Row i of f/g starts at i*rowstride, of out at i*outstride (out may equal
 * f).  Replaces Convolution::convolveRaw (convolve.cc:7513-7575) for the
 * device multipliers. */
int fftwpp_amd_conv_rows(const fftwpp_amd_dim *d, int dtype, int mult,
                         const void *f, const void *g, size_t rows,
                         size_t rowstride, void *out, size_t outstride,
                         double scale, void *stream);

/* The y-z sub-convolution of every x-plane of a 3D convolution in one
 * persistent cooperative launch: planes of Ly x lz complex words (y rows of
 * lz contiguous words, plane p at p*Ly*lz) of Xf and Xg are convolved over
 * their y (dy) and z (dz) dimensions -- forward residues, `mult`, backward
 * residues summed -- and written to Xh (which may equal Xf), unscaled.  This
 * is the x-plane loop of Convolution3::convolveRaw (convolve.h:1812-1879)
 * with its inner Convolution2, for dy and dz complex with m = 256, q = 2,
 * L = 256 and mult multNone, multBinary or multcorrelation.  The padded
 * y-transformed planes stay in the L2 cache of the XCD whose blocks share
 * them; `work` holds fftwpp_amd_plane_work(dtype) bytes (exchange areas and
 * a control block) and must not be shared by launches that overlap; the
 * launch is not capturable into a graph (its control block is cleared by a
 * memset).  fftwpp_amd_plane_supported() says whether the geometry and the
 * device take the kernel (8 XCDs of 32 CUs; the first call on a device
 * synchronises it); fftwpp_amd_plane_status() synchronises `stream` and
 * returns the launch's abort code (0 = completed, 1 = a team wait timed
 * out, 2 = an XCD held another number of blocks than the kernel needs). */
size_t fftwpp_amd_plane_work(int dtype);
int fftwpp_amd_plane_supported(const fftwpp_amd_dim *dy, const fftwpp_amd_dim *dz);
int fftwpp_amd_plane_conv(const fftwpp_amd_dim *dy, const fftwpp_amd_dim *dz,
                          int dtype, int mult, const void *Xf, const void *Xg,
                          void *Xh, size_t planes, size_t Ly, size_t lz,
                          void *work, void *stream);
int fftwpp_amd_plane_status(void *work, void *stream);

/* Hermitian rows longer than one block of m (H = ceil(L/2) > m): the
 * fused row convolution of fftwpp_amd_conv_rows with realMultBinary, for
 * power-of-two m in [16, 4096].  Each residue pair (fftPadHermitian's D = 2
 * passes, convolve.cc:4517-5012) of each row runs as an independent work
 * unit writing its contribution into `part` (fftwpp_amd_hconv_rows_long_part
 * words, 0 when d is not such a dimension); a second launch adds the passes
 * in order into out (out may equal f).  Replaces the residue loop of
 * Convolution::convolveRaw (convolve.cc:7513-7575) for these rows. */
size_t fftwpp_amd_hconv_rows_long_part(const fftwpp_amd_dim *d, size_t rows);
int fftwpp_amd_hconv_rows_long(const fftwpp_amd_dim *d, int dtype, const void *f,
                               const void *g, size_t rows, size_t rowstride,
                               void *out, size_t outstride, void *part,
                               double scale, void *stream);

/* Fused convolutions of `count` independent items that share the geometry
 * (m, q) of d.  items is a device array of {int64 off, int64 L, int64 poff}
 * triples: item i convolves f[off..off+L) with g[off..off+L) into
 * out[off..off+L), which is exact when N = m q >= 2L - 1.  Tiles of TR items
 * per block, `threads` per block (a multiple of 64, at most 1024), `nblocks`
 * blocks looping over the work units; lacc > 0 accumulates the residues in
 * LDS rows of lacc >= max L words, lacc = 0 in the destination.  D residues
 * per work unit: D >= q writes out directly; D < q writes G = ceil(q/D)
 * partial sums into part (G slices of ptotal words, item i at poff) and adds
 * them into out in a second launch; D = 0 (register path only) makes each
 * residue a work unit that writes its inverse-transformed sequence (q m
 * words per item in part, item i at i q m), and the second launch folds the
 * sequences back onto the item's blocks.  Each item follows the residue loop of
 * Convolution::convolveRaw (convolve.cc:7513-7575); the batching itself has
 * no reference counterpart.  fftwpp_amd_batch_rows_lds gives the dynamic
 * LDS bytes of a tile, 0 when it does not fit. */
size_t fftwpp_amd_batch_rows_lds(const fftwpp_amd_dim *d, int dtype,
                                 size_t TR, size_t lacc);
/* Threads per sequence of the register-resident FFT path, taken for
 * power-of-two m in [16, 4096] (TR and lacc are then ignored, `threads` is
 * rounded to a multiple of it, at most 512, one sequence per block when it
 * exceeds 64); 0 when the LDS path runs. */
size_t fftwpp_amd_batch_rows_tps(const fftwpp_amd_dim *d, int dtype);
int fftwpp_amd_batch_rows(const fftwpp_amd_dim *d, int dtype, int mult,
                          const void *items, size_t count, const void *f,
                          const void *g, void *out, size_t TR, size_t threads,
                          size_t nblocks, size_t lacc, size_t D, void *part,
                          size_t ptotal, double scale, void *stream);
/* fftwpp_amd_batch_rows with its partial sums (or sequences) in a fixed
 * device array of the library instead of a caller's `part` buffer.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc); it calls no allocating routine.  The array holds
 * FFTWPP_AMD_NOALLOC_SCRATCH_BYTES (16 MiB) per GPU: launches needing more
 * (G ptotal words for D < q, count q m words for D = 0, times 16 bytes in
 * FP64 or 8 in FP32) are refused with FFTWPP_AMD_ERR_ARG before anything is
 * enqueued.  Every call on a GPU shares its array, so calls that overlap on
 * one GPU (two streams, two threads) race; serialise them.  Nothing in the
 * array needs clearing between calls (every word read was written by the
 * same launch). */
#define FFTWPP_AMD_NOALLOC_SCRATCH_BYTES ((size_t) 16 << 20)
int fftwpp_amd_batch_rows_noalloc(const fftwpp_amd_dim *d, int dtype, int mult,
                                  const void *items, size_t count, const void *f,
                                  const void *g, void *out, size_t TR,
                                  size_t threads, size_t nblocks, size_t lacc,
                                  size_t D, size_t ptotal, double scale,
                                  void *stream);

/* Real convolutions (fftPadReal, convolve.h:805-980; tests/hybridconvr*.cc):
 * z[i] = f0[i] + i 2^e f1[i] for n real words of f0 and f1, and out[i] =
 * scale 2^-e Im z[i]; between the two, a complex self-convolution of z gives
 * 2 i 2^e f0 * f1 in its imaginary part.  The exponent e = round(log2(||f0|| /
 * ||f1||)) (0 when either input is zero) balances the norms so that the
 * rounding error stays relative to ||f0|| ||f1||; fftwpp_amd_real_balance
 * computes it on the device into `bal` (fftwpp_amd_real_balance_bytes()
 * bytes of device memory), which pack and unpack read (bal = NULL: e = 0).
 * The scaling by a power of two is exact; when either input is zero, unpack
 * writes exact zeros. */
size_t fftwpp_amd_real_balance_bytes(void);
int fftwpp_amd_real_balance(int dtype, const void *f0, const void *f1,
                            size_t n, void *bal, void *stream);
int fftwpp_amd_real_pack(int dtype, const void *f0, const void *f1, void *z,
                         size_t n, const void *bal, void *stream);
int fftwpp_amd_real_unpack(int dtype, const void *z, void *out, size_t n,
                           double scale, const void *bal, void *stream);

/* The multipliers on device arrays (convolve.cc:26-110): n counts complex
 * words for multBinary/multcorrelation and real words for realMultBinary. */
int fftwpp_amd_multiply(int dtype, int mult, void *F0, const void *F1,
                        size_t n, void *stream);

int fftwpp_amd_scale(int dtype, void *f, size_t n, double s, void *stream);

/* HermitianSymmetrizeX / XY on device data (convolve.h:1177-1256). */
int fftwpp_amd_symmetrize_x(size_t Hx, size_t Hy, size_t x0, void *f,
                            size_t Sx, void *stream);
int fftwpp_amd_symmetrize_xy(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                             size_t y0, void *f, size_t Sx, size_t Sy,
                             void *stream);

/* Operations on the self-conjugate plane of a centred Hermitian array (the
 * kz = 0 plane in 3D, the ky = 0 line in 2D with Hy = 1, k = 0 in 1D with
 * Hx = Hy = 1): element (i, j), |i| < Hx, |j| < Hy, lives at
 * f[(x0+i)*Sx + (y0+j)*Sy] and pairs with (-i, -j); "upper" elements have
 * j > 0, or j = 0 and i > 0.  dtype selects complex128 or complex64.
 *   op 0: lower <- conj(upper), origin <- Re origin: HermitianSymmetrizeX/XY
 *         (convolve.h:1177-1256) without the Nyquist-row zeroing;
 *   op 1: w(k) <- w(k) + conj w(-k) for both members, origin <- 2 Re origin;
 *   op 2: upper kept, lower zeroed, origin halved (the adjoint of op 0).
 * Used by the Hermitian adjoint; no reference counterpart for ops 1-2. */
int fftwpp_amd_hermitian_plane(int dtype, int op, size_t Hx, size_t Hy,
                               size_t x0, size_t y0, void *f, size_t Sx,
                               size_t Sy, void *stream);

/* Device memory and stream plumbing (host code includes no HIP header). */
int fftwpp_amd_malloc(void **p, size_t bytes);
int fftwpp_amd_free(void *p);
int fftwpp_amd_memcpy(void *dst, const void *src, size_t bytes, void *stream);
int fftwpp_amd_memset(void *dst, size_t bytes, void *stream);
int fftwpp_amd_synchronize(void *stream);
int fftwpp_amd_is_device_pointer(const void *p);
int fftwpp_amd_device_count(void);

/* HIP events (timing of launches on a stream). */
int fftwpp_amd_event_create(void **event);
int fftwpp_amd_event_record(void *event, void *stream);
int fftwpp_amd_event_elapsed(float *ms, void *start, void *stop);
int fftwpp_amd_event_destroy(void *event);

/* ---------------------------------------------------------------------- */
/* 2. Convolution layer                                                    */
/* ---------------------------------------------------------------------- */

typedef struct fftwpp_amd_conv fftwpp_amd_conv;

/* A whole dealiased convolution of `dim` dimensions (1..3).
 * hermitian = 0: complex non-centred (HybridConvolution{,2,3},
 *   wrappers/HybridConvolution.h:12-79,113-147), data L[0] x ... contiguous.
 * hermitian = 2: centred complex, dim = 1 (fftPadCentered + Convolution,
 *   tests/hybridconv.cc:44-46 with -c); M defaults to the 3/2 rule.
 * hermitian = 1: centred Hermitian (HybridConvolutionHermitian{,2,3},
 *   HybridConvolution.h:33-52,80-111,149-182); L[] are the full lengths and
 *   the last stored dimension holds ceil(L/2) modes.
 * hermitian = 3: real data (fftPadReal as the outer or only dimension,
 *   tests/hybridconvr{,2,3}.cc), arrays of L[0] x ... real words (float64
 *   or float32); multBinary only, no adjoint.
 * M[d] = 0 selects the reference default (2L-1, or 3ceil(L/2)-2(L mod 2));
 * m[d] = 0 lets the library choose the sub-FFT size (autotuned plan).
 * The result is normalised and written over `a`. */
int fftwpp_amd_conv_create(fftwpp_amd_conv **h, int dim, int hermitian,
                           const size_t *L, const size_t *M,
                           const size_t *m, int dtype);
int fftwpp_amd_conv_destroy(fftwpp_amd_conv *h);

/* fftwpp_amd_conv_create with timed sub-FFT sizes: the GPU counterpart of the
 * reference's optimizer, which times candidate parameters on the CPU when a
 * plan is made without forced ones (fftBase::OptBase, convolve.h:164-236, convolve.cc:239-470).
 * The library's own choice is timed first; then, per dimension, the sizes of
 * every split of L into p <= 8 blocks (smooth sizes) and the covering powers
 * of two that a plan of that dimension accepts, the shortest padded lengths
 * first (within 1.5 times the shortest; 6, 4 or 3 per dimension for 1D, 2D,
 * 3D); every combination runs on device data and the fastest average wins.
 * The result is kept in the tuning file of the batch planner (see
 * fftwpp_amd_batch_create) under a key of (dim, kind, L, M, precision,
 * device architecture and CU count), so a later creation of the same plan
 * takes no timing unless flags has FFTWPP_AMD_TUNE_RETIME.  chosen (may be
 * NULL) receives the sub-FFT size of each dimension.  Tuning allocates and
 * runs on the current device, synchronously; with FFTWPP_AMD_TUNE_VERBOSE set
 * in the environment every candidate and its time go to stderr. */
#define FFTWPP_AMD_TUNE_RETIME 1
int fftwpp_amd_conv_create_tuned(fftwpp_amd_conv **h, int dim, int hermitian,
                                 const size_t *L, const size_t *M, int dtype,
                                 int flags, size_t *chosen);

/* Memory-bounded execution of 2D and 3D plans (the reference bounds its
 * buffers by processing D residues per pass, convolve.h:445-451, and loops
 * over the outer residues, convolve.h:1475-1535, 1812-1879).  The plan's
 * device work buffers stay within `bytes` (0 = no limit): a 3D plan whose
 * y-z stage the plane kernel takes (y and z complex with m = 256, q = 2,
 * L = 256) runs it when the separate y, z and y stages would need more (two
 * x-transformed arrays plus fftwpp_amd_plane_work() bytes instead of 2 + 2 q_y
 * arrays); otherwise the outer (x) residues run in the largest blocks that
 * fit, each block's backward transform added to the output, with a copy of
 * the input an in-place call overwrites.  Real plans run in one pass.  An
 * execution that fits none of these fails with FFTWPP_AMD_ERR_UNSUPPORTED.
 * Results are the same in every case.  fftwpp_amd_conv_work_bytes reports the
 * bytes the next execution in the plan's precision holds and whether it takes
 * the plane kernel. */
int fftwpp_amd_conv_set_memory_limit(fftwpp_amd_conv *h, size_t bytes);
int fftwpp_amd_conv_work_bytes(const fftwpp_amd_conv *h, size_t *bytes, int *plane);

/* Device pointers: a <- a (*) b on `stream`. */
int fftwpp_amd_conv_execute(fftwpp_amd_conv *h, void *a, const void *b,
                            void *stream);

/* Adjoint of the bilinear map (a,b) -> a (*) b: given the cotangent w of the
 * output, grad_a = w (x) b and grad_b = w (x) a (dealiased correlations,
 * multcorrelation of convolve.cc:85-110 on the same padded residues).
 * Real plans take it when their outer dimension runs on the real residue
 * transforms (fftwpp_amd_real_planes > 0; else FFTWPP_AMD_ERR_UNSUPPORTED).
 * Pointers are device pointers; either gradient pointer may be NULL. */
int fftwpp_amd_conv_adjoint(fftwpp_amd_conv *h, const void *w, const void *a,
                            const void *b, void *grad_a, void *grad_b,
                            void *stream);

/* Out-of-place: out <- a (*) b; a and b are only read (out may equal a). */
int fftwpp_amd_conv_execute_out(fftwpp_amd_conv *h, void *out, const void *a,
                                const void *b, void *stream);

/* conv_execute_out with a HIP event recorded after every kernel launch of
 * the plan: stage_ms[i] is the duration of stage i (waits for completion).
 * fftwpp_amd_conv_stage_name names stage i of the last timed execution. */
int fftwpp_amd_conv_execute_timed(fftwpp_amd_conv *h, void *out,
                                  const void *a, const void *b, void *stream,
                                  float *stage_ms, int max_stages,
                                  int *nstages);
const char *fftwpp_amd_conv_stage_name(const fftwpp_amd_conv *h, int i);

/* Padded sizes N[d] = m*q actually used and the sub-FFT sizes m[d], q[d]. */
int fftwpp_amd_conv_geometry(const fftwpp_amd_conv *h, size_t *N,
                             size_t *m, size_t *q);

/* ------------------------------------------------------------------------
 * User multipliers: A inputs, B outputs and a DEVICE multiplier, the
 * multiplier/Indices contract of convolve.h:48-82 (Convolution::operate,
 * convolve.h:1120-1135; padded index Indices::fft->index(r,i),
 * convolve.h:297-326) on the GPU.
 *
 * The reference calls a host multiplier once per residue block.  Here every
 * dimension's q residues are expanded and the multiplier is called ONCE per
 * convolution on the whole padded arrays, which are device arrays:
 *   F[a], a < max(A,B), holds n = N[0] ... N[dims-1] complex words (complex
 *   double for FFTWPP_AMD_F64, complex float for FFTWPP_AMD_F32).  F[0..A-1]
 *   hold the transformed inputs; the multiplier leaves the B products in
 *   F[0..B-1] (pointwise products on the padded grid).
 *   The word at offset sum_d j[d]*stride[d] (j[d] < N[d] = m[d]*q[d]) is the
 *   value at padded grid point x[d] = FFTWPP_AMD_GRID_POINT(m[d],q[d],j[d]),
 *   i.e. sum_k f_k exp(2 pi i k x / N) per dimension, the index q s + r of
 *   the reference's residue r at sub-FFT position s.  For centred data k
 *   runs over -floor(L/2) ... ceil(L/2)-1.
 * The multiplier enqueues its work on `stream` and returns; it must not free
 * F or synchronise other streams.  `user` is passed through. */
typedef struct fftwpp_amd_layout {
  int dims;                  /* 1..3 */
  int dtype;                 /* FFTWPP_AMD_F64 or FFTWPP_AMD_F32 */
  size_t A, B;               /* inputs, outputs */
  size_t N[3], m[3], q[3];   /* padded length, sub-FFT size, residues */
  size_t stride[3];          /* word stride of position j[d] */
} fftwpp_amd_layout;

#define FFTWPP_AMD_GRID_POINT(m, q, j) ((q) * ((j) % (m)) + (j) / (m))

typedef void (*fftwpp_amd_multiplier)(void **F, size_t n,
                                      const fftwpp_amd_layout *layout,
                                      void *stream, void *user);

/* A dealiased convolution of A inputs into B outputs with a user device
 * multiplier (Application(A,B,mult) with fftPad or fftPadCentered in every
 * dimension, convolve.h:84-121,510-699).  Data are L[0] x ... contiguous
 * complex arrays; centred = 1 makes every dimension centred (default M the
 * 3/2 rule, else 2L-1).  1 <= A, B <= 8. */
int fftwpp_amd_conv_create_general(fftwpp_amd_conv **h, int dim, int centred,
                                   const size_t *L, const size_t *M,
                                   const size_t *m, int dtype, size_t A,
                                   size_t B, fftwpp_amd_multiplier mult,
                                   void *user);

/* Device pointers f[0..max(A,B)-1]: inputs f[0..A-1], normalised results
 * written to f[0..B-1] (Convolution::convolve, convolve.cc:7577-7599). */
int fftwpp_amd_conv_execute_general(fftwpp_amd_conv *h, void *const *f,
                                    void *stream);

/* ------------------------------------------------------------------------
 * Fused user multipliers: the multiplier runs inside the fused innermost-row
 * kernels, as the library multipliers do (Convolution::operate inside the
 * residue loop, convolve.h:1120-1135), instead of on expanded arrays.  The
 * multiplier is device code compiled into those kernels by the caller
 * (include/fftwpp_amd_fused.h, hipcc); the library runs every other stage of
 * a plan made by fftwpp_amd_conv_create and hands the row stage to `hook`.
 *
 * Row i of the stage (f, g, out at i*rowstride, i*outstride; out may equal
 * f) lies at the outer padded grid point (x_0, ..., x_{nouter-1}):
 * i = sum_d j_d ostride[d] with j_d < ocount[d], and
 * x_d = oq[d] (j_d mod om[d]) + j_d / om[d] (FFTWPP_AMD_GRID_POINT), except
 * for a real outer dimension (oreal[d] = 1), whose j_d indexes the kept planes
 * of fftwpp_amd_real_planes (x_d of plane base(r) + s is q s + r).  Within a
 * row the multiplier sees padded index l = q s + r of the row dimension.
 * The hook returns 0 after enqueueing the stage on `stream`, or nonzero (the
 * row geometry is outside the register row kernels, or a HIP error). */
typedef struct fftwpp_amd_rows_args {
  const fftwpp_amd_dim *dim;   /* the row (innermost) dimension */
  int dtype;                   /* FFTWPP_AMD_F64 or FFTWPP_AMD_F32 */
  int hermitian;               /* Hermitian rows: real residue values */
  const void *f, *g;           /* transformed inputs of the row stage */
  void *out;
  size_t rows, rowstride, outstride;
  double scale;                /* applied to the row results */
  void *stream;
  int nouter;                  /* 0 (1D), 1 (2D), 2 (3D) */
  size_t om[2], oq[2], ocount[2], ostride[2];
  int oreal[2];
} fftwpp_amd_rows_args;

typedef int (*fftwpp_amd_rows_hook)(const fftwpp_amd_rows_args *args, void *user);

/* conv_execute_out with the row stage handed to `hook` (plans of
 * fftwpp_amd_conv_create; a 1D real plan has no row stage and is refused;
 * the 3D plane kernel is not used). */
int fftwpp_amd_conv_execute_fused(fftwpp_amd_conv *h, void *out, const void *a,
                                  const void *b, fftwpp_amd_rows_hook hook,
                                  void *user, void *stream);

/* The kernel arguments of a dimension (a fftwpp_amd::DimArgs of
 * fftwpp_amd/csrc/common.h, owned by d) for code that launches the row
 * kernels itself; NULL for a bad dimension. */
const void *fftwpp_amd_dim_kernel_args(const fftwpp_amd_dim *d, int dtype);

/* ------------------------------------------------------------------------
 * Batches of independent 1D complex convolutions of uneven lengths.  No
 * reference counterpart: FFTW++ convolves one set of arrays per Convolution
 * object and call (convolve.h:982-1166); a batch of n items replaces n such
 * objects (each with A = 2, B = 1 and the given multiplier) and n calls.
 *
 * Items are packed: item i occupies [off_i, off_i + L_i) of the device arrays
 * f, g and out, off_0 = 0, off_{i+1} = off_i + L_i.  M[i] is the minimum
 * padded length of item i (M = NULL: 2 L_i - 1, as HybridConvolution's
 * default, wrappers/HybridConvolution.h:19).  The result is normalised like
 * Convolution::convolve.
 * ------------------------------------------------------------------------ */
typedef struct fftwpp_amd_batch fftwpp_amd_batch;

/* time candidate sub-FFT sizes and tile shapes per length bucket at
 * creation (in place of the reference's CPU timing optimizer) */
#define FFTWPP_AMD_BATCH_TUNE 1
/* record execute() into a hipGraph per (out, f, g) and replay it */
#define FFTWPP_AMD_BATCH_GRAPH 2

int fftwpp_amd_batch_create(fftwpp_amd_batch **h, size_t n, const size_t *L,
                            const size_t *M, int dtype, int mult, int flags);
int fftwpp_amd_batch_destroy(fftwpp_amd_batch *h);
/* total packed length (sum of L) and the n+1 offsets */
size_t fftwpp_amd_batch_size(const fftwpp_amd_batch *h);
int fftwpp_amd_batch_offsets(const fftwpp_amd_batch *h, size_t *off);
/* sub-FFT size, residue count and padded length chosen for item i */
int fftwpp_amd_batch_geometry(const fftwpp_amd_batch *h, size_t i, size_t *m,
                              size_t *q, size_t *N);
/* one line per launch class: m, q, items, tile shape */
int fftwpp_amd_batch_describe(const fftwpp_amd_batch *h, char *buf,
                              size_t len);
/* out <- f (*) g item by item (device pointers; out distinct from f, g) on the
 * HIP stream; with FFTWPP_AMD_BATCH_GRAPH the first call on a given
 * (out, f, g) runs directly and records the graph that later calls replay
 * (up to 16 graphs per handle, the least recently used one replaced; the
 * graphs hold the addresses, so arrays freed and reallocated at the same
 * addresses replay the same graph, which is correct since the plan is
 * immutable).  A handle is single-stream: its calls must not overlap, on one
 * stream or several (they share the partial-sum buffers, the fork/join
 * events and the side streams). */
int fftwpp_amd_batch_execute(fftwpp_amd_batch *h, void *out, const void *f,
                             const void *g, void *stream);

/* ------------------------------------------------------------------------
 * 3. Standalone FFTs: the transforms of the fft1d, mfft1d, fft2d, fft3d,
 * rcfft1d, crfft1d, mrcfft1d and mcrfft1d classes (fftw++.h:556-1444) on
 * device arrays, unnormalised unless `scale` says otherwise:
 *   out_k = scale * sum_j in_j exp(sign 2 pi i j k / n)  (FFTW's sign).
 * kind FFTWPP_AMD_FFT_C2C: any sign; dims 1..3 of n[0] x ... contiguous
 *   words, or (dims = 1) `howmany` transforms with element j of transform t
 *   at j*istride + t*idist (output: ostride, odist), as mfft1d
 *   (fftw++.h:827-896).
 * kind FFTWPP_AMD_FFT_R2C: sign -1: n real words in (istride, idist in
 *   reals), n/2+1 complex words out (ostride, odist), as rcfft1d/mrcfft1d
 *   (fftw++.h:917-955,1023-1072); dims 2, 3: one transform of row-major
 *   n[0] x ... x n[d-1] reals to n[0] x ... x (n[d-1]/2+1) complex words,
 *   as rcfft2d/rcfft3d (fftw++.h:1249-1316,1436-1520).
 * kind FFTWPP_AMD_FFT_C2R: sign +1, the inverse layouts (a Hermitian
 *   spectrum in, reals out), as crfft1d/mcrfft1d/crfft2d/crfft3d
 *   (fftw++.h:957-975,1122-1171,1318-1388,1522-1585).
 * In place (in == out), real 2D/3D rows are padded to 2 (n[d-1]/2+1) reals
 * as FFTW stores them.
 * A length n = m q runs as q residue passes of m-point sub-FFTs (m the
 * largest divisor of n the sub-FFT layer takes, at most 4096 here), writing
 * natural order through strides: a length with such a divisor and small q
 * is one or two passes, longer contiguous 1D transforms take four steps
 * (columns of q, twiddles, rows of m).  A 1D length without a divisor in
 * [n/4096, 4096] (up to 2^23) runs through Bluestein's algorithm; such a
 * length as an axis of a 2D/3D transform is FFTWPP_AMD_ERR_UNSUPPORTED. */
#define FFTWPP_AMD_FFT_C2C 0
#define FFTWPP_AMD_FFT_R2C 1
#define FFTWPP_AMD_FFT_C2R 2

typedef struct fftwpp_amd_fft fftwpp_amd_fft;

int fftwpp_amd_fft_create(fftwpp_amd_fft **h, int kind, int dims,
                          const size_t *n, int sign, int dtype,
                          size_t howmany, size_t istride, size_t ostride,
                          size_t idist, size_t odist);
int fftwpp_amd_fft_destroy(fftwpp_amd_fft *h);
/* Device pointers; in may equal out (in place). */
int fftwpp_amd_fft_execute(fftwpp_amd_fft *h, const void *in, void *out,
                           double scale, void *stream);
/* n[0]*...*n[dims-1]: the factor fftNormalized divides by. */
size_t fftwpp_amd_fft_normalization(const fftwpp_amd_fft *h);
/* out <- out / normalization over the plan's output layout (contiguous
 * output rows for real 2D/3D), as fftw::Normalize (fftw++.h:357-434). */
int fftwpp_amd_fft_normalize(const fftwpp_amd_fft *h, void *out, void *stream);

/* Origin-shifted real 2D/3D transforms (FP64 device arrays).  op 0: the
 * Shift of fftw++.h:144-222, negating the reals of odd rows i (2D, nx x
 * rowlen) or of rows (i, j) with i + j odd (3D, nx x ny x rowlen), rowlen =
 * ny or nz, or 2 (n/2+1) for padded in-place rows; nx (and ny in 3D) even.
 * op 1 / 2: deNyquist of a shifted 2D (nx x ny/2+1) / 3D (nx x ny x nz/2+1)
 * spectrum, fftw++.h:1342-1360,1466-1495. */
int fftwpp_amd_fft_shift(int op, double *data, size_t nx, size_t ny, size_t nz,
                         size_t rowlen, void *stream);

/* out[c][r] = in[r][c] for a rows x cols matrix of elements of `words`
 * doubles (device arrays, in != out): the Transpose class of fftw++.h:436-497
 * (an FFTW rank-0 guru plan there). */
int fftwpp_amd_transpose(size_t rows, size_t cols, size_t words, const void *in,
                         void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif
