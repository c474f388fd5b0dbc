/* cfftw++.h -- drop-in C wrapper of the hybrid dealiased convolutions,
 * exporting the symbol set that wrappers/cfftw++.cc:22-166 defines (and that
 * wrappers/fftwpp.py:26-82 binds through ctypes), implemented on the GPU by
 * libfftwpp_amd.so.
 *
 * Semantics follow the reference (wrappers/HybridConvolution.h):
 *   - create_* take the unpadded lengths; M defaults to 2L-1 (complex) or
 *     3*ceil(L/2)-2*(L mod 2) (centred Hermitian);
 *   - conv1d: L; conv2d: Lx x Ly; conv3d: Lx x Ly x Lz, row-major, contiguous;
 *     hconv1d: ceil(L/2); hconv2d: Lx x ceil(Ly/2); hconv3d: Lx x Ly x
 *     ceil(Lz/2) -- the hconv creators take the full (odd) lengths;
 *   - *_convolve(h, a, b) writes the normalised convolution over a; b is not
 *     modified.  a and b may be host pointers (staged through device memory)
 *     or device pointers (used in place);
 *   - Hermitian inputs must already be symmetrised (fftwpp_HermitianSymmetrize*).
 *
 * The reference header also declares *_dot, *_dotf, *_work and *_work_dot
 * variants and create_doublealign; wrappers/cfftw++.cc defines none of them,
 * so no existing caller can link against them and they are not provided.
 * Failures print the reason to stderr and return NULL / leave `a` untouched
 * where the reference would have called exit(); fftwpp_amd_last_error() holds
 * the message.
 */
#ifndef FFTWPP_AMD_CFFTWPP_H
#define FFTWPP_AMD_CFFTWPP_H

#include <stddef.h>

#ifdef __cplusplus
namespace fftwpp { extern "C" {
#define FFTWPP_AMD_DCOMPLEX double
#else
#define FFTWPP_AMD_DCOMPLEX double __complex__
#endif

/* aligned host arrays (cfftw++.cc:27-41) */
double *create_doubleAlign(size_t n);
void delete_doubleAlign(double *p);
FFTWPP_AMD_DCOMPLEX *create_complexAlign(size_t n);
void delete_complexAlign(FFTWPP_AMD_DCOMPLEX *p);

/* thread count (cfftw++.cc:43-50); stored for compatibility, the device
 * path does not use host threads */
size_t get_fftwpp_maxthreads(void);
void set_fftwpp_maxthreads(size_t nthreads);

/* 1d complex non-centred convolution (cfftw++.cc:52-65) */
typedef struct HybridConvolution HybridConvolution;
HybridConvolution *fftwpp_create_conv1d(size_t L);
void fftwpp_conv1d_convolve(HybridConvolution *conv, FFTWPP_AMD_DCOMPLEX *a,
                            FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_conv1d_delete(HybridConvolution *conv);

/* 1d Hermitian convolution (cfftw++.cc:67-84) */
typedef struct HybridConvolutionHermitian HybridConvolutionHermitian;
HybridConvolutionHermitian *fftwpp_create_hconv1d(size_t L);
void fftwpp_hconv1d_convolve(HybridConvolutionHermitian *conv,
                             FFTWPP_AMD_DCOMPLEX *a, FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_hconv1d_delete(HybridConvolutionHermitian *conv);

/* 2d complex non-centred convolution (cfftw++.cc:86-106) */
typedef struct HybridConvolution2 HybridConvolution2;
HybridConvolution2 *fftwpp_create_conv2d(size_t Lx, size_t Ly);
void fftwpp_conv2d_convolve(HybridConvolution2 *conv, FFTWPP_AMD_DCOMPLEX *a,
                            FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_conv2d_delete(HybridConvolution2 *conv);

/* 2d centred Hermitian convolution (cfftw++.cc:108-122) */
typedef struct HybridConvolutionHermitian2 HybridConvolutionHermitian2;
HybridConvolutionHermitian2 *fftwpp_create_hconv2d(size_t Lx, size_t Ly);
void fftwpp_hconv2d_convolve(HybridConvolutionHermitian2 *conv,
                             FFTWPP_AMD_DCOMPLEX *a, FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_hconv2d_delete(HybridConvolutionHermitian2 *conv);

/* 3d complex non-centred convolution (cfftw++.cc:124-146) */
typedef struct HybridConvolution3 HybridConvolution3;
HybridConvolution3 *fftwpp_create_conv3d(size_t Lx, size_t Ly, size_t Lz);
void fftwpp_conv3d_convolve(HybridConvolution3 *conv, FFTWPP_AMD_DCOMPLEX *a,
                            FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_conv3d_delete(HybridConvolution3 *conv);

/* 3d centred Hermitian convolution (cfftw++.cc:148-162) */
typedef struct HybridConvolutionHermitian3 HybridConvolutionHermitian3;
HybridConvolutionHermitian3 *fftwpp_create_hconv3d(size_t Lx, size_t Ly,
                                                   size_t Lz);
void fftwpp_hconv3d_convolve(HybridConvolutionHermitian3 *conv,
                             FFTWPP_AMD_DCOMPLEX *a, FFTWPP_AMD_DCOMPLEX *b);
void fftwpp_hconv3d_delete(HybridConvolutionHermitian3 *conv);

/* Hermitian symmetrisers (cfftw++.cc:76-78,98-101,134-139) */
void fftwpp_HermitianSymmetrize(FFTWPP_AMD_DCOMPLEX *f);
void fftwpp_HermitianSymmetrizeX(size_t Hx, size_t Hy, size_t x0,
                                 FFTWPP_AMD_DCOMPLEX *f);
void fftwpp_HermitianSymmetrizeXY(size_t Hx, size_t Hy, size_t Hz, size_t x0,
                                  size_t y0, FFTWPP_AMD_DCOMPLEX *f);

/* ------------------------------------------------------------------------
 * Extensions with no reference counterpart: the same handles on device
 * pointers and a HIP stream (hipStream_t passed as void *, NULL = default
 * stream), asynchronous, returning 0 or a negative fftwpp_amd status
 * (FFTWPP_AMD_ERR_*; fftwpp_amd_last_error() holds the message).  a <- a (*) b
 * normalised, as *_convolve; b is only read.
 * ------------------------------------------------------------------------ */
int fftwpp_conv1d_convolve_device(HybridConvolution *conv, FFTWPP_AMD_DCOMPLEX *a,
                                  const FFTWPP_AMD_DCOMPLEX *b, void *stream);
int fftwpp_hconv1d_convolve_device(HybridConvolutionHermitian *conv,
                                   FFTWPP_AMD_DCOMPLEX *a, const FFTWPP_AMD_DCOMPLEX *b,
                                   void *stream);
int fftwpp_conv2d_convolve_device(HybridConvolution2 *conv, FFTWPP_AMD_DCOMPLEX *a,
                                  const FFTWPP_AMD_DCOMPLEX *b, void *stream);
int fftwpp_hconv2d_convolve_device(HybridConvolutionHermitian2 *conv,
                                   FFTWPP_AMD_DCOMPLEX *a, const FFTWPP_AMD_DCOMPLEX *b,
                                   void *stream);
int fftwpp_conv3d_convolve_device(HybridConvolution3 *conv, FFTWPP_AMD_DCOMPLEX *a,
                                  const FFTWPP_AMD_DCOMPLEX *b, void *stream);
int fftwpp_hconv3d_convolve_device(HybridConvolutionHermitian3 *conv,
                                   FFTWPP_AMD_DCOMPLEX *a, const FFTWPP_AMD_DCOMPLEX *b,
                                   void *stream);

#ifdef __cplusplus
} }
#endif

#undef FFTWPP_AMD_DCOMPLEX

#endif
