/* fftwpp_amd_fused.h -- fused user multipliers (C++ for hipcc, gfx950).
 *
 * The reference calls a user multiplier inside its residue loop
 * (Convolution::operate, convolve.h:1120-1135, the multiplier typedef at
 * convolve.h:78-82), so the multiplier runs where the transformed data are.
 * Here a user multiplier is a device functor compiled into the library's
 * fused innermost-row kernels (the register FFT kernels of rowsfused.h): the
 * library runs every other stage of a plan made by fftwpp_amd_conv_create
 * and the caller's translation unit launches the row stage with its own
 * functor, so the innermost dimension never leaves the chip.
 *
 * Build: hipcc --offload-arch=gfx950 -I <repo>/include -I <repo>/fftwpp_amd/csrc
 *        (with -Xclang -target-feature -Xclang -packed-fp32-ops, as the
 *        library), linked against libfftwpp_amd.so.
 *
 * The functor (trivially copyable; it is a kernel argument):
 *
 *   struct Op {
 *     // complex and centred rows: the product at padded point l of row `row`
 *     template<typename T> __device__ fftwpp_amd::cplx<T>
 *     operator()(fftwpp_amd::cplx<T> F, fftwpp_amd::cplx<T> G, int64_t row, int l) const;
 *     // Hermitian rows: the transforms are real (realMultBinary's values)
 *     template<typename T> __device__ T operator()(T F, T G, int64_t row, int l) const;
 *   };
 *
 * l = q s + r is the padded index of the row dimension (0 <= l < N).  A
 * functor derived from fftwpp_amd::RowLayout also gets the outer padded grid
 * point of each row: outer(row, x) fills x[0..nouter-1] (the layout of
 * fftwpp_amd_rows_args).  Work that depends on the row alone belongs in an
 * optional member `__device__ R at(int64_t row) const`: the kernels then call
 * at(row) once per row and R's operator()(F, G, l) (complex and real forms)
 * for every element.  Multipliers on Hermitian or real plans must keep
 * the conjugate symmetry of the data (a product weighted by a function of
 * |k| in every dimension does).  Row geometries outside the register row
 * kernels (power-of-two m in [16, 4096] covering the row) are refused with
 * FFTWPP_AMD_ERR_UNSUPPORTED. */
#ifndef FFTWPP_AMD_FUSED_H
#define FFTWPP_AMD_FUSED_H

#include <type_traits>

#include "fftwpp_amd.h"
#include "rowsfused.h"

namespace fftwpp_amd {

/* The outer layout of the rows of a row stage (fftwpp_amd_rows_args). */
struct RowLayout {
  int nouter=0;
  int wide=0;                // rows or strides beyond 32 bits
  int oreal[2]={0,0};
  int64_t om[2]={1,1}, oq[2]={1,1}, ocount[2]={1,1}, ostride[2]={1,1};

  void setLayout(const fftwpp_amd_rows_args &a)
  {
    nouter=a.nouter;
    for(int d=0; d < 2; ++d) {
      om[d]=(int64_t) a.om[d];
      oq[d]=(int64_t) a.oq[d];
      ocount[d]=a.ocount[d] ? (int64_t) a.ocount[d] : 1;
      ostride[d]=a.ostride[d] ? (int64_t) a.ostride[d] : 1;
      oreal[d]=a.oreal[d];
    }
    wide=a.rows >= ((size_t) 1 << 32) || ocount[0]*ostride[0] >= ((int64_t) 1 << 32);
  }

  // j -> padded grid point of outer dimension d
  template<typename I>
  __host__ __device__ int64_t point(int d, I j) const
  {
    const I m=(I) om[d];
    if(oreal[d]) {
      // kept planes of real data: residue 0 at s <= m/2, then whole residues
      if(j <= m/2) return oq[d]*(int64_t) j;
      const I t=j-(m/2+1);
      return oq[d]*(int64_t) (t % m)+1+(int64_t) (t/m);
    }
    return oq[d]*(int64_t) (j % m)+(int64_t) (j/m);
  }

  // x[d] = the padded grid point of row `row` in outer dimension d (32-bit
  // divisions unless the stage is wider)
  __host__ __device__ void outer(int64_t row, int64_t *x) const
  {
#pragma unroll
    for(int d=0; d < 2; ++d)
      if(d < nouter) {
        if(wide)
          x[d]=point<int64_t>(d,(row/ostride[d]) % ocount[d]);
        else
          x[d]=point<uint32_t>(d,((uint32_t) row/(uint32_t) ostride[d]) % (uint32_t) ocount[d]);
      }
  }
};

/* What a translation unit compiles for one functor: row sizes (bit i of
 * sizes = m of 16 << i, ROWS_ALL_SIZES for 16 ... 4096), precisions (bit 0
 * FP64, bit 1 FP32) and row kinds (bit 0 complex or centred, bit 1
 * Hermitian).  Every kernel is instantiated per functor, so narrowing these to
 * the plans a program runs shortens its build; other geometries are refused. */
template<unsigned SIZES=ROWS_ALL_SIZES, unsigned PRECISIONS=3u, unsigned KINDS=3u>
struct FusedSet {};

/* Launch the row stage described by `a` with functor op: 0 after the launch,
 * FFTWPP_AMD_ERR_UNSUPPORTED when no compiled register row kernel takes the
 * geometry, FFTWPP_AMD_ERR_HIP on a launch failure. */
template<class Op, unsigned SIZES=ROWS_ALL_SIZES, unsigned PRECISIONS=3u, unsigned KINDS=3u>
inline int fused_rows(const fftwpp_amd_rows_args *a, Op op,
                      FusedSet<SIZES,PRECISIONS,KINDS> = FusedSet<SIZES,PRECISIONS,KINDS>())
{
  if constexpr(std::is_base_of<RowLayout,Op>::value) op.setLayout(*a);
  const DimArgs *d=(const DimArgs *) fftwpp_amd_dim_kernel_args(a->dim,a->dtype);
  if(!d) return FFTWPP_AMD_ERR_ARG;
  hipStream_t st=(hipStream_t) a->stream;
  const bool herm=a->hermitian != 0;
  int rc=1;
  if(a->dtype == FFTWPP_AMD_F32) {
    if constexpr((PRECISIONS & 2u) != 0)
      rc=rows_fused<float,Op,SIZES,KINDS>(*d,herm,a->f,a->g,a->rows,a->rowstride,a->out,
                                           a->outstride,a->scale,op,st);
  } else {
    if constexpr((PRECISIONS & 1u) != 0)
      rc=rows_fused<double,Op,SIZES,KINDS>(*d,herm,a->f,a->g,a->rows,a->rowstride,a->out,
                                            a->outstride,a->scale,op,st);
  }
  return rc == 0 ? FFTWPP_AMD_OK : (rc == 1 ? FFTWPP_AMD_ERR_UNSUPPORTED : FFTWPP_AMD_ERR_HIP);
}

/* out <- a (*) b on device pointers with the product op in the row stage
 * (normalised, like fftwpp_amd_conv_execute_out). */
template<class Op, unsigned SIZES=ROWS_ALL_SIZES, unsigned PRECISIONS=3u, unsigned KINDS=3u>
inline int conv_execute_fused(fftwpp_amd_conv *h, void *out, const void *a, const void *b,
                              const Op &op, void *stream=nullptr,
                              FusedSet<SIZES,PRECISIONS,KINDS> = FusedSet<SIZES,PRECISIONS,KINDS>())
{
  fftwpp_amd_rows_hook hook=[](const fftwpp_amd_rows_args *args, void *user) -> int {
    return fused_rows(args,*static_cast<const Op *>(user),FusedSet<SIZES,PRECISIONS,KINDS>());
  };
  return fftwpp_amd_conv_execute_fused(h,out,a,b,hook,const_cast<Op *>(&op),stream);
}

} // namespace fftwpp_amd

#endif
