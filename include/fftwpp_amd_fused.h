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
 * fftwpp_amd_rows_args).  Multipliers on Hermitian or real plans must keep
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
  }

  // x[d] = the padded grid point of row `row` in outer dimension d
  __host__ __device__ void outer(int64_t row, int64_t *x) const
  {
    for(int d=0; d < nouter; ++d) {
      const int64_t j=(row/ostride[d]) % ocount[d], m=om[d];
      if(oreal[d]) {
        // kept planes of real data: residue 0 at s <= m/2, then whole residues
        int64_t r=0, s=j;
        if(j > m/2) {
          const int64_t t=j-(m/2+1);
          r=1+t/m;
          s=t % m;
        }
        x[d]=oq[d]*s+r;
      } else
        x[d]=oq[d]*(j % m)+j/m;
    }
  }
};

/* Launch the row stage described by `a` with functor op: 0 after the launch,
 * FFTWPP_AMD_ERR_UNSUPPORTED when no register row kernel takes the geometry,
 * FFTWPP_AMD_ERR_HIP on a launch failure. */
template<class Op>
inline int fused_rows(const fftwpp_amd_rows_args *a, Op op)
{
  if constexpr(std::is_base_of<RowLayout,Op>::value) op.setLayout(*a);
  const DimArgs *d=(const DimArgs *) fftwpp_amd_dim_kernel_args(a->dim,a->dtype);
  if(!d) return FFTWPP_AMD_ERR_ARG;
  hipStream_t st=(hipStream_t) a->stream;
  const bool herm=a->hermitian != 0;
  const int rc=a->dtype == FFTWPP_AMD_F32 ?
    rows_fused<float>(*d,herm,a->f,a->g,a->rows,a->rowstride,a->out,a->outstride,a->scale,op,st) :
    rows_fused<double>(*d,herm,a->f,a->g,a->rows,a->rowstride,a->out,a->outstride,a->scale,op,st);
  return rc == 0 ? FFTWPP_AMD_OK : (rc == 1 ? FFTWPP_AMD_ERR_UNSUPPORTED : FFTWPP_AMD_ERR_HIP);
}

/* out <- a (*) b on device pointers with the product op in the row stage
 * (normalised, like fftwpp_amd_conv_execute_out). */
template<class Op>
inline int conv_execute_fused(fftwpp_amd_conv *h, void *out, const void *a, const void *b,
                              const Op &op, void *stream=nullptr)
{
  fftwpp_amd_rows_hook hook=[](const fftwpp_amd_rows_args *args, void *user) -> int {
    return fused_rows(args,*static_cast<const Op *>(user));
  };
  return fftwpp_amd_conv_execute_fused(h,out,a,b,hook,const_cast<Op *>(&op),stream);
}

} // namespace fftwpp_amd

#endif
