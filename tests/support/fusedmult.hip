// fusedmult.hip -- fused user multipliers for the GPU tests: device functors
// compiled into the library's fused row kernels through
// include/fftwpp_amd_fused.h, exported as C entry points for ctypes.
#include "fftwpp_amd_fused.h"

namespace {

// F G / (1 + c |k|^2): k the signed frequency of the padded grid point in every
// dimension (l or l - N), so the weight is even in each k and keeps the
// conjugate symmetry of Hermitian and real data.  N[0..nouter-1] are the outer
// padded lengths, N[nouter] the row dimension's.
struct Weighted : fftwpp_amd::RowLayout {
  int64_t N[3];
  double c;

  __device__ double weight(int64_t row, int l) const
  {
    int64_t x[2]={0,0};
    outer(row,x);
    double k2=0.0;
    for(int d=0; d < nouter; ++d) {
      const double k=(double) (2*x[d] <= N[d] ? x[d] : x[d]-N[d]);
      k2 += k*k;
    }
    const int64_t n=N[nouter];
    const double k=(double) (2*l <= n ? l : l-n);
    return 1.0/(1.0+c*(k2+k*k));
  }
  template<typename T>
  __device__ fftwpp_amd::cplx<T> operator()(fftwpp_amd::cplx<T> F, fftwpp_amd::cplx<T> G,
                                           int64_t row, int l) const
  {
    return (T) weight(row,l)*(F*G);
  }
  template<typename T>
  __device__ T operator()(T F, T G, int64_t row, int l) const
  {
    return (T) weight(row,l)*(F*G);
  }
};

} // namespace

// out <- a (*) b with the product weighted by 1/(1 + c |k|^2); N holds the
// padded lengths of the plan's dimensions (outer first).
extern "C" int fusedmult_weighted(fftwpp_amd_conv *h, void *out, const void *a,
                                  const void *b, const int64_t *N, int nd, double c,
                                  void *stream)
{
  Weighted op;
  for(int d=0; d < 3; ++d) op.N[d]=d < nd ? N[d] : 1;
  op.c=c;
  return fftwpp_amd::conv_execute_fused(h,out,a,b,op,stream);
}
