// fusedmult.hip -- fused user multipliers for the GPU tests: device functors
// compiled into the library's fused row kernels through
// include/fftwpp_amd_fused.h, exported as C entry points for ctypes.
#include "fftwpp_amd_fused.h"

namespace {

// F G / (1 + c |k|^2): k the signed frequency of the padded grid point in every
// dimension (l or l - N), so the weight is even in each k and keeps the
// conjugate symmetry of Hermitian and real data.  N[0..nouter-1] are the outer
// padded lengths, N[nouter] the row dimension's.  at(row) does the outer part
// of |k|^2 once per row.
struct Weighted : fftwpp_amd::RowLayout {
  int64_t N[3];
  double c;

  struct Row {
    double k2, c;
    int64_t n;
    __device__ double w(int l) const
    {
      const double k=(double) (2*l <= n ? l : l-n);
      return 1.0/(1.0+c*(k2+k*k));
    }
    template<typename T>
    __device__ fftwpp_amd::cplx<T> operator()(fftwpp_amd::cplx<T> F, fftwpp_amd::cplx<T> G,
                                             int l) const
    {
      return (T) w(l)*(F*G);
    }
    template<typename T>
    __device__ T operator()(T F, T G, int l) const { return (T) w(l)*(F*G); }
  };

  __device__ Row at(int64_t row) const
  {
    int64_t x[2]={0,0};
    outer(row,x);
    double k2=0.0;
    for(int d=0; d < nouter; ++d) {
      const double k=(double) (2*x[d] <= N[d] ? x[d] : x[d]-N[d]);
      k2 += k*k;
    }
    return Row{k2,c,N[nouter]};
  }
  template<typename T>
  __device__ fftwpp_amd::cplx<T> operator()(fftwpp_amd::cplx<T> F, fftwpp_amd::cplx<T> G,
                                           int64_t row, int l) const
  {
    return at(row)(F,G,l);
  }
  template<typename T>
  __device__ T operator()(T F, T G, int64_t row, int l) const { return at(row)(F,G,l); }
};

// The plain product F G as a caller's functor (the user multiplier of
// multipliers.hip, testmult_binary, fused).
struct Binary {
  template<typename T>
  __device__ fftwpp_amd::cplx<T> operator()(fftwpp_amd::cplx<T> F, fftwpp_amd::cplx<T> G,
                                           int64_t, int) const
  {
    return F*G;
  }
  template<typename T>
  __device__ T operator()(T F, T G, int64_t, int) const { return F*G; }
};

} // namespace

extern "C" int fusedmult_binary(fftwpp_amd_conv *h, void *out, const void *a,
                                const void *b, void *stream)
{
  // FP64 complex rows of m = 256 (the 3D 256^3 -> 512^3 workload)
  return fftwpp_amd::conv_execute_fused(h,out,a,b,Binary(),stream,
                                        fftwpp_amd::FusedSet<1u << 4,1u,1u>());
}

// out <- a (*) b with the product weighted by 1/(1 + c |k|^2); N holds the
// padded lengths of the plan's dimensions (outer first).
extern "C" int fusedmult_weighted(fftwpp_amd_conv *h, void *out, const void *a,
                                  const void *b, const int64_t *N, int nd, double c,
                                  void *stream)
{
  Weighted op;
  for(int d=0; d < 3; ++d) op.N[d]=d < nd ? N[d] : 1;
  op.c=c;
  // the row sizes of the GPU tests: m = 16 ... 256 and 1024
  return fftwpp_amd::conv_execute_fused(h,out,a,b,op,stream,
                                        fftwpp_amd::FusedSet<0x5Fu>());
}
