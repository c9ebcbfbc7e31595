// L_p norms of the PCFs of a device set, ||f||_p = (int_0^MAX |f|^p)^(1/p):
// l1_norm / lp_norm of include/sbear/algorithms/functional/apply_functional.hpp:63-100
// over iterate_segments (iterate_rectangles.hpp:106-155). One lane per PCF
// walks its records (zero-length pieces are already dropped: they add
// fma(0, |v|, acc) == acc); the last record's segment runs to the domain
// end MAX in the time type, as the reference's does. FP64 sets accumulate
// with one FMA per segment (the reference's operation sequence); FP32C sets
// with the Kahan-compensated FP32 sum.
#include "walk_common.hpp"

namespace sbamd
{
  template <typename TT, typename TV, bool L1, bool KAHAN>
  __global__ __launch_bounds__(256) void norm_kernel(NormParams P)
  {
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= P.n)
      return;
    const RecT* r = static_cast<const RecT*>(P.rec) + P.roff[q];
    const int64_t m = P.roff[q + 1] - P.roff[q];
    const TV p = static_cast<TV>(P.p);
    using Acc = typename std::conditional<KAHAN, KahanAcc, PlainAcc<TV>>::type;
    Acc acc;
    TT t = TT(0);
    for (int64_t k = 0; k + 1 < m; ++k)
    {
      const RecT x = r[k];
      const TV a = fabs(x.v);
      acc.add(static_cast<TV>(x.t - t), L1 ? a : pow(a, p));
      t = x.t;
    }
    const TV a = fabs(r[m - 1].v);
    double res = acc.finish(L::max() - t, L1 ? a : pow(a, p));
    if (!L1)
      res = pow(res, 1.0 / P.p);
    const int64_t o = P.perm[q];
    if (P.out_f64)
      static_cast<double*>(P.out)[o] = res;
    else
      static_cast<float*>(P.out)[o] = static_cast<float>(res);
  }

  template <typename TT, typename TV, bool KAHAN>
  static void launch_norm_typed(const NormParams& prm, hipStream_t stream)
  {
    const unsigned grid = static_cast<unsigned>((prm.n + 255) / 256);
    if (prm.p == 1.0)
      hipLaunchKernelGGL((norm_kernel<TT, TV, true, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
    else
      hipLaunchKernelGGL((norm_kernel<TT, TV, false, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
    SBAMD_HIP(hipGetLastError());
  }

  void launch_norm(const NormParams& prm, int tt_f64, int tv_f64, hipStream_t stream)
  {
    if (prm.n <= 0)
      return;
    if (tt_f64 && tv_f64)
      launch_norm_typed<double, double, false>(prm, stream);
    else if (tt_f64)
      launch_norm_typed<double, float, true>(prm, stream);
    else
      launch_norm_typed<float, float, true>(prm, stream);
  }
}
