// L_p norms of the PCFs of a device set, ||f||_p = (int_0^MAX |f|^p)^(1/p):
// l1_norm / lp_norm of include/sbear/algorithms/functional/apply_functional.hpp:63-100
// over iterate_segments (iterate_rectangles.hpp:106-155). One lane per PCF
// walks its records (zero-length pieces are already dropped: they add
// fma(0, |v|, acc) == acc); the last record's segment runs to the domain
// end MAX in the time type, as the reference's does. FP64 sets accumulate
// with one FMA per segment (the reference's operation sequence); FP32C sets
// with the Kahan-compensated FP32 sum. p = 2 uses v * v and sqrt for the
// reference's pow(|v|, 2) and pow(x, 1/2), both correctly rounded.
#include "walk_common.hpp"

namespace sbamd
{
  // |v|^p per segment: p = 1 and p = 2 without pow (as the distance kernels
  // take d and d * d), any other p with pow.
  template <int PK, typename TV>
  __device__ __forceinline__ TV norm_term(TV a, TV p)
  {
    if (PK == 1)
      return a;
    if (PK == 2)
      return a * a;
    return pow(a, p);
  }

  template <typename TT, typename TV, int PK, bool KAHAN>
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
    // Records come in batches of NORM_BATCH loads issued together, so a lane
    // waits on memory once per batch, not once per record; the sum keeps the
    // sequential order.
    constexpr int NORM_BATCH = 8;
    int64_t k = 0;
    for (; k + NORM_BATCH < m; k += NORM_BATCH)
    {
      RecT x[NORM_BATCH];
#pragma unroll
      for (int u = 0; u < NORM_BATCH; ++u)
        x[u] = r[k + u];
#pragma unroll
      for (int u = 0; u < NORM_BATCH; ++u)
      {
        acc.add(static_cast<TV>(x[u].t - t), norm_term<PK>(fabs(x[u].v), p));
        t = x[u].t;
      }
    }
    for (; k + 1 < m; ++k)
    {
      const RecT x = r[k];
      acc.add(static_cast<TV>(x.t - t), norm_term<PK>(fabs(x.v), p));
      t = x.t;
    }
    double res = acc.finish(L::max() - t, norm_term<PK>(fabs(r[m - 1].v), p));
    if (PK == 2)
      res = sqrt(res);
    else if (PK != 1)
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
      hipLaunchKernelGGL((norm_kernel<TT, TV, 1, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
    else if (prm.p == 2.0)
      hipLaunchKernelGGL((norm_kernel<TT, TV, 2, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
    else
      hipLaunchKernelGGL((norm_kernel<TT, TV, 0, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
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
