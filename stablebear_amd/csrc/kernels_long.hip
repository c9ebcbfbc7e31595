// Long-pair forward: one wavefront per pair, the pair's union segments split
// across the 64 lanes.
//
// For few pairs with many segments each (a handful of long PCFs, or one
// lp_distance between two long PCFs) the tile kernel, one lane per pair,
// leaves most of the device idle. Here the E interior breakpoints of a pair
// (m_f - 1 of f and m_g - 1 of g, merged) are cut into 64 equal runs: lane l
// finds the time T_l of merged breakpoint E*l/64 by a merge-path binary
// search over the two sorted lists, takes T_{l+1} from lane l + 1 by a
// shuffle, and walks [T_l, T_{l+1}) with the same sorted-pair step as the
// tile kernel, reading records from global memory (sequential per lane, so
// they hit in L1 / L2). Each lane integrates its span exactly (the last
// segment clipped at T_{l+1}; the last lane adds the final rectangle to
// numeric_limits<T>::max() as the reference does), and the 64 partial sums
// are added in FP64 by a butterfly of shuffles. Lanes cut at a time, not at
// an event count, so ties (equal breakpoint times, which only make
// zero-length segments) need no common rule across lanes.
//
// The per-segment terms are those of the sequential walk; the summation order
// is not (64 partial sums), so the FP64 form is not bit-identical to the
// reference here and the planner uses this kernel for FP32C plans unless
// asked to.
#include "walk_common.hpp"

namespace sbamd
{
  // Pair q of a plan in sorted positions: the strict lower triangle (pdist),
  // the lower triangle with diagonal (kernel) or the full grid (cdist).
  __device__ __forceinline__ void pair_of(int64_t q, int32_t out_kind, int64_t n_cols, int64_t& a,
                                          int64_t& b)
  {
    if (out_kind == SBAMD_OUT_DENSE)
    {
      a = q / n_cols;
      b = q - a * n_cols;
      return;
    }
    const bool diag = out_kind == SBAMD_OUT_SYMMAT;
    // a(a-1)/2 <= q < a(a+1)/2 (pdist, a > b) or a(a+1)/2 <= q < (a+1)(a+2)/2 (kernel)
    const double qd = static_cast<double>(q);
    int64_t r = static_cast<int64_t>(diag ? (sqrt(8.0 * qd + 1.0) - 1.0) * 0.5
                                          : (sqrt(8.0 * qd + 1.0) + 1.0) * 0.5);
    auto first = [&](int64_t x) { return diag ? x * (x + 1) / 2 : x * (x - 1) / 2; };
    while (r > 0 && first(r) > q)
      --r;
    while (first(r + 1) <= q)
      ++r;
    a = r;
    b = q - first(r);
  }

  template <typename TT, typename TV, int OP, bool KAHAN>
  __global__ __launch_bounds__(256) void long_pair_kernel(LongParams P)
  {
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    const TV p = static_cast<TV>(P.p);
    const RecT* __restrict__ rrec = static_cast<const RecT*>(P.r_rec);
    const RecT* __restrict__ crec = static_cast<const RecT*>(P.c_rec);

    for (int64_t q = P.q0 + wave; q < P.q1; q += nwaves)
    {
      int64_t a, b;
      pair_of(q, P.out_kind, P.n_cols, a, b);
      const RecT* f = rrec + P.r_off[a];
      const RecT* g = crec + P.c_off[b];
      const int64_t ef = P.r_off[a + 1] - P.r_off[a] - 1; // interior breakpoints of f
      const int64_t eg = P.c_off[b + 1] - P.c_off[b] - 1;
      const int64_t E = ef + eg;

      // T_l: time of merged breakpoint d = E*l/64 (0 for d = 0). Merge path:
      // i = f breakpoints among the first d, f before g on equal times.
      const int64_t d = E * lane / 64;
      TT t0 = TT(0);
      if (d > 0)
      {
        int64_t lo = d - eg > 0 ? d - eg : 0, hi = d < ef ? d : ef;
        while (lo < hi)
        {
          const int64_t mid = (lo + hi) >> 1;
          if (f[mid].t <= g[d - mid - 1].t)
            lo = mid + 1;
          else
            hi = mid;
        }
        const int64_t i = lo, j = d - lo;
        const TT ti = i > 0 ? f[i - 1].t : TT(0);
        const TT tj = j > 0 ? g[j - 1].t : TT(0);
        t0 = tmax(ti, tj);
      }
      // T_{l+1} from the next lane; the last lane runs to the end of the domain
      const TT t1n = __shfl_down(t0, 1, 64);
      const bool last = lane == 63;
      const TT t_end = last ? L::max() : t1n;

      // start state at t0: every breakpoint <= t0 consumed (upper bounds)
      auto upper = [&](const RecT* x, int64_t n) {
        int64_t lo = 0, hi = n;
        while (lo < hi)
        {
          const int64_t mid = (lo + hi) >> 1;
          if (x[mid].t <= t0)
            lo = mid + 1;
          else
            hi = mid;
        }
        return lo;
      };
      using Acc = typename std::conditional<KAHAN, KahanAcc, PlainAcc<TV>>::type;
      Acc acc;
      double part = 0.0;
      if (t0 < t_end)
      {
        const int64_t is = upper(f, ef), js = upper(g, eg);
        const RecT rf = f[is], rg = g[js];
        const bool k0 = rf.t <= rg.t;
        TT lo_t = tmin(rf.t, rg.t), hi_t = tmax(rf.t, rg.t);
        TV hi_v = k0 ? rg.v : rf.v;
        TV x = pair_pre<OP, TV>(rf.v, rg.v);
        const RecT* lo_p = k0 ? f + is + 1 : g + js + 1;
        const RecT* hi_p = k0 ? g + js + 1 : f + is + 1;
        TT t = t0;
        while (lo_t < t_end)
        {
          acc.add(static_cast<TV>(lo_t - t), pair_post<OP, TV>(x, p));
          t = lo_t;
          const RecT r = *lo_p;
          const bool k = r.t <= hi_t;
          x = pair_pre<OP, TV>(r.v, hi_v);
          lo_t = tmin(r.t, hi_t);
          hi_t = tmax(r.t, hi_t);
          const RecT* nlp = k ? lo_p + 1 : hi_p;
          hi_p = k ? hi_p : lo_p + 1;
          lo_p = nlp;
          hi_v = k ? hi_v : r.v;
        }
        if (last)
        {
          part = acc.finish(L::max() - t, pair_post<OP, TV>(x, p)); // final rectangle [t, MAX)
        }
        else
        {
          acc.add(static_cast<TV>(t_end - t), pair_post<OP, TV>(x, p));
          part = acc.finish(TT(0), TV(0));
        }
      }
      // FP64 butterfly over the wavefront
      for (int off = 32; off > 0; off >>= 1)
        part += __shfl_xor(part, off, 64);
      if (lane == 0)
      {
        const double res = finalize<OP>(part, P.p);
        const int64_t i = P.r_perm[a];
        const int64_t j = P.c_perm[b];
        int64_t idx;
        if (P.out_kind == SBAMD_OUT_DISTMAT)
          idx = distmat_index(i, j);
        else if (P.out_kind == SBAMD_OUT_SYMMAT)
          idx = symmat_index(i, j);
        else
          idx = i * P.n_cols + j;
        if (P.out_f64)
          static_cast<double*>(P.out)[idx] = res;
        else
          static_cast<float*>(P.out)[idx] = static_cast<float>(res);
        if (P.row_sum)
          atomicAdd(&P.row_sum[i], res);
      }
    }
  }

  template <typename TT, typename TV, bool KAHAN>
  static void launch_long_typed(const LongParams& prm, int op, int cu_count, hipStream_t stream)
  {
    const int64_t pairs = prm.q1 - prm.q0;
    if (pairs <= 0)
      return;
    const int64_t want = (pairs + 3) / 4; // four wavefronts per workgroup
    const int64_t cap = static_cast<int64_t>(cu_count > 0 ? cu_count : 1) * 8;
    const unsigned grid = static_cast<unsigned>(want < cap ? want : cap);
    switch (op)
    {
    case OPK_L1:
      hipLaunchKernelGGL((long_pair_kernel<TT, TV, OPK_L1, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
      break;
    case OPK_L2:
      hipLaunchKernelGGL((long_pair_kernel<TT, TV, OPK_L2, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
      break;
    case OPK_LP:
      hipLaunchKernelGGL((long_pair_kernel<TT, TV, OPK_LP, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
      break;
    default:
      hipLaunchKernelGGL((long_pair_kernel<TT, TV, OPK_DOT, KAHAN>), dim3(grid), dim3(256), 0, stream, prm);
      break;
    }
    SBAMD_HIP(hipGetLastError());
  }

  void launch_long(const LongParams& prm, int tt_f64, int tv_f64, int op, int kahan, int cu_count,
                   hipStream_t stream)
  {
    if (tt_f64 && tv_f64)
      launch_long_typed<double, double, false>(prm, op, cu_count, stream);
    else if (tt_f64)
      launch_long_typed<double, float, true>(prm, op, cu_count, stream);
    else
      launch_long_typed<float, float, true>(prm, op, cu_count, stream);
    (void)kahan;
  }
}
