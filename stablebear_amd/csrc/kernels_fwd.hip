// Forward pairwise integration: one workgroup per pair tile, one lane per pair.
//
// Replaces the reference's one-thread-per-pair kernel over global memory
// (include/sbear/cuda/cuda_pcf_kernel.cuh:36-136). Here a workgroup owns an
// R x C tile of pairs from PCFs sorted by record count (so the lanes of a
// wave walk pairs of near-equal length), and stages the R + C PCFs of the tile
// in LDS in time phases: every staged record serves all R (or C) pairs that
// use it. A phase ends at the time Tb where the first tile PCF would overflow
// `cap` staged records; each lane walks its pair up to Tb, then the next
// records are staged. Records are stored slot-interleaved ([record][slot]) so
// lanes reading different PCFs hit different LDS banks.
#include "walk_common.hpp"

namespace sbamd
{
  template <typename TT, typename TV, int OP, bool KAHAN>
  __global__ __launch_bounds__(256) void fwd_tile_kernel(FwdParams P)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using L = Lim<TT>;
    using Bits = typename L::Bits;

    const int NS = P.R + P.C;
    const int stride = P.stride;
    const int cap = P.cap;
    const LdsLayout lay = lds_layout(sizeof(TT), sizeof(TV), NS, stride, cap, 0);
    TT* s_tend = reinterpret_cast<TT*>(smem + lay.tend_off);
    TV* s_val = reinterpret_cast<TV*>(smem + lay.val_off);
    int64_t* s_base = reinterpret_cast<int64_t*>(smem + lay.base_off);
    int64_t* s_m = reinterpret_cast<int64_t*>(smem + lay.m_off);
    int64_t* s_cur = reinterpret_cast<int64_t*>(smem + lay.cur_off);
    int64_t* s_next = reinterpret_cast<int64_t*>(smem + lay.next_off);
    TT* s_tstart = reinterpret_cast<TT*>(smem + lay.tstart_off);
    Bits* s_tb = reinterpret_cast<Bits*>(smem + lay.tb_off);

    const TT* __restrict__ g_rt = static_cast<const TT*>(P.r_tend);
    const TV* __restrict__ g_rv = static_cast<const TV*>(P.r_val);
    const TT* __restrict__ g_ct = static_cast<const TT*>(P.c_tend);
    const TV* __restrict__ g_cv = static_cast<const TV*>(P.c_val);

    const int tid = threadIdx.x;
    const TV p = static_cast<TV>(P.p);

    for (int64_t ti = blockIdx.x; ti < P.ntiles; ti += gridDim.x)
    {
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];
      const bool fold = (tl.flags & TILE_FOLD) != 0;

      if (tid < NS)
      {
        const bool isrow = tid < P.R;
        const int k = isrow ? tid : tid - P.R;
        const int lim = isrow ? tl.nr : tl.nc;
        const int64_t* off = isrow ? P.r_off : P.c_off;
        if (k < lim)
        {
          const int64_t g = isrow ? pr.row_base + tl.r0 + k : pr.col_base + tl.c0 + k;
          const int64_t b = off[g];
          s_base[tid] = b;
          s_m[tid] = off[g + 1] - b;
        }
        else
        {
          s_base[tid] = 0;
          s_m[tid] = 0;
        }
        s_next[tid] = 0;
        s_tstart[tid] = TT(0);
      }
      if (tid == 0)
      {
        s_tb[0] = L::bits(L::inf());
        s_tb[1] = L::bits(L::inf());
      }
      __syncthreads();

      // Pair of this lane (sorted positions a = r0 + lr, b = c0 + lc).
      const int lr = tid / P.C;
      const int lc = tid - lr * P.C;
      bool valid = tid < P.R * P.C && lr < tl.nr && lc < tl.nc;
      if (valid && (tl.flags & TILE_DIAG))
        valid = P.include_diag ? (lr >= lc) : (lr > lc);
      const int slot_f = lr;
      const int slot_g = P.R + lc;

      TT t = TT(0);
      TT tf = L::inf(), tg = L::inf();
      TV fv = TV(0), gv = TV(0);
      int64_t nf = 1, ng = 1; // global record index (within the PCF) of the next record
      using Acc = typename std::conditional<KAHAN, KahanAcc, PlainAcc<TV>>::type;
      Acc acc;

      for (int phase = 0;; ++phase)
      {
        // ---- stage the records that start before this phase's end time ----
        if (tid < NS)
        {
          const int64_t cur = s_next[tid];
          s_cur[tid] = cur;
          const int64_t rem = s_m[tid] - cur;
          TT tbs = L::inf();
          if (rem > cap)
          {
            const TT* gt = tid < P.R ? g_rt : g_ct;
            tbs = gt[s_base[tid] + cur + cap - 1];
          }
          atomicMin(&s_tb[phase & 1], L::bits(tbs));
        }
        for (int idx = tid; idx < NS * cap; idx += blockDim.x)
        {
          const int s = idx / cap;
          const int j = idx - s * cap;
          const int64_t k = s_next[s] + j;
          if (k < s_m[s])
          {
            const bool isrow = s < P.R;
            const int64_t gk = s_base[s] + k;
            s_tend[j * stride + s] = isrow ? g_rt[gk] : g_ct[gk];
            s_val[j * stride + s] = isrow ? g_rv[gk] : g_cv[gk];
          }
        }
        if (tid == 0)
          s_tb[(phase + 1) & 1] = L::bits(L::inf());
        __syncthreads();

        const TT Tb = L::from_bits(s_tb[phase & 1]);

        // ---- next cursor of each slot: first staged record starting at or after Tb ----
        if (tid < NS)
        {
          const int64_t cur = s_cur[tid];
          const int64_t rem = s_m[tid] - cur;
          const int lim = rem < cap ? static_cast<int>(rem) : cap;
          int lo = 0, hi = lim;
          while (lo < hi)
          {
            const int mid = (lo + hi) >> 1;
            const TT ts = mid == 0 ? s_tstart[tid] : s_tend[(mid - 1) * stride + tid];
            if (ts < Tb)
              lo = mid + 1;
            else
              hi = mid;
          }
          s_next[tid] = cur + lo;
          if (lo > 0)
            s_tstart[tid] = s_tend[(lo - 1) * stride + tid];
        }

        // ---- walk this lane's pair up to Tb ----
        if (valid)
        {
          const int64_t cf = s_cur[slot_f];
          const int64_t cg = s_cur[slot_g];
          if (phase == 0)
          {
            tf = s_tend[slot_f];
            fv = s_val[slot_f];
            tg = s_tend[slot_g];
            gv = s_val[slot_g];
          }
          int lf = static_cast<int>(nf - cf);
          int lg = static_cast<int>(ng - cg);
          for (;;)
          {
            const bool af = tf <= tg;
            const TT tn = af ? tf : tg;
            if (!(tn < Tb))
              break;
            const TV dt = static_cast<TV>(tn - t);
            acc.add(dt, pair_op<OP, TV>(fv, gv, p));
            t = tn;
            const int j = af ? lf : lg;
            const int s = af ? slot_f : slot_g;
            const TT tnew = s_tend[j * stride + s];
            const TV vnew = s_val[j * stride + s];
            if (af)
            {
              tf = tnew;
              fv = vnew;
              ++lf;
            }
            else
            {
              tg = tnew;
              gv = vnew;
              ++lg;
            }
          }
          nf = cf + lf;
          ng = cg + lg;
          if (fold)
            acc.fold();
        }
        __syncthreads();
        if (!(Tb < L::inf()))
          break;
      }

      if (valid)
      {
        const TV w = pair_op<OP, TV>(fv, gv, p);
        const TT tail = L::max() - t;
        const double raw = acc.finish(tail, w);
        const double res = finalize<OP>(raw, P.p);
        const int64_t i = P.r_perm[pr.row_base + tl.r0 + lr];
        const int64_t j = P.c_perm[pr.col_base + tl.c0 + lc];
        int64_t idx;
        if (P.out_kind == SBAMD_OUT_DISTMAT)
          idx = distmat_index(i, j);
        else if (P.out_kind == SBAMD_OUT_SYMMAT)
          idx = symmat_index(i, j);
        else
          idx = i * P.out_cols + j;
        idx += pr.out_base;
        if (P.out_f64)
          static_cast<double*>(P.out)[idx] = res;
        else
          static_cast<float*>(P.out)[idx] = static_cast<float>(res);
        if (P.row_sum)
          atomicAdd(&P.row_sum[i], res);
      }
      __syncthreads();
    }
  }

  size_t fwd_lds_bytes(int tt_f64, int tv_f64, int R, int C, int cap)
  {
    const int NS = R + C;
    return static_cast<size_t>(lds_layout(tt_f64 ? 8 : 4, tv_f64 ? 8 : 4, NS, NS, cap, 0).total);
  }

  template <typename TT, typename TV, bool KAHAN>
  static void launch_typed(const FwdParams& prm, int op, hipStream_t stream)
  {
    const int NS = prm.R + prm.C;
    const size_t lds = static_cast<size_t>(
        lds_layout(sizeof(TT), sizeof(TV), NS, prm.stride, prm.cap, 0).total);
    const unsigned grid = static_cast<unsigned>(prm.ntiles < 0x7fffffff ? prm.ntiles : 0x7fffffff);
    if (grid == 0)
      return;
    const dim3 block(tile_block_threads(prm.R, prm.C));
    switch (op)
    {
    case OPK_L1:
      hipLaunchKernelGGL((fwd_tile_kernel<TT, TV, OPK_L1, KAHAN>), dim3(grid), block, lds, stream, prm);
      break;
    case OPK_L2:
      hipLaunchKernelGGL((fwd_tile_kernel<TT, TV, OPK_L2, KAHAN>), dim3(grid), block, lds, stream, prm);
      break;
    case OPK_LP:
      hipLaunchKernelGGL((fwd_tile_kernel<TT, TV, OPK_LP, KAHAN>), dim3(grid), block, lds, stream, prm);
      break;
    default:
      hipLaunchKernelGGL((fwd_tile_kernel<TT, TV, OPK_DOT, KAHAN>), dim3(grid), block, lds, stream, prm);
      break;
    }
    SBAMD_HIP(hipGetLastError());
  }

  void launch_forward(const FwdParams& prm, int tt_f64, int tv_f64, int op, int kahan,
                      hipStream_t stream)
  {
    if (tt_f64 && tv_f64 && !kahan)
      launch_typed<double, double, false>(prm, op, stream);
    else if (tt_f64 && !tv_f64 && kahan)
      launch_typed<double, float, true>(prm, op, stream);
    else if (!tt_f64 && !tv_f64 && kahan)
      launch_typed<float, float, true>(prm, op, stream);
    else
      fail(SBAMD_ERR_INVALID, "unsupported record layout / precision combination");
  }
}
