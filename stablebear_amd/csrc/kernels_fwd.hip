// Forward pairwise integration: one workgroup per pair tile, one lane per pair.
//
// Replaces the reference's one-thread-per-pair kernel over global memory
// (include/sbear/cuda/cuda_pcf_kernel.cuh:36-136). Here a workgroup owns an
// R x C tile of pairs from PCFs sorted by record count (so the lanes of a
// wave walk pairs of near-equal length), and stages the R + C PCFs of the tile
// in LDS in time phases: every staged record serves all R (or C) pairs that
// use it. A phase ends at the time Tb where the first tile PCF would overflow
// `cap` staged records; each lane walks its pair up to Tb, then the next
// records are staged.
//
// The walk step is branch-free: one compare picks the side whose record ends
// first, one LDS load fetches that side's next record, and selects update the
// state, so a wavefront issues the same instruction stream for every lane.
#include "walk_common.hpp"

namespace sbamd
{
  template <typename TT, typename TV, int OP, bool KAHAN>
  // 8 waves per SIMD (at most 64 VGPRs): the walk is latency-bound on its
  // dependent LDS loads, so residency matters more than registers. The
  // general-p operation carries an inlined pow, which would spill at 64 VGPRs.
  __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OP == OPK_LP ? 4 : 6, 8))) void fwd_tile_kernel(FwdParams P)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;

    const LdsLayout lay = lds_layout(sizeof(TT), P.R, P.C, P.cap, 0);
    const TileLds<TT, TV> S(smem, lay);
    constexpr uint32_t SF = sizeof(RecT); // record stride within a slot

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const TV p = static_cast<TV>(P.p);

    // Tiles come either one per workgroup (grid = tiles, the dispatcher hands
    // them out in launch order) or, for persistent workgroups, from an atomic
    // counter over the longest-first tile list.
    uint32_t* const s_next_tile = reinterpret_cast<uint32_t*>(smem + lay.misc_off);
    // row sums (when requested): the tile's R partial sums, one global atomic
    // per row and tile instead of one per pair
    double* const s_rows = reinterpret_cast<double*>(smem + lay.total);
    int64_t ti = blockIdx.x;
    for (;; ti += gridDim.x)
    {
      if (P.work_counter)
      {
        if (tid == 0)
          *s_next_tile = atomicAdd(P.work_counter, 1u);
        __syncthreads();
        ti = *s_next_tile;
      }
      if (ti >= P.ntiles)
        break;
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];
      const bool fold = (tl.flags & TILE_FOLD) != 0;

      tile_begin(S, P, tl, pr, tid);
      if (P.row_sum && tid < P.R)
        s_rows[tid] = 0.0;
      __syncthreads();

      // Pair of this lane (sorted positions a = r0 + lr, b = c0 + lc).
      const int lr = tid / P.C;
      const int lc = tid - lr * P.C;
      bool valid = tid < P.R * P.C && lr < tl.nr && lc < tl.nc;
      if (valid && (tl.flags & TILE_DIAG))
        valid = P.include_diag ? (lr >= lc) : (lr > lc);
      const int slot_f = valid ? lr : 0;
      const int slot_g = valid ? P.R + lc : P.R;

      // The two current records of the pair, ordered by end time: the next
      // union breakpoint is always lo.t, and the side that owns lo advances.
      // (Which PCF lo belongs to does not matter to a symmetric pair op.)
      // lo_p / hi_p: LDS address of the owning PCF's next record; between
      // phases they are kept as (slot, record index) instead. x is the first
      // half of the pair op of the current segment (pair_pre); lo's value is
      // not kept, since after a step the two current values are {new
      // record's value, hi_v} whichever side advanced, and every pair op is
      // symmetric in its operands. len is the length of the current segment
      // [previous breakpoint, lo_t), formed in the time type when the
      // segment starts and rounded to the value type.
      TT lo_t = L::inf(), hi_t = L::inf();
      TV len = TV(0), x = TV(0), hi_v = TV(0);
      int lo_s = slot_f, hi_s = slot_g;  // owning slots
      int64_t lo_k = 1, hi_k = 1;        // record index of each owner's next record
      using Acc = typename std::conditional<KAHAN, KahanAcc, PlainAcc<TV>>::type;
      Acc acc;

      for (int phase = 0;; ++phase)
      {
        phase_stage(S, P, phase, tid, nthreads);
        __syncthreads();
        const TT Tb = L::from_bits(S.tb[phase & 1]);
        phase_advance_cursor(S, P, Tb, tid);
        __syncthreads();
        // Steps of this lane's walk in the phase: every record of either PCF
        // that ends before Tb is consumed, so the walk leaves each side's
        // next-record index at that slot's new cursor. The wavefront runs the
        // smallest count over its lanes under a scalar loop counter (no
        // per-step time compare), then the rest under the compare.
        uint32_t n_walk = 0x7fffffffu;
        if (valid)
          n_walk = static_cast<uint32_t>(S.next[slot_f] + S.next[slot_g] - (lo_k + hi_k));
        const int n_uniform = static_cast<int>(wave_min(n_walk));

        if (valid)
        {
          if (phase == 0)
          {
            const RecT rf = S.rec(S.lbase[slot_f]);
            const RecT rg = S.rec(S.lbase[slot_g]);
            const bool k = rf.t <= rg.t;
            lo_t = k ? rf.t : rg.t;
            hi_t = k ? rg.t : rf.t;
            hi_v = k ? rg.v : rf.v;
            x = pair_pre<OP, TV>(rf.v, rg.v);
            len = static_cast<TV>(lo_t - TT(0));
            lo_s = k ? slot_f : slot_g;
            hi_s = k ? slot_g : slot_f;
          }
          const int64_t lo_c = S.cur[lo_s], hi_c = S.cur[hi_s];
          const RecT* const recs = reinterpret_cast<const RecT*>(smem);
          const RecT* lo_p = recs + (S.lbase[lo_s] / SF + static_cast<uint32_t>(lo_k - lo_c));
          const RecT* hi_p = recs + (S.lbase[hi_s] / SF + static_cast<uint32_t>(hi_k - hi_c));
          // One merge step: accumulate the segment ending at lo.t, then
          // replace lo by the next record of its PCF and restore the order.
          // Two steps per trip let registers alternate roles.
          auto step = [&]() {
            acc.add(len, pair_post<OP, TV>(x, p));
            const RecT r = lds_rec(lo_p);
            const RecT* np = lo_p + 1;
            const bool k = r.t <= hi_t;
            x = pair_pre<OP, TV>(r.v, hi_v);
            const TT nlt = tmin(r.t, hi_t);
            len = static_cast<TV>(nlt - lo_t);
            lo_t = nlt;
            hi_t = tmax(r.t, hi_t);
            const RecT* nlp = k ? np : hi_p;
            hi_p = k ? hi_p : np;
            lo_p = nlp;
            hi_v = k ? hi_v : r.v;
          };
          int n = 0;
          for (; n + 2 <= n_uniform; n += 2)
          {
            step();
            step();
          }
          if (n < n_uniform)
            step();
          while (lo_t < Tb)
          {
            step();
            if (!(lo_t < Tb))
              break;
            step();
          }
          // back to (slot, record index). A next-record pointer lies within
          // local 0 .. cap of its slot's window (slot_recs > cap), and the
          // row slot precedes the column slot, so one compare with the
          // column window's start tells the owners apart.
          const uint32_t lo_i = static_cast<uint32_t>(lo_p - recs);
          const uint32_t hi_i = static_cast<uint32_t>(hi_p - recs);
          const bool lo_in_g = lo_i >= S.lbase[slot_g] / SF;
          lo_s = lo_in_g ? slot_g : slot_f;
          hi_s = lo_in_g ? slot_f : slot_g;
          lo_k = S.cur[lo_s] + (lo_i - S.lbase[lo_s] / SF);
          hi_k = S.cur[hi_s] + (hi_i - S.lbase[hi_s] / SF);
          if (fold)
            acc.fold();
        }
        __syncthreads();
        if (!(Tb < L::inf()))
          break;
      }

      if (valid)
      {
        // final rectangle [t, MAX) from the last union breakpoint t: the
        // later of the two PCFs' last record starts
        auto last_start = [&](int s, const void* rec) {
          const int64_t m = S.m[s];
          return m >= 2 ? static_cast<const RecT*>(rec)[S.base[s] + m - 2].t : TT(0);
        };
        const TT t = tmax(last_start(slot_f, P.r_rec), last_start(slot_g, P.c_rec));
        const TT tail = L::max() - t;
        const double raw = acc.finish(tail, pair_post<OP, TV>(x, p));
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
          atomicAdd(&s_rows[lr], res);
      }
      __syncthreads();
      if (P.row_sum && tid < tl.nr)
        atomicAdd(&P.row_sum[P.r_perm[pr.row_base + tl.r0 + tid]], s_rows[tid]);
    }
  }

  // LDS of the forward kernel: the staging layout plus R row-sum partials.
  size_t fwd_lds_bytes(int tt_f64, int tv_f64, int R, int C, int cap)
  {
    (void)tv_f64;
    return static_cast<size_t>(lds_layout(tt_f64 ? 8 : 4, R, C, cap, 0).total) + 8 * static_cast<size_t>(R);
  }

  template <typename TT, typename TV, bool KAHAN>
  static void launch_typed(const FwdParams& prm, int op, hipStream_t stream)
  {
    const size_t lds = fwd_lds_bytes(sizeof(TT) == 8, 0, prm.R, prm.C, prm.cap);
    if (prm.ntiles <= 0)
      return;
    const unsigned threads = tile_block_threads(prm.R, prm.C);
    void (*kern)(FwdParams) = nullptr;
    switch (op)
    {
    case OPK_L1:
      kern = fwd_tile_kernel<TT, TV, OPK_L1, KAHAN>;
      break;
    case OPK_L2:
      kern = fwd_tile_kernel<TT, TV, OPK_L2, KAHAN>;
      break;
    case OPK_LP:
      kern = fwd_tile_kernel<TT, TV, OPK_LP, KAHAN>;
      break;
    default:
      kern = fwd_tile_kernel<TT, TV, OPK_DOT, KAHAN>;
      break;
    }
    int64_t grid = prm.ntiles;
    if (prm.work_counter)
    {
      // persistent: as many workgroups as can be resident at once
      int per_cu = 0, dev = 0, cus = 0;
      SBAMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, static_cast<int>(threads), lds));
      SBAMD_HIP(hipGetDevice(&dev));
      SBAMD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const int64_t resident = static_cast<int64_t>(per_cu > 0 ? per_cu : 1) * cus;
      grid = grid < resident ? grid : resident;
    }
    if (grid > 0x7fffffff)
      grid = 0x7fffffff;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(threads), lds, stream, prm);
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
