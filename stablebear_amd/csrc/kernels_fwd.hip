// Forward pairwise integration: one workgroup per pair tile, one or two pairs
// per lane (the NCH template parameter).
//
// Replaces the reference's one-thread-per-pair kernel over global memory
// (include/sbear/cuda/cuda_pcf_kernel.cuh:36-136). Here a workgroup owns an
// R x C tile of pairs from PCFs sorted by record count (so the lanes of a
// wave walk pairs of near-equal length), and stages the R + C PCFs of the tile
// in LDS in time phases: every staged record serves all R (or C) pairs that
// use it. A phase ends at the time Tb where the first tile PCF would overflow
// `cap` staged records; each lane walks its pairs up to Tb, then the next
// records are staged. A wavefront waits for its slowest pair to reach Tb;
// meanwhile lanes whose pairs got there first walk on into the records that
// are already staged past Tb (their budget), so the next phase has less left
// for them, and the next windows start at the phase's cursors with the
// pointers of such pairs moved along.
//
// The walk step is branch-free: one compare picks the side whose record ends
// first, one LDS load fetches that side's next record, and selects update the
// state, so a wavefront issues the same instruction stream for every lane.
// Each step waits on its own LDS load; with NCH = 2 a lane walks two pairs,
// (r, c) and (r, c + ceil(C/2)), and issues both pairs' loads before it
// consumes either, so two independent chains per lane keep the SIMDs issuing
// while loads are in flight (the 32 x 32 shapes the tuner picks for long
// PCFs). With NCH = 1, more workgroups per CU hide the same latency.
#include "walk_common.hpp"

namespace sbamd
{
  // Register budgets: one pair per lane fits 80 VGPRs (6 waves per SIMD);
  // two pairs per lane need about 100 (4 waves per SIMD, so two 32 x 32
  // workgroups of 8 waves per CU). The general-p operation carries an
  // inlined pow and gets twice the registers.
  template <int OP, int NCH>
  constexpr int fwd_min_waves() { return (OP == OPK_LP ? 2 : 4) + (NCH == 1 ? 2 : 0); }

  template <typename TT, typename TV, int OP, bool KAHAN, int NCH>
  __global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(fwd_min_waves<OP, NCH>(), 8))) void fwd_tile_kernel(FwdParams P)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    constexpr uint32_t RB = sizeof(RecT);

    const LdsLayout lay = lds_layout(sizeof(TT), P.R, P.C, P.cap);
    const TileLds<TT, TV> S(smem, lay);

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const TV p = static_cast<TV>(P.p);
    const int half = tile_lane_cols(P.C, NCH);

    // Tiles come either one per workgroup (grid = tiles, the dispatcher hands
    // them out in launch order) or, for persistent workgroups, from per-XCD
    // queues: the dispatcher places workgroup w on XCD w mod 8, which takes
    // tiles from its own queue (tiles that share their PCFs, so they stay in
    // that XCD's L2) through an atomic counter, then from the other queues.
    uint32_t* const s_next_tile = reinterpret_cast<uint32_t*>(smem + lay.misc_off);
    uint32_t queue = blockIdx.x & (DEVICE_QUEUES - 1);
    // row sums (when requested): the tile's R partial sums, one global atomic
    // per row and tile instead of one per pair
    double* const s_rows = reinterpret_cast<double*>(smem + lay.rows_off);
    int64_t ti = blockIdx.x;
    for (;; ti += gridDim.x)
    {
      if (P.work_counter)
      {
        if (tid == 0)
        {
          uint32_t next = 0xffffffffu; // every queue is empty
          for (int tries = 0; tries < DEVICE_QUEUES; ++tries)
          {
            const int64_t qb = P.q_tab ? P.q_tab[queue] : 0;
            const int64_t ql = P.q_tab ? P.q_tab[DEVICE_QUEUES + queue] : (queue == 0 ? P.ntiles : 0);
            const uint32_t t = atomicAdd(P.work_counter + 16 * queue, 1u);
            if (static_cast<int64_t>(t) < ql)
            {
              next = static_cast<uint32_t>(qb) + t;
              break;
            }
            queue = (queue + 1) & (DEVICE_QUEUES - 1);
          }
          *s_next_tile = next;
        }
        __syncthreads();
        const uint32_t nt = *s_next_tile;
        ti = nt == 0xffffffffu ? P.ntiles : static_cast<int64_t>(nt);
      }
      if (ti >= P.ntiles)
        break;
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];
      const bool fold = (tl.flags & TILE_FOLD) != 0;

      tile_begin(S, P, tl, pr, tid);
      if (P.row_sum && tid < P.R)
        s_rows[tid] = 0.0;

      // The lane's pairs: row lr against columns lc + q ceil(C / NCH). A pair
      // outside the tile (edge tiles, the upper half of diagonal tiles) walks
      // the pair of slots 0 and R instead, which every tile has, and stores
      // nothing.
      const int lr = tid / half;
      const int lc = tid - lr * half;
      bool valid[NCH];
      int slot_f[NCH], slot_g[NCH];
      uint32_t wb_f[NCH], wb_g[NCH];
#pragma unroll
      for (int q = 0; q < NCH; ++q)
      {
        const int c = lc + q * half;
        bool v = lr < tl.nr && c < tl.nc;
        // tiles that reach the diagonal of a symmetric plan: only pairs
        // below it (or on it, for kernel matrices), in sorted positions
        if (v && (tl.flags & TILE_DIAG))
          v = P.include_diag ? (tl.r0 + lr >= tl.c0 + c) : (tl.r0 + lr > tl.c0 + c);
        valid[q] = v;
        slot_f[q] = v ? lr : 0;
        slot_g[q] = v ? P.R + c : P.R;
        wb_f[q] = S.win(slot_f[q]);
        wb_g[q] = S.win(slot_g[q]);
      }

      // Per pair: the two current records ordered by end time. The next union
      // breakpoint is always lo.t, and the side that owns lo advances (which
      // PCF lo belongs to does not matter to a symmetric pair op). lo_p /
      // hi_p: LDS address of the owning PCF's next record; which window a
      // pointer lies in tells its side (row windows precede column windows).
      // x is the first half of the pair op of the current segment
      // (pair_pre); lo's value is not kept, since after a step the two
      // current values are {new record's value, hi_v} whichever side
      // advanced, and every pair op is symmetric in its operands. len is the
      // length of the current segment [previous breakpoint, lo_t), formed in
      // the time type when the segment starts and rounded to the value type.
      TT lo_t[NCH], hi_t[NCH];
      TV len[NCH], x[NCH], hi_v[NCH];
      uint32_t lo_p[NCH], hi_p[NCH];
      using Acc = typename std::conditional<KAHAN, BlockKahanAcc, PlainAcc<TV>>::type;
      Acc acc[NCH];
#pragma unroll
      for (int q = 0; q < NCH; ++q)
      {
        lo_t[q] = hi_t[q] = L::inf();
        len[q] = x[q] = hi_v[q] = TV(0);
        lo_p[q] = hi_p[q] = 0;
      }
      __syncthreads();

      for (int phase = 0;; ++phase)
      {
        phase_stage(S, P, tid, nthreads);
        __syncthreads();
        const TT Tb = L::from_bits(*S.tb);
        phase_cursors(S, P, Tb, tid);
        __syncthreads();

        // Steps of each pair to reach the phase end: every record of either
        // PCF that starts before Tb and is not yet consumed (the slots'
        // cursor counts less the records the pair took past the previous
        // phase end; a pair is either before Tb on both sides or past it on
        // both); in the first phase the two first records are read before
        // the walk. A pair may go on past Tb while the next record of its
        // walk is staged, up to the earlier of its two slots' bounds, so
        // lanes whose pairs reach Tb early keep walking while the wavefront
        // waits for its slowest pair. Its budget counts the steps it can
        // surely take: those to Tb and then the rest of the staged records of
        // the side whose bound is earlier, each of which starts before that
        // bound.
        uint32_t n_need[NCH], n_cap[NCH];
#pragma unroll
        for (int q = 0; q < NCH; ++q)
        {
          if (phase == 0)
          {
            const RecT rf = lds_rec(&S.rec(wb_f[q]));
            const RecT rg = lds_rec(&S.rec(wb_g[q]));
            const bool k = rf.t <= rg.t;
            lo_t[q] = k ? rf.t : rg.t;
            hi_t[q] = k ? rg.t : rf.t;
            hi_v[q] = k ? rg.v : rf.v;
            x[q] = pair_pre<OP, TV>(rf.v, rg.v);
            len[q] = static_cast<TV>(lo_t[q] - TT(0));
            lo_p[q] = (k ? wb_f[q] : wb_g[q]) + RB;
            hi_p[q] = (k ? wb_g[q] : wb_f[q]) + RB;
          }
          const bool lg = lo_p[q] >= wb_g[q];
          const int of = static_cast<int>(((lg ? hi_p[q] : lo_p[q]) - wb_f[q]) / RB);
          const int og = static_cast<int>(((lg ? lo_p[q] : hi_p[q]) - wb_g[q]) / RB);
          const uint32_t tf = S.cnt[slot_f[q]], tg = S.cnt[slot_g[q]];
          const int cf = static_cast<int>(tf & 0xffffu), cg = static_cast<int>(tg & 0xffffu);
          const int n = cf + cg - of - og;
          n_need[q] = n > 0 ? static_cast<uint32_t>(n) : 0u;
          using Bits = typename L::Bits;
          const bool fb = L::from_bits(static_cast<Bits>(S.bnd[slot_f[q]])) <=
                          L::from_bits(static_cast<Bits>(S.bnd[slot_g[q]]));
          const int ns_b = static_cast<int>((fb ? tf : tg) >> 16);
          const int c_b = fb ? cf : cg;
          const int o_b = fb ? of : og;
          n_cap[q] = n_need[q] + static_cast<uint32_t>(ns_b - (o_b > c_b ? o_b : c_b));
        }
        // One merge step of pair q on the record r that lo_p[q] pointed at:
        // accumulate the segment ending at lo.t, then replace lo by r and
        // restore the order.
        auto step = [&](int q, const RecT& r) {
          acc[q].add(len[q], pair_post<OP, TV>(x[q], p));
          const uint32_t np = lo_p[q] + RB;
          const bool k = r.t <= hi_t[q];
          x[q] = pair_pre<OP, TV>(r.v, hi_v[q]);
          const TT nlt = tmin(r.t, hi_t[q]);
          len[q] = static_cast<TV>(nlt - lo_t[q]);
          lo_t[q] = nlt;
          hi_t[q] = tmax(r.t, hi_t[q]);
          const uint32_t nlp = k ? np : hi_p[q];
          hi_p[q] = k ? hi_p[q] : np;
          lo_p[q] = nlp;
          hi_v[q] = k ? hi_v[q] : r.v;
        };
        // all pairs of the lane, loads first (NCH independent chains)
        auto step_all = [&]() {
          RecT r[NCH];
#pragma unroll
          for (int q = 0; q < NCH; ++q)
            r[q] = lds_rec(&S.rec(lo_p[q]));
#pragma unroll
          for (int q = 0; q < NCH; ++q)
            step(q, r[q]);
        };
        auto flush_all = [&]() {
#pragma unroll
          for (int q = 0; q < NCH; ++q)
            acc[q].flush();
        };
        auto step_one = [&](int q) { step(q, lds_rec(&S.rec(lo_p[q]))); };

        // The wavefront walks until its slowest pair reaches Tb (the largest
        // count), each lane at most its budget: (1) every pair of every lane
        // up to the smallest of those limits over the wavefront, eight steps
        // per compensated add; then (2) each pair up to the largest count,
        // lanes past their budget masked off, four steps per compensated add.
        uint32_t n_top[NCH], c_min = 0x7fffffffu;
#pragma unroll
        for (int q = 0; q < NCH; ++q)
        {
          n_top[q] = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(wave_max(n_need[q]))));
          c_min = n_cap[q] < c_min ? n_cap[q] : c_min;
        }
        int n_all = __builtin_amdgcn_readfirstlane(static_cast<int>(wave_min(c_min)));
#pragma unroll
        for (int q = 0; q < NCH; ++q)
          n_all = n_all < static_cast<int>(n_top[q]) ? n_all : static_cast<int>(n_top[q]);
        int done = 0;
        // sixteen steps per trip (two compensated adds): 0.4 % faster on
        // L1 pdists than eight, the loop's branch and counter halved
        for (; done + 16 <= n_all; done += 16)
        {
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          flush_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          flush_all();
        }
        for (; done + 8 <= n_all; done += 8)
        {
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          step_all();
          flush_all();
        }
        if (done + 4 <= n_all)
        {
          step_all();
          step_all();
          step_all();
          step_all();
          flush_all();
          done += 4;
        }
        if (done + 2 <= n_all)
        {
          step_all();
          step_all();
          done += 2;
        }
        if (done < n_all)
        {
          step_all();
          ++done;
        }
#pragma unroll
        for (int q = 0; q < NCH; ++q)
        {
          const int nt = static_cast<int>(n_top[q]);
          // steps left to this lane in the phase
          const int r = static_cast<int>(n_cap[q]) - done;
          int d = 0;
          for (; done + d + 4 <= nt; d += 4)
          {
            if (d < r)
              step_one(q);
            if (d + 1 < r)
              step_one(q);
            if (d + 2 < r)
              step_one(q);
            if (d + 3 < r)
              step_one(q);
            acc[q].flush();
          }
          for (; done + d < nt; ++d)
            if (d < r)
              step_one(q);
          acc[q].flush();
          if (fold)
            acc[q].fold();
        }
        const bool last = !(Tb < L::inf());
        if (!last)
        {
          // the next windows start at this phase's cursors: move the
          // pointers of pairs that went past Tb along with their records
#pragma unroll
          for (int q = 0; q < NCH; ++q)
          {
            const bool lg = lo_p[q] >= wb_g[q];
            const uint32_t sf = (S.cnt[slot_f[q]] & 0xffffu) * RB;
            const uint32_t sg = (S.cnt[slot_g[q]] & 0xffffu) * RB;
            lo_p[q] -= lg ? sg : sf;
            hi_p[q] -= lg ? sf : sg;
          }
        }
        __syncthreads();
        if (last)
          break;
      }

      // final rectangle [t, MAX) from the last union breakpoint t: the
      // later of the two PCFs' last record starts
      auto last_start = [&](bool row, int64_t g) {
        const int64_t* off = row ? P.r_off : P.c_off;
        const RecT* rec = static_cast<const RecT*>(row ? P.r_rec : P.c_rec);
        const int64_t b = off[g], m = off[g + 1] - b;
        return m >= 2 ? rec[b + m - 2].t : TT(0);
      };
#pragma unroll
      for (int q = 0; q < NCH; ++q)
      {
        if (!valid[q])
          continue;
        const int c = lc + q * half;
        const int64_t gr = pr.row_base + tl.r0 + lr;
        const int64_t gc = pr.col_base + tl.c0 + c;
        const TT t = tmax(last_start(true, gr), last_start(false, gc));
        const TT tail = L::max() - t;
        const double raw = acc[q].finish(tail, pair_post<OP, TV>(x[q], p));
        const double res = finalize<OP>(raw, P.p);
        const int64_t i = P.r_perm[gr];
        const int64_t j = P.c_perm[gc];
        int64_t idx;
        if (P.out_kind == SBAMD_OUT_DISTMAT)
          idx = distmat_index(i, j);
        else if (P.out_kind == SBAMD_OUT_SYMMAT)
          idx = symmat_index(i, j);
        else
          idx = i * P.out_cols + j;
        idx += pr.out_base;
        // the outputs are scattered through the permutation: streaming
        // stores keep them from taking L2 lines that the staged records need
        if (P.out_f64)
          __builtin_nontemporal_store(res, static_cast<double*>(P.out) + idx);
        else
          __builtin_nontemporal_store(static_cast<float>(res), static_cast<float*>(P.out) + idx);
        if (P.row_sum)
          atomicAdd(&s_rows[lr], res);
      }
      __syncthreads();
      if (P.row_sum && tid < tl.nr)
        atomicAdd(&P.row_sum[P.r_perm[pr.row_base + tl.r0 + tid]], s_rows[tid]);
    }
  }

  // LDS of the forward kernel (staging layout, tables, R row-sum partials).
  size_t fwd_lds_bytes(int tt_f64, int tv_f64, int R, int C, int cap)
  {
    (void)tv_f64;
    return static_cast<size_t>(lds_layout(tt_f64 ? 8 : 4, R, C, cap).total);
  }

  template <typename TT, typename TV, bool KAHAN, int NCH>
  static void launch_typed(const FwdParams& prm, int op, hipStream_t stream)
  {
    const size_t lds = fwd_lds_bytes(sizeof(TT) == 8, 0, prm.R, prm.C, prm.cap);
    if (prm.ntiles <= 0)
      return;
    if (!tile_shape_ok(prm.R, prm.C, prm.cap, NCH))
      fail(SBAMD_ERR_INVALID, "tile shape not supported by the tile kernel");
    const unsigned threads = static_cast<unsigned>(tile_block_threads(prm.R, prm.C, NCH));
    void (*kern)(FwdParams) = nullptr;
    switch (op)
    {
    case OPK_L1:
      kern = fwd_tile_kernel<TT, TV, OPK_L1, KAHAN, NCH>;
      break;
    case OPK_L2:
      kern = fwd_tile_kernel<TT, TV, OPK_L2, KAHAN, NCH>;
      break;
    case OPK_LP:
      kern = fwd_tile_kernel<TT, TV, OPK_LP, KAHAN, NCH>;
      break;
    default:
      kern = fwd_tile_kernel<TT, TV, OPK_DOT, KAHAN, NCH>;
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

  template <typename TT, typename TV, bool KAHAN>
  static void launch_ppl(const FwdParams& prm, int op, hipStream_t stream)
  {
    if (prm.ppl == 1)
      launch_typed<TT, TV, KAHAN, 1>(prm, op, stream);
    else
      launch_typed<TT, TV, KAHAN, 2>(prm, op, stream);
  }

  void launch_forward(const FwdParams& prm, int tt_f64, int tv_f64, int op, int kahan,
                      hipStream_t stream)
  {
    if (tt_f64 && tv_f64 && !kahan)
      launch_ppl<double, double, false>(prm, op, stream);
    else if (tt_f64 && !tv_f64 && kahan)
      launch_ppl<double, float, true>(prm, op, stream);
    else if (!tt_f64 && !tv_f64 && kahan)
      launch_ppl<float, float, true>(prm, op, stream);
    else
      fail(SBAMD_ERR_INVALID, "unsupported record layout / precision combination");
  }
}
