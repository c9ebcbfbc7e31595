// Backward of L = sum_e G_e * X_e over the pair outputs X of a pdist (p = 2)
// or l2_kernel plan. No reference counterpart (the reference has no autograd).
//
// For a PCF f with records r (value v_r on [t_start(r), t_end(r))) and its
// partners g_j with pair weights w_j (G / (2 D) for the L2 distance
// D = sqrt(int (f - g)^2), G for the inner product K = int f g):
//   dL/dv_r  (L2)  = 2 (v_r len_r W - Q_r)        (dot) = Q_r
//   dL/dt_r  (L2)  = (v_{r-1} - v_r) ((v_{r-1} + v_r) W - 2 A_{r-1})
//            (dot) = (v_{r-1} - v_r) A_{r-1}
// with W = sum_j w_j, Q_r = sum_j w_j int_{record r} g_j, and A_r = sum_j w_j
// g_j(t_end(r)) (the partner value where record r ends). The tile kernel
// walks every pair once (the forward walk) accumulating Q and A per record:
// w times the integral of the partner over the current record in a
// register, added to the record's LDS slot together with w times the partner
// value when the walk leaves the record; the slots go to global memory once
// per phase. A small finalize kernel applies the formulas above.
//
// Ties: where f and g end a record at the same time, the side whose record
// the walk already holds as the earlier-ending one leaves first and sees the
// partner's left-limit value; the other side leaves second and sees the
// first side's new, right-limit value. The derivative does not exist there;
// for a self pair (kernel diagonal) the two halves add up to the exact
// derivative whichever side leaves first.
#include "walk_common.hpp"

namespace sbamd
{
  template <typename TV, int OP>
  __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void bwd_tile_kernel(BwdParams B)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using TT = double;
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    const FwdParams& P = B.f;

    // 16 bytes of accumulators (Q, A) per staged record, then the W per slot
    const LdsLayout lay = lds_layout(sizeof(TT), P.R, P.C, P.cap, 16);
    const TileLds<TT, TV> S(smem, lay);
    static_assert(sizeof(RecT) == 16, "accumulator slots mirror 16-byte records");
    constexpr uint32_t REC = sizeof(RecT);
    // Q and A are two arrays of one FP64 per staged record (8-byte stride,
    // so the lanes' atomics of one instruction spread over all LDS banks);
    // record at LDS byte address x has its slots at QOFF / AOFF + x / 2.
    const uint32_t QOFF = static_cast<uint32_t>(lay.extra_off);
    const uint32_t AOFF = QOFF + static_cast<uint32_t>(lay.rec_end / 2);
    double* const s_w = reinterpret_cast<double*>(smem + lay.total);

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int NS = P.R + P.C;
    const uint32_t SR = static_cast<uint32_t>(slot_recs(P.cap));

    for (int64_t ti = blockIdx.x; ti < P.ntiles; ti += gridDim.x)
    {
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];

      tile_begin(S, P, tl, pr, tid);
      if (tid < NS)
        s_w[tid] = 0.0;
      __syncthreads();

      // Lanes take the tile's pairs along diagonals, (r, (r + k) mod C) with
      // r = tid mod R and k = tid / R, so a wavefront holds each row and each
      // column only a few times: the lanes of one PCF that leave the same
      // record in one instruction add to the same LDS word, and the
      // hardware serialises those adds.
      const int lr = tid % P.R;
      const int lc = (lr + tid / P.R) % P.C;
      bool valid = tid < P.R * P.C && lr < tl.nr && lc < tl.nc;
      if (valid && (tl.flags & TILE_DIAG))
        valid = P.include_diag ? (lr >= lc) : (lr > lc);

      double w = 0.0;
      if (valid)
      {
        const int64_t i = P.r_perm[pr.row_base + tl.r0 + lr];
        const int64_t j = P.c_perm[pr.col_base + tl.c0 + lc];
        const int64_t idx = (P.out_kind == SBAMD_OUT_DISTMAT ? distmat_index(i, j) : symmat_index(i, j)) + pr.out_base;
        const double G = B.in_f64 ? static_cast<const double*>(B.grad_out)[idx]
                                  : static_cast<double>(static_cast<const float*>(B.grad_out)[idx]);
        if constexpr (OP == OPK_DOT)
        {
          w = G;
        }
        else
        {
          const double D = B.in_f64 ? static_cast<const double*>(B.fwd_out)[idx]
                                    : static_cast<double>(static_cast<const float*>(B.fwd_out)[idx]);
          w = D > 0.0 ? G / (2.0 * D) : 0.0;
        }
        if (w == 0.0)
          valid = false;
      }
      const int slot_f = valid ? lr : 0;
      const int slot_g = valid ? P.R + lc : P.R;
      if (valid)
      {
        atomicAdd(&s_w[slot_f], w);
        atomicAdd(&s_w[slot_g], w);
      }

      // Walk state, values pre-multiplied by the pair weight w: lo is the
      // side whose current record ends first (at lo_t), hi the other side.
      TT t = TT(0);
      TT lo_t = L::inf(), hi_t = L::inf();
      double lo_w = 0.0, hi_w = 0.0;     // w * current value of each side
      double lo_acc = 0.0, hi_acc = 0.0; // w * integral of the partner over the current record
      int lo_s = slot_f, hi_s = slot_g;
      int64_t lo_k = 1, hi_k = 1; // record index of each side's next record

      for (int phase = 0;; ++phase)
      {
        phase_stage<TT, TV, 1>(S, P, phase, tid, nthreads);
        for (int k = tid; k < lay.rec_end / 16; k += nthreads)
          *reinterpret_cast<double2*>(smem + QOFF + 16 * k) = make_double2(0.0, 0.0);
        __syncthreads();
        const TT Tb = L::from_bits(S.tb[phase & 1]);
        phase_advance_cursor<TT, TV, 1>(S, P, Tb, tid);

        if (valid)
        {
          if (phase == 0)
          {
            // record 0 sits at local index 1 (index 0 is the lead slot)
            const RecT rf = S.rec(S.lbase[slot_f] + REC);
            const RecT rg = S.rec(S.lbase[slot_g] + REC);
            const bool k = rf.t <= rg.t;
            lo_t = k ? rf.t : rg.t;
            lo_w = w * static_cast<double>(k ? rf.v : rg.v);
            hi_t = k ? rg.t : rf.t;
            hi_w = w * static_cast<double>(k ? rg.v : rf.v);
            lo_s = k ? slot_f : slot_g;
            hi_s = k ? slot_g : slot_f;
          }
          // window: record k at local k - cur + 1
          uint32_t lo_p = S.lbase[lo_s] + static_cast<uint32_t>(lo_k - S.cur[lo_s] + 1) * REC;
          uint32_t hi_p = S.lbase[hi_s] + static_cast<uint32_t>(hi_k - S.cur[hi_s] + 1) * REC;
          while (lo_t < Tb)
          {
            const double len = lo_t - t;
            lo_acc = fma(hi_w, len, lo_acc);
            hi_acc = fma(lo_w, len, hi_acc);
            t = lo_t;
            // lo leaves its current record (the one before lo_p)
            const uint32_t half = (lo_p - REC) >> 1;
            atomicAdd(reinterpret_cast<double*>(smem + QOFF + half), lo_acc);
            atomicAdd(reinterpret_cast<double*>(smem + AOFF + half), hi_w);
            const RecT r = S.rec(lo_p);
            const double rw = w * static_cast<double>(r.v);
            const uint32_t np = lo_p + REC;
            const bool k = r.t <= hi_t;
            const TT nlt = tmin(r.t, hi_t);
            const double nlw = k ? rw : hi_w;
            const uint32_t nlp = k ? np : hi_p;
            const double nla = k ? 0.0 : hi_acc;
            hi_t = tmax(r.t, hi_t);
            hi_w = k ? hi_w : rw;
            hi_p = k ? hi_p : np;
            hi_acc = k ? hi_acc : 0.0;
            lo_t = nlt;
            lo_w = nlw;
            lo_p = nlp;
            lo_acc = nla;
          }
          // a next-record pointer is at local 1 .. cap + 1 (slot_recs >= cap + 1),
          // so local - 1 identifies the slot even at the end of a full window
          const uint32_t lo_i = lo_p / REC, hi_i = hi_p / REC;
          lo_s = static_cast<int>((lo_i - 1) / SR);
          hi_s = static_cast<int>((hi_i - 1) / SR);
          lo_k = S.cur[lo_s] - 1 + (lo_i - S.lbase[lo_s] / REC);
          hi_k = S.cur[hi_s] - 1 + (hi_i - S.lbase[hi_s] / REC);
        }
        __syncthreads();

        // this phase's accumulators: records cur - 1 .. next - 1 of each slot
        for (int idx = tid; idx < NS * (P.cap + 1); idx += nthreads)
        {
          const int s = idx / (P.cap + 1);
          const int j = idx - s * (P.cap + 1);
          const int64_t k = S.cur[s] - 1 + j;
          if (k >= 0 && k < S.next[s])
          {
            const uint32_t half = (S.lbase[s] + static_cast<uint32_t>(j) * REC) >> 1;
            const double q = *reinterpret_cast<const double*>(smem + QOFF + half);
            const double a = *reinterpret_cast<const double*>(smem + AOFF + half);
            const int64_t g = S.base[s] + k;
            if (q != 0.0)
              atomicAdd(&B.q[g], q);
            if (a != 0.0)
              atomicAdd(&B.a[g], a);
          }
        }
        __syncthreads();
        if (!(Tb < L::inf()))
          break;
      }

      if (valid)
      {
        // final rectangle [t, MAX), then the last records' integrals
        const double len = L::max() - t;
        if (hi_w != 0.0)
          lo_acc = fma(hi_w, len, lo_acc);
        if (lo_w != 0.0)
          hi_acc = fma(lo_w, len, hi_acc);
        atomicAdd(&B.q[S.base[lo_s] + lo_k - 1], lo_acc);
        atomicAdd(&B.q[S.base[hi_s] + hi_k - 1], hi_acc);
      }
      if (tid < NS)
      {
        const bool isrow = tid < P.R;
        const int kk = isrow ? tid : tid - P.R;
        if (kk < (isrow ? tl.nr : tl.nc) && s_w[tid] != 0.0)
          atomicAdd(&B.wsum[isrow ? pr.row_base + tl.r0 + kk : pr.col_base + tl.c0 + kk], s_w[tid]);
      }
      __syncthreads();
    }
  }

  // One thread per record: the formulas in the file comment.
  __global__ void finalize_grad_kernel(FinalizeParams F)
  {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= F.records)
      return;
    int64_t lo = 0, hi = F.n; // PCF q with roff[q] <= r < roff[q + 1]
    while (hi - lo > 1)
    {
      const int64_t mid = (lo + hi) >> 1;
      if (F.roff[mid] <= r)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t q = lo;
    const int64_t first = F.roff[q];
    // 16-byte records: FP64 end time, then the value (FP64 or FP32)
    const unsigned char* rec = static_cast<const unsigned char*>(F.rec);
    auto tend = [&](int64_t k) { return *reinterpret_cast<const double*>(rec + 16 * k); };
    auto val = [&](int64_t k) {
      return F.val_f64 ? *reinterpret_cast<const double*>(rec + 16 * k + 8)
                       : static_cast<double>(*reinterpret_cast<const float*>(rec + 16 * k + 8));
    };
    const double v = val(r);
    const double W = F.wsum[q];
    double gv, gt = 0.0;
    if (F.op == OPK_DOT)
    {
      gv = F.q[r];
      if (r > first)
        gt = (val(r - 1) - v) * F.a[r - 1];
    }
    else
    {
      // the last record ends where the domain does, at DBL_MAX
      const double end = tend(r) < __builtin_inf() ? tend(r) : DBL_MAX;
      const double len = end - (r > first ? tend(r - 1) : 0.0);
      gv = 2.0 * ((v != 0.0 ? v * len * W : 0.0) - F.q[r]);
      if (r > first)
      {
        const double vp = val(r - 1);
        gt = (vp - v) * ((vp + v) * W - 2.0 * F.a[r - 1]);
      }
    }
    const int64_t pt = F.src[r];
    if (pt < 0) // the zero record standing in for an empty PCF
      return;
    if (F.grad_f64)
    {
      static_cast<double*>(F.grad_v)[pt] = gv;
      static_cast<double*>(F.grad_t)[pt] = gt;
    }
    else
    {
      static_cast<float*>(F.grad_v)[pt] = static_cast<float>(gv);
      static_cast<float*>(F.grad_t)[pt] = static_cast<float>(gt);
    }
  }

  size_t bwd_lds_bytes(int tv_f64, int R, int C, int cap)
  {
    (void)tv_f64;
    return static_cast<size_t>(lds_layout(8, R, C, cap, 16).total) + 8 * static_cast<size_t>(R + C);
  }

  void launch_backward(const BwdParams& prm, int op, hipStream_t stream)
  {
    const FwdParams& P = prm.f;
    const unsigned grid = static_cast<unsigned>(P.ntiles < 0x7fffffff ? P.ntiles : 0x7fffffff);
    if (grid == 0)
      return;
    const size_t lds = bwd_lds_bytes(prm.val_f64, P.R, P.C, P.cap);
    const dim3 block(tile_block_threads(P.R, P.C));
    if (prm.val_f64)
    {
      if (op == OPK_DOT)
        hipLaunchKernelGGL((bwd_tile_kernel<double, OPK_DOT>), dim3(grid), block, lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_tile_kernel<double, OPK_L2>), dim3(grid), block, lds, stream, prm);
    }
    else
    {
      if (op == OPK_DOT)
        hipLaunchKernelGGL((bwd_tile_kernel<float, OPK_DOT>), dim3(grid), block, lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_tile_kernel<float, OPK_L2>), dim3(grid), block, lds, stream, prm);
    }
    SBAMD_HIP(hipGetLastError());
  }

  void launch_finalize_grad(const FinalizeParams& prm, hipStream_t stream)
  {
    if (prm.records <= 0)
      return;
    const unsigned grid = static_cast<unsigned>((prm.records + 255) / 256);
    hipLaunchKernelGGL(finalize_grad_kernel, dim3(grid), dim3(256), 0, stream, prm);
    SBAMD_HIP(hipGetLastError());
  }
}
