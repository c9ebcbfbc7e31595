// Backward of L = sum_e G_e * X_e over the pair outputs X of a pdist (p = 2)
// or l2_kernel plan. No reference counterpart (the reference has no autograd).
//
// Per pair (f, g) with weight w (G / (2 D) for the L2 distance D = sqrt(S),
// S = int (f - g)^2; G for the inner product K = int f g), walking the union
// segments [t, tn) with values (fv, gv):
//   dS/dv_f(record) += 2 (fv - gv) len,   dS/dv_g(record) -= 2 (fv - gv) len
//   dK/dv_f(record) += gv len,            dK/dv_g(record) += fv len
// and at each crossing of a breakpoint of f (the start time of its next
// record, with the partner value c):
//   dS/dt = (v_old - c)^2 - (v_new - c)^2,  dK/dt = (v_old - v_new) c
// f advances first on ties, so f's breakpoint sees g's left-limit value and g's
// breakpoint sees f's right-limit value; for a self pair (kernel diagonal) the
// two halves add up to the exact derivative v_old^2 - v_new^2.
//
// Value gradients of the current record accumulate in registers and are
// flushed when the record is left; flushes and time gradients go to FP64 LDS
// slots beside the staged records, flushed to global memory once per staged
// record and phase. A record staged in an earlier phase is flushed straight
// to global memory.
#include "walk_common.hpp"

namespace sbamd
{
  __device__ __forceinline__ void gadd(void* base, int64_t idx, double x, int f64)
  {
    if (f64)
      atomicAdd(static_cast<double*>(base) + idx, x);
    else
      atomicAdd(static_cast<float*>(base) + idx, static_cast<float>(x));
  }

  template <typename TV, int OP>
  __global__ __launch_bounds__(256) void bwd_tile_kernel(BwdParams B)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using TT = double;
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    const FwdParams& P = B.f;

    // 16 bytes of gradient slots (value, time) per staged record
    const LdsLayout lay = lds_layout(sizeof(TT), P.R, P.C, P.cap, 16);
    const TileLds<TT, TV> S(smem, lay);
    static_assert(sizeof(RecT) == 16, "gradient slots mirror 16-byte records");
    constexpr uint32_t SF = sizeof(RecT); // record stride of every slot
    constexpr uint32_t SG = SF;
    const uint32_t GOFF = static_cast<uint32_t>(lay.extra_off);
    const int64_t* __restrict__ src = B.r_src;

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int NS = P.R + P.C;

    for (int64_t ti = blockIdx.x; ti < P.ntiles; ti += gridDim.x)
    {
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];

      tile_begin(S, P, tl, pr, tid);
      __syncthreads();

      const int lr = tid / P.C;
      const int lc = tid - lr * P.C;
      bool valid = tid < P.R * P.C && lr < tl.nr && lc < tl.nc;
      if (valid && (tl.flags & TILE_DIAG))
        valid = P.include_diag ? (lr >= lc) : (lr > lc);

      // Pair weight.
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
      const double scale = OP == OPK_DOT ? w : 2.0 * w;

      TT t = TT(0);
      TT tf = L::inf(), tg = L::inf();
      TV fv = TV(0), gv = TV(0);
      int64_t nf = 1, ng = 1;
      double accf = 0.0, accg = 0.0; // unscaled value-gradient sums of the current records

      for (int phase = 0;; ++phase)
      {
        phase_stage(S, P, phase, tid, nthreads);
        for (int k = tid; k < lay.rec_end / 16; k += nthreads)
          *reinterpret_cast<double2*>(smem + GOFF + 16 * k) = make_double2(0.0, 0.0);
        __syncthreads();
        const TT Tb = L::from_bits(S.tb[phase & 1]);
        phase_advance_cursor(S, P, Tb, tid);

        if (valid)
        {
          const int64_t cf = S.cur[slot_f];
          const int64_t cg = S.cur[slot_g];
          const uint32_t bf = S.lbase[slot_f];
          const uint32_t bg = S.lbase[slot_g];
          const int64_t basef = S.base[slot_f];
          const int64_t baseg = S.base[slot_g];
          if (phase == 0)
          {
            const RecT rf = S.rec(bf);
            const RecT rg = S.rec(bg);
            tf = rf.t;
            fv = rf.v;
            tg = rg.t;
            gv = rg.v;
          }
          uint32_t af_addr = bf + static_cast<uint32_t>(nf - cf) * SF;
          uint32_t ag_addr = bg + static_cast<uint32_t>(ng - cg) * SG;
          for (;;)
          {
            const bool af = tf <= tg;
            const TT tn = af ? tf : tg;
            if (!(tn < Tb))
              break;
            const double len = tn - t;
            if constexpr (OP == OPK_DOT)
            {
              accf += static_cast<double>(gv) * len;
              accg += static_cast<double>(fv) * len;
            }
            else
            {
              const double d = static_cast<double>(fv) - static_cast<double>(gv);
              accf += d * len;
              accg -= d * len;
            }
            t = tn;
            const uint32_t a = af ? af_addr : ag_addr;
            const uint32_t st = af ? SF : SG;
            const uint32_t b0 = af ? bf : bg;
            const RecT r = S.rec(a);
            // flush the value gradient of the record being left
            const double gflush = (af ? accf : accg) * scale;
            if (a - b0 >= st)
              atomicAdd(reinterpret_cast<double*>(smem + GOFF + a - st), gflush);
            else
              gadd(B.grad_v, src[(af ? basef + cf : baseg + cg) - 1], gflush, B.grad_f64);
            // time gradient of the breakpoint just crossed (start of record r)
            const double vold = af ? static_cast<double>(fv) : static_cast<double>(gv);
            const double c = af ? static_cast<double>(gv) : static_cast<double>(fv);
            const double vn = static_cast<double>(r.v);
            double dt_grad;
            if constexpr (OP == OPK_DOT)
              dt_grad = (vold - vn) * c;
            else
              dt_grad = (vold - c) * (vold - c) - (vn - c) * (vn - c);
            atomicAdd(reinterpret_cast<double*>(smem + GOFF + a + 8), dt_grad * w);
            af_addr = af ? af_addr + SF : af_addr;
            ag_addr = af ? ag_addr : ag_addr + SG;
            tf = af ? r.t : tf;
            fv = af ? r.v : fv;
            tg = af ? tg : r.t;
            gv = af ? gv : r.v;
            accf = af ? 0.0 : accf;
            accg = af ? accg : 0.0;
          }
          nf = cf + (af_addr - bf) / SF;
          ng = cg + (ag_addr - bg) / SG;
        }
        __syncthreads();

        // flush this phase's record slots
        for (int idx = tid; idx < NS * P.cap; idx += nthreads)
        {
          const int s = idx / P.cap;
          const int j = idx - s * P.cap;
          const int64_t k = S.cur[s] + j;
          if (k < S.next[s])
          {
            const uint32_t a = S.lbase[s] + static_cast<uint32_t>(j) * S.lstride[s];
            const double2 g2 = *reinterpret_cast<const double2*>(smem + GOFF + a);
            const int64_t pt = src[S.base[s] + k];
            if (g2.x != 0.0)
              gadd(B.grad_v, pt, g2.x, B.grad_f64);
            if (g2.y != 0.0 && k > 0)
              gadd(B.grad_t, pt, g2.y, B.grad_f64);
          }
        }
        __syncthreads();
        if (!(Tb < L::inf()))
          break;
      }

      if (valid)
      {
        // final rectangle [t, MAX) and flush of the last records
        const double len = L::max() - t;
        if constexpr (OP == OPK_DOT)
        {
          if (gv != TV(0))
            accf += static_cast<double>(gv) * len;
          if (fv != TV(0))
            accg += static_cast<double>(fv) * len;
        }
        else
        {
          const double d = static_cast<double>(fv) - static_cast<double>(gv);
          if (d != 0.0)
          {
            accf += d * len;
            accg -= d * len;
          }
        }
        gadd(B.grad_v, src[S.base[slot_f] + nf - 1], accf * scale, B.grad_f64);
        gadd(B.grad_v, src[S.base[slot_g] + ng - 1], accg * scale, B.grad_f64);
      }
      __syncthreads();
    }
  }

  size_t bwd_lds_bytes(int tv_f64, int R, int C, int cap)
  {
    (void)tv_f64;
    return static_cast<size_t>(lds_layout(8, R, C, cap, 16).total);
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
}
