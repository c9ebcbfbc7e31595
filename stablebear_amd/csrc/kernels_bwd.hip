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
// slots of the staged records, which are flushed to global memory (one atomic
// per staged record and phase). Records staged in an earlier phase are
// flushed straight to global memory.
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

  template <typename TT, typename TV, int OP>
  __global__ __launch_bounds__(256) void bwd_tile_kernel(BwdParams B)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    using L = Lim<TT>;
    using Bits = typename L::Bits;
    const FwdParams& P = B.f;

    const int NS = P.R + P.C;
    const int stride = P.stride;
    const int cap = P.cap;
    const LdsLayout lay = lds_layout(sizeof(TT), sizeof(TV), NS, stride, cap, 16);
    TT* s_tend = reinterpret_cast<TT*>(smem + lay.tend_off);
    TV* s_val = reinterpret_cast<TV*>(smem + lay.val_off);
    int64_t* s_base = reinterpret_cast<int64_t*>(smem + lay.base_off);
    int64_t* s_m = reinterpret_cast<int64_t*>(smem + lay.m_off);
    int64_t* s_cur = reinterpret_cast<int64_t*>(smem + lay.cur_off);
    int64_t* s_next = reinterpret_cast<int64_t*>(smem + lay.next_off);
    TT* s_tstart = reinterpret_cast<TT*>(smem + lay.tstart_off);
    Bits* s_tb = reinterpret_cast<Bits*>(smem + lay.tb_off);
    double* s_gv = reinterpret_cast<double*>(smem + lay.extra_off);
    double* s_gt = s_gv + cap * stride;

    const TT* __restrict__ g_t = static_cast<const TT*>(P.r_tend);
    const TV* __restrict__ g_v = static_cast<const TV*>(P.r_val);
    const int64_t* __restrict__ src = B.r_src;

    const int tid = threadIdx.x;

    for (int64_t ti = blockIdx.x; ti < P.ntiles; ti += gridDim.x)
    {
      const Tile tl = P.tiles[ti];
      const Problem pr = P.probs[tl.prob];

      if (tid < NS)
      {
        const bool isrow = tid < P.R;
        const int k = isrow ? tid : tid - P.R;
        const int lim = isrow ? tl.nr : tl.nc;
        if (k < lim)
        {
          const int64_t g = isrow ? pr.row_base + tl.r0 + k : pr.col_base + tl.c0 + k;
          const int64_t b = P.r_off[g];
          s_base[tid] = b;
          s_m[tid] = P.r_off[g + 1] - b;
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

      const int lr = tid / P.C;
      const int lc = tid - lr * P.C;
      bool valid = tid < P.R * P.C && lr < tl.nr && lc < tl.nc;
      if (valid && (tl.flags & TILE_DIAG))
        valid = P.include_diag ? (lr >= lc) : (lr > lc);
      const int slot_f = lr;
      const int slot_g = P.R + lc;

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

      TT t = TT(0);
      TT tf = L::inf(), tg = L::inf();
      TV fv = TV(0), gv = TV(0);
      int64_t nf = 1, ng = 1;
      double accf = 0.0, accg = 0.0; // unscaled value-gradient sums of the current records

      for (int phase = 0;; ++phase)
      {
        if (tid < NS)
        {
          const int64_t cur = s_next[tid];
          s_cur[tid] = cur;
          const int64_t rem = s_m[tid] - cur;
          TT tbs = L::inf();
          if (rem > cap)
            tbs = g_t[s_base[tid] + cur + cap - 1];
          atomicMin(&s_tb[phase & 1], L::bits(tbs));
        }
        for (int idx = tid; idx < NS * cap; idx += blockDim.x)
        {
          const int s = idx / cap;
          const int j = idx - s * cap;
          const int64_t k = s_next[s] + j;
          if (k < s_m[s])
          {
            const int64_t gk = s_base[s] + k;
            s_tend[j * stride + s] = g_t[gk];
            s_val[j * stride + s] = g_v[gk];
          }
          s_gv[j * stride + s] = 0.0;
          s_gt[j * stride + s] = 0.0;
        }
        if (tid == 0)
          s_tb[(phase + 1) & 1] = L::bits(L::inf());
        __syncthreads();

        const TT Tb = L::from_bits(s_tb[phase & 1]);

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
          const double scale = OP == OPK_DOT ? w : 2.0 * w;
          for (;;)
          {
            const bool af = tf <= tg;
            const TT tn = af ? tf : tg;
            if (!(tn < Tb))
              break;
            const double len = static_cast<double>(tn - t);
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
            const int s = af ? slot_f : slot_g;
            const int lj = af ? lf : lg;
            // flush the value gradient of the record being left (local lj - 1)
            const double gflush = (af ? accf : accg) * scale;
            if (lj - 1 >= 0)
              atomicAdd(&s_gv[(lj - 1) * stride + s], gflush);
            else
              gadd(B.grad_v, src[s_base[s] + (af ? cf : cg) + lj - 1], gflush, B.grad_f64);
            const TT tnew = s_tend[lj * stride + s];
            const TV vnew = s_val[lj * stride + s];
            const double vold = af ? static_cast<double>(fv) : static_cast<double>(gv);
            const double c = af ? static_cast<double>(gv) : static_cast<double>(fv);
            const double vn = static_cast<double>(vnew);
            double dt_grad;
            if constexpr (OP == OPK_DOT)
              dt_grad = (vold - vn) * c;
            else
              dt_grad = (vold - c) * (vold - c) - (vn - c) * (vn - c);
            atomicAdd(&s_gt[lj * stride + s], dt_grad * w);
            if (af)
            {
              tf = tnew;
              fv = vnew;
              ++lf;
              accf = 0.0;
            }
            else
            {
              tg = tnew;
              gv = vnew;
              ++lg;
              accg = 0.0;
            }
          }
          nf = cf + lf;
          ng = cg + lg;
        }
        __syncthreads();

        // flush this phase's record slots
        for (int idx = tid; idx < NS * cap; idx += blockDim.x)
        {
          const int s = idx / cap;
          const int j = idx - s * cap;
          const int64_t k = s_cur[s] + j;
          if (k < s_next[s])
          {
            const int64_t pt = src[s_base[s] + k];
            const double a = s_gv[j * stride + s];
            const double b2 = s_gt[j * stride + s];
            if (a != 0.0)
              gadd(B.grad_v, pt, a, B.grad_f64);
            if (b2 != 0.0 && k > 0)
              gadd(B.grad_t, pt, b2, B.grad_f64);
          }
        }
        __syncthreads();
        if (!(Tb < L::inf()))
          break;
      }

      if (valid)
      {
        // final rectangle [t, MAX) and flush of the last records
        const double len = static_cast<double>(L::max() - t);
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
        const double scale = OP == OPK_DOT ? w : 2.0 * w;
        gadd(B.grad_v, src[s_base[slot_f] + nf - 1], accf * scale, B.grad_f64);
        gadd(B.grad_v, src[s_base[slot_g] + ng - 1], accg * scale, B.grad_f64);
      }
      __syncthreads();
    }
  }

  size_t bwd_lds_bytes(int tv_f64, int R, int C, int cap)
  {
    const int NS = R + C;
    return static_cast<size_t>(lds_layout(8, tv_f64 ? 8 : 4, NS, NS, cap, 16).total);
  }

  void launch_backward(const BwdParams& prm, int op, hipStream_t stream)
  {
    const FwdParams& P = prm.f;
    const int NS = P.R + P.C;
    const unsigned grid = static_cast<unsigned>(P.ntiles < 0x7fffffff ? P.ntiles : 0x7fffffff);
    if (grid == 0)
      return;
    const int tt = 8;
    if (prm.val_f64)
    {
      const size_t lds = static_cast<size_t>(lds_layout(tt, 8, NS, P.stride, P.cap, 16).total);
      if (op == OPK_DOT)
        hipLaunchKernelGGL((bwd_tile_kernel<double, double, OPK_DOT>), dim3(grid), dim3(tile_block_threads(P.R, P.C)), lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_tile_kernel<double, double, OPK_L2>), dim3(grid), dim3(tile_block_threads(P.R, P.C)), lds, stream, prm);
    }
    else
    {
      const size_t lds = static_cast<size_t>(lds_layout(tt, 4, NS, P.stride, P.cap, 16).total);
      if (op == OPK_DOT)
        hipLaunchKernelGGL((bwd_tile_kernel<double, float, OPK_DOT>), dim3(grid), dim3(tile_block_threads(P.R, P.C)), lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_tile_kernel<double, float, OPK_L2>), dim3(grid), dim3(tile_block_threads(P.R, P.C)), lds, stream, prm);
    }
    SBAMD_HIP(hipGetLastError());
  }
}
