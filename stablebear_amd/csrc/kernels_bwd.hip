// Backward of L = sum_e G_e * X_e over the pair outputs X of a pdist (p = 2)
// or l2_kernel plan. No reference counterpart (the reference has no autograd).
//
// For a PCF f = f_i with records r (value v_r on [t_start(r), t_end(r))) and
// its partners g_j with pair weights w_ij (G / (2 D) for the L2 distance
// D = sqrt(int (f - g)^2), G for the inner product K = int f g, 2 G for the
// kernel's diagonal entry int f^2):
//   dL/dv_r  (L2)  = 2 (v_r len_r W - Q_r)        (dot) = Q_r
//   dL/dt_r  (L2)  = (v_{r-1} - v_r) ((v_{r-1} + v_r) W - 2 A_{r-1})
//            (dot) = (v_{r-1} - v_r) A_{r-1}
// with W = sum_j w_ij, Q_r = integral of H_i over record r, A_r = H_i at
// t_end(r), and H_i = sum_j w_ij g_j.
//
// With F_j(t) = integral of g_j over [0, t) and S_i(t) = sum_j w_ij F_j(t),
// the integral of H_i from 0, Q_r = S_i(t_end(r)) - S_i(t_start(r)). So the
// backward needs S_i and H_i at every breakpoint of every PCF, each a sum
// over the partners. On a record of g_j, F_j(t) = P_j + f_j t (f_j its
// value, P_j = F_j(t_start) - f_j t_start); a partner breakpoint at time t_e
// that moves f_j by d moves P_j by -d t_e.
//
// PCFs are taken in groups of 64 ("bands") and of CH partners ("chunks": 64
// with FP32 pair weights, 32 with FP64). A workgroup handles one band against
// a group of two chunks. Its wavefronts split the band's breakpoints (one
// time-ordered list) into slices and take them 64 at a time, one breakpoint
// t_b (of row i_b) per lane. For a block whose first breakpoint is at T0, a
// lane starts from its row's sums over the group's partners at T0,
//   X = sum_j w_ij f_j(T0),  Y = -sum_j w_ij P_j(T0),
// and walks the partner breakpoints from T0 to the block's last breakpoint:
// each one with t_e < t_b adds w d to X and w d t_e to Y, so that at the end
// X = H_i(t_b^-) and S_i(t_b) = t_b X - Y. The walk reads the events through
// the scalar cache (they are the same for every lane), four at a time with
// the next four in flight, and reads the lane's w_ij for the four events'
// partners from the workgroup's weights in LDS. A chunk's partners come as
// two sublists (two event lists) of 32 (FP64 weights: 16). The row sums at T0
// live in the lanes, row r in lane r: formed once per slice from the
// partners' state at its start (f_j and P_j, in LDS), then moved along by
// every event of every window (unmasked, with the lane's own row's weight),
// so that at a block's end they hold the next block's T0 values; a block's
// lanes read their rows' sums from those lanes.
//
// Each breakpoint adds S_i(t_b) (summed over the group's partners) to q at
// the record it starts and H_i(t_b) to a at the record it ends; the finalize
// kernel takes Q_r = q[r + 1] - q[r], with S = 0 at the start of a first
// record and, for a last record, S at the end of the domain (sum_j w_ij
// F_j(DBL_MAX), kept in a[] at that record, where no breakpoint writes). The
// chunk groups of a band add their parts with global atomics.
//
// All arrays here are in the original PCF order ("position" p of record k of
// PCF i is ooff[i] + k); a layout kernel copies them out of the sorted
// device set, and the group event lists come from a merge sort by time.
//
// Ties: where f and a partner have a breakpoint at the same time, A takes the
// mean of H_i's left and right limits there (the mean of the partner's values
// on both sides). For a self pair (kernel diagonal) that is the exact
// derivative; for two PCFs jumping together the two one-sided results add up
// to the derivative of moving the shared time. S is continuous there. A
// plan's first backward also sorts the whole set by time and flags a time
// that two PCFs share; without one, the walk skips its tie test and each
// lane adds its own PCF's term (the only tie left) directly.
#include "walk_common.hpp"

namespace sbamd
{
  namespace
  {
    __device__ __forceinline__ double load_out(const void* p, int64_t idx, int f64)
    {
      return f64 ? static_cast<const double*>(p)[idx]
                 : static_cast<double>(static_cast<const float*>(p)[idx]);
    }

    // w_ij of the loss for rows i and partner j (original indices).
    template <int OP>
    __device__ __forceinline__ double pair_weight(const BwdParams& B, int64_t i, int64_t j)
    {
      if (OP == OPK_DOT)
      {
        const double G = load_out(B.grad_out, symmat_index(i, j), B.in_f64);
        return i == j ? 2.0 * G : G;
      }
      if (i == j)
        return 0.0;
      const int64_t idx = distmat_index(i, j);
      const double D = load_out(B.fwd_out, idx, B.in_f64);
      return D > 0.0 ? load_out(B.grad_out, idx, B.in_f64) / (2.0 * D) : 0.0;
    }

    // Number of keys in [lo, hi) below `key` (strict) or not above it.
    __device__ __forceinline__ int64_t count_before(const double* keys, int64_t lo, int64_t hi,
                                                    double key, bool inclusive)
    {
      while (lo < hi)
      {
        const int64_t mid = lo + ((hi - lo) >> 1);
        const double k = keys[mid];
        if (inclusive ? k <= key : k < key)
          lo = mid + 1;
        else
          hi = mid;
      }
      return lo;
    }

    // Wave-uniform loads through the scalar cache: the pointer is cast to the
    // constant address space, which the compiler reads with s_load.
    template <typename T>
    __device__ __forceinline__ T uniform_load(const T* p, int64_t i)
    {
      return ((const __attribute__((address_space(4))) T*)(p))[i];
    }

    // Records at positions [a, b) of one PCF that start before t. Eight
    // independent probes per round (the keys are sorted), so a search takes
    // about log8 of the length in dependent loads.
    __device__ __forceinline__ int64_t starts_below(const double* key0, int64_t a, int64_t b, double t)
    {
      int64_t lo = a, hi = b; // the count (relative to a at the end) is in [lo, hi]
      while (hi - lo > 8)
      {
        const int64_t step = (hi - lo + 7) / 8;
        int below = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
          const int64_t idx = lo + k * step;
          below += idx < hi && key0[idx] < t;
        }
        // the probes are in key order: the first `below` of them lie below t
        const int64_t nlo = below > 0 ? lo + (below - 1) * step + 1 : lo;
        const int64_t nhi = below > 0 ? min(hi, lo + below * step) : lo;
        lo = nlo;
        hi = nhi;
      }
      int below = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        below += lo + k < hi && key0[lo + k] < t;
      return lo + below - a;
    }

    // First index of a time-ordered event list of n events whose time is
    // >= t, found by the whole wavefront (every lane active): 64 probes per
    // round, so two or three dependent loads for a list of thousands.
    __device__ __forceinline__ int32_t lower_bound_time_wave(const BwdEvent* p, int32_t n, double t)
    {
      const int lane = __lane_id();
      int32_t lo = 0, hi = n; // the answer is in [lo, hi]
      while (hi - lo > 64)
      {
        const int32_t step = (hi - lo + 63) / 64;
        const int32_t idx = lo + lane * step;
        const bool below = idx < hi && p[idx].t < t;
        const int32_t cnt = static_cast<int32_t>(__popcll(__ballot(below)));
        const int32_t nlo = cnt > 0 ? lo + (cnt - 1) * step + 1 : lo;
        const int32_t nhi = cnt > 0 ? min(hi, lo + cnt * step) : lo;
        lo = nlo;
        hi = nhi;
      }
      const bool below = lo + lane < hi && p[lo + lane].t < t;
      return __builtin_amdgcn_readfirstlane(lo + static_cast<int32_t>(__popcll(__ballot(below))));
    }

    // 1 if a < b, for times held in scalar registers as their bit patterns
    // (non-negative doubles and +inf order like unsigned integers): the
    // borrow of a 64-bit subtract, on the scalar unit.
    __device__ __forceinline__ int32_t bits_before(uint64_t a, uint64_t b)
    {
      uint32_t tmp, r;
      asm volatile("s_sub_u32 %0, %2, %4\n\t"
                   "s_subb_u32 %0, %3, %5\n\t"
                   "s_cselect_b32 %1, 1, 0"
                   : "=&s"(tmp), "=s"(r)
                   : "s"(static_cast<uint32_t>(a)), "s"(static_cast<uint32_t>(a >> 32)),
                     "s"(static_cast<uint32_t>(b)), "s"(static_cast<uint32_t>(b >> 32))
                   : "scc");
      return static_cast<int32_t>(r);
    }

    // Lane u's double, as a wavefront-uniform value.
    __device__ __forceinline__ double readlane_f64(double x, int u)
    {
      const uint64_t b = __builtin_bit_cast(uint64_t, x);
      const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), u));
      const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), u));
      return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
    }

    // A partner event as the walk reads it: one 32-byte scalar load.
    struct WalkEv
    {
      double t, d, dt;
      int32_t col, pos;
    };
    static_assert(sizeof(WalkEv) == sizeof(BwdEvent), "the walk reads whole events");

    __device__ __forceinline__ WalkEv walk_load(const BwdEvent* cl, int32_t i)
    {
      // one 32-byte scalar load per event
      typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
      const __attribute__((address_space(4))) u32x8* c = (const __attribute__((address_space(4))) u32x8*)(cl);
      const u32x8 w = c[i];
      const uint32_t w0 = w.s0, w1 = w.s1, w2 = w.s2, w3 = w.s3, w4 = w.s4, w5 = w.s5, w6 = w.s6, w7 = w.s7;
      WalkEv e;
      e.t = __builtin_bit_cast(double, (static_cast<uint64_t>(w1) << 32) | w0);
      e.d = __builtin_bit_cast(double, (static_cast<uint64_t>(w3) << 32) | w2);
      e.dt = __builtin_bit_cast(double, (static_cast<uint64_t>(w5) << 32) | w4);
      e.col = static_cast<int32_t>(w6);
      e.pos = static_cast<int32_t>(w7);
      return e;
    }

    // One partner event (time t, jump d, dt = d t, in scalar registers) for
    // the lane's breakpoint tb and pair weight w: a lane with t < tb adds
    // w d to X and w dt to Y; lanes with t == tb are set in `tie`. Written
    // out because the compiler turns the masked update into selects of both
    // FP64 results (four instructions more per event).
    __device__ __forceinline__ void walk_step(double& X, double& Y, uint64_t& tie, double tb, float w,
                                              double t, double d, double dt)
    {
      float wm;
      double wd;
      asm volatile("v_cmp_lt_f64 vcc, %[t], %[tb]\n\t"
                   "s_nop 1\n\t"
                   "v_cndmask_b32 %[wm], 0, %[w], vcc\n\t"
                   "v_cvt_f64_f32 %[wd], %[wm]\n\t"
                   "v_cmp_eq_f64 vcc, %[t], %[tb]\n\t"
                   "v_fmac_f64 %[X], %[d], %[wd]\n\t"
                   "v_fmac_f64 %[Y], %[dt], %[wd]\n\t"
                   "s_or_b64 %[tie], %[tie], vcc"
                   : [X] "+v"(X), [Y] "+v"(Y), [tie] "+s"(tie), [wm] "=&v"(wm), [wd] "=&v"(wd)
                   : [t] "s"(t), [tb] "v"(tb), [w] "v"(w), [d] "s"(d), [dt] "s"(dt)
                   : "vcc", "scc");
    }

    __device__ __forceinline__ void walk_step(double& X, double& Y, uint64_t& tie, double tb, double w,
                                              double t, double d, double dt)
    {
      const bool lt = t < tb;
      const double wm = lt ? w : 0.0;
      X = fma(wm, d, X);
      Y = fma(wm, dt, Y);
      tie |= __ballot(t == tb);
    }

    // The same without the tie flags, for lists in which no two PCFs share a
    // breakpoint time (the only ties left are a PCF's with itself, which the
    // caller adds on its own).
    __device__ __forceinline__ void walk_step_untied(double& X, double& Y, double tb, float w, double t, double d,
                                                     double dt)
    {
      float wm;
      double wd;
      asm volatile("v_cmp_lt_f64 vcc, %[t], %[tb]\n\t"
                   "s_nop 1\n\t"
                   "v_cndmask_b32 %[wm], 0, %[w], vcc\n\t"
                   "v_cvt_f64_f32 %[wd], %[wm]\n\t"
                   "v_fmac_f64 %[X], %[d], %[wd]\n\t"
                   "v_fmac_f64 %[Y], %[dt], %[wd]"
                   : [X] "+v"(X), [Y] "+v"(Y), [wm] "=&v"(wm), [wd] "=&v"(wd)
                   : [t] "s"(t), [tb] "v"(tb), [w] "v"(w), [d] "s"(d), [dt] "s"(dt)
                   : "vcc");
    }

    __device__ __forceinline__ void walk_step_untied(double& X, double& Y, double tb, double w, double t, double d,
                                                     double dt)
    {
      const double wm = t < tb ? w : 0.0;
      X = fma(wm, d, X);
      Y = fma(wm, dt, Y);
    }

    template <typename WT>
    struct WalkVec;
    template <>
    struct WalkVec<float>
    {
      static constexpr int VW = 32;
    };
    template <>
    struct WalkVec<double>
    {
      static constexpr int VW = 16;
    };

    // Walks the partner events of one sublist from index e while their time
    // (bits) is below `lim` (bits), for the lane's breakpoint tb; returns the
    // first index not taken. Events come four at a time through the scalar
    // cache from one base address (the event arrays have four events of
    // padding, so a group may run past the list); the next four are
    // requested after the current ones have been waited for, so that they
    // arrive while the current four are applied.
    template <typename WT, bool TIES>
    __device__ __forceinline__ int32_t walk_window(const BwdEvent* cl, int32_t n, int32_t e, uint64_t lim,
                                                   const WT* w_row, const WT* w_own, double tb, double& X,
                                                   double& Y, double& XA, double& YA, uint64_t& tie)
    {
      constexpr int VW = WalkVec<WT>::VW;
      if (e >= n)
        return e;
      WalkEv A[4], Bq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        A[u] = walk_load(cl + e, u);
      auto half = [&](WalkEv(&cur)[4], WalkEv(&nxt)[4]) -> bool {
        // the lane's weights for the four events' partners, requested
        // before they are needed (an event past the list reads a weight of
        // the sublist too, and is not applied)
        WT w[4], wo[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
          w[u] = w_row[(cur[u].col & (VW - 1)) * 64];
          wo[u] = w_own[(cur[u].col & (VW - 1)) * 64];
        }
        // events are in time order: all four are in the window when the
        // fourth is (and the list has four more), else count them
        int32_t k;
        if (bits_before(static_cast<uint64_t>(__double_as_longlong(cur[3].t)), lim) && n - e >= 4)
          k = 4;
        else
        {
          // only events of this list count (those past its end belong to
          // another list or are padding)
          const int32_t nv = n - e;
          k = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u)
            k += u < nv ? bits_before(static_cast<uint64_t>(__double_as_longlong(cur[u].t)), lim) : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          nxt[u] = walk_load(cl + e + 4, u);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (u < k)
          {
            if constexpr (TIES)
              walk_step(X, Y, tie, tb, w[u], cur[u].t, cur[u].d, cur[u].dt);
            else
              walk_step_untied(X, Y, tb, w[u], cur[u].t, cur[u].d, cur[u].dt);
            // the lane's own row (row = lane) takes every event of the window
            const double wd = static_cast<double>(wo[u]);
            XA = fma(wd, cur[u].d, XA);
            YA = fma(wd, cur[u].dt, YA);
          }
        e += k;
        return k == 4;
      };
      while (half(A, Bq) && half(Bq, A))
      {
      }
      return e;
    }
  }

  // Wavefronts per backward workgroup (slices of one task's breakpoints):
  // with 59 vector registers a lane and 34.5 KB of LDS, four 8-wave
  // workgroups share a CU, so one's set-up (weights, searches) overlaps the
  // others' walks.
  // (With the weights in vector registers instead of LDS, 91 registers and
  // two workgroups per CU: 1.39-1.42 ms against 1.32-1.36 on 4096 PCFs.)
  constexpr int BWD_WAVES = 8;

  // Chunks of partners per workgroup: a breakpoint's two atomics are paid
  // once per group of chunks (measured on 4096 PCFs: 1, 2 and 4 chunks take
  // 1.83, 1.61 and 1.73 ms; four leave too few workgroups).
  constexpr int BWD_CHUNKS = 2;

  // LDS of the gather kernel, per chunk of the group: pair weights (ch
  // partners x 64 rows), the partners' record end positions (ch ints), and
  // the partners' state (f, P) at the end of the domain (ch x 16 bytes).
  __host__ __device__ constexpr size_t bwd_lds_weights(size_t wbytes, int ch)
  {
    return static_cast<size_t>(ch) * 64 * wbytes;
  }
  __host__ __device__ constexpr size_t bwd_lds_bytes(size_t wbytes, int ch)
  {
    return BWD_CHUNKS * (bwd_lds_weights(wbytes, ch) + static_cast<size_t>(ch) * 4 +
                         static_cast<size_t>(ch) * 16);
  }

  // Thread per sorted record: copies start time, value, owner and source
  // point into the original-order position layout and seeds the sort keys.
  __global__ void bwd_layout_kernel(BwdParams B)
  {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= B.records)
      return;
    int64_t lo = 0, hi = B.n; // sorted PCF s with roff[s] <= r < roff[s + 1]
    while (hi - lo > 1)
    {
      const int64_t mid = (lo + hi) >> 1;
      if (B.roff[mid] <= r)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t k = r - B.roff[lo];
    const int64_t i = B.perm[lo];
    const int64_t p = B.ooff[i] + k;
    const unsigned char* rec = static_cast<const unsigned char*>(B.rec);
    const double start = k ? *reinterpret_cast<const double*>(rec + 16 * (r - 1)) : 0.0;
    const double v = B.val_f64 ? *reinterpret_cast<const double*>(rec + 16 * r + 8)
                               : static_cast<double>(*reinterpret_cast<const float*>(rec + 16 * r + 8));
    B.key0[p] = start;
    B.val0[p] = v;
    B.col0[p] = static_cast<int32_t>(i);
    B.src0[p] = B.src[r];
    B.keys[0][p] = start;
    B.ids[0][p] = static_cast<int32_t>(p);
  }

  // One level of the per-group merge sort: runs of `width` consecutive PCFs
  // (each PCF's records are already in time order) merge pairwise; equal
  // times keep the left run first, so the order is (time, PCF, record).
  __global__ void bwd_merge_kernel(BwdParams B, int32_t width, int32_t from, int32_t group)
  {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= B.records)
      return;
    const double* kin = B.keys[from];
    const int32_t* iin = B.ids[from];
    const int64_t i = B.col0[e];
    const int64_t cb = (i / group) * group;
    const int64_t ce = min(cb + group, B.n);
    const int64_t run = (i - cb) / width;
    const int64_t sib = run ^ 1;
    const int64_t own0 = cb + run * width;
    const int64_t sib0 = cb + sib * width;
    const double key = kin[e];
    int64_t rank = 0;
    if (sib0 < ce)
    {
      const int64_t lo = B.ooff[sib0], hi = B.ooff[min(sib0 + width, ce)];
      rank = count_before(kin, lo, hi, key, (run & 1) != 0) - lo;
    }
    const int64_t pos = B.ooff[cb + (run & ~int64_t(1)) * width] + (e - B.ooff[own0]) + rank;
    B.keys[from ^ 1][pos] = key;
    B.ids[from ^ 1][pos] = iin[e];
  }

  // Thread per position of the sorted group lists (groups of `group` PCFs):
  // the event record (time, jump, jump x time, PCF within the group, record).
  __global__ void bwd_events_kernel(BwdParams B, int32_t from, int32_t group, BwdEvent* out)
  {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= B.records)
      return;
    const int64_t cb = (static_cast<int64_t>(B.col0[p]) / group) * group;
    const double t = B.keys[from][p];
    const int64_t id = B.ids[from][p];
    const int64_t c = B.col0[id];
    const bool first = id == B.ooff[c];
    BwdEvent ev;
    ev.t = t;
    ev.d = B.val0[id] - (first ? 0.0 : B.val0[id - 1]);
    ev.dt = ev.d * t;
    ev.col = static_cast<int32_t>(c - cb) | (first ? BWD_FIRST : 0);
    ev.pos = static_cast<int32_t>(id);
    out[p] = ev;
  }

  // Thread per event of both lists (the bands' lists, then the sublists):
  // the value jump at the event and its product with the time, from the
  // current values. The lists' order, times and records depend on the
  // breakpoint times only, so a backward over a set whose lists were already
  // sorted refreshes these two fields instead of sorting again; the results
  // are those of bwd_events_kernel bit for bit.
  __global__ void bwd_refresh_kernel(BwdParams B)
  {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= 2 * B.records)
      return;
    BwdEvent* const ev = e < B.records ? B.events + e : B.events2 + (e - B.records);
    const int32_t id = ev->pos;
    const bool first = (ev->col & BWD_FIRST) != 0;
    const double d = B.val0[id] - (first ? 0.0 : B.val0[id - 1]);
    ev->d = d;
    ev->dt = d * ev->t;
  }

  // Wavefront per PCF: pst0[p] = F(t_start(p)) - v_p t_start(p), with F the
  // PCF's integral from 0 (an inclusive scan of the record integrals over
  // the lanes, 64 records at a time).
  __global__ void bwd_prefix_kernel(BwdParams B)
  {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= B.n)
      return;
    const int lane = threadIdx.x & 63;
    const int64_t a = B.ooff[i], b = B.ooff[i + 1];
    double carry = 0.0;
    for (int64_t base = a; base < b; base += 64)
    {
      const int64_t p = base + lane;
      const bool in = p < b;
      const double k = in ? B.key0[p] : 0.0;
      const double v = in ? B.val0[p] : 0.0;
      const double kn = p + 1 < b ? B.key0[p + 1] : k;
      const double c = v * (kn - k);
      double incl = c;
      for (int o = 1; o < 64; o <<= 1)
      {
        const double u = __shfl_up(incl, o);
        if (lane >= o)
          incl += u;
      }
      const double excl = incl - c;
      if (in)
        B.pst0[p] = (carry + excl) - v * k;
      carry += __shfl(incl, 63);
    }
  }

  // One workgroup per (band of 64 rows, group of BWD_CHUNKS chunks of CH
  // partners); the design is in the file comment. A lane's breakpoint
  // collects X, Y and the tie jumps over all chunks of the group before its
  // two atomics, so a breakpoint costs two atomics per group rather than per
  // chunk. Workgroups are dealt to the 8 XCDs round-robin; all bands of a
  // chunk group run on one XCD so the group's lists stay in that XCD's L2,
  // and groups go to XCDs in snake order (0..7, 7..0, ...).
  template <typename WT, int OP>
  __global__ __launch_bounds__(64 * BWD_WAVES) void bwd_gather_kernel(BwdParams B)
  {
    constexpr int VW = WalkVec<WT>::VW;
    constexpr int CH = 2 * VW;
    constexpr int G = BWD_CHUNKS;
    extern __shared__ __align__(16) unsigned char smem[];
    // s_w[g][jj * 64 + ii] = w(row ii, partner jj of chunk g)
    WT* const s_w = reinterpret_cast<WT*>(smem);
    int32_t* const s_oend = reinterpret_cast<int32_t*>(smem + G * bwd_lds_weights(sizeof(WT), CH));
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // the partners' states (f, P) at the end of the domain, CH per chunk
    double2* const s_fp = reinterpret_cast<double2*>(smem + G * bwd_lds_weights(sizeof(WT), CH) + G * CH * 4);
    const int64_t nbands = (B.n + BWD_GROUP - 1) / BWD_GROUP;
    const int64_t nchunks = (B.n + CH - 1) / CH;
    const int64_t ngroups = (nchunks + G - 1) / G;
    const int64_t x = blockIdx.x;
    const int64_t slot = x >> 3;
    const int64_t gslot = slot / nbands;
    const int64_t cr = ((gslot & 1) ? 7 - (x & 7) : (x & 7)) + 8 * gslot;
    if (cr >= ngroups)
      return;
    const int64_t grp = ngroups - 1 - cr;
    const bool ties_possible = __builtin_amdgcn_readfirstlane(B.cross_ties ? *B.cross_ties : 1) != 0;
    const int64_t band = nbands - 1 - slot % nbands;
    const int nrow = static_cast<int>(min(static_cast<int64_t>(BWD_GROUP), B.n - band * BWD_GROUP));

    // pair weights, partner-major, and the partners' record ends
    for (int k = tid; k < G * CH * 64; k += 64 * BWD_WAVES)
    {
      const int g = k / (CH * 64);
      const int jj = (k >> 6) % CH, ii = k & 63;
      const int64_t jg = (grp * G + g) * CH + jj;
      const double w = jg < B.n && ii < nrow ? pair_weight<OP>(B, band * BWD_GROUP + ii, jg) : 0.0;
      s_w[k] = static_cast<WT>(w);
    }
    if (tid < G * CH)
    {
      const int64_t jg = grp * G * CH + tid;
      s_oend[tid] = jg < B.n ? static_cast<int32_t>(B.ooff[jg + 1]) : 0;
    }
    __syncthreads();
    if (wv == 0 && lane < nrow)
    {
      double ws = 0.0;
      for (int k = 0; k < G * CH; ++k)
        ws += static_cast<double>(s_w[k * 64 + lane]);
      if (ws != 0.0)
        atomicAdd(&B.wsum[band * BWD_GROUP + lane], ws);
    }
    if (wv == 1)
    {
      // each row's S at the end of the domain: sum_j w_ij (P_j + f_j DBL_MAX)
      // over the partners' last records, kept at the row's last record
      for (int k = lane; k < G * CH; k += 64)
      {
        double f = 0.0, P = 0.0;
        if (grp * G * CH + k < B.n)
        {
          const int64_t lp = s_oend[k] - 1;
          f = B.val0[lp];
          P = B.pst0[lp];
        }
        s_fp[k] = make_double2(f, P);
      }
      __builtin_amdgcn_wave_barrier();
      if (lane < nrow)
      {
        double A = 0.0, Hend = 0.0;
        for (int k = 0; k < G * CH; ++k)
        {
          const double w = static_cast<double>(s_w[k * 64 + lane]);
          const double2 fp = s_fp[k];
          A = fma(w, fp.y, A);
          Hend = fma(w, fp.x, Hend);
        }
        const double Send = Hend != 0.0 ? fma(Hend, DBL_MAX, A) : A;
        if (Send != 0.0)
          atomicAdd(&B.a[B.ooff[band * BWD_GROUP + lane + 1] - 1], Send);
      }
      __builtin_amdgcn_wave_barrier();
    }

    // the band's breakpoints; the records starting at 0 (one per row, first
    // in the list) are not breakpoints
    const int64_t r0 = uniform_load(B.ooff, band * BWD_GROUP);
    const int32_t nr = static_cast<int32_t>(uniform_load(B.ooff, band * BWD_GROUP + nrow) - r0);
    const BwdEvent* const rl = B.events + r0;
    int32_t skip;
    {
      const bool first = lane < nr && (rl[lane].col & BWD_FIRST) != 0;
      const uint64_t m = ~__ballot(first);
      skip = m ? static_cast<int32_t>(__builtin_ctzll(m)) : 64;
      skip = __builtin_amdgcn_readfirstlane(min(skip, nr));
    }
    const int32_t span = nr - skip;
    const int32_t b_lo = skip + static_cast<int32_t>((static_cast<int64_t>(span) * wv) / BWD_WAVES);
    const int32_t b_hi = skip + static_cast<int32_t>((static_cast<int64_t>(span) * (wv + 1)) / BWD_WAVES);
    if (b_lo >= b_hi)
      return;

    // chunk g's sublist s of partner events: cl(g, s)[0, cn(g, s))
    auto list_of = [&](int g, int s, const BwdEvent*& lst, int32_t& cnt) {
      const int64_t p0 = (grp * G + g) * CH + s * VW;
      if (p0 < B.n)
      {
        const int64_t base = uniform_load(B.ooff, p0);
        lst = B.events2 + base;
        cnt = static_cast<int32_t>(uniform_load(B.ooff, min(p0 + VW, B.n)) - base);
      }
      else
      {
        lst = B.events2;
        cnt = 0;
      }
    };

    const double T0 = uniform_load(&rl[b_lo].t, 0);
    // the walk's position in each list: lane 2 g + s of one vector register
    int32_t ecv = 0;
    for (int k = 0; k < 2 * G; ++k)
    {
      const BwdEvent* lst;
      int32_t cnt;
      list_of(k >> 1, k & 1, lst, cnt);
      const int32_t e = cnt > 0 ? lower_bound_time_wave(lst, cnt, T0) : 0;
      ecv = lane == k ? e : ecv;
    }
    __builtin_amdgcn_wave_barrier();
    // Row `lane`'s sums over the group's partners at the current block start,
    // X = sum_j w f_j and Y = -sum_j w P_j: formed once from the partners'
    // state just before the slice's first breakpoint (lane l finds partner
    // k0 + l's, every lane takes it from there), then moved along by every
    // event the walk reads; a block's lanes take their rows' sums from these
    // lanes.
    double XA = 0.0, YA = 0.0;
    for (int g = 0; g < G; ++g)
    {
      if ((grp * G + g) * CH >= B.n)
        break;
      const WT* const w_g = s_w + g * CH * 64;
      for (int k0 = 0; k0 < CH; k0 += 64)
      {
        double f = 0.0, P = 0.0;
        const int k = k0 + lane;
        const int64_t jg = (grp * G + g) * CH + k;
        if (k < CH && jg < B.n)
        {
          const int64_t a = B.ooff[jg];
          const int64_t c = starts_below(B.key0, a, s_oend[g * CH + k], T0);
          if (c > 0)
          {
            f = B.val0[a + c - 1];
            P = B.pst0[a + c - 1];
          }
        }
        const int nk = min(64, CH - k0);
        for (int u = 0; u < nk; ++u)
        {
          const double w = static_cast<double>(w_g[(k0 + u) * 64 + lane]);
          XA = fma(w, readlane_f64(f, u), XA);
          YA = fma(-w, readlane_f64(P, u), YA);
        }
      }
    }

    for (int32_t blk = b_lo; blk < b_hi; blk += 64)
    {
      const int32_t bb = blk + lane;
      const bool valid = bb < b_hi;
      const BwdEvent& E = rl[min(bb, b_hi - 1)];
      const double tb = valid ? E.t : -1.0; // an invalid lane takes no event
      const int32_t ecol = E.col;
      const int32_t pos = E.pos;
      const int ib = ecol & 63;
      const bool emit = valid && !(ecol & BWD_FIRST);
      const bool lastblk = blk + 64 >= b_hi;
      const double Tlast = uniform_load(&rl[min(blk + 63, b_hi - 1)].t, 0);
      const double Tend = lastblk ? Tlast : uniform_load(&rl[blk + 64].t, 0);
      // the walk takes events with t < Tend, or t <= Tlast in a slice's last block
      uint64_t lim = static_cast<uint64_t>(__double_as_longlong(Tend)) + (lastblk ? 1u : 0u);
      // (uniform by construction; stated for the compiler, whose analysis
      // loses it through the loop)
      lim = (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(lim >> 32))) << 32) |
            static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(lim)));
      const bool edge = !lastblk && __ballot(valid && tb == Tend) != 0;
      // without ties between PCFs (and none at this block's end) the walk
      // skips the tie flags; a block with an edge takes the checked walk
      const bool untied = !ties_possible && !edge;

      // the row's sums at the block start (over the group's partners)
      double X = __shfl(XA, ib);
      double Y = __shfl(YA, ib);
      double J = 0.0;
#pragma unroll 1
      for (int g = 0; g < G; ++g)
      {
        if ((grp * G + g) * CH >= B.n)
          break;
        const WT* const w_g = s_w + g * CH * 64;
        const BwdEvent *cl0, *cl1;
        int32_t cn0, cn1;
        list_of(g, 0, cl0, cn0);
        list_of(g, 1, cl1, cn1);
        // the lane's row of weights of each sublist (partner c at [64 c]),
        // and its own row's (row = lane), read from LDS event by event
        const WT* const w_row0 = w_g + ib;
        const WT* const w_row1 = w_g + VW * 64 + ib;
        const WT* const w_own0 = w_g + lane;
        const WT* const w_own1 = w_g + VW * 64 + lane;

        uint64_t tie = 0;
        const int32_t ws0 = __builtin_amdgcn_readlane(ecv, 2 * g);
        const int32_t ws1 = __builtin_amdgcn_readlane(ecv, 2 * g + 1);
        int32_t we0, we1;
        if (untied)
        {
          we0 = walk_window<WT, false>(cl0, cn0, ws0, lim, w_row0, w_own0, tb, X, Y, XA, YA, tie);
          we1 = walk_window<WT, false>(cl1, cn1, ws1, lim, w_row1, w_own1, tb, X, Y, XA, YA, tie);
          // the one tie left: the lane's own PCF among the partners, whose
          // event at tb is the lane's breakpoint itself
          const int64_t js = band * BWD_GROUP + ib - (grp * G + g) * CH;
          if (js >= 0 && js < CH)
            J = fma(static_cast<double>(w_g[js * 64 + ib]), E.d, J);
        }
        else
        {
          we0 = walk_window<WT, true>(cl0, cn0, ws0, lim, w_row0, w_own0, tb, X, Y, XA, YA, tie);
          we1 = walk_window<WT, true>(cl1, cn1, ws1, lim, w_row1, w_own1, tb, X, Y, XA, YA, tie);
        }
        ecv = lane == 2 * g ? we0 : (lane == 2 * g + 1 ? we1 : ecv);

        if (tie != 0 || edge)
        {
          // events inside the window at tb, and (edge) those at Tend just past it
          for (int32_t e = ws0; e < cn0; ++e)
          {
            const WalkEv ev = walk_load(cl0, e);
            if (e >= we0 && !(edge && ev.t == Tend))
              break;
            if (ev.t == tb)
              J = fma(static_cast<double>(w_row0[(ev.col & (VW - 1)) * 64]), ev.d, J);
          }
          for (int32_t e = ws1; e < cn1; ++e)
          {
            const WalkEv ev = walk_load(cl1, e);
            if (e >= we1 && !(edge && ev.t == Tend))
              break;
            if (ev.t == tb)
              J = fma(static_cast<double>(w_row1[(ev.col & (VW - 1)) * 64]), ev.d, J);
          }
        }
      }
      if (emit)
      {
        atomicAdd(&B.q[pos], fma(tb, X, -Y));
        atomicAdd(&B.a[pos - 1], fma(0.5, J, X));
      }
      __builtin_amdgcn_wave_barrier();
    }
  }

  // Thread per position: the formulas in the file comment.
  __global__ void finalize_grad_kernel(BwdParams B)
  {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= B.records)
      return;
    const int64_t pt = B.src0[p];
    if (pt < 0) // the zero record standing in for an empty PCF
      return;
    const int64_t i = B.col0[p];
    const int64_t first = B.ooff[i];
    const bool last = p + 1 == B.ooff[i + 1];
    const double v = B.val0[p];
    // integral of H_i over the record: S at its end minus S at its start
    const double Q = (last ? B.a[p] : B.q[p + 1]) - (p > first ? B.q[p] : 0.0);
    double gv, gt = 0.0;
    if (B.op == OPK_DOT)
    {
      gv = Q;
      if (p > first)
        gt = (B.val0[p - 1] - v) * B.a[p - 1];
    }
    else
    {
      const double W = B.wsum[i];
      // the last record ends where the domain does, at DBL_MAX
      const double len = (last ? DBL_MAX : B.key0[p + 1]) - B.key0[p];
      gv = 2.0 * ((v != 0.0 ? v * len * W : 0.0) - Q);
      if (p > first)
      {
        const double vp = B.val0[p - 1];
        gt = (vp - v) * ((vp + v) * W - 2.0 * B.a[p - 1]);
      }
    }
    if (B.grad_f64)
    {
      static_cast<double*>(B.grad_v)[pt] = gv;
      static_cast<double*>(B.grad_t)[pt] = gt;
    }
    else
    {
      static_cast<float*>(B.grad_v)[pt] = static_cast<float>(gv);
      static_cast<float*>(B.grad_t)[pt] = static_cast<float>(gt);
    }
  }

  // Thread per position of the whole set's time order: flags a breakpoint
  // time (> 0) that two different PCFs share.
  __global__ void bwd_ties_kernel(BwdParams B, int32_t from)
  {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p + 1 >= B.records)
      return;
    const double t = B.keys[from][p];
    if (t > 0.0 && t == B.keys[from][p + 1] && B.col0[B.ids[from][p]] != B.col0[B.ids[from][p + 1]])
      atomicOr(B.cross_ties, 1);
  }

  void launch_backward(const BwdParams& prm, bool lists_ready, hipStream_t stream)
  {
    if (prm.records <= 0 || prm.n <= 0)
      return;
    const int32_t vw = prm.w_f64 ? WalkVec<double>::VW : WalkVec<float>::VW;
    if (prm.ch != 2 * vw)
      fail(SBAMD_ERR_RUNTIME, "backward: chunks of 64 partners (FP32 weights) or 32 (FP64)");
    const unsigned rgrid = static_cast<unsigned>((prm.records + 255) / 256);
    hipLaunchKernelGGL(bwd_layout_kernel, dim3(rgrid), dim3(256), 0, stream, prm);
    SBAMD_HIP(hipGetLastError());
    // merge levels up to groups of 64 (the bands' lists); the sublists of
    // vw partners are read off at their level
    int from = 0;
    for (int32_t width = 1; width < BWD_GROUP; width *= 2)
    {
      if (width == vw && !lists_ready)
      {
        hipLaunchKernelGGL(bwd_events_kernel, dim3(rgrid), dim3(256), 0, stream, prm, from, vw, prm.events2);
        SBAMD_HIP(hipGetLastError());
      }
      if (!lists_ready)
      {
        hipLaunchKernelGGL(bwd_merge_kernel, dim3(rgrid), dim3(256), 0, stream, prm, width, from, BWD_GROUP);
        SBAMD_HIP(hipGetLastError());
      }
      from ^= 1;
    }
    if (lists_ready)
      hipLaunchKernelGGL(bwd_refresh_kernel, dim3(2 * rgrid), dim3(256), 0, stream, prm);
    else
      hipLaunchKernelGGL(bwd_events_kernel, dim3(rgrid), dim3(256), 0, stream, prm, from, BWD_GROUP, prm.events);
    SBAMD_HIP(hipGetLastError());
    if (!lists_ready && prm.cross_ties)
    {
      // the whole set's time order (merge levels over all PCFs, kept with the
      // lists' flag), then a check for a breakpoint time that two PCFs share:
      // without one the gather walk skips its tie flags
      int32_t all = 1;
      while (all < prm.n)
        all *= 2;
      int f = from;
      for (int32_t width = BWD_GROUP; width < prm.n; width *= 2)
      {
        hipLaunchKernelGGL(bwd_merge_kernel, dim3(rgrid), dim3(256), 0, stream, prm, width, f, all);
        SBAMD_HIP(hipGetLastError());
        f ^= 1;
      }
      launch_fill_zero(prm.cross_ties, sizeof(int32_t), stream);
      hipLaunchKernelGGL(bwd_ties_kernel, dim3(rgrid), dim3(256), 0, stream, prm, f);
      SBAMD_HIP(hipGetLastError());
    }
    // the sort keys are free now: pst0 takes the other buffer
    BwdParams B = prm;
    B.pst0 = prm.keys[from ^ 1];
    const unsigned pgrid = static_cast<unsigned>((prm.n + 3) / 4);
    hipLaunchKernelGGL(bwd_prefix_kernel, dim3(pgrid), dim3(256), 0, stream, B);
    SBAMD_HIP(hipGetLastError());

    // (band, chunk group) workgroups: every band of 64 rows against every
    // group of BWD_CHUNKS chunks of ch partners, group-major per XCD
    const int64_t nbands = (prm.n + BWD_GROUP - 1) / BWD_GROUP;
    const int64_t nchunks = (prm.n + prm.ch - 1) / prm.ch;
    const int64_t ngroups = (nchunks + BWD_CHUNKS - 1) / BWD_CHUNKS;
    const int64_t grid = 8 * nbands * ((ngroups + 7) / 8);
    if (grid > 0x7fffffff)
      fail(SBAMD_ERR_INVALID, "backward: too many PCFs for one launch");
    const size_t lds = bwd_lds_bytes(prm.w_f64 ? 8 : 4, prm.ch);
    const dim3 gr(static_cast<unsigned>(grid));
    const dim3 blk(64 * BWD_WAVES);
    if (prm.w_f64)
    {
      if (prm.op == OPK_DOT)
        hipLaunchKernelGGL((bwd_gather_kernel<double, OPK_DOT>), gr, blk, lds, stream, B);
      else
        hipLaunchKernelGGL((bwd_gather_kernel<double, OPK_L2>), gr, blk, lds, stream, B);
    }
    else
    {
      if (prm.op == OPK_DOT)
        hipLaunchKernelGGL((bwd_gather_kernel<float, OPK_DOT>), gr, blk, lds, stream, B);
      else
        hipLaunchKernelGGL((bwd_gather_kernel<float, OPK_L2>), gr, blk, lds, stream, B);
    }
    SBAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(finalize_grad_kernel, dim3(rgrid), dim3(256), 0, stream, B);
    SBAMD_HIP(hipGetLastError());
  }
}
