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
// H_i is piecewise constant and changes only where a partner has a
// breakpoint, by w_ij times that partner's jump. PCFs are taken in groups of
// 64 consecutive ones; each group's record starts, merged into one
// time-ordered event list, serve twice. A wavefront handles a band of 64
// rows (one lane per row) against a chunk of 64 partners: it walks the
// chunk's list, where every event moves all lanes' H_i (w_ij from LDS), and
// before each event every lane emits the Q and A of its own records that
// end before that event's time, all lanes at once (the band's list, sorted
// into per-lane queues in LDS, supplies each lane's breakpoints). The cost
// is the chunk list's length plus one emission round per distinct gap
// holding breakpoints; the chunks of a band add their Q and A with global
// atomics, 64 records at a time.
//
// All arrays here are in the original PCF order ("position" p of record k of
// PCF i is ooff[i] + k); a layout kernel copies them out of the sorted
// device set, and the group event lists come from a merge sort by time.
//
// Ties: where f and a partner have a breakpoint at the same time, A takes the
// mean of H_i's left and right limits there (the mean of the partner's values
// on both sides). For a self pair (kernel diagonal) that is the exact
// derivative; for two PCFs jumping together the two one-sided results add up
// to the derivative of moving the shared time.
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

    __device__ __forceinline__ BwdEvent uniform_event(const BwdEvent* p, int64_t i)
    {
      const __attribute__((address_space(4))) BwdEvent* c =
          (const __attribute__((address_space(4))) BwdEvent*)(p);
      BwdEvent e;
      e.t = c[i].t;
      e.dt = c[i].dt;
      e.d = c[i].d;
      e.col = c[i].col;
      e.pos = c[i].pos;
      return e;
    }

    __device__ __forceinline__ double uniform_time(const BwdEvent* p, int32_t i)
    {
      return ((const __attribute__((address_space(4))) BwdEvent*)(p))[i].t;
    }

    // First index of a time-ordered event list of n events whose time is >= t.
    __device__ __forceinline__ int32_t lower_bound_time(const BwdEvent* p, int32_t n, double t)
    {
      int32_t lo = 0, hi = n;
      while (lo < hi)
      {
        const int32_t mid = (lo + hi) >> 1;
        if (uniform_time(p, mid) < t)
          lo = mid + 1;
        else
          hi = mid;
      }
      return lo;
    }

    // Records at positions [a, b) of one PCF that start before t.
    __device__ __forceinline__ int64_t starts_below(const double* key0, int64_t a, int64_t b, double t)
    {
      return count_before(key0, a, b, t, false) - a;
    }
  }

  constexpr int BWD_WAVES = 16; // wavefronts per backward workgroup (time slices of one task)

  // Per wave LDS: four arrays of 64 8-byte words (row queue times, pending
  // Q, pending A, row masks) and three of 64 ints (row queue positions,
  // pending records, queue offsets). The partners' start values (ch FP64,
  // needed only before the main loop) share the first two 8-byte arrays.
  constexpr size_t BWD_WAVE_LDS = 4 * 64 * sizeof(double) + 3 * 64 * sizeof(int32_t);
  static_assert(BWD_WAVE_LDS >= 128 * sizeof(double), "start values of 128 partners must fit");

  // LDS bytes of the pair weights (ch partners x 64 rows) and of the whole
  // workgroup.
  __host__ __device__ constexpr size_t bwd_lds_weights(size_t wbytes, int ch)
  {
    return static_cast<size_t>(ch) * 64 * wbytes;
  }
  __host__ __device__ constexpr size_t bwd_lds_bytes(size_t wbytes, int ch)
  {
    return bwd_lds_weights(wbytes, ch) + BWD_WAVES * BWD_WAVE_LDS;
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

  // One level of the per-chunk merge sort: runs of `width` consecutive PCFs
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

  // Thread per position of the sorted chunk lists: the event record.
  __global__ void bwd_events_kernel(BwdParams B, int32_t from, int32_t group, BwdEvent* out)
  {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= B.records)
      return;
    const int64_t cb = (static_cast<int64_t>(B.col0[p]) / group) * group;
    const double t = B.keys[from][p];
    const double tp = p > B.ooff[cb] ? B.keys[from][p - 1] : 0.0;
    const int64_t id = B.ids[from][p];
    const int64_t c = B.col0[id];
    const bool first = id == B.ooff[c];
    BwdEvent ev;
    ev.t = t;
    ev.dt = t - tp;
    ev.d = B.val0[id] - (first ? 0.0 : B.val0[id - 1]);
    ev.col = static_cast<int32_t>(c - cb) | (first ? BWD_FIRST : 0);
    ev.pos = static_cast<int32_t>(id);
    out[p] = ev;
  }

  // One workgroup per (band of 64 rows, chunk of CH partners), BWD_WAVES
  // wavefronts splitting the task's time axis: wave k takes the events in
  // [T_k, T_{k+1}), with T_k the time of the chunk's event at k/BWD_WAVES of
  // its list. A wave starts from H_i just before T_k (from the partners'
  // values there) and from the row record open there; the part of that
  // record's integral before T_k was added by wave k - 1. Each wave holds
  // blocks of 64 events of both lists in vector registers (lane l holds
  // event l) and reads them with readlane. A breakpoint event costs one lane's work and a partner event
  // all lanes', so wider chunks (CH = 128 with FP32 weights) spend a smaller
  // share of the events on breakpoints. Workgroups are dealt to the 8 XCDs
  // round-robin; all bands of a chunk run on one XCD so the chunk's list
  // stays in that XCD's L2. Chunks go to XCDs in snake order (0..7, 7..0,
  // ...) so that every XCD gets a like share of light and heavy chunks.
  template <typename WT, int OP, int CH>
  __global__ __launch_bounds__(64 * BWD_WAVES) void bwd_stream_kernel(BwdParams B)
  {
    extern __shared__ __align__(16) unsigned char smem[];
    WT* const s_w = reinterpret_cast<WT*>(smem); // s_w[jj * 64 + ii] = w(row ii, partner jj)
    unsigned char* const s_wave = smem + bwd_lds_weights(sizeof(WT), CH);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // the wave index, as a wave-uniform value, so that everything derived
    // from it (slice bounds, list indices) stays in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    double* const s_qt = reinterpret_cast<double*>(s_wave + wv * BWD_WAVE_LDS); // row queue: times
    double* const s_eq = s_qt + 64;                                              // pending Q
    double* const s_ea = s_eq + 64;                                              // pending A
    unsigned long long* const s_m = reinterpret_cast<unsigned long long*>(s_ea + 64); // row masks
    int32_t* const s_qp = reinterpret_cast<int32_t*>(s_m + 64);                  // row queue: positions
    int32_t* const s_ep = s_qp + 64;                                             // pending records
    int32_t* const s_off = s_ep + 64;                                            // queue offsets
    double* const s_g = s_qt; // partners' values at tlo (before the main loop only)
    const int64_t nbands = (B.n + BWD_GROUP - 1) / BWD_GROUP;
    const int64_t nchunks = (B.n + CH - 1) / CH;
    const int64_t x = blockIdx.x;
    const int64_t slot = x >> 3;
    const int64_t g = slot / nbands;
    const int64_t cr = ((g & 1) ? 7 - (x & 7) : (x & 7)) + 8 * g;
    if (cr >= nchunks)
      return;
    // PCFs are sorted by record count, so the last chunks and bands are the
    // heaviest tasks: they go first
    const int64_t c = nchunks - 1 - cr;
    const int64_t band = nbands - 1 - slot % nbands;
    const int64_t i = band * BWD_GROUP + lane;
    const bool act = i < B.n;
    const int64_t j0 = c * CH;
    const int ncol = static_cast<int>(min(static_cast<int64_t>(CH), B.n - j0));
    const int nrow = static_cast<int>(min(static_cast<int64_t>(BWD_GROUP), B.n - band * BWD_GROUP));

    // pair weights, partner-major: the lanes read one row of 64 consecutive
    // words per event
    for (int k = tid; k < CH * 64; k += 64 * BWD_WAVES)
    {
      const int jj = k >> 6, ii = k & 63;
      const double w = jj < ncol && ii < nrow ? pair_weight<OP>(B, band * BWD_GROUP + ii, j0 + jj) : 0.0;
      s_w[k] = static_cast<WT>(w);
    }
    __syncthreads();
    if (wv == 0 && act)
    {
      double ws = 0.0;
      for (int jj = 0; jj < ncol; ++jj)
        ws += static_cast<double>(s_w[jj * 64 + lane]);
      if (ws != 0.0)
        atomicAdd(&B.wsum[i], ws);
    }

    // this wave's slice of both lists
    const int64_t r0 = uniform_load(B.ooff, band * BWD_GROUP);
    const int32_t nr = static_cast<int32_t>(uniform_load(B.ooff, band * BWD_GROUP + nrow) - r0);
    const int64_t c0 = uniform_load(B.ooff, j0);
    const int32_t ncv = static_cast<int32_t>(uniform_load(B.ooff, j0 + ncol) - c0);
    const BwdEvent* const rl = B.events + r0;
    const BwdEvent* const cl = B.events2 + c0;
    const bool last = wv == BWD_WAVES - 1;
    double tlo = 0.0, thi = __builtin_inf();
    int32_t cb = 0, ce = ncv, rb = 0, re = nr;
    if (wv > 0)
    {
      tlo = uniform_time(cl, static_cast<int32_t>((static_cast<int64_t>(wv) * ncv) / BWD_WAVES));
      cb = lower_bound_time(cl, ncv, tlo);
      rb = lower_bound_time(rl, nr, tlo);
    }
    if (!last)
    {
      thi = uniform_time(cl, static_cast<int32_t>((static_cast<int64_t>(wv + 1) * ncv) / BWD_WAVES));
      ce = lower_bound_time(cl, ncv, thi);
      re = lower_bound_time(rl, nr, thi);
    }

    // H_i just before tlo, from the partners' values there (lane = partner)
    double H = 0.0;
    if (wv > 0)
    {
      for (int jb = 0; jb < CH; jb += 64)
      {
        const int jj = jb + lane;
        double gv = 0.0;
        if (jj < ncol)
        {
          const int64_t a = B.ooff[j0 + jj];
          const int64_t k = starts_below(B.key0, a, B.ooff[j0 + jj + 1], tlo);
          if (k > 0)
            gv = B.val0[a + k - 1];
        }
        s_g[jj] = gv;
      }
      for (int jj = 0; jj < ncol; ++jj)
        H = fma(static_cast<double>(s_w[jj * 64 + lane]), s_g[jj], H);
    }
    // the row record open just before tlo
    int32_t open = 0; // positions are below 2^31 (checked when the plan is made)
    if (act)
    {
      const int64_t a = B.ooff[i];
      const int64_t k = wv > 0 ? starts_below(B.key0, a, B.ooff[i + 1], tlo) : 0;
      open = static_cast<int32_t>(a + (k > 0 ? k - 1 : 0));
    }
    // past the chunk's last event H_i is sum_j w_ij (last value of g_j),
    // taken exactly there rather than as the running sum of the jumps
    double Hend = 0.0;
    if (last)
      for (int jj = 0; jj < ncol; ++jj)
        Hend = fma(static_cast<double>(s_w[jj * 64 + lane]),
                   B.val0[uniform_load(B.ooff, j0 + jj + 1) - 1], Hend);

    // A block of 64 events of a list in vector registers: lane l holds
    // event s + l (index clamped into the list, so that every register is
    // defined). Only the fields used are loaded: a load whose destination is
    // never read leaves a register that the compiler reuses, and that reuse
    // waits for the load.
    struct PartnerBlk
    {
      double t, d;
      int32_t col;
    };
    auto load_partners = [&](int32_t s0) -> PartnerBlk {
      const BwdEvent* const ev = cl + max(min(s0 + lane, ncv - 1), 0);
      PartnerBlk b;
      b.t = ev->t;
      b.d = ev->d;
      b.col = ev->col;
      return b;
    };
    auto lane_d = [](double v, int l) -> double {
      return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                              __builtin_amdgcn_readlane(__double2loint(v), l));
    };

    // Row breakpoints are taken by their own lanes, all lanes at once. The
    // wave holds a block of up to 64 events of the band's list, sorted by
    // lane into per-lane queues in LDS (s_qt / s_qp, lane L's events at
    // [qptr, qend) in time order): a 64-bit mask per lane (bit e set when
    // event e of the block ends one of its records, built with LDS atomic
    // ORs) gives each event its rank in its lane's queue. nb is the lane's
    // next breakpoint and npos the position of the record that starts
    // there; a lane with no events left in the block has nb = the block's
    // last time, so that a partner event past that time brings in the next
    // block. The row block is loaded when it is needed, not ahead: a block
    // held in registers across the partner loop costs register copies there
    // that wait on every outstanding memory operation.
    int32_t rs = rb;
    int32_t qptr = 0, qend = 0;
    double nb = 0.0, blast = 0.0;
    int32_t npos = 0;
    auto next_own = [&]() {
      // qptr = 64 (a lane that had all 64 events) reads s_eq[0] / s_ep[0],
      // still this wave's LDS, and the value is discarded
      const double t = s_qt[qptr];
      npos = s_qp[qptr];
      nb = qptr < qend ? t : blast;
    };
    auto open_block = [&]() {
      const BwdEvent* const ev = rl + max(min(rs + lane, nr - 1), 0);
      const double t = ev->t;
      const int32_t col = ev->col & 63, pos = ev->pos;
      const bool valid = rs + lane < re && !(ev->col & BWD_FIRST);
      s_m[lane] = 0ull;
      __builtin_amdgcn_wave_barrier();
      if (valid)
        atomicOr(&s_m[col], 1ull << lane);
      __builtin_amdgcn_wave_barrier();
      // queue offsets: exclusive prefix sum of the lanes' event counts
      const int32_t cnt = __popcll(s_m[lane]);
      int32_t incl = cnt;
      for (int o = 1; o < 64; o <<= 1)
      {
        const int32_t v = __shfl_up(incl, o);
        if (lane >= o)
          incl += v;
      }
      qptr = incl - cnt;
      qend = incl;
      s_off[lane] = qptr;
      __builtin_amdgcn_wave_barrier();
      if (valid)
      {
        const int32_t slot = s_off[col] + static_cast<int32_t>(__popcll(s_m[col] & ((1ull << lane) - 1ull)));
        s_qt[slot] = t;
        s_qp[slot] = pos;
      }
      __builtin_amdgcn_wave_barrier();
      blast = rs + 64 < re ? lane_d(t, 63) : __builtin_inf();
      next_own();
    };
    open_block();

    // Pending emissions (record, Q, A), added to q and a 64 at a time.
    int32_t nem = 0;
    auto flush = [&]() {
      __builtin_amdgcn_wave_barrier();
      if (lane < nem)
      {
        atomicAdd(&B.q[s_ep[lane]], s_eq[lane]);
        atomicAdd(&B.a[s_ep[lane]], s_ea[lane]);
      }
      nem = 0;
    };

    // next chunk event; blocks start at cb + 64 k, so event ic is lane
    // (ic - cb) & 63 of its block
    int32_t ic = cb;
    PartnerBlk Cc = load_partners(cb);
    __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0): see the block loads below
    double acc = 0.0;   // integral of H_i from the open record's start (or tlo) to tprev
    double tprev = tlo; // time of the chunk's last event taken

    // Sum over the chunk's events at time tprev (all taken) of w_ij * jump:
    // at a breakpoint there, A is the mean of H_i's two limits, H - J / 2.
    auto jumps_at_tprev = [&](int32_t ic, double tprev) -> double {
      double J = 0.0;
      for (int32_t e = ic - 1; e >= cb; --e)
      {
        const BwdEvent T = uniform_event(cl, e);
        if (__double_as_longlong(T.t) != __double_as_longlong(tprev))
          break;
        J = fma(static_cast<double>(s_w[(T.col & (CH - 1)) * 64 + lane]), T.d, J);
      }
      return J;
    };
    // Emission rounds (every lane with a breakpoint before T emits one
    // record per round) and new row blocks, until no lane is below T.
    auto drain = [&](double T) {
      while (__ballot(nb < T))
      {
        const bool go = qptr < qend && nb < T;
        const unsigned long long gom = __ballot(go);
        if (gom)
        {
          const int32_t ngo = __popcll(gom);
          if (nem + ngo > 64)
            flush();
          if (go)
          {
            // record npos - 1 ends at nb
            double J = 0.0;
            if (__double_as_longlong(nb) == __double_as_longlong(tprev))
              J = jumps_at_tprev(ic, tprev);
            const double d = nb - tprev;
            const int32_t slot = nem + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(gom >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(gom), 0u));
            s_eq[slot] = fma(H, d, acc);
            s_ea[slot] = fma(-0.5, J, H);
            s_ep[slot] = npos - 1;
            acc = -H * d;
            open = npos;
            ++qptr;
            next_own();
          }
          nem += ngo;
        }
        else
        {
          // the lanes below T are done with this block, which ends before T
          rs += 64;
          open_block();
        }
      }
    };
    double Ct = lane_d(Cc.t, 0);
    while (ic < ce)
    {
      // breakpoints before the chunk's event ic see H before it (at equal
      // times the chunk's events come first)
      if (__ballot(nb < Ct))
        drain(Ct);
      const int l = (ic - cb) & 63;
      const double Cd = lane_d(Cc.d, l);
      const int32_t Ccol = __builtin_amdgcn_readlane(Cc.col, l) & (CH - 1);
      const double w = static_cast<double>(s_w[Ccol * 64 + lane]);
      acc = fma(H, Ct - tprev, acc);
      H = fma(w, Cd, H);
      tprev = Ct;
      if (++ic < ce)
      {
        if (l == 63)
        {
          // The next block is loaded when needed and waited for here, once
          // per block: a block still in flight where the paths of this loop
          // meet would make the compiler wait on it (and on every atomic
          // issued before it) at each event. 0x0F70 = vmcnt(0) alone.
          Cc = load_partners(ic);
          __builtin_amdgcn_s_waitcnt(0x0F70);
        }
        Ct = lane_d(Cc.t, (ic - cb) & 63);
      }
    }
    // past the chunk's last event H_i is exact
    if (last)
      H = Hend;
    drain(__builtin_inf());
    flush();
    if (act)
    {
      // the open record's integral up to thi; the last record runs to the
      // end of the domain, DBL_MAX
      double qv;
      if (!last)
        qv = fma(H, thi - tprev, acc);
      else
        qv = H != 0.0 ? fma(H, DBL_MAX - tprev, acc) : acc;
      atomicAdd(&B.q[open], qv);
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
    double gv, gt = 0.0;
    if (B.op == OPK_DOT)
    {
      gv = B.q[p];
      if (p > first)
        gt = (B.val0[p - 1] - v) * B.a[p - 1];
    }
    else
    {
      const double W = B.wsum[i];
      // the last record ends where the domain does, at DBL_MAX
      const double len = (last ? DBL_MAX : B.key0[p + 1]) - B.key0[p];
      gv = 2.0 * ((v != 0.0 ? v * len * W : 0.0) - B.q[p]);
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

  void launch_backward(const BwdParams& prm, hipStream_t stream)
  {
    if (prm.records <= 0 || prm.n <= 0)
      return;
    if (prm.ch != BWD_GROUP && !(prm.ch == 2 * BWD_GROUP && !prm.w_f64))
      fail(SBAMD_ERR_RUNTIME, "backward: chunks of 64 partners, or 128 with FP32 weights");
    const unsigned rgrid = static_cast<unsigned>((prm.records + 255) / 256);
    hipLaunchKernelGGL(bwd_layout_kernel, dim3(rgrid), dim3(256), 0, stream, prm);
    SBAMD_HIP(hipGetLastError());
    // merge levels up to groups of 64 (the bands' lists), then one more
    // level for chunks of 128
    int from = 0;
    for (int32_t width = 1; width < prm.ch; width *= 2)
    {
      const int32_t group = width < BWD_GROUP ? BWD_GROUP : prm.ch;
      if (width == BWD_GROUP)
      {
        hipLaunchKernelGGL(bwd_events_kernel, dim3(rgrid), dim3(256), 0, stream, prm, from, BWD_GROUP,
                           prm.events);
        SBAMD_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(bwd_merge_kernel, dim3(rgrid), dim3(256), 0, stream, prm, width, from, group);
      SBAMD_HIP(hipGetLastError());
      from ^= 1;
    }
    hipLaunchKernelGGL(bwd_events_kernel, dim3(rgrid), dim3(256), 0, stream, prm, from, prm.ch,
                       prm.events2);
    SBAMD_HIP(hipGetLastError());

    // (band, chunk) workgroups: every band of 64 rows against every chunk of
    // ch partners, chunk-major per XCD
    const int64_t nbands = (prm.n + BWD_GROUP - 1) / BWD_GROUP;
    const int64_t nchunks = (prm.n + prm.ch - 1) / prm.ch;
    const int64_t grid = 8 * nbands * ((nchunks + 7) / 8);
    if (grid > 0x7fffffff)
      fail(SBAMD_ERR_INVALID, "backward: too many PCFs for one launch");
    const size_t lds = bwd_lds_bytes(prm.w_f64 ? 8 : 4, prm.ch);
    const dim3 g(static_cast<unsigned>(grid));
    const dim3 blk(64 * BWD_WAVES);
    if (prm.w_f64)
    {
      if (prm.op == OPK_DOT)
        hipLaunchKernelGGL((bwd_stream_kernel<double, OPK_DOT, BWD_GROUP>), g, blk, lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_stream_kernel<double, OPK_L2, BWD_GROUP>), g, blk, lds, stream, prm);
    }
    else if (prm.ch == 2 * BWD_GROUP)
    {
      if (prm.op == OPK_DOT)
        hipLaunchKernelGGL((bwd_stream_kernel<float, OPK_DOT, 2 * BWD_GROUP>), g, blk, lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_stream_kernel<float, OPK_L2, 2 * BWD_GROUP>), g, blk, lds, stream, prm);
    }
    else
    {
      if (prm.op == OPK_DOT)
        hipLaunchKernelGGL((bwd_stream_kernel<float, OPK_DOT, BWD_GROUP>), g, blk, lds, stream, prm);
      else
        hipLaunchKernelGGL((bwd_stream_kernel<float, OPK_L2, BWD_GROUP>), g, blk, lds, stream, prm);
    }
    SBAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(finalize_grad_kernel, dim3(rgrid), dim3(256), 0, stream, prm);
    SBAMD_HIP(hipGetLastError());
  }
}
