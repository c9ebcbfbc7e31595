// Device-side building blocks shared by the forward and backward tile kernels.
//
// Layout of one PCF on the device (a "record" per constant piece): record k
// holds the value on [t_start(k), t_end(k)) and t_end(k); the last record has
// t_end = +inf. Zero-length pieces (duplicate breakpoint times) are dropped
// on the host, so t_end is strictly increasing within a PCF and t_start(0) = 0.
//
// The walk over a pair advances exactly one side per step (the side whose
// record ends first, f on ties), so a pair takes m_f + m_g - 2 steps plus a
// final rectangle to the end of the domain. A tie of the two sides becomes
// two steps, the second of zero length; the reference walk
// (include/sbear/algorithms/functional/iterate_rectangles.hpp:54-103) takes
// one step there. A zero-length step adds fma(0, w, acc) == acc, so the FP64
// sums are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "sbamd_internal.hpp"

namespace sbamd
{
  template <typename T>
  struct Lim;

  template <>
  struct Lim<double>
  {
    __device__ static double max() { return DBL_MAX; }
    __device__ static double inf() { return __builtin_inf(); }
    using Bits = unsigned long long;
    __device__ static Bits bits(double x) { return static_cast<Bits>(__double_as_longlong(x)); }
    __device__ static double from_bits(Bits b) { return __longlong_as_double(static_cast<long long>(b)); }
  };

  template <>
  struct Lim<float>
  {
    __device__ static float max() { return FLT_MAX; }
    __device__ static float inf() { return __builtin_inff(); }
    using Bits = unsigned int;
    __device__ static Bits bits(float x) { return __float_as_uint(x); }
    __device__ static float from_bits(Bits b) { return __uint_as_float(b); }
  };

  // Minimum and maximum of two breakpoint times, one v_min / v_max each.
  // Times are never NaN (inputs are validated finite; +inf is the end
  // sentinel), so the operand canonicalisation the compiler wraps around
  // fmin / fmax for IEEE-mode NaN quieting is dead weight in the walk.
  __device__ __forceinline__ double tmin(double a, double b)
  {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  __device__ __forceinline__ double tmax(double a, double b)
  {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  __device__ __forceinline__ float tmin(float a, float b)
  {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  __device__ __forceinline__ float tmax(float a, float b)
  {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }

  // Pair operations: operations.cuh:19-65. OPK_L2 is the Lp operation at
  // p == 2 (|d|^2 computed as d * d, finalize sqrt).
  template <int OP, typename TV>
  __device__ __forceinline__ TV pair_op(TV f, TV g, TV p)
  {
    if constexpr (OP == OPK_L1)
      return fabs(f - g);
    else if constexpr (OP == OPK_L2)
    {
      TV d = f - g;
      return d * d;
    }
    else if constexpr (OP == OPK_LP)
      return pow(fabs(f - g), p);
    else
      return f * g;
  }

  // The same operations split in two, for a walk that keeps the pair's
  // current segment as loop state: pair_pre(f, g) is the part formed when the
  // segment starts (f - g, or f * g), pair_post(x, p) the rest, applied where
  // the segment is accumulated, so that |x| folds into the consuming FMA as an
  // operand modifier. pair_post(pair_pre(f, g), p) == pair_op(f, g, p) bit for
  // bit, and pair_pre(f, g) == +-pair_pre(g, f), which pair_post erases.
  template <int OP, typename TV>
  __device__ __forceinline__ TV pair_pre(TV f, TV g)
  {
    if constexpr (OP == OPK_DOT)
      return f * g;
    else
      return f - g;
  }

  template <int OP, typename TV>
  __device__ __forceinline__ TV pair_post(TV x, TV p)
  {
    if constexpr (OP == OPK_L1)
      return fabs(x);
    else if constexpr (OP == OPK_L2)
      return x * x;
    else if constexpr (OP == OPK_LP)
      return pow(fabs(x), p);
    else
      return x;
  }

  template <int OP>
  __device__ __forceinline__ double finalize(double x, double p)
  {
    if constexpr (OP == OPK_L2)
      return sqrt(x);
    else if constexpr (OP == OPK_LP)
      return pow(x, 1.0 / p);
    else
      return x;
  }

  // Plain accumulator: acc = fma(dt, w, acc), one rounding per segment, the
  // operation the reference's integrate() performs (matrix_integrate.hpp:25-27).
  template <typename TV>
  struct PlainAcc
  {
    TV acc = TV(0);
    __device__ __forceinline__ void add(TV dt, TV w) { acc = fma(dt, w, acc); }
    __device__ __forceinline__ void flush() {}
    __device__ __forceinline__ void fold() {}
    template <typename TT>
    __device__ __forceinline__ double finish(TT tail, TV w)
    {
      // final rectangle [t, MAX) in the working type, as the reference does
      acc = fma(static_cast<TV>(tail), w, acc);
      return static_cast<double>(acc);
    }
  };

  // Kahan-compensated FP32 accumulator with an FP64 fold for long pairs.
  struct KahanAcc
  {
    float s = 0.f;
    float c = 0.f;
    double big = 0.0;
    __device__ __forceinline__ void add(float dt, float w)
    {
      float y = fmaf(dt, w, -c);
      float t = s + y;
      c = (t - s) - y;
      s = t;
    }
    __device__ __forceinline__ void fold()
    {
      big += static_cast<double>(s) - static_cast<double>(c);
      s = 0.f;
      c = 0.f;
    }
    template <typename TT>
    __device__ __forceinline__ double finish(TT tail, float w)
    {
      // The final rectangle [t, MAX) is formed in the time type, whose MAX
      // bounds the domain (FLT_MAX or DBL_MAX), so a non-zero tail overflows
      // exactly as the reference's working type does; a zero tail adds 0.
      const double r = big + (static_cast<double>(s) - static_cast<double>(c));
      const TT term = tail * static_cast<TT>(w);
      return r + static_cast<double>(term);
    }
  };

  // Kahan-compensated FP32 sum over blocks of segments (the forward tile
  // walk). add() chains a block's products into one FP32 partial with FMAs;
  // flush() adds the partial to the compensated sum, which the walk does
  // every 8 segments in its common steps and every 4 in its tails. A block of
  // b non-negative terms carries a relative error of at most b u (u = 2^-24)
  // and the compensated sum of the blocks at most 2u more, so an L1 / Lp sum
  // stays within 10u = 6.0e-7 of the exact sum of its FP32 terms, at a
  // fraction of the per-segment work of a compensated add per segment. fold()
  // moves the sum into the FP64 accumulator (long pairs, every phase).
  struct BlockKahanAcc
  {
    float s = 0.f;
    float c = 0.f;
    float part = 0.f;
    double big = 0.0;
    __device__ __forceinline__ void add(float dt, float w) { part = fmaf(dt, w, part); }
    __device__ __forceinline__ void flush()
    {
      const float y = part - c;
      const float t = s + y;
      c = (t - s) - y;
      s = t;
      part = 0.f;
    }
    __device__ __forceinline__ void fold()
    {
      flush();
      big += static_cast<double>(s) - static_cast<double>(c);
      s = 0.f;
      c = 0.f;
    }
    template <typename TT>
    __device__ __forceinline__ double finish(TT tail, float w)
    {
      flush();
      const double r = big + (static_cast<double>(s) - static_cast<double>(c));
      const TT term = tail * static_cast<TT>(w);
      return r + static_cast<double>(term);
    }
  };

  __device__ __forceinline__ int64_t distmat_index(int64_t i, int64_t j)
  {
    int64_t hi = i > j ? i : j;
    int64_t lo = i > j ? j : i;
    return hi * (hi - 1) / 2 + lo;
  }

  __device__ __forceinline__ int64_t symmat_index(int64_t i, int64_t j)
  {
    int64_t hi = i > j ? i : j;
    int64_t lo = i > j ? j : i;
    return hi * (hi + 1) / 2 + lo;
  }

  // One staged record in LDS: end time and value of a constant piece. 16 bytes
  // for FP64 times (so one record is one aligned 16-byte access), 8 for FP32.
  template <typename TT, typename TV>
  struct alignas(sizeof(TT) == 8 ? 16 : 8) Rec
  {
    TT t;
    TV v;
  };

  // One LDS read of a staged record. An FP64-time / FP32-value record is
  // read as one vector load: field by field it becomes a 64-bit and a 32-bit
  // LDS read, two LDS instructions per walk step instead of one.
  template <typename TT, typename TV>
  __device__ __forceinline__ Rec<TT, TV> lds_rec(const Rec<TT, TV>* p)
  {
    if constexpr (sizeof(TT) == 8 && sizeof(TV) == 4)
    {
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 w = *reinterpret_cast<const u32x4*>(p);
      // the lanes go through scalars first: this clang lowers
      // __builtin_bit_cast of a vector element expression (w.z) to element 0
      const uint32_t t_lo = w.x, t_hi = w.y, v_bits = w.z;
      Rec<TT, TV> r;
      r.t = __builtin_bit_cast(double, (static_cast<uint64_t>(t_hi) << 32) | t_lo);
      r.v = __builtin_bit_cast(float, v_bits);
      return r;
    }
    else
      return *p;
  }

  // ---------------------------------------------------------------------
  // LDS of a tile kernel.
  //
  // Slot s (rows 0..R-1, then columns R..R+C-1) has a window of sr =
  // slot_recs(cap) records; record j of slot s sits at
  //   win_off + (s * sr + j) * RB
  // with RB the record size (16 bytes for FP64 times, 8 for FP32 times).
  // Every window has the same stride, so a walk pointer advances by RB
  // whichever side it belongs to, and a pointer is known to lie in a
  // column window when it is at or past the first column window (rows
  // precede columns). sr is 1 modulo 16 so that 16-byte records of
  // different slots at the same position fall in different bank groups.
  // Then per-slot tables and the phase end time.
  // ---------------------------------------------------------------------
  __host__ __device__ inline int slot_recs(int cap) { return (cap + 15) / 16 * 16 + 1; }

  __host__ __device__ inline int align_up(int x, int a) { return (x + a - 1) / a * a; }

  struct LdsLayout
  {
    int rb;         // record bytes
    int sr;         // records per slot window
    int win_off;    // slot 0's window
    int src_off;    // u64 per slot: global byte address of the next window's first record
    int tstart_off; // time type per slot: start time of the next window's first record
    int rem_off;    // i32 per slot: records from the next window's first to the PCF's end
    int cnt_off;    // u32 per slot: records consumed in the current phase (low 16 bits) and records staged (high 16)
    int bnd_off;    // time type (bits, 8 bytes) per slot: start of the first record not staged
    int tb_off;     // phase end time (bits)
    int misc_off;   // next tile of a persistent workgroup
    int rows_off;   // R FP64 row-sum partials
    int total;
  };

  __host__ __device__ inline LdsLayout lds_layout(int tt_bytes, int R, int C, int cap)
  {
    LdsLayout L;
    const int NS = R + C;
    L.rb = tt_bytes == 8 ? 16 : 8;
    L.sr = slot_recs(cap);
    int o = 0;
    L.win_off = o;
    o += NS * L.sr * L.rb;
    L.src_off = o;
    o += 8 * NS;
    L.tstart_off = o;
    o += 8 * NS;
    L.rem_off = o;
    o += 4 * NS;
    L.cnt_off = o;
    o += 4 * NS;
    o = align_up(o, 8);
    L.bnd_off = o;
    o += 8 * NS;
    L.tb_off = o;
    o += 8;
    L.misc_off = o;
    o += 8;
    L.rows_off = o;
    o += 8 * R;
    L.total = align_up(o, 16);
    return L;
  }

  template <typename TT, typename TV>
  struct TileLds
  {
    unsigned char* smem;
    LdsLayout lay;
    uint32_t base; // LDS address of smem
    unsigned long long* src;
    TT* tstart;
    int32_t* rem;
    uint32_t* cnt;
    unsigned long long* bnd;
    typename Lim<TT>::Bits* tb;

    __device__ TileLds(unsigned char* s, const LdsLayout& L) : smem(s), lay(L)
    {
#if __HIP_DEVICE_COMPILE__
      base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) unsigned char*)s));
#else
      base = 0;
#endif
      src = reinterpret_cast<unsigned long long*>(s + L.src_off);
      tstart = reinterpret_cast<TT*>(s + L.tstart_off);
      rem = reinterpret_cast<int32_t*>(s + L.rem_off);
      cnt = reinterpret_cast<uint32_t*>(s + L.cnt_off);
      bnd = reinterpret_cast<unsigned long long*>(s + L.bnd_off);
      tb = reinterpret_cast<typename Lim<TT>::Bits*>(s + L.tb_off);
    }

    // LDS address of record 0 of slot s's window. Walk pointers are LDS
    // addresses (not offsets from smem), so a record read takes the pointer
    // as its address with no add of the dynamic-LDS base in the loop.
    __device__ __forceinline__ uint32_t win(int s) const
    {
      return base + static_cast<uint32_t>(lay.win_off) +
             static_cast<uint32_t>(s) * static_cast<uint32_t>(lay.sr * lay.rb);
    }

    __device__ __forceinline__ Rec<TT, TV>& rec(uint32_t addr) const
    {
#if __HIP_DEVICE_COMPILE__
      using P3 = __attribute__((address_space(3))) unsigned char*;
      return *reinterpret_cast<Rec<TT, TV>*>((unsigned char*)(P3)(uintptr_t)addr);
#else
      return *reinterpret_cast<Rec<TT, TV>*>(smem + (addr - base));
#endif
    }
  };

  // One 16-byte global -> LDS load (global_load_lds_dwordx4). lds_addr is the
  // wavefront-uniform LDS address of lane 0's 16 bytes; lane i's record lands
  // at lds_addr + 16 i (lanes that do not call leave their slot untouched).
  // The builtin exists only in the device compilation.
  __device__ __forceinline__ void lds_dma16(const void* src, uint32_t lds_addr)
  {
#if __HIP_DEVICE_COMPILE__
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(uintptr_t)lds_addr, 16, 0, 0);
#else
    (void)src;
    (void)lds_addr;
#endif
  }

  // Minimum over the 64 lanes of a wavefront (all lanes active), as a
  // wavefront-uniform value. Within each row of 16 lanes, shifts right by 1,
  // 2, 4 and 8 (DPP row_shr; a lane without a source keeps its own value)
  // leave the row's minimum in the row's last lane; the four row minima are
  // read from lanes 15, 31, 47 and 63. No LDS round trips.
  template <int SHIFT>
  __device__ __forceinline__ uint32_t dpp_shr(uint32_t v)
  {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x110 + SHIFT, 0xf, 0xf, false));
  }

  __device__ __forceinline__ uint32_t wave_min(uint32_t v)
  {
    v = min(v, dpp_shr<1>(v));
    v = min(v, dpp_shr<2>(v));
    v = min(v, dpp_shr<4>(v));
    v = min(v, dpp_shr<8>(v));
    const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 15));
    const uint32_t b = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 31));
    const uint32_t c = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 47));
    const uint32_t d = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
    return min(min(a, b), min(c, d));
  }

  // Maximum over the 64 lanes of a wavefront, as wave_min.
  __device__ __forceinline__ uint32_t wave_max(uint32_t v)
  {
    v = max(v, dpp_shr<1>(v));
    v = max(v, dpp_shr<2>(v));
    v = max(v, dpp_shr<4>(v));
    v = max(v, dpp_shr<8>(v));
    const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 15));
    const uint32_t b = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 31));
    const uint32_t c = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 47));
    const uint32_t d = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
    return max(max(a, b), max(c, d));
  }

  template <int SHIFT>
  __device__ __forceinline__ unsigned long long dpp_shr(unsigned long long v)
  {
    const uint32_t lo = dpp_shr<SHIFT>(static_cast<uint32_t>(v));
    const uint32_t hi = dpp_shr<SHIFT>(static_cast<uint32_t>(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
  }

  __device__ __forceinline__ unsigned long long wave_min(unsigned long long v)
  {
    unsigned long long u = dpp_shr<1>(v);
    v = u < v ? u : v;
    u = dpp_shr<2>(v);
    v = u < v ? u : v;
    u = dpp_shr<4>(v);
    v = u < v ? u : v;
    u = dpp_shr<8>(v);
    v = u < v ? u : v;
    unsigned long long r = ~0ull;
#pragma unroll
    for (int row = 0; row < 4; ++row)
    {
      const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 16 * row + 15));
      const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), 16 * row + 15));
      const unsigned long long x = (static_cast<unsigned long long>(hi) << 32) | lo;
      r = x < r ? x : r;
    }
    return r;
  }

  // Tile start: slot s < R is row PCF r0 + s, slot R + c is column c0 + c;
  // slots past the tile's extent hold no records.
  template <typename TT, typename TV>
  __device__ void tile_begin(const TileLds<TT, TV>& S, const FwdParams& P, const Tile& tl,
                             const Problem& pr, int tid)
  {
    const int NS = P.R + P.C;
    if (tid < NS)
    {
      const bool isrow = tid < P.R;
      const int k = isrow ? tid : tid - P.R;
      const int lim = isrow ? tl.nr : tl.nc;
      const int64_t* off = isrow ? P.r_off : P.c_off;
      const unsigned char* recs = static_cast<const unsigned char*>(isrow ? P.r_rec : P.c_rec);
      int64_t b = 0, m = 0;
      if (k < lim)
      {
        const int64_t g = isrow ? pr.row_base + tl.r0 + k : pr.col_base + tl.c0 + k;
        b = off[g];
        m = off[g + 1] - b;
      }
      S.src[tid] = reinterpret_cast<unsigned long long>(recs + b * S.lay.rb);
      S.rem[tid] = static_cast<int32_t>(m);
      S.tstart[tid] = TT(0);
      S.cnt[tid] = 0;
    }
  }

  // First part of a phase (before the barrier): every slot's next window
  // (up to cap records from its cursor) is copied into LDS, one wavefront
  // per slot, each lane moving whole records; with 16-byte records the
  // copies go global -> LDS directly (global_load_lds_dwordx4), so one
  // vmcnt wait covers them all. Wavefront 0 also finds the phase end: each
  // slot with more than cap records left proposes the end time of the last
  // record it stages, and the phase ends at the earliest proposal (+inf
  // when every slot stages all it has left). Each slot's proposal is kept as
  // its bound: a pair may walk on past the phase end while the next record
  // of either side starts before the earlier of its two slots' bounds.
  template <typename TT, typename TV>
  __device__ void phase_stage(const TileLds<TT, TV>& S, const FwdParams& P, int tid, int nthreads)
  {
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    constexpr uint32_t RB = sizeof(RecT);
    const int NS = P.R + P.C;
    const int cap = P.cap;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nwaves = nthreads >> 6;
    typename L::Bits prop = L::bits(L::inf());
    if (w == 0 && lane < NS && S.rem[lane] > cap)
    {
      const RecT* last = reinterpret_cast<const RecT*>(S.src[lane]) + (cap - 1);
      prop = L::bits(last->t);
    }
    for (int s = w; s < NS; s += nwaves)
    {
      const int rem = S.rem[s];
      const int n = rem < cap ? rem : cap;
      const unsigned char* src = reinterpret_cast<const unsigned char*>(S.src[s]);
      const uint32_t dst = S.win(s);
      for (int c0 = 0; c0 < n; c0 += 64)
      {
        const int j = c0 + lane;
        if (j < n)
        {
          if constexpr (RB == 16)
            lds_dma16(src + static_cast<size_t>(j) * RB, dst + RB * static_cast<uint32_t>(c0));
          else
            S.rec(dst + RB * static_cast<uint32_t>(j)) = reinterpret_cast<const RecT*>(src)[j];
        }
      }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (w == 0)
    {
      if (lane < NS)
        S.bnd[lane] = static_cast<unsigned long long>(prop);
      const typename L::Bits tb = wave_min(prop);
      if (lane == 0)
        *S.tb = tb;
    }
  }

  // Second part (after the barrier): the cursor of every slot -- the number
  // of staged records that start before the phase end Tb, found by a binary
  // search over the window (starts increase along it), kept with the number
  // of staged records -- and the slot's next window: its first record, start
  // time and remaining count. One lane of wavefront 0 per slot, all slots at
  // once.
  template <typename TT, typename TV>
  __device__ void phase_cursors(const TileLds<TT, TV>& S, const FwdParams& P, TT Tb, int tid)
  {
    using RecT = Rec<TT, TV>;
    constexpr uint32_t RB = sizeof(RecT);
    const int NS = P.R + P.C;
    if (tid < NS)
    {
      const int rem = S.rem[tid];
      const int n = rem < P.cap ? rem : P.cap;
      const uint32_t wb = S.win(tid);
      const TT t0 = S.tstart[tid];
      // first j in [0, n) whose record starts at or after Tb (n if none);
      // record j starts where record j - 1 ends, record 0 at t0
      int lo = 0, hi = n;
      while (lo < hi)
      {
        const int mid = (lo + hi) >> 1;
        const TT st = mid == 0 ? t0 : S.rec(wb + RB * static_cast<uint32_t>(mid - 1)).t;
        if (st < Tb)
          lo = mid + 1;
        else
          hi = mid;
      }
      S.cnt[tid] = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(n) << 16);
      if (lo > 0)
      {
        S.tstart[tid] = S.rec(wb + RB * static_cast<uint32_t>(lo - 1)).t;
        S.rem[tid] = rem - lo;
        S.src[tid] += static_cast<unsigned long long>(lo) * RB;
      }
    }
  }
}
