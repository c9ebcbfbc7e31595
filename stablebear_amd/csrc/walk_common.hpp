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

  // Byte layout of the dynamic LDS block of a tile kernel:
  //   slot s (rows 0..R-1, then columns), record j at
  //     rows_off + (s * slot_recs(cap) + j) * REC
  //   Every slot is contiguous with the same record stride, so advancing a
  //   walk pointer is one add whichever side it belongs to; slot_recs is 1
  //   modulo 16 so 16-byte records of different slots at the same position,
  //   and of one slot at nearby positions, fall in different bank groups.
  //   Then per-slot tables, the phase end time, and optional per-record extra
  //   space (backward gradient slots) mirroring the record geometry.
  constexpr int ROW_PAD = 4;

  __host__ __device__ inline int slot_recs(int cap) { return (cap + 15) / 16 * 16 + 1; }

  struct LdsLayout
  {
    int rec_bytes;
    int rows_off, cols_off, rec_end;
    int base_off, m_off, cur_off, next_off, tstart_off, lbase_off, lstride_off, tb_off, misc_off;
    int extra_off, total;
  };

  __host__ __device__ inline int align_up(int x, int a) { return (x + a - 1) / a * a; }

  // extra_per_rec: bytes per staged record of the optional extra region
  // (a multiple of rec_bytes; placed at extra_off + rec_offset * extra_per_rec / rec_bytes).
  __host__ __device__ inline LdsLayout lds_layout(int tt_bytes, int R, int C, int cap, int extra_per_rec)
  {
    LdsLayout L;
    const int NS = R + C;
    L.rec_bytes = tt_bytes == 8 ? 16 : 8;
    int o = 0;
    L.rows_off = o;
    o += R * slot_recs(cap) * L.rec_bytes;
    L.cols_off = o;
    o += C * slot_recs(cap) * L.rec_bytes;
    L.rec_end = o;
    L.base_off = o;
    o += 8 * NS;
    L.m_off = o;
    o += 8 * NS;
    L.cur_off = o;
    o += 8 * NS;
    L.next_off = o;
    o += 8 * NS;
    L.tstart_off = o;
    o += 8 * NS;
    L.lbase_off = o;
    o += 4 * NS;
    L.lstride_off = o;
    o += 4 * NS;
    o = align_up(o, 8);
    L.tb_off = o;
    o += 16;
    L.misc_off = o; // next tile index of a persistent workgroup
    o += 16;
    o = align_up(o, 16);
    L.extra_off = o;
    o += (L.rec_end / L.rec_bytes) * extra_per_rec;
    L.total = align_up(o, 16);
    return L;
  }

  // Workgroup size of a tile kernel: one lane per pair, whole wavefronts.
  inline unsigned tile_block_threads(int R, int C)
  {
    const int lanes = R * C;
    return static_cast<unsigned>(lanes <= 64 ? 64 : (lanes + 63) / 64 * 64);
  }

  // ---------------------------------------------------------------------
  // Phase staging shared by the forward and backward tile kernels.
  // ---------------------------------------------------------------------
  template <typename TT, typename TV>
  struct TileLds
  {
    unsigned char* smem;
    LdsLayout lay;
    int64_t* base;   // global record offset of each slot's PCF
    int64_t* m;      // records of each slot's PCF (0 for an unused slot)
    int64_t* cur;    // first record staged this phase
    int64_t* next;   // first record starting at or after this phase's end
    TT* tstart;      // start time of record `next` (phase bookkeeping)
    uint32_t* lbase; // LDS byte address of record 0 of the slot's phase window
    uint32_t* lstride;
    typename Lim<TT>::Bits* tb;

    __device__ TileLds(unsigned char* s, const LdsLayout& L) : smem(s), lay(L)
    {
      base = reinterpret_cast<int64_t*>(s + L.base_off);
      m = reinterpret_cast<int64_t*>(s + L.m_off);
      cur = reinterpret_cast<int64_t*>(s + L.cur_off);
      next = reinterpret_cast<int64_t*>(s + L.next_off);
      tstart = reinterpret_cast<TT*>(s + L.tstart_off);
      lbase = reinterpret_cast<uint32_t*>(s + L.lbase_off);
      lstride = reinterpret_cast<uint32_t*>(s + L.lstride_off);
      tb = reinterpret_cast<typename Lim<TT>::Bits*>(s + L.tb_off);
    }

    __device__ __forceinline__ Rec<TT, TV>& rec(uint32_t addr) const
    {
      return *reinterpret_cast<Rec<TT, TV>*>(smem + addr);
    }
  };

  // Tile start: slot s < R is row PCF r0 + s, slot R + c is column c0 + c.
  template <typename TT, typename TV>
  __device__ void tile_begin(const TileLds<TT, TV>& S, const FwdParams& P, const Tile& tl,
                             const Problem& pr, int tid)
  {
    using L = Lim<TT>;
    const int NS = P.R + P.C;
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
        S.base[tid] = b;
        S.m[tid] = off[g + 1] - b;
      }
      else
      {
        S.base[tid] = 0;
        S.m[tid] = 0;
      }
      S.next[tid] = 0;
      S.tstart[tid] = TT(0);
      const uint32_t rb = static_cast<uint32_t>(S.lay.rec_bytes);
      S.lbase[tid] = static_cast<uint32_t>(S.lay.rows_off) +
                     static_cast<uint32_t>(tid * slot_recs(P.cap)) * rb;
      S.lstride[tid] = rb;
    }
    if (tid == 0)
    {
      S.tb[0] = L::bits(L::inf());
      S.tb[1] = L::bits(L::inf());
    }
  }

  // One 16-byte global -> LDS load (global_load_lds_dwordx4). lds_base is the
  // wavefront-uniform LDS address of lane 0; lane i's record lands at
  // lds_base + 16 i (lanes that do not call leave their slot untouched).
  // The builtin exists only in the device compilation.
  __device__ __forceinline__ void lds_dma16(const void* src, unsigned char* lds_base)
  {
#if __HIP_DEVICE_COMPILE__
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_base, 16, 0, 0);
#else
    (void)src;
    (void)lds_base;
#endif
  }

  // A value every lane of the wavefront holds (read from one LDS address),
  // moved to a scalar register so that the address arithmetic on it is
  // scalar.
  __device__ __forceinline__ int64_t wave_uniform(int64_t x)
  {
    const uint64_t u = static_cast<uint64_t>(x);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(u)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
  }
  __device__ __forceinline__ uint32_t wave_uniform(uint32_t x)
  {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
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

  // First half of a phase (before the barrier): each slot proposes the phase
  // end (start time of its (cap+1)-th unconsumed record), and the workgroup
  // copies up to cap records of every slot into LDS, one wavefront per slot
  // (a slot's window is contiguous in global memory and in LDS, so each lane
  // moves whole 16- or 8-byte records). With LEAD = 1 the window also holds
  // the record before the cursor (the record a walk is inside when the phase
  // starts) at local index 0, and record cur + j at local j + 1.
  template <typename TT, typename TV, int LEAD = 0>
  __device__ void phase_stage(const TileLds<TT, TV>& S, const FwdParams& P, int phase, int tid,
                              int nthreads)
  {
    using L = Lim<TT>;
    using RecT = Rec<TT, TV>;
    const int NS = P.R + P.C;
    const int cap = P.cap;
    const RecT* __restrict__ g_r = static_cast<const RecT*>(P.r_rec);
    const RecT* __restrict__ g_c = static_cast<const RecT*>(P.c_rec);
    if ((tid & ~63) < NS)
    {
      // the proposals of a wavefront's slot threads are reduced in registers
      // (times are >= 0, so their bit patterns order as the times do) and
      // entered with one LDS atomic per wavefront
      typename L::Bits tbs = L::bits(L::inf());
      if (tid < NS)
      {
        const int64_t c = S.next[tid];
        S.cur[tid] = c;
        if (S.m[tid] - c > cap)
          tbs = L::bits((tid < P.R ? g_r : g_c)[S.base[tid] + c + cap - 1].t);
      }
      tbs = wave_min(tbs);
      if ((tid & 63) == 0)
        atomicMin(&S.tb[phase & 1], tbs);
    }
    const int lane = tid & 63;
    const int nwaves = nthreads >> 6;
    if constexpr (sizeof(RecT) == 16)
    {
      // Direct global -> LDS loads (global_load_lds_dwordx4): lane i of a
      // 64-record chunk lands at chunk base + 16 i, which is exactly the
      // window layout, so records bypass registers and every slot's loads
      // are in flight together; one vmcnt wait below covers them. The slot
      // tables are read per lane (one broadcast LDS read each) and the
      // window address comes from the slot index, so a slot costs a few
      // vector operations and no readfirstlane round trips.
      const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
      const uint32_t slot_bytes = static_cast<uint32_t>(slot_recs(cap)) * 16u;
      for (int s = w; s < NS; s += nwaves)
      {
        const int64_t m = S.m[s];
        const int64_t first = S.next[s] - LEAD; // window record 0
        // window records j in [jlo, jhi): record first + j exists
        const int jlo = first < 0 ? 1 : 0;
        const int jhi = static_cast<int>(m - first < cap + LEAD ? m - first : cap + LEAD);
        const RecT* src = (s < P.R ? g_r : g_c) + (S.base[s] + first);
        const uint32_t dst = static_cast<uint32_t>(S.lay.rows_off) + static_cast<uint32_t>(s) * slot_bytes;
        for (int c0 = 0; c0 < cap + LEAD; c0 += 64)
        {
          const int j = c0 + lane;
          if (j >= jlo && j < jhi)
            lds_dma16(src + j, S.smem + dst + 16u * static_cast<uint32_t>(c0));
        }
      }
    }
    else
    {
      for (int s = tid >> 6; s < NS; s += nwaves)
      {
        const int64_t m = wave_uniform(S.m[s]);
        const int64_t first = wave_uniform(S.next[s]) - LEAD; // window record 0
        const int jlo = first < 0 ? 1 : 0;
        const int jhi = static_cast<int>(m - first < cap + LEAD ? m - first : cap + LEAD);
        const RecT* src = (s < P.R ? g_r : g_c) + wave_uniform(S.base[s]);
        const uint32_t dst = wave_uniform(S.lbase[s]);
        for (int j = lane; j < jhi; j += 64)
          if (j >= jlo)
            S.rec(dst + static_cast<uint32_t>(j) * static_cast<uint32_t>(sizeof(RecT))) = src[first + j];
      }
    }
    if constexpr (sizeof(RecT) == 16)
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (tid == 0)
      S.tb[(phase + 1) & 1] = L::bits(L::inf());
  }

  // After the barrier: the slot threads find the first staged record that
  // starts at or after Tb (binary search) -- the next phase's cursor.
  template <typename TT, typename TV, int LEAD = 0>
  __device__ void phase_advance_cursor(const TileLds<TT, TV>& S, const FwdParams& P, TT Tb, int tid)
  {
    const int NS = P.R + P.C;
    if (tid < NS)
    {
      const int64_t c = S.cur[tid];
      const int64_t rem = S.m[tid] - c;
      const int lim = rem < P.cap ? static_cast<int>(rem) : P.cap;
      const uint32_t lb = S.lbase[tid], ls = S.lstride[tid];
      int lo = 0, hi = lim;
      while (lo < hi)
      {
        const int mid = (lo + hi) >> 1;
        const TT ts = mid == 0 ? S.tstart[tid] : S.rec(lb + static_cast<uint32_t>(mid - 1 + LEAD) * ls).t;
        if (ts < Tb)
          lo = mid + 1;
        else
          hi = mid;
      }
      S.next[tid] = c + lo;
      if (lo > 0)
        S.tstart[tid] = S.rec(lb + static_cast<uint32_t>(lo - 1 + LEAD) * ls).t;
    }
  }
}
