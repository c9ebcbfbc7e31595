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

  // Byte layout of the dynamic LDS block of a tile kernel.
  struct LdsLayout
  {
    int tend_off, val_off, base_off, m_off, cur_off, next_off, tstart_off, tb_off, extra_off, total;
  };

  __host__ __device__ inline int align_up(int x, int a) { return (x + a - 1) / a * a; }

  // Workgroup size of a tile kernel: one lane per pair, whole wavefronts.
  inline unsigned tile_block_threads(int R, int C)
  {
    const int lanes = R * C;
    return static_cast<unsigned>(lanes <= 64 ? 64 : (lanes + 63) / 64 * 64);
  }

  // extra_per_slot_rec: bytes per staged record for backward gradient slots.
  __host__ __device__ inline LdsLayout lds_layout(int tt_bytes, int tv_bytes, int NS, int stride, int cap,
                                                  int extra_per_rec)
  {
    LdsLayout L;
    int o = 0;
    L.tend_off = o;
    o += tt_bytes * cap * stride;
    o = align_up(o, 8);
    L.val_off = o;
    o += tv_bytes * cap * stride;
    o = align_up(o, 8);
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
    L.tb_off = o;
    o += 16;
    o = align_up(o, 8);
    L.extra_off = o;
    o += extra_per_rec * cap * stride;
    L.total = align_up(o, 16);
    return L;
  }
}
