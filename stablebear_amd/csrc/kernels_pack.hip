// Device-side packing of a CSR PCF set into the record layout, the same
// records bit for bit as the host packer (pack.cpp): the reference's
// breakpoint rules (src/python/pypcf_support.hpp:44-87) checked on every
// point, zero-length pieces dropped, record end times and values converted
// to the set's types. Replaces the host flattening of the reference
// (init_pcf_data, include/sbear/cuda/cuda_pcf_data_manager.cuh:13-23;
// OffsetDataManager::init, offset_data_manager.hpp:42-68), which copies every
// point on one host thread.
//
// A wavefront owns a PCF; a point is kept when the next point has a later
// time or it is the last one, and its record index is the number of kept
// points before it. Pass 1 only counts and walks 256 points per step, four
// consecutive ones per lane with all of the lane's loads in flight together;
// pass 2 walks 64 points per step, one per lane, so its record stores are
// contiguous across the wavefront, and ranks with a ballot and mbcnt. Inputs
// are strided views, so the same kernels read SoA arrays (stride = element
// size) and AoS {t, v} point arrays (stride = 2 x element size); with
// per-PCF byte bases they also read a file image as it lies, where each
// PCF's points follow its own point count.
#include "sbamd_internal.hpp"

#include <cmath>

namespace sbamd
{
  namespace
  {
    constexpr int PACK_WAVES = 4; // wavefronts (PCFs) per workgroup
    constexpr int PACK_RUN = 4;   // consecutive points per lane and step

    template <typename T>
    __device__ __forceinline__ double load_at(const unsigned char* base, int64_t k, int64_t stride)
    {
      return static_cast<double>(*reinterpret_cast<const T*>(base + k * stride));
    }
  }

  // Pass 1: records per PCF and rule violations (the lowest offending PCF
  // index goes to *bad).
  template <typename T>
  __global__ __launch_bounds__(64 * PACK_WAVES) void pack_count_kernel(PackParams P)
  {
    const int lane = threadIdx.x & 63;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * PACK_WAVES + (threadIdx.x >> 6);
    if (i >= P.n)
      return;
    const int64_t b = P.off[i], e = P.off[i + 1];
    if (e == b)
    {
      if (lane == 0)
        P.m[i] = 1; // the zero PCF: one record
      return;
    }
    const unsigned char* const tp = P.t + (P.pbase ? P.pbase[i] : 0);
    int64_t cnt = 0;
    bool bad = false;
    // PACK_RUN consecutive points per lane and step: the lane's loads
    // (its points and their two neighbours) are all in flight together
    for (int64_t c = b; c < e; c += 64 * PACK_RUN)
    {
      const int64_t k0 = c + PACK_RUN * lane;
      double tt[PACK_RUN + 2]; // times of points k0 - 1 .. k0 + PACK_RUN
#pragma unroll
      for (int u = 0; u < PACK_RUN + 2; ++u)
      {
        const int64_t k = k0 - 1 + u;
        tt[u] = k >= b && k < e ? load_at<T>(tp, k, P.t_stride) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < PACK_RUN; ++u)
      {
        const int64_t k = k0 + u;
        const bool in = k < e;
        const double t = tt[u + 1];
        if (in && !isfinite(t))
          bad = true;
        if (in && k == b && t != 0.0)
          bad = true;
        if (in && k > b && t < tt[u])
          bad = true;
        const bool keep = in && (k + 1 == e || t < tt[u + 2]);
        cnt += __popcll(__ballot(keep));
      }
    }
    if (__ballot(bad) && lane == 0)
      atomicMin(P.bad, static_cast<unsigned long long>(i));
    if (lane == 0)
      P.m[i] = cnt;
  }

  // Pass 2: the records of PCF i at its sorted position.
  template <typename T>
  __global__ __launch_bounds__(64 * PACK_WAVES) void pack_records_kernel(PackParams P)
  {
    const int lane = threadIdx.x & 63;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * PACK_WAVES + (threadIdx.x >> 6);
    if (i >= P.n)
      return;
    const int64_t b = P.off[i], e = P.off[i + 1];
    const int64_t r0 = P.roff[P.qinv[i]];
    const int64_t pb = P.pbase ? P.pbase[i] : 0;
    const unsigned char* const tp = P.t + pb;
    const unsigned char* const vp = P.v + pb;
    const int rb = P.tt_f64 ? 16 : 8;
    unsigned char* const rec = static_cast<unsigned char*>(P.rec);
    // record r = {end time, value} in the set's types; the value sits at
    // byte rb / 2 (FP32 values pad to 16 bytes behind FP64 times)
    auto put = [&](int64_t r, double tend, double v) {
      unsigned char* const p = rec + r * rb;
      if (P.tt_f64)
        *reinterpret_cast<double*>(p) = tend;
      else
        *reinterpret_cast<float*>(p) = static_cast<float>(tend);
      if (P.tv_f64)
        *reinterpret_cast<double*>(p + rb / 2) = v;
      else if (P.tt_f64)
      {
        // FP32 value and 4 zero bytes of padding, as one 8-byte store
        const float2 vz = make_float2(static_cast<float>(v), 0.0f);
        *reinterpret_cast<float2*>(p + rb / 2) = vz;
      }
      else
        *reinterpret_cast<float*>(p + rb / 2) = static_cast<float>(v);
    };
    if (e == b)
    {
      if (lane == 0)
      {
        put(r0, __builtin_inf(), 0.0);
        P.src[r0] = -1;
      }
      return;
    }
    int64_t base = 0;
    for (int64_t c = b; c < e; c += 64)
    {
      const int64_t k = c + lane;
      const bool in = k < e;
      const bool last = k + 1 == e;
      double tend = __builtin_inf();
      bool keep = false;
      if (in)
      {
        const double t = load_at<T>(tp, k, P.t_stride);
        if (!last)
        {
          const double tn = load_at<T>(tp, k + 1, P.t_stride);
          keep = t < tn;
          tend = tn + 0.0; // + 0.0 maps -0.0 to +0.0
        }
        else
          keep = true;
      }
      const unsigned long long mask = __ballot(keep);
      if (keep)
      {
        const int64_t r = r0 + base +
                          __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
        put(r, tend, load_at<T>(vp, k, P.v_stride));
        P.src[r] = k;
      }
      base += __popcll(mask);
    }
  }

  static unsigned pack_blocks(int64_t n)
  {
    return static_cast<unsigned>((n + PACK_WAVES - 1) / PACK_WAVES);
  }

  void launch_pack_count(const PackParams& P, int f64, hipStream_t stream)
  {
    if (P.n <= 0)
      return;
    if (f64)
      hipLaunchKernelGGL(pack_count_kernel<double>, dim3(pack_blocks(P.n)), dim3(64 * PACK_WAVES), 0, stream, P);
    else
      hipLaunchKernelGGL(pack_count_kernel<float>, dim3(pack_blocks(P.n)), dim3(64 * PACK_WAVES), 0, stream, P);
    SBAMD_HIP(hipGetLastError());
  }

  void launch_pack_records(const PackParams& P, int f64, hipStream_t stream)
  {
    if (P.n <= 0)
      return;
    if (f64)
      hipLaunchKernelGGL(pack_records_kernel<double>, dim3(pack_blocks(P.n)), dim3(64 * PACK_WAVES), 0, stream, P);
    else
      hipLaunchKernelGGL(pack_records_kernel<float>, dim3(pack_blocks(P.n)), dim3(64 * PACK_WAVES), 0, stream, P);
    SBAMD_HIP(hipGetLastError());
  }
}
