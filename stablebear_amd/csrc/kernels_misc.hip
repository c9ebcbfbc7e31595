// Small kernels: the general bounded walk that lists rectangles, value
// refresh of a staged set, and a zero fill (used instead of hipMemsetAsync so
// that captured graphs replay correctly).
#include "walk_common.hpp"

namespace sbamd
{
  // Last index with t <= a, clamped to 0 (iterate_rectangles.hpp:23-35).
  template <typename T>
  __device__ int64_t start_index(const T* t, int64_t m, T a)
  {
    int64_t lo = 0, hi = m;
    while (lo < hi)
    {
      const int64_t mid = lo + (hi - lo) / 2;
      if (a < t[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
    return lo == 0 ? 0 : lo - 1;
  }

  // The bounded walk of iterate_rectangles.hpp:37-104 on raw points (not the
  // dropped-duplicate records), one thread: it reproduces every rectangle the
  // reference emits, zero-length ones included.
  template <typename T>
  __global__ void rectangles_kernel(const T* ft, const T* fvals, int64_t mf, const T* gt,
                                    const T* gvals, int64_t mg, T a, T b, T* out, int64_t cap,
                                    int64_t* count)
  {
    if (blockIdx.x != 0 || threadIdx.x != 0)
      return;
    int64_t fi = start_index(ft, mf, a);
    int64_t gi = start_index(gt, mg, a);
    T t = a;
    T tprev = a;
    int64_t n = 0;
    while (t < b)
    {
      tprev = t;
      const T fv = fvals[fi];
      const T gv = gvals[gi];
      const bool fnext = fi + 1 < mf;
      const bool gnext = gi + 1 < mg;
      bool done = false;
      if (fnext && gnext)
      {
        const T delta = ft[fi + 1] - gt[gi + 1];
        if (delta <= 0)
          ++fi;
        if (delta >= 0)
          ++gi;
      }
      else if (fnext)
        ++fi;
      else if (gnext)
        ++gi;
      else
        done = true;
      T r;
      if (done)
        r = b;
      else
      {
        const T mx = ft[fi] > gt[gi] ? ft[fi] : gt[gi];
        r = mx < b ? mx : b;
        t = r;
      }
      if (n < cap)
      {
        out[4 * n + 0] = tprev;
        out[4 * n + 1] = r;
        out[4 * n + 2] = fv;
        out[4 * n + 3] = gv;
      }
      ++n;
      if (done)
        break;
    }
    *count = n;
  }

  void launch_rectangles(const double* ft, const double* fv, int64_t mf, const double* gt,
                         const double* gv, int64_t mg, double a, double b, double* out,
                         int64_t cap, int64_t* count, hipStream_t stream)
  {
    hipLaunchKernelGGL((rectangles_kernel<double>), dim3(1), dim3(64), 0, stream, ft, fv, mf, gt, gv,
                       mg, a, b, out, cap, count);
    SBAMD_HIP(hipGetLastError());
  }

  void launch_rectangles_f32(const float* ft, const float* fv, int64_t mf, const float* gt,
                             const float* gv, int64_t mg, float a, float b, float* out,
                             int64_t cap, int64_t* count, hipStream_t stream)
  {
    hipLaunchKernelGGL((rectangles_kernel<float>), dim3(1), dim3(64), 0, stream, ft, fv, mf, gt, gv,
                       mg, a, b, out, cap, count);
    SBAMD_HIP(hipGetLastError());
  }

  // rec: record array of rec_bytes-byte records, value at byte rec_bytes / 2.
  template <typename TI, typename TO>
  __global__ void gather_values_kernel(const TI* __restrict__ v, const int64_t* __restrict__ src,
                                       int64_t records, unsigned char* __restrict__ rec,
                                       int rec_bytes)
  {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < records)
      *reinterpret_cast<TO*>(rec + i * rec_bytes + rec_bytes / 2) = static_cast<TO>(v[src[i]]);
  }

  void launch_update_values(const void* v, const int64_t* src, int64_t records, void* rec,
                            int rec_bytes, int in_f64, int out_f64, hipStream_t stream)
  {
    if (records <= 0)
      return;
    const unsigned grid = static_cast<unsigned>((records + 255) / 256);
    unsigned char* r = static_cast<unsigned char*>(rec);
    if (in_f64 && out_f64)
      hipLaunchKernelGGL((gather_values_kernel<double, double>), dim3(grid), dim3(256), 0, stream,
                         static_cast<const double*>(v), src, records, r, rec_bytes);
    else if (in_f64)
      hipLaunchKernelGGL((gather_values_kernel<double, float>), dim3(grid), dim3(256), 0, stream,
                         static_cast<const double*>(v), src, records, r, rec_bytes);
    else if (out_f64)
      hipLaunchKernelGGL((gather_values_kernel<float, double>), dim3(grid), dim3(256), 0, stream,
                         static_cast<const float*>(v), src, records, r, rec_bytes);
    else
      hipLaunchKernelGGL((gather_values_kernel<float, float>), dim3(grid), dim3(256), 0, stream,
                         static_cast<const float*>(v), src, records, r, rec_bytes);
    SBAMD_HIP(hipGetLastError());
  }

  __global__ void fill_zero_kernel(uint32_t* __restrict__ p, size_t words)
  {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
      p[i] = 0u;
  }

  void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream)
  {
    if (bytes == 0)
      return;
    if (bytes % 4 != 0)
      fail(SBAMD_ERR_INVALID, "fill size must be a multiple of 4 bytes");
    const size_t words = bytes / 4;
    size_t blocks = (words + 255) / 256;
    if (blocks > 65536)
      blocks = 65536;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                       static_cast<uint32_t*>(ptr), words);
    SBAMD_HIP(hipGetLastError());
  }

  // Several buffers cleared by one launch (a fill kernel, not a memset, so
  // that a captured graph replays it correctly).
  struct FillList
  {
    uint32_t* ptr[FILL_MAX];
    size_t words[FILL_MAX];
    int n;
  };

  __global__ void fill_zero_multi_kernel(FillList L)
  {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t i0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int b = 0; b < L.n; ++b)
      for (size_t i = i0; i < L.words[b]; i += stride)
        L.ptr[b][i] = 0u;
  }

  void launch_fill_zero_multi(void* const* ptrs, const size_t* bytes, int n, hipStream_t stream)
  {
    if (n < 0 || n > FILL_MAX)
      fail(SBAMD_ERR_INVALID, "too many buffers for one fill");
    FillList L{};
    size_t most = 0;
    for (int b = 0; b < n; ++b)
    {
      if (bytes[b] % 4 != 0)
        fail(SBAMD_ERR_INVALID, "fill size must be a multiple of 4 bytes");
      if (bytes[b] == 0)
        continue;
      L.ptr[L.n] = static_cast<uint32_t*>(ptrs[b]);
      L.words[L.n] = bytes[b] / 4;
      most = L.words[L.n] > most ? L.words[L.n] : most;
      ++L.n;
    }
    if (L.n == 0)
      return;
    size_t blocks = (most + 255) / 256;
    if (blocks > 65536)
      blocks = 65536;
    hipLaunchKernelGGL(fill_zero_multi_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, L);
    SBAMD_HIP(hipGetLastError());
  }

  // Thread per record of a subset set: the subset's PCF p is PCF sel[p] of
  // the full set (sorted positions, kept in the full set's order), so record
  // k of p is record roff[sel[p]] + k of the full set.
  template <typename W>
  __global__ void gather_subset_kernel(const int64_t* __restrict__ roff, const W* __restrict__ rec,
                                       const int64_t* __restrict__ sel, const int64_t* __restrict__ droff,
                                       int64_t n, int64_t records, W* __restrict__ drec)
  {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= records)
      return;
    int64_t lo = 0, hi = n; // droff[lo] <= r < droff[lo + 1]
    while (hi - lo > 1)
    {
      const int64_t mid = (lo + hi) >> 1;
      if (droff[mid] <= r)
        lo = mid;
      else
        hi = mid;
    }
    drec[r] = rec[roff[sel[lo]] + (r - droff[lo])];
  }

  void launch_gather_subset(const int64_t* roff, const void* rec, int rec_bytes, const int64_t* sel,
                            const int64_t* droff, int64_t n, int64_t records, void* drec, hipStream_t stream)
  {
    if (records <= 0)
      return;
    const unsigned grid = static_cast<unsigned>((records + 255) / 256);
    if (rec_bytes == 16)
      hipLaunchKernelGGL(gather_subset_kernel<uint4>, dim3(grid), dim3(256), 0, stream, roff,
                         static_cast<const uint4*>(rec), sel, droff, n, records, static_cast<uint4*>(drec));
    else
      hipLaunchKernelGGL(gather_subset_kernel<uint2>, dim3(grid), dim3(256), 0, stream, roff,
                         static_cast<const uint2*>(rec), sel, droff, n, records, static_cast<uint2*>(drec));
    SBAMD_HIP(hipGetLastError());
  }
}
