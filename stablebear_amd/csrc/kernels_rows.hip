// Row reductions of a pairwise output already in device memory: per row i
// the sum (and so the mean), minimum and maximum, and the k nearest columns
// (smallest values, ties to the lower column index). Generalises the fused
// row means of the cross-distance path to any reduction a caller of the
// distance matrix applies to its rows (D.mean(axis=1) in the reference's
// benchmark of that workload; k-nearest neighbours).
//
// Matrices: dense row-major n_rows x n_cols (cdist), or the compressed lower
// triangle of a distance matrix (distance_matrix.hpp:120-146: D(i, j), i > j,
// at i(i-1)/2 + j) whose row i is every j != i: the j < i part is contiguous,
// the j > i part one element per later row (consecutive rows run together, so
// those lines are shared in L2).
//
// A workgroup owns a row; each thread scans the row with a stride of the
// workgroup size, so a wavefront's loads are coalesced. The k nearest: a
// first pass bounds the row's k-th smallest value by the k-th smallest of
// the threads' minima; in the second pass each thread keeps its k smallest
// (value, column) pairs below that bound sorted in registers, then the
// workgroup's lists merge pairwise in LDS, halving the list count per step.
// HBM-bound: the sums read every element once, the k nearest twice.
#include "sbamd_internal.hpp"

#include <cfloat>

namespace sbamd
{
  namespace
  {
    constexpr int ROWS_THREADS = 256;

    template <typename T>
    struct RowView
    {
      const T* d;
      int64_t n_rows, n_cols;
      int32_t compressed;

      __device__ __forceinline__ int64_t count(int64_t i) const { return compressed ? n_rows - 1 : n_cols; }
      // column of the row's c-th element (the diagonal skipped)
      __device__ __forceinline__ int64_t col(int64_t i, int64_t c) const
      {
        return compressed ? (c < i ? c : c + 1) : c;
      }
      __device__ __forceinline__ T at(int64_t i, int64_t j) const
      {
        if (!compressed)
          return d[i * n_cols + j];
        return j < i ? d[i * (i - 1) / 2 + j] : d[j * (j - 1) / 2 + i];
      }
    };

    // (value, column) order: value, then the lower column
    template <typename T>
    __device__ __forceinline__ bool before(T a, int32_t ia, T b, int32_t ib)
    {
      return a < b || (a == b && ia < ib);
    }
  }

  template <typename T>
  __global__ __launch_bounds__(ROWS_THREADS) void rows_reduce_kernel(RowView<T> V, double* sum, double* mn,
                                                                     double* mx)
  {
    __shared__ double s_sum[ROWS_THREADS / 64], s_min[ROWS_THREADS / 64], s_max[ROWS_THREADS / 64];
    const int64_t i = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t cnt = V.count(i);
    double a = 0.0, lo = __builtin_inf(), hi = -__builtin_inf();
    for (int64_t c = tid; c < cnt; c += ROWS_THREADS)
    {
      const double x = static_cast<double>(V.at(i, V.col(i, c)));
      a += x;
      lo = fmin(lo, x);
      hi = fmax(hi, x);
    }
    for (int o = 32; o > 0; o >>= 1)
    {
      a += __shfl_xor(a, o);
      lo = fmin(lo, __shfl_xor(lo, o));
      hi = fmax(hi, __shfl_xor(hi, o));
    }
    if ((tid & 63) == 0)
    {
      s_sum[tid >> 6] = a;
      s_min[tid >> 6] = lo;
      s_max[tid >> 6] = hi;
    }
    __syncthreads();
    if (tid == 0)
    {
      for (int w = 1; w < ROWS_THREADS / 64; ++w)
      {
        a += s_sum[w];
        lo = fmin(lo, s_min[w]);
        hi = fmax(hi, s_max[w]);
      }
      if (sum)
        sum[i] = a;
      if (mn)
        mn[i] = lo;
      if (mx)
        mx[i] = hi;
    }
  }

  // Threads of a k-nearest workgroup: one wavefront per row. (Measured on a
  // 50,000 x 50,000 matrix, k = 16: 9.4 ms with 64 threads, 10.3 ms with
  // 256, whose bound step and merge tree cost more than the wider scan
  // gains.)
  template <int K>
  constexpr int knn_threads()
  {
    return 64;
  }

  template <typename T, int K>
  __global__ __launch_bounds__(knn_threads<K>()) void rows_knn_kernel(RowView<T> V, int32_t k, T* out_v,
                                                                       int64_t* out_i)
  {
    constexpr int NT = knn_threads<K>();
    __shared__ T s_v[NT * K];
    __shared__ int32_t s_i[NT * K];
    const int64_t i = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t cnt = V.count(i);
    const T inf = static_cast<T>(__builtin_inf());
    T tv[K];
    int32_t ti[K];
#pragma unroll
    for (int p = 0; p < K; ++p)
    {
      tv[p] = inf;
      ti[p] = INT32_MAX; // no column yet
    }
    // Pass 1: a bound on the row's K-th smallest value. The K smallest of
    // the threads' minima are K of the row's values, so the row's K-th
    // smallest is at most the K-th of those minima; pass 2 inserts only
    // values up to that bound (a handful per row instead of a few per
    // thread and K).
    T tmin = inf;
    for (int64_t c = tid; c < cnt; c += NT)
      tmin = min(tmin, V.at(i, V.col(i, c)));
    s_v[tid] = tmin;
    __syncthreads();
    int32_t rank = 0; // position of this thread's minimum, ties by thread
    for (int u = 0; u < NT; ++u)
    {
      const T o = s_v[u];
      rank += o < tmin || (o == tmin && u < tid);
    }
    __syncthreads();
    if (rank == K - 1)
      s_v[0] = tmin;
    __syncthreads();
    const T bound = s_v[0];
    __syncthreads();
    for (int64_t c = tid; c < cnt; c += NT)
    {
      const int32_t j = static_cast<int32_t>(V.col(i, c));
      const T x = V.at(i, j);
      if (!(x <= bound) || !before(x, j, tv[K - 1], ti[K - 1]))
        continue;
      // insert (x, j) into the ascending list; the last entry drops out
#pragma unroll
      for (int p = K - 1; p >= 0; --p)
      {
        const bool shift = p > 0 && before(x, j, tv[p - 1], ti[p - 1]);
        const bool here = !shift && before(x, j, tv[p], ti[p]);
        if (shift)
        {
          tv[p] = tv[p - 1];
          ti[p] = ti[p - 1];
        }
        else if (here)
        {
          tv[p] = x;
          ti[p] = j;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < K; ++p)
    {
      s_v[tid * K + p] = tv[p];
      s_i[tid * K + p] = ti[p];
    }
    __syncthreads();
    // pairwise merges of sorted lists: list t absorbs list t + h
    for (int h = NT / 2; h > 0; h >>= 1)
    {
      if (tid < h)
      {
        const T* av = s_v + tid * K;
        const int32_t* ai = s_i + tid * K;
        const T* bv = s_v + (tid + h) * K;
        const int32_t* bi = s_i + (tid + h) * K;
        int pa = 0, pb = 0; // pa + pb = p < K: both stay inside their lists
#pragma unroll
        for (int p = 0; p < K; ++p)
        {
          if (before(bv[pb], bi[pb], av[pa], ai[pa]))
          {
            tv[p] = bv[pb];
            ti[p] = bi[pb];
            ++pb;
          }
          else
          {
            tv[p] = av[pa];
            ti[p] = ai[pa];
            ++pa;
          }
        }
      }
      __syncthreads();
      if (tid < h)
      {
#pragma unroll
        for (int p = 0; p < K; ++p)
        {
          s_v[tid * K + p] = tv[p];
          s_i[tid * K + p] = ti[p];
        }
      }
      __syncthreads();
    }
    for (int p = tid; p < k; p += NT)
    {
      const bool real = s_i[p] != INT32_MAX;
      out_v[i * k + p] = real ? s_v[p] : inf;
      out_i[i * k + p] = real ? static_cast<int64_t>(s_i[p]) : -1;
    }
  }

  template <typename T>
  static RowView<T> row_view(const RowsParams& P)
  {
    RowView<T> V;
    V.d = static_cast<const T*>(P.d);
    V.n_rows = P.n_rows;
    V.n_cols = P.n_cols;
    V.compressed = P.compressed;
    return V;
  }

  void launch_rows_reduce(const RowsParams& P, hipStream_t stream)
  {
    if (P.n_rows <= 0)
      return;
    const dim3 grid(static_cast<unsigned>(P.n_rows)), block(ROWS_THREADS);
    if (P.f64)
      hipLaunchKernelGGL(rows_reduce_kernel<double>, grid, block, 0, stream, row_view<double>(P), P.sum, P.min,
                         P.max);
    else
      hipLaunchKernelGGL(rows_reduce_kernel<float>, grid, block, 0, stream, row_view<float>(P), P.sum, P.min,
                         P.max);
    SBAMD_HIP(hipGetLastError());
  }

  template <typename T>
  static void knn_dispatch(const RowsParams& P, hipStream_t stream)
  {
    const dim3 grid(static_cast<unsigned>(P.n_rows));
    const RowView<T> V = row_view<T>(P);
    T* ov = static_cast<T*>(P.knn_v);
    if (P.k <= 1)
      hipLaunchKernelGGL((rows_knn_kernel<T, 1>), grid, dim3(knn_threads<1>()), 0, stream, V, P.k, ov, P.knn_i);
    else if (P.k <= 4)
      hipLaunchKernelGGL((rows_knn_kernel<T, 4>), grid, dim3(knn_threads<4>()), 0, stream, V, P.k, ov, P.knn_i);
    else if (P.k <= 8)
      hipLaunchKernelGGL((rows_knn_kernel<T, 8>), grid, dim3(knn_threads<8>()), 0, stream, V, P.k, ov, P.knn_i);
    else if (P.k <= 16)
      hipLaunchKernelGGL((rows_knn_kernel<T, 16>), grid, dim3(knn_threads<16>()), 0, stream, V, P.k, ov,
                         P.knn_i);
    else
      hipLaunchKernelGGL((rows_knn_kernel<T, 32>), grid, dim3(knn_threads<32>()), 0, stream, V, P.k, ov,
                         P.knn_i);
    SBAMD_HIP(hipGetLastError());
  }

  void launch_rows_knn(const RowsParams& P, hipStream_t stream)
  {
    if (P.n_rows <= 0)
      return;
    if (P.f64)
      knn_dispatch<double>(P, stream);
    else
      knn_dispatch<float>(P, stream);
  }
}
