// C entry points over the reference's own header-only walk, compiled from the
// sources where they lie (see oracle/Makefile). TEST INFRASTRUCTURE ONLY: used
// to validate the C oracle and as the CPU baseline; never part of the product.
//
// Compiled reference headers:
//   include/sbear/functional/time_point.hpp        (TimePoint, infinity)
//   include/sbear/functional/rectangle.hpp         (Rectangle)
//   include/sbear/algorithms/functional/iterate_rectangles.hpp (the walk)
//   include/sbear/functional/operations.cuh        (L1 / Lp / inner-product ops)
// include/sbear/algorithms/functional/matrix_integrate.hpp needs taskflow (an
// empty submodule here), so its 3-line integrate() accumulation and the
// row-parallel CpuPairwiseIntegrationTask loop (matrix_integrate.hpp:18-29,
// 80-90) and CdistTask loop (py_distance.cpp:116-126) are driven below with
// std::thread and a dynamic row counter.

#include <sbear/functional/time_point.hpp>
#include <sbear/functional/rectangle.hpp>
#include <sbear/algorithms/functional/iterate_rectangles.hpp>
#include <sbear/functional/operations.cuh>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

namespace
{
  template <typename T>
  using Points = std::vector<sb::TimePoint<T, T>>;

  template <typename T>
  Points<T> make_points(const T* t, const T* v, long m)
  {
    Points<T> pts(static_cast<size_t>(m));
    for (long k = 0; k < m; ++k)
    {
      pts[k].t = t[k];
      pts[k].v = v[k];
    }
    return pts;
  }

  template <typename T>
  std::vector<Points<T>> make_set(long n, const int64_t* off, const T* t, const T* v)
  {
    std::vector<Points<T>> set(static_cast<size_t>(n));
    for (long i = 0; i < n; ++i)
      set[i] = make_points(t + off[i], v + off[i], static_cast<long>(off[i + 1] - off[i]));
    return set;
  }

  template <typename T, typename Op>
  T integrate(const Points<T>& f, const Points<T>& g, Op op)
  {
    using rect_t = sb::Rectangle<T, T>;
    T val = T(0);
    sb::iterate_rectangles(f, g, [&val, &op](const rect_t& rect) -> void {
      val += (rect.right - rect.left) * op(rect.f_value, rect.g_value);
    }, T(0), std::numeric_limits<T>::max());
    return val;
  }

  // l1_norm / lp_norm (apply_functional.hpp:63-100, which needs taskflow):
  // their accumulation over the reference's own iterate_segments.
  template <typename T>
  T lp_norm(const Points<T>& f, T p)
  {
    using seg_t = sb::Segment<T, T>;
    T out = T(0);
    if (p == T(1))
    {
      sb::iterate_segments(f.cbegin(), f.cend(), T(0), std::numeric_limits<T>::infinity(),
                           [&out](const seg_t& seg) { out += (seg.right - seg.left) * std::abs(seg.value); });
      return out;
    }
    sb::iterate_segments(f.cbegin(), f.cend(), T(0), std::numeric_limits<T>::infinity(),
                         [&out, p](const seg_t& seg) {
                           out += (seg.right - seg.left) * std::pow(std::abs(seg.value), p);
                         });
    return std::pow(out, T(1) / p);
  }

  template <typename T, typename Op>
  void run_rows(Op op, int kind, const std::vector<Points<T>>& xs, const std::vector<Points<T>>& ys,
                long row_begin, long row_end, T* out, int nthreads)
  {
    std::atomic<long> next{row_begin};
    long n = static_cast<long>(xs.size());
    long n2 = static_cast<long>(ys.size());
    auto worker = [&]() {
      for (;;)
      {
        long i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= row_end)
          break;
        if (kind == 2)
        {
          for (long j = 0; j < n2; ++j)
            out[i * n2 + j] = op(integrate<T>(xs[i], ys[j], op));
          continue;
        }
        long j0 = kind == 1 ? i : i + 1;
        for (long j = j0; j < n; ++j)
        {
          long idx = kind == 1 ? j * (j + 1) / 2 + i : j * (j - 1) / 2 + i;
          out[idx] = op(integrate<T>(xs[i], xs[j], op));
        }
      }
    };
    if (nthreads < 1)
      nthreads = 1;
    std::vector<std::thread> th;
    for (int k = 1; k < nthreads; ++k)
      th.emplace_back(worker);
    worker();
    for (auto& x : th)
      x.join();
  }

  template <typename T>
  int pairwise(int op, T p, int kind, long n, const int64_t* off, const T* t, const T* v,
               long n2, const int64_t* off2, const T* t2, const T* v2,
               long row_begin, long row_end, T* out, int nthreads)
  {
    auto xs = make_set<T>(n, off, t, v);
    std::vector<Points<T>> ys;
    if (kind == 2)
      ys = make_set<T>(n2, off2, t2, v2);
    if (row_end > n)
      row_end = n;
    if (op == 0)
      run_rows<T>(sb::OperationL1Dist<T, T>{}, kind, xs, ys, row_begin, row_end, out, nthreads);
    else if (op == 1)
      run_rows<T>(sb::OperationLpDist<T, T>(p), kind, xs, ys, row_begin, row_end, out, nthreads);
    else
      run_rows<T>(sb::OperationL2InnerProduct<T, T>{}, kind, xs, ys, row_begin, row_end, out, nthreads);
    return 0;
  }

  template <typename T>
  long rects(const T* ft, const T* fv, long mf, const T* gt, const T* gv, long mg, T a, T b,
             T* out, long cap)
  {
    auto f = make_points(ft, fv, mf);
    auto g = make_points(gt, gv, mg);
    long count = 0;
    sb::iterate_rectangles(f, g, [&](const sb::Rectangle<T, T>& r) {
      if (count < cap)
      {
        out[4 * count + 0] = r.left;
        out[4 * count + 1] = r.right;
        out[4 * count + 2] = r.f_value;
        out[4 * count + 3] = r.g_value;
      }
      ++count;
    }, a, b);
    return count;
  }
}

extern "C"
{
  long ref_rects_f64(const double* ft, const double* fv, long mf, const double* gt, const double* gv,
                     long mg, double a, double b, double* out, long cap)
  {
    return rects<double>(ft, fv, mf, gt, gv, mg, a, b, out, cap);
  }

  long ref_rects_f32(const float* ft, const float* fv, long mf, const float* gt, const float* gv,
                     long mg, float a, float b, float* out, long cap)
  {
    return rects<float>(ft, fv, mf, gt, gv, mg, a, b, out, cap);
  }

  int ref_pairwise_f64(int op, double p, int kind, long n, const int64_t* off, const double* t,
                       const double* v, long n2, const int64_t* off2, const double* t2,
                       const double* v2, double* out, int nthreads)
  {
    return pairwise<double>(op, p, kind, n, off, t, v, n2, off2, t2, v2, 0, n, out, nthreads);
  }

  int ref_pairwise_rows_f64(int op, double p, int kind, long n, const int64_t* off, const double* t,
                            const double* v, long n2, const int64_t* off2, const double* t2,
                            const double* v2, long row_begin, long row_end, double* out,
                            int nthreads)
  {
    return pairwise<double>(op, p, kind, n, off, t, v, n2, off2, t2, v2, row_begin, row_end, out,
                            nthreads);
  }

  int ref_pairwise_f32(int op, float p, int kind, long n, const int64_t* off, const float* t,
                       const float* v, long n2, const int64_t* off2, const float* t2,
                       const float* v2, float* out, int nthreads)
  {
    return pairwise<float>(op, p, kind, n, off, t, v, n2, off2, t2, v2, 0, n, out, nthreads);
  }

  void ref_lp_norms_f64(double p, long n, const int64_t* off, const double* t, const double* v,
                        double* out)
  {
    for (long i = 0; i < n; ++i)
      out[i] = lp_norm<double>(make_points(t + off[i], v + off[i], static_cast<long>(off[i + 1] - off[i])), p);
  }

  void ref_lp_norms_f32(float p, long n, const int64_t* off, const float* t, const float* v,
                        float* out)
  {
    for (long i = 0; i < n; ++i)
      out[i] = lp_norm<float>(make_points(t + off[i], v + off[i], static_cast<long>(off[i + 1] - off[i])), p);
  }
}
