/*
 * CPU oracle for the pairwise PCF integration path (pdist / l2_kernel / cdist).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product
 * (stablebear_amd) never links, loads or calls it.
 *
 * It restates, in plain C, the reference algorithm:
 *   - the merge walk            include/sbear/algorithms/functional/iterate_rectangles.hpp:23-104
 *   - integrate()               include/sbear/algorithms/functional/matrix_integrate.hpp:18-29
 *   - the pair ops + finalizers include/sbear/functional/operations.cuh:19-65
 *   - CpuPairwiseIntegrationTask include/sbear/algorithms/functional/matrix_integrate.hpp:35-95
 *   - CdistTask                 src/python/functional/py_distance.cpp:101-136
 *   - output index maps         include/sbear/distance_matrix.hpp:120-146,
 *                               include/sbear/symmetric_matrix.hpp:80-103
 * Accumulation is `val = fma(r - l, op(fv, gv), val)`: the reference builds
 * with -std=gnu++20 -march=x86-64-v3 (CMakeLists.txt:368, 489), where GCC's
 * default -ffp-contract=fast contracts `val += (r - l) * op(...)` into an FMA.
 *
 * Gradients (no reference counterpart): analytic dL/dv and dL/dt of
 * L = sum_pairs G_ij * D_ij (p = 2 distance) or sum_pairs G_ij * K_ij (inner
 * product), following the same walk. Tie convention: when f and g have a
 * breakpoint at the same time, each side's breakpoint sees the mean of the
 * other side's left and right limits (this makes self-pairs exact, and the
 * two sides add up to the derivative of moving the shared time).
 *
 * Pinning (what checks each function): tests/test_oracle.py.
 */
#ifndef SBAMD_ORACLE_H
#define SBAMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OP_L1 = 0, ORC_OP_LP = 1, ORC_OP_DOT = 2 };
enum { ORC_KIND_PDIST = 0, ORC_KIND_KERNEL = 1, ORC_KIND_CDIST = 2 };

/* Rectangles of the walk over [a, b): out holds 4 values per rectangle
 * (left, right, f_value, g_value). Returns the rectangle count (may exceed cap;
 * only the first cap are written). */
long orc_rects_f64(const double* ft, const double* fv, long mf,
                   const double* gt, const double* gv, long mg,
                   double a, double b, double* out, long cap);
long orc_rects_f32(const float* ft, const float* fv, long mf,
                   const float* gt, const float* gv, long mg,
                   float a, float b, float* out, long cap);

/* Raw integral sum_rect (r - l) * op(fv, gv) over [a, b) (before finalize). */
double orc_integrate_f64(int op, double p,
                         const double* ft, const double* fv, long mf,
                         const double* gt, const double* gv, long mg,
                         double a, double b);
float orc_integrate_f32(int op, float p,
                        const float* ft, const float* fv, long mf,
                        const float* gt, const float* gv, long mg,
                        float a, float b);

/* L_p norms of each PCF of a CSR set over [0, MAX) (apply_functional.hpp:63-100). */
void orc_lp_norms_f64(double p, long n, const int64_t* off, const double* t, const double* v,
                      double* out);
void orc_lp_norms_f32(float p, long n, const int64_t* off, const float* t, const float* v,
                      float* out);

double orc_finalize_f64(int op, double p, double x);
float orc_finalize_f32(int op, float p, float x);

/* Whole-matrix drivers, row-parallel over nthreads with a dynamic row counter.
 * kind PDIST: out[j(j-1)/2 + i] for i < j  (n(n-1)/2 entries)
 * kind KERNEL: out[j(j+1)/2 + i] for i <= j (n(n+1)/2 entries)
 * kind CDIST: out[i * n2 + j] (n * n2 entries; second set given by off2/t2/v2)
 * Returns 0 on success. */
int orc_pairwise_f64(int op, double p, int kind,
                     long n, const int64_t* off, const double* t, const double* v,
                     long n2, const int64_t* off2, const double* t2, const double* v2,
                     double* out, int nthreads);
int orc_pairwise_f32(int op, float p, int kind,
                     long n, const int64_t* off, const float* t, const float* v,
                     long n2, const int64_t* off2, const float* t2, const float* v2,
                     float* out, int nthreads);

/* Rows [row_begin, row_end) of a pdist/kernel/cdist matrix only; same layout as
 * above, entries of other rows untouched. Used to time a bounded sample. */
int orc_pairwise_rows_f64(int op, double p, int kind,
                          long n, const int64_t* off, const double* t, const double* v,
                          long n2, const int64_t* off2, const double* t2, const double* v2,
                          long row_begin, long row_end, double* out, int nthreads);

/* Analytic gradient of L = sum_{pairs} G[pair] * X[pair] where X is the L2
 * distance (op = ORC_OP_LP with p = 2, kind PDIST: pairs i < j, compressed
 * index j(j-1)/2 + i, D = forward distances) or the inner product (op =
 * ORC_OP_DOT, kind KERNEL: pairs i <= j, compressed index j(j+1)/2 + i; D is
 * ignored). grad_v and grad_t are indexed like v and t and are overwritten. */
int orc_grad_f64(int op, int kind, long n, const int64_t* off,
                 const double* t, const double* v,
                 const double* G, const double* D,
                 double* grad_v, double* grad_t, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
