// The plain FP64 GPU kernel: one lane per pair, both PCFs read from global
// memory, PCFs in their original order. It is the same design as the
// reference's CUDA kernel (cuda_pcf_kernel.cuh:95-136: a thread per pair,
// no shared memory, the merge walk of iterate_rectangles.hpp:37-104 over the
// raw breakpoints) and is kept as the measured GPU baseline of the tiled
// kernels. Accumulation is fma(r - l, op, acc) per rectangle, the reference
// CPU integral's operation, so L1 / inner-product results are bit-exact.
#include "walk_common.hpp"

namespace sbamd
{
  struct PlainParams
  {
    const int64_t* off;
    const double* t;
    const double* v;
    int64_t n;
    int64_t col_blocks;
    double* out;
    double p;
    int32_t out_kind;
  };

  template <int OP>
  __global__ __launch_bounds__(64) void plain_pair_kernel(PlainParams P)
  {
    const int64_t i = static_cast<int64_t>(blockIdx.x) / P.col_blocks;
    const int64_t jb = static_cast<int64_t>(blockIdx.x) - i * P.col_blocks;
    const int64_t j = jb * 64 + threadIdx.x;
    const bool diag = P.out_kind == SBAMD_OUT_SYMMAT;
    if (i >= P.n || j >= P.n || (diag ? j > i : j >= i))
      return;
    const double* ft = P.t + P.off[i];
    const double* fvs = P.v + P.off[i];
    const int64_t mf = P.off[i + 1] - P.off[i];
    const double* gt = P.t + P.off[j];
    const double* gvs = P.v + P.off[j];
    const int64_t mg = P.off[j + 1] - P.off[j];
    const double zero = 0.0;
    // empty PCFs are the zero PCF [(0, 0)]
    const double* fT = mf ? ft : &zero;
    const double* fV = mf ? fvs : &zero;
    const double* gT = mg ? gt : &zero;
    const double* gV = mg ? gvs : &zero;
    const int64_t nf = mf ? mf : 1, ng = mg ? mg : 1;
    // start at the last point with t <= 0
    int64_t fi = 0, gi = 0;
    while (fi + 1 < nf && fT[fi + 1] <= 0.0)
      ++fi;
    while (gi + 1 < ng && gT[gi + 1] <= 0.0)
      ++gi;
    const double b = DBL_MAX;
    const double pp = P.p;
    double t = 0.0, acc = 0.0;
    while (t < b)
    {
      const double tprev = t;
      const double fv = fV[fi], gv = gV[gi];
      const bool fn = fi + 1 < nf, gn = gi + 1 < ng;
      if (fn && gn)
      {
        const double delta = fT[fi + 1] - gT[gi + 1];
        if (delta <= 0.0)
          ++fi;
        if (delta >= 0.0)
          ++gi;
      }
      else if (fn)
        ++fi;
      else if (gn)
        ++gi;
      else
      {
        acc = fma(b - tprev, pair_op<OP, double>(fv, gv, pp), acc);
        break;
      }
      const double mx = fT[fi] > gT[gi] ? fT[fi] : gT[gi];
      t = mx < b ? mx : b;
      acc = fma(t - tprev, pair_op<OP, double>(fv, gv, pp), acc);
    }
    const int64_t idx = diag ? symmat_index(i, j) : distmat_index(i, j);
    P.out[idx] = finalize<OP>(acc, pp);
  }

  void launch_plain(const int64_t* off, const double* t, const double* v, int64_t n, int op,
                    double p, int out_kind, double* out, hipStream_t stream)
  {
    if (n <= 0)
      return;
    PlainParams P;
    P.off = off;
    P.t = t;
    P.v = v;
    P.n = n;
    P.col_blocks = (n + 63) / 64;
    P.out = out;
    P.p = p;
    P.out_kind = out_kind;
    const int64_t blocks = n * P.col_blocks;
    if (blocks > 0x7fffffff)
      fail(SBAMD_ERR_INVALID, "plain kernel grid too large");
    const dim3 grid(static_cast<unsigned>(blocks));
    switch (op)
    {
    case OPK_L1:
      hipLaunchKernelGGL(plain_pair_kernel<OPK_L1>, grid, dim3(64), 0, stream, P);
      break;
    case OPK_L2:
      hipLaunchKernelGGL(plain_pair_kernel<OPK_L2>, grid, dim3(64), 0, stream, P);
      break;
    case OPK_LP:
      hipLaunchKernelGGL(plain_pair_kernel<OPK_LP>, grid, dim3(64), 0, stream, P);
      break;
    default:
      hipLaunchKernelGGL(plain_pair_kernel<OPK_DOT>, grid, dim3(64), 0, stream, P);
      break;
    }
    SBAMD_HIP(hipGetLastError());
  }
}
