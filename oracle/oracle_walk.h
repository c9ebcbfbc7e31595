/*
 * Type-generic body of the oracle walk; included by oracle.c once per floating
 * type with T, SUF(name), FMA, FABS, POW and TMAX defined. Test infrastructure
 * only (see oracle.h).
 */

/* Index of the last point with t <= a, clamped to 0.
 * Reference: max_time_iterator_prior_to, iterate_rectangles.hpp:23-35
 * (std::upper_bound then std::prev, begin if none). */
static long SUF(start_index)(const T* pt, long m, T a)
{
  long lo = 0, hi = m; /* first index with pt[i] > a */
  while (lo < hi)
  {
    long mid = lo + (hi - lo) / 2;
    if (a < pt[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo == 0 ? 0 : lo - 1;
}

/* The reference walk, iterate_rectangles.hpp:37-104, calling emit() per
 * rectangle. The callback receives (left, right, fv, gv, fi_after, gi_after,
 * fi_before, gi_before) so gradients can be attributed to points. */
typedef void (*SUF(emit_fn))(void* ctx, T l, T r, T fv, T gv, long fi0, long gi0, long fi1, long gi1);

static void SUF(walk)(const T* ft, const T* fvals, long mf,
                      const T* gt, const T* gvals, long mg,
                      T a, T b, SUF(emit_fn) emit, void* ctx)
{
  long fi = SUF(start_index)(ft, mf, a);
  long gi = SUF(start_index)(gt, mg, a);
  T t = a;
  T tprev = a;
  while (t < b)
  {
    tprev = t;
    T fv = fvals[fi];
    T gv = gvals[gi];
    long fi0 = fi, gi0 = gi;
    int fnext = fi + 1 < mf;
    int gnext = gi + 1 < mg;
    if (fnext && gnext)
    {
      T delta = ft[fi + 1] - gt[gi + 1];
      if (delta <= 0)
        ++fi;
      if (delta >= 0)
        ++gi;
    }
    else if (fnext)
    {
      ++fi;
    }
    else if (gnext)
    {
      ++gi;
    }
    else
    {
      emit(ctx, tprev, b, fv, gv, fi0, gi0, fi, gi);
      return;
    }
    T mx = ft[fi] > gt[gi] ? ft[fi] : gt[gi];
    t = mx < b ? mx : b;
    emit(ctx, tprev, t, fv, gv, fi0, gi0, fi, gi);
  }
}

typedef struct
{
  T* out;
  long cap;
  long count;
} SUF(rect_ctx);

static void SUF(rect_emit)(void* vctx, T l, T r, T fv, T gv, long fi0, long gi0, long fi1, long gi1)
{
  (void)fi0; (void)gi0; (void)fi1; (void)gi1;
  SUF(rect_ctx)* c = (SUF(rect_ctx)*)vctx;
  if (c->count < c->cap)
  {
    c->out[4 * c->count + 0] = l;
    c->out[4 * c->count + 1] = r;
    c->out[4 * c->count + 2] = fv;
    c->out[4 * c->count + 3] = gv;
  }
  c->count++;
}

long SUF(orc_rects)(const T* ft, const T* fv, long mf,
                    const T* gt, const T* gv, long mg,
                    T a, T b, T* out, long cap)
{
  SUF(rect_ctx) c = {out, cap, 0};
  SUF(walk)(ft, fv, mf, gt, gv, mg, a, b, SUF(rect_emit), &c);
  return c.count;
}

/* Norm of one PCF over [0, MAX): l1_norm / lp_norm of
 * include/sbear/algorithms/functional/apply_functional.hpp:63-100 over
 * iterate_segments (iterate_rectangles.hpp:106-155): from the last point at
 * or before 0, segments [t_k, min(t_{k+1}, MAX)) with value v_k, then the
 * last point's segment to MAX; out = fma(r - l, |v| (p = 1) or |v|^p, out);
 * the result is out (p = 1) or out^(1/p). An empty PCF has norm 0. */
T SUF(orc_lp_norm)(T p, const T* t, const T* v, long m)
{
  if (m == 0)
    return (T)0;
  const T b = TMAX;
  const int l1 = p == (T)1;
  long k = SUF(start_index)(t, m, (T)0);
  T tprev = 0;
  T vprev = v[k];
  T out = 0;
  for (long i = k + 1; i < m; ++i)
  {
    T tnext = t[i] < b ? t[i] : b;
    out = FMA(tnext - tprev, l1 ? FABS(vprev) : POW(FABS(vprev), p), out);
    if (tnext >= b)
      return l1 ? out : POW(out, (T)1 / p);
    tprev = t[i];
    vprev = v[i];
  }
  out = FMA(b - tprev, l1 ? FABS(vprev) : POW(FABS(vprev), p), out);
  return l1 ? out : POW(out, (T)1 / p);
}

/* Norms of a CSR set: out[i] = norm of PCF i. */
void SUF(orc_lp_norms)(T p, long n, const int64_t* off, const T* t, const T* v, T* out)
{
  for (long i = 0; i < n; ++i)
    out[i] = SUF(orc_lp_norm)(p, t + off[i], v + off[i], (long)(off[i + 1] - off[i]));
}

/* Pair ops, operations.cuh:19-65. */
static inline T SUF(pair_op)(int op, T p, T x, T y)
{
  if (op == ORC_OP_L1)
    return FABS(x - y);
  if (op == ORC_OP_LP)
    return POW(FABS(x - y), p);
  return x * y;
}

T SUF(orc_finalize)(int op, T p, T x)
{
  if (op == ORC_OP_LP)
    return POW(x, (T)1 / p);
  return x;
}

/* integrate(), matrix_integrate.hpp:18-29: val += (r - l) * op(fv, gv), which
 * the reference's build contracts to one FMA per rectangle. The walk is
 * inlined here (same steps as SUF(walk)) to keep the driver fast. */
T SUF(orc_integrate)(int op, T p,
                     const T* ft, const T* fvals, long mf,
                     const T* gt, const T* gvals, long mg,
                     T a, T b)
{
  long fi = SUF(start_index)(ft, mf, a);
  long gi = SUF(start_index)(gt, mg, a);
  T t = a;
  T tprev = a;
  T val = 0;
  while (t < b)
  {
    tprev = t;
    T fv = fvals[fi];
    T gv = gvals[gi];
    int fnext = fi + 1 < mf;
    int gnext = gi + 1 < mg;
    if (fnext && gnext)
    {
      T delta = ft[fi + 1] - gt[gi + 1];
      if (delta <= 0)
        ++fi;
      if (delta >= 0)
        ++gi;
    }
    else if (fnext)
    {
      ++fi;
    }
    else if (gnext)
    {
      ++gi;
    }
    else
    {
      val = FMA(b - tprev, SUF(pair_op)(op, p, fv, gv), val);
      return val;
    }
    T mx = ft[fi] > gt[gi] ? ft[fi] : gt[gi];
    t = mx < b ? mx : b;
    val = FMA(t - tprev, SUF(pair_op)(op, p, fv, gv), val);
  }
  return val;
}

typedef struct
{
  int op;
  T p;
  int kind;
  long n;
  const int64_t* off;
  const T* t;
  const T* v;
  long n2;
  const int64_t* off2;
  const T* t2;
  const T* v2;
  T* out;
  long row_end;
  long next_row; /* accessed atomically */
} SUF(pair_job);

static void SUF(compute_row)(SUF(pair_job)* jb, long i)
{
  const int64_t* off = jb->off;
  const T* ft = jb->t + off[i];
  const T* fv = jb->v + off[i];
  long mf = (long)(off[i + 1] - off[i]);
  if (jb->kind == ORC_KIND_CDIST)
  {
    for (long j = 0; j < jb->n2; ++j)
    {
      const int64_t* o2 = jb->off2;
      T r = SUF(orc_integrate)(jb->op, jb->p, ft, fv, mf,
                               jb->t2 + o2[j], jb->v2 + o2[j], (long)(o2[j + 1] - o2[j]),
                               (T)0, TMAX);
      jb->out[i * jb->n2 + j] = SUF(orc_finalize)(jb->op, jb->p, r);
    }
    return;
  }
  long j0 = jb->kind == ORC_KIND_KERNEL ? i : i + 1;
  for (long j = j0; j < jb->n; ++j)
  {
    T r = SUF(orc_integrate)(jb->op, jb->p, ft, fv, mf,
                             jb->t + off[j], jb->v + off[j], (long)(off[j + 1] - off[j]),
                             (T)0, TMAX);
    long idx = jb->kind == ORC_KIND_KERNEL ? j * (j + 1) / 2 + i : j * (j - 1) / 2 + i;
    jb->out[idx] = SUF(orc_finalize)(jb->op, jb->p, r);
  }
}

static void* SUF(pair_worker)(void* arg)
{
  SUF(pair_job)* jb = (SUF(pair_job)*)arg;
  for (;;)
  {
    long i = __atomic_fetch_add(&jb->next_row, 1, __ATOMIC_RELAXED);
    if (i >= jb->row_end)
      break;
    SUF(compute_row)(jb, i);
  }
  return NULL;
}

static int SUF(run_pairwise)(SUF(pair_job)* jb, long row_begin, int nthreads)
{
  jb->next_row = row_begin;
  if (nthreads < 1)
    nthreads = 1;
  if (nthreads == 1)
  {
    SUF(pair_worker)(jb);
    return 0;
  }
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  if (!th)
    return -1;
  int started = 0;
  for (int k = 0; k < nthreads; ++k)
  {
    if (pthread_create(&th[k], NULL, SUF(pair_worker), jb) != 0)
      break;
    started++;
  }
  if (started == 0)
    SUF(pair_worker)(jb);
  for (int k = 0; k < started; ++k)
    pthread_join(th[k], NULL);
  free(th);
  return 0;
}
