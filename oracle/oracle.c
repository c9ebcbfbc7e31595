/*
 * CPU oracle for the pairwise PCF integration path. TEST INFRASTRUCTURE ONLY:
 * the checker loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; never linked or called by the product. See oracle.h for
 * the reference files each function follows.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- double instantiation ---- */
#define T double
#define SUF(name) name##_f64
#define FMA fma
#define FABS fabs
#define POW pow
#define TMAX DBL_MAX
#include "oracle_walk.h"
#undef T
#undef SUF
#undef FMA
#undef FABS
#undef POW
#undef TMAX

/* ---- float instantiation ---- */
#define T float
#define SUF(name) name##_f32
#define FMA fmaf
#define FABS fabsf
#define POW powf
#define TMAX FLT_MAX
#include "oracle_walk.h"
#undef T
#undef SUF
#undef FMA
#undef FABS
#undef POW
#undef TMAX

int orc_pairwise_f64(int op, double p, int kind,
                     long n, const int64_t* off, const double* t, const double* v,
                     long n2, const int64_t* off2, const double* t2, const double* v2,
                     double* out, int nthreads)
{
  return orc_pairwise_rows_f64(op, p, kind, n, off, t, v, n2, off2, t2, v2, 0, n, out, nthreads);
}

int orc_pairwise_rows_f64(int op, double p, int kind,
                          long n, const int64_t* off, const double* t, const double* v,
                          long n2, const int64_t* off2, const double* t2, const double* v2,
                          long row_begin, long row_end, double* out, int nthreads)
{
  pair_job_f64 jb;
  memset(&jb, 0, sizeof(jb));
  jb.op = op; jb.p = p; jb.kind = kind;
  jb.n = n; jb.off = off; jb.t = t; jb.v = v;
  jb.n2 = n2; jb.off2 = off2; jb.t2 = t2; jb.v2 = v2;
  jb.out = out;
  jb.row_end = row_end < n ? row_end : n;
  return run_pairwise_f64(&jb, row_begin, nthreads);
}

int orc_pairwise_f32(int op, float p, int kind,
                     long n, const int64_t* off, const float* t, const float* v,
                     long n2, const int64_t* off2, const float* t2, const float* v2,
                     float* out, int nthreads)
{
  pair_job_f32 jb;
  memset(&jb, 0, sizeof(jb));
  jb.op = op; jb.p = p; jb.kind = kind;
  jb.n = n; jb.off = off; jb.t = t; jb.v = v;
  jb.n2 = n2; jb.off2 = off2; jb.t2 = t2; jb.v2 = v2;
  jb.out = out;
  jb.row_end = n;
  return run_pairwise_f32(&jb, 0, nthreads);
}

/* ---- gradients ---- */

typedef struct
{
  int op;
  double w;            /* dL/dS for L2 (G / (2 D)), or G for the inner product */
  const double* fv;    /* values of f (the row PCF) */
  const double* gv;    /* values of g */
  double* gvf;         /* grad_v slice of f */
  double* gtf;         /* grad_t slice of f */
  double* gvg;
  double* gtg;
} grad_ctx;

/* One rectangle [l, r) with pre-step values (a_f, a_g) at point indices
 * (fi0, gi0); (fi1, gi1) are the indices after the step. */
static void grad_emit(void* vctx, double l, double r, double fval, double gval,
                      long fi0, long gi0, long fi1, long gi1)
{
  grad_ctx* c = (grad_ctx*)vctx;
  double len = r - l;
  if (c->op == ORC_OP_DOT)
  {
    c->gvf[fi0] += c->w * gval * len;
    c->gvg[gi0] += c->w * fval * len;
  }
  else
  {
    double d = fval - gval;
    c->gvf[fi0] += c->w * 2.0 * d * len;
    c->gvg[gi0] -= c->w * 2.0 * d * len;
  }
  if (fi1 != fi0)
  {
    /* f crossed its breakpoint fi1; the partner value is the mean of g's
     * values before and after the step (they differ when g crosses too). */
    double a = c->fv[fi0], b = c->fv[fi1], s = 0.5 * (gval + c->gv[gi1]);
    if (c->op == ORC_OP_DOT)
      c->gtf[fi1] += c->w * (a - b) * s;
    else
      c->gtf[fi1] += c->w * ((a - s) * (a - s) - (b - s) * (b - s));
  }
  if (gi1 != gi0)
  {
    /* g crossed its breakpoint gi1; the partner value is the mean of f's
     * values before and after the step. */
    double a = c->gv[gi0], b = c->gv[gi1], s = 0.5 * (fval + c->fv[fi1]);
    if (c->op == ORC_OP_DOT)
      c->gtg[gi1] += c->w * (a - b) * s;
    else
      c->gtg[gi1] += c->w * ((a - s) * (a - s) - (b - s) * (b - s));
  }
}

typedef struct
{
  int op;
  int kind;
  long n;
  const int64_t* off;
  const double* t;
  const double* v;
  const double* G;
  const double* D;
  long npts;
  long next_row;
  pthread_mutex_t mu;
  double* grad_v;
  double* grad_t;
} grad_job;

static void* grad_worker(void* arg)
{
  grad_job* jb = (grad_job*)arg;
  double* gv = (double*)calloc((size_t)jb->npts, sizeof(double));
  double* gt = (double*)calloc((size_t)jb->npts, sizeof(double));
  if (!gv || !gt)
  {
    free(gv);
    free(gt);
    return (void*)1;
  }
  const int64_t* off = jb->off;
  for (;;)
  {
    long i = __atomic_fetch_add(&jb->next_row, 1, __ATOMIC_RELAXED);
    if (i >= jb->n)
      break;
    long j0 = jb->kind == ORC_KIND_KERNEL ? i : i + 1;
    for (long j = j0; j < jb->n; ++j)
    {
      long idx = jb->kind == ORC_KIND_KERNEL ? j * (j + 1) / 2 + i : j * (j - 1) / 2 + i;
      grad_ctx c;
      c.op = jb->op;
      if (jb->op == ORC_OP_DOT)
      {
        c.w = jb->G[idx];
      }
      else
      {
        double D = jb->D ? jb->D[idx]
                         : sqrt(orc_integrate_f64(ORC_OP_LP, 2.0, jb->t + off[i], jb->v + off[i],
                                                  (long)(off[i + 1] - off[i]), jb->t + off[j],
                                                  jb->v + off[j], (long)(off[j + 1] - off[j]),
                                                  0.0, DBL_MAX));
        c.w = D > 0.0 ? jb->G[idx] / (2.0 * D) : 0.0;
      }
      c.fv = jb->v + off[i];
      c.gv = jb->v + off[j];
      c.gvf = gv + off[i];
      c.gtf = gt + off[i];
      c.gvg = gv + off[j];
      c.gtg = gt + off[j];
      walk_f64(jb->t + off[i], jb->v + off[i], (long)(off[i + 1] - off[i]),
               jb->t + off[j], jb->v + off[j], (long)(off[j + 1] - off[j]),
               0.0, DBL_MAX, grad_emit, &c);
    }
  }
  pthread_mutex_lock(&jb->mu);
  for (long k = 0; k < jb->npts; ++k)
  {
    jb->grad_v[k] += gv[k];
    jb->grad_t[k] += gt[k];
  }
  pthread_mutex_unlock(&jb->mu);
  free(gv);
  free(gt);
  return NULL;
}

int orc_grad_f64(int op, int kind, long n, const int64_t* off,
                 const double* t, const double* v,
                 const double* G, const double* D,
                 double* grad_v, double* grad_t, int nthreads)
{
  grad_job jb;
  memset(&jb, 0, sizeof(jb));
  jb.op = op; jb.kind = kind; jb.n = n; jb.off = off; jb.t = t; jb.v = v;
  jb.G = G; jb.D = D;
  jb.npts = n > 0 ? (long)off[n] : 0;
  jb.grad_v = grad_v; jb.grad_t = grad_t;
  pthread_mutex_init(&jb.mu, NULL);
  memset(grad_v, 0, sizeof(double) * (size_t)jb.npts);
  memset(grad_t, 0, sizeof(double) * (size_t)jb.npts);
  if (nthreads < 1)
    nthreads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  if (!th)
    return -1;
  int started = 0;
  int rc = 0;
  for (int k = 0; k < nthreads; ++k)
  {
    if (pthread_create(&th[k], NULL, grad_worker, &jb) != 0)
      break;
    started++;
  }
  if (started == 0)
    rc = grad_worker(&jb) ? -1 : 0;
  for (int k = 0; k < started; ++k)
  {
    void* res = NULL;
    pthread_join(th[k], &res);
    if (res)
      rc = -1;
  }
  free(th);
  pthread_mutex_destroy(&jb.mu);
  /* t_0 = 0 is fixed by the data model (the first breakpoint must be 0). */
  for (long i = 0; i < n; ++i)
    if (off[i + 1] > off[i])
      grad_t[off[i]] = 0.0;
  return rc;
}
