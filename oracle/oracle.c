/*
 * oracle.c -- CPU restatement of AHMED's H-matrix hot path (test checker).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.  See oracle.h for the
 * parity pin of every entry point.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */

/* BLAS / LAPACK (Fortran ABI, MKL).  Same routines the reference wraps in
 * Include/blas.h:176-560.                                                   */
extern double ddot_(const int*, const double*, const int*, const double*, const int*);
extern void   daxpy_(const int*, const double*, const double*, const int*, double*, const int*);
extern double dnrm2_(const int*, const double*, const int*);
extern void   dscal_(const int*, const double*, double*, const int*);
extern void   dcopy_(const int*, const double*, const int*, double*, const int*);
extern void   dgemv_(const char*, const int*, const int*, const double*, const double*,
                     const int*, const double*, const int*, const double*, double*, const int*);
extern void   dspmv_(const char*, const int*, const double*, const double*, const double*,
                     const int*, const double*, double*, const int*);
extern void   dgemm_(const char*, const char*, const int*, const int*, const int*,
                     const double*, const double*, const int*, const double*, const int*,
                     const double*, double*, const int*);
extern void   dgeqrf_(const int*, const int*, double*, const int*, double*, double*,
                      const int*, int*);
extern void   dgesvd_(const char*, const char*, const int*, const int*, double*, const int*,
                      double*, double*, const int*, double*, const int*, double*,
                      const int*, int*);
extern void   dormqr_(const char*, const char*, const int*, const int*, const int*,
                      const double*, const int*, const double*, double*, const int*,
                      double*, const int*, int*);
extern void   dtrtrs_(const char*, const char*, const char*, const int*, const int*,
                      const double*, const int*, double*, const int*, int*);
extern int    MKL_CBWR_Set(int);
extern void   MKL_Get_Version_String(char*, int);

static const int I1 = 1;
static const double DONE = 1.0, DZERO = 0.0;

static void blas_init(void)
{
  static int done = 0;
  if (!done) { MKL_CBWR_Set(2 /* MKL_CBWR_COMPATIBLE */); done = 1; }
}

void or_blas_version(char* buf, int len) { MKL_Get_Version_String(buf, len); }

static double scpr(unsigned n, const double* a, const double* b)
{ int nn = (int)n; return ddot_(&nn, a, &I1, b, &I1); }
static void axpy(unsigned n, double d, const double* x, double* y)
{ int nn = (int)n; daxpy_(&nn, &d, x, &I1, y, &I1); }
static double nrm2(unsigned n, const double* x)
{ int nn = (int)n; return dnrm2_(&nn, x, &I1); }
static void scal(unsigned n, double d, double* x)
{ int nn = (int)n; dscal_(&nn, &d, x, &I1); }
static void copyv(unsigned n, const double* x, double* y)
{ int nn = (int)n; dcopy_(&nn, x, &I1, y, &I1); }
/* y += d*A*x, A m x n col-major ld m  (Include/blas.h:1172-1175) */
static void gemva(unsigned m, unsigned n, double d, const double* A,
                  const double* x, double* y)
{ int mm = (int)m, nn = (int)n; dgemv_("N", &mm, &nn, &d, A, &mm, x, &I1, &DONE, y, &I1); }

static void* xmalloc(size_t n)
{
  void* p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n); abort(); }
  return p;
}

/* ======================================================================== */
/* Cluster trees                                                            */
/* ======================================================================== */

typedef struct {
  unsigned nbeg, nend;
  int son[2];
  double bb[6];     /* bbx: xmin,ymin,zmin,xmax,ymax,zmax */
  double com[3];    /* pca: centre of mass                */
  double diam2;
  unsigned maindir;
  double cntrdir;
  unsigned icom;
} or_cl;

struct or_tree {
  int kind;
  int sym;          /* symmetric block tree: leaves on and above the diagonal */
  unsigned N;
  double* pts;
  unsigned *op, *po;
  or_cl* cl;
  unsigned ncl, capcl;
  unsigned nl, capl;
  unsigned *b1, *n1, *b2, *n2, *rcl;
  unsigned char* adm;
  double eta2;
  /* symmetric trees: the leaf ranges in the order mltaHeHVec_ visits them
     (H/mltaGeHVec.cpp:420-446): op 0 = y += A x over leaves [l0, l1),
     op 1 = y += A^T x over [l0, l1) */
  unsigned nsched, capsched;
  unsigned* sched;   /* triples (op, l0, l1) */
};

static const double* P(const or_tree* t, unsigned orig) { return t->pts + 3 * (size_t)orig; }

static unsigned new_cluster(or_tree* t, unsigned k, unsigned l)
{
  if (t->ncl == t->capcl) {
    t->capcl = t->capcl ? 2 * t->capcl : 64;
    t->cl = (or_cl*)realloc(t->cl, sizeof(or_cl) * t->capcl);
    if (!t->cl) abort();
  }
  or_cl* c = &t->cl[t->ncl];
  memset(c, 0, sizeof(*c));
  c->nbeg = k; c->nend = l; c->son[0] = c->son[1] = -1;
  return t->ncl++;
}

/* cluster_bbx::init, Include/cluster_bbx.h:52-98 (points have radius 0) */
static void bbx_init(or_tree* t, unsigned ci)
{
  or_cl* c = &t->cl[ci];
  const double* v = P(t, t->op[c->nbeg]);
  for (int i = 0; i < 3; ++i) { c->bb[i] = v[i]; c->bb[i + 3] = v[i]; }
  for (unsigned j = c->nbeg + 1; j < c->nend; ++j) {
    v = P(t, t->op[j]);
    for (int i = 0; i < 3; ++i) {
      const double x = v[i];
      if (x < c->bb[i]) c->bb[i] = x;
      else if (x > c->bb[i + 3]) c->bb[i + 3] = x;
    }
  }
  c->diam2 = 0.0;
  unsigned imax = 0;
  double mx = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double e = c->bb[i + 3] - c->bb[i];
    c->diam2 += e * e;
    if (e > mx) { mx = e; imax = (unsigned)i; }
  }
  c->maindir = (c->bb[3] - c->bb[0] > 0.8 * mx) ? 0u : imax;
  c->cntrdir = 0.5 * (c->bb[c->maindir] + c->bb[c->maindir + 3]);
}

static void swapu(unsigned* a, unsigned* b) { unsigned t = *a; *a = *b; *b = t; }

/* cluster_bbx::createClusterTree, Include/cluster_bbx.h:100-128 */
static void bbx_split(or_tree* t, unsigned ci, unsigned bmin)
{
  const unsigned nbeg = t->cl[ci].nbeg, nend = t->cl[ci].nend;
  const unsigned md = t->cl[ci].maindir;
  const double cd = t->cl[ci].cntrdir;
  unsigned nnbeg = nend;
  for (unsigned i = nbeg; i < nnbeg;) {
    if (P(t, t->op[i])[md] < cd) ++i;
    else {
      --nnbeg;
      swapu(&t->op[i], &t->op[nnbeg]);
      t->po[t->op[i]] = i;
      t->po[t->op[nnbeg]] = nnbeg;
    }
  }
  if (nnbeg - nbeg > bmin && nend - nnbeg > bmin) {
    unsigned s0 = new_cluster(t, nbeg, nnbeg);
    bbx_init(t, s0);
    bbx_split(t, s0, bmin);
    unsigned s1 = new_cluster(t, nnbeg, nend);
    bbx_init(t, s1);
    bbx_split(t, s1, bmin);
    t->cl[ci].son[0] = (int)s0;
    t->cl[ci].son[1] = (int)s1;
  }
}

/* cluster_pca::init, Include/cluster_pca.h:82-115 */
static void pca_init(or_tree* t, unsigned ci)
{
  or_cl* c = &t->cl[ci];
  for (int j = 0; j < 3; ++j) c->com[j] = 0.0;
  for (unsigned i = c->nbeg; i < c->nend; ++i)
    for (int j = 0; j < 3; ++j) c->com[j] += P(t, t->op[i])[j];
  const double d = 1.0 / (c->nend - c->nbeg);
  for (int j = 0; j < 3; ++j) c->com[j] *= d;
  c->diam2 = 0.0;
  for (unsigned i = c->nbeg; i < c->nend; ++i) {
    double d2 = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double d1 = 2 * (P(t, t->op[i])[j] - c->com[j]);
      d2 += d1 * d1;
    }
    if (d2 > c->diam2) c->diam2 = d2;
  }
  /* bemcluster icom, Include/bemcluster.h:25-56: scans dofs[i] UNPERMUTED */
  c->icom = c->nbeg;
  double smin = 0.0;
  for (int j = 0; j < 3; ++j) {
    const double e = c->com[j] - P(t, c->nbeg)[j];
    smin += e * e;
  }
  for (unsigned i = c->nbeg + 1; i < c->nend; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = c->com[j] - P(t, i)[j];
      s += e * e;
    }
    if (s < smin) { c->icom = i; smin = s; }
  }
}

/* cluster_pca::subdivide + createClusterTree, Include/cluster_pca.h:118-217 */
static void pca_split(or_tree* t, unsigned ci, unsigned bmin)
{
  const unsigned nbeg = t->cl[ci].nbeg, nend = t->cl[ci].nend;
  double com[3];
  memcpy(com, t->cl[ci].com, sizeof com);
  double c[6] = {0, 0, 0, 0, 0, 0}, d[3], u[3];
  for (unsigned i = nbeg; i < nend; ++i) {
    for (int j = 0; j < 3; ++j) d[j] = P(t, t->op[i])[j] - com[j];
    int l = 0;
    for (int j = 0; j < 3; ++j)
      for (int k = j; k < 3; ++k) c[l++] += d[j] * d[k];
  }
  for (int i = 0; i < 3; ++i) u[i] = 1.0 / sqrt(3.0);
  unsigned it = 0;
  do {
    for (int j = 0; j < 3; ++j) d[j] = u[j];
    for (int j = 0; j < 3; ++j) {
      int k;
      u[j] = 0.0;
      for (k = 0; k < j; ++k) u[j] += c[k * 3 - k * (k + 1) / 2 + j] * d[k];
      for (; k < 3; ++k) u[j] += c[j * 3 - j * (j + 1) / 2 + k] * d[k];
    }
    double modu = 0.0;
    for (int j = 0; j < 3; ++j) modu += u[j] * u[j];
    modu = 1.0 / sqrt(modu);
    for (int j = 0; j < 3; ++j) u[j] *= modu;
  } while (++it < 10);
  double mdtcl = 0.0;
  for (int i = 0; i < 3; ++i) mdtcl += u[i] * com[i];
  unsigned nsep = nend;
  for (unsigned i = nbeg; i < nsep;) {
    const double* v = P(t, t->op[i]);
    double skp = 0.0;
    for (int j = 0; j < 3; ++j) skp += v[j] * u[j];
    if (skp < mdtcl) ++i;
    else {
      --nsep;
      swapu(&t->op[i], &t->op[nsep]);
      t->po[t->op[i]] = i;
      t->po[t->op[nsep]] = nsep;
    }
  }
  if (nsep == nbeg || nsep == nend) nsep = (nend + nbeg) / 2;
  if (nsep >= nbeg + bmin && nend >= nsep + bmin) {
    unsigned s0 = new_cluster(t, nbeg, nsep);
    pca_init(t, s0);
    pca_split(t, s0, bmin);
    unsigned s1 = new_cluster(t, nsep, nend);
    pca_init(t, s1);
    pca_split(t, s1, bmin);
    t->cl[ci].son[0] = (int)s0;
    t->cl[ci].son[1] = (int)s1;
  }
}

/* ACA start row for bbx clusters (the bbx tree has no icom in the reference):
 * permuted position whose point is closest to the cluster's centre of mass,
 * com summed in position order, first minimum wins.                        */
static void bbx_icom(or_tree* t, unsigned ci)
{
  or_cl* c = &t->cl[ci];
  double com[3] = {0, 0, 0};
  for (unsigned i = c->nbeg; i < c->nend; ++i)
    for (int j = 0; j < 3; ++j) com[j] += P(t, t->op[i])[j];
  const double d = 1.0 / (c->nend - c->nbeg);
  for (int j = 0; j < 3; ++j) com[j] *= d;
  double smin = INFINITY;
  c->icom = c->nbeg;
  for (unsigned i = c->nbeg; i < c->nend; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = P(t, t->op[i])[j] - com[j];
      s += e * e;
    }
    if (s < smin) { smin = s; c->icom = i; }
  }
}

/* cluster3d_bbx::isadm, Include/cluster_bbx.h:245-297 (dist2 :33-47) */
static int bbx_isadm(const or_cl* a, const or_cl* b, double eta2)
{
  double dd = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double e = b->bb[i] - a->bb[i + 3];
    if (e > 0.0) dd += e * e;
    else {
      const double e1 = a->bb[i] - b->bb[i + 3];
      if (e1 > 0.0) dd += e1 * e1;
    }
  }
  const double d2 = a->diam2 < b->diam2 ? a->diam2 : b->diam2;
  if (d2 < eta2 * dd) return 1;
  /* weak admissibility on exactly touching boxes */
  const int yc = (a->bb[4] == b->bb[1]) || (b->bb[4] == a->bb[1]);
  const int zc = (a->bb[5] == b->bb[2]) || (b->bb[5] == a->bb[2]);
  if (a->bb[3] == b->bb[0] && yc && zc) return 1;
  if (b->bb[3] == a->bb[0] && yc && zc) return 1;
  return 0;
}

/* cluster_pca::isadm, Include/cluster_pca.h:54-80 */
static int pca_isadm(const or_cl* a, const or_cl* b, double eta2)
{
  double d2c = 0.0;
  for (int j = 0; j < 3; ++j) {
    const double d = a->com[j] - b->com[j];
    d2c += d * d;
  }
  const double d2 = a->diam2 < b->diam2 ? a->diam2 : b->diam2;
  return d2 < eta2 * d2c;
}

static void add_leaf(or_tree* t, unsigned c1, unsigned c2, int adm)
{
  if (t->nl == t->capl) {
    t->capl = t->capl ? 2 * t->capl : 1024;
    t->b1 = (unsigned*)realloc(t->b1, sizeof(unsigned) * t->capl);
    t->n1 = (unsigned*)realloc(t->n1, sizeof(unsigned) * t->capl);
    t->b2 = (unsigned*)realloc(t->b2, sizeof(unsigned) * t->capl);
    t->n2 = (unsigned*)realloc(t->n2, sizeof(unsigned) * t->capl);
    t->rcl = (unsigned*)realloc(t->rcl, sizeof(unsigned) * t->capl);
    t->adm = (unsigned char*)realloc(t->adm, t->capl);
    if (!t->b1 || !t->n1 || !t->b2 || !t->n2 || !t->rcl || !t->adm) abort();
  }
  const or_cl *a = &t->cl[c1], *b = &t->cl[c2];
  t->b1[t->nl] = a->nbeg; t->n1[t->nl] = a->nend - a->nbeg;
  t->b2[t->nl] = b->nbeg; t->n2[t->nl] = b->nend - b->nbeg;
  t->rcl[t->nl] = c1;
  t->adm[t->nl] = (unsigned char)adm;
  t->nl++;
}

/* blcluster::subdivide_, cluster/blcluster.cpp:40-66: leaves numbered DFS */
static void subdivide(or_tree* t, unsigned c1, unsigned c2, double eta2)
{
  const or_cl *a = &t->cl[c1], *b = &t->cl[c2];
  int adm = t->kind == OR_TREE_BBX ? bbx_isadm(a, b, eta2) : pca_isadm(a, b, eta2);
  if (adm) { add_leaf(t, c1, c2, 1); return; }
  if (a->son[0] >= 0 && b->son[0] >= 0) {
    const int s1[2] = {a->son[0], a->son[1]}, s2[2] = {b->son[0], b->son[1]};
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) subdivide(t, (unsigned)s1[i], (unsigned)s2[j], eta2);
  } else add_leaf(t, c1, c2, 0);
}

/* blcluster::subdivide_sym_, cluster/blcluster.cpp:90-113: diagonal blocks
   are never admissible; below the diagonal nothing, diagonal sons recurse
   symmetrically, the sons above the diagonal subdivide as above, in
   row-major son order */
static void push_sched(or_tree* t, unsigned op, unsigned l0, unsigned l1)
{
  if (t->nsched == t->capsched) {
    t->capsched = t->capsched ? 2 * t->capsched : 1024;
    t->sched = (unsigned*)realloc(t->sched, sizeof(unsigned) * 3 * t->capsched);
    if (!t->sched) abort();
  }
  unsigned* e = t->sched + 3 * (size_t)t->nsched++;
  e[0] = op; e[1] = l0; e[2] = l1;
}

static void subdivide_sym(or_tree* t, unsigned c)
{
  const or_cl* a = &t->cl[c];
  if (a->son[0] >= 0) {
    const int s[2] = {a->son[0], a->son[1]};
    for (int i = 0; i < 2; ++i) {
      subdivide_sym(t, (unsigned)s[i]);
      for (int j = i + 1; j < 2; ++j) {
        const unsigned l0 = t->nl;
        subdivide(t, (unsigned)s[i], (unsigned)s[j], t->eta2);
        push_sched(t, 0, l0, t->nl);
        push_sched(t, 1, l0, t->nl);
      }
    }
  } else {
    push_sched(t, 0, t->nl, t->nl + 1);
    add_leaf(t, c, c, 0);
  }
}

or_tree* or_tree_build(const double* pts, unsigned N, unsigned bmin, double eta2, int kind)
{
  if (N == 0) return NULL;
  or_tree* t = (or_tree*)calloc(1, sizeof(or_tree));
  t->sym = (kind & OR_TREE_SYM) != 0;
  kind &= 0xff;
  t->eta2 = eta2;
  t->kind = kind;
  t->N = N;
  t->pts = (double*)xmalloc(sizeof(double) * 3 * (size_t)N);
  memcpy(t->pts, pts, sizeof(double) * 3 * (size_t)N);
  t->op = (unsigned*)xmalloc(sizeof(unsigned) * N);
  t->po = (unsigned*)xmalloc(sizeof(unsigned) * N);
  for (unsigned i = 0; i < N; ++i) t->op[i] = t->po[i] = i;
  unsigned r = new_cluster(t, 0, N);
  if (kind == OR_TREE_BBX) {
    bbx_init(t, r);
    bbx_split(t, r, bmin);
    for (unsigned c = 0; c < t->ncl; ++c) bbx_icom(t, c);
  } else {
    pca_init(t, r);
    pca_split(t, r, bmin);
  }
  if (t->sym) subdivide_sym(t, r);
  else subdivide(t, r, r, eta2);
  return t;
}

void or_tree_free(or_tree* t)
{
  if (!t) return;
  free(t->pts); free(t->op); free(t->po); free(t->cl);
  free(t->b1); free(t->n1); free(t->b2); free(t->n2); free(t->rcl); free(t->adm);
  free(t->sched);
  free(t);
}

unsigned or_tree_nleaves(const or_tree* t) { return t->nl; }
unsigned or_tree_nclusters(const or_tree* t) { return t->ncl; }

void or_tree_perm(const or_tree* t, unsigned* op, unsigned* po)
{
  memcpy(op, t->op, sizeof(unsigned) * t->N);
  memcpy(po, t->po, sizeof(unsigned) * t->N);
}

void or_tree_leaves(const or_tree* t, unsigned* b1, unsigned* n1, unsigned* b2,
                    unsigned* n2, unsigned char* adm, unsigned* i0)
{
  for (unsigned l = 0; l < t->nl; ++l) {
    b1[l] = t->b1[l]; n1[l] = t->n1[l]; b2[l] = t->b2[l]; n2[l] = t->n2[l];
    adm[l] = t->adm[l];
    i0[l] = t->cl[t->rcl[l]].icom - t->b1[l];
  }
}

void or_tree_clusters(const or_tree* t, unsigned* nbeg, unsigned* nend, int* son0,
                      int* son1, unsigned* icom)
{
  for (unsigned c = 0; c < t->ncl; ++c) {
    nbeg[c] = t->cl[c].nbeg; nend[c] = t->cl[c].nend;
    son0[c] = t->cl[c].son[0]; son1[c] = t->cl[c].son[1];
    icom[c] = t->cl[c].icom;
  }
}

/* ======================================================================== */
/* Kernel generators (MATGEN_T plugin, contract Include/matgen_sqntl.h:21-29) */
/* ======================================================================== */

#define ONE_OVER_4PI 0.079577471545947667884

/* weight w scales the off-diagonal entries (a quadrature weight; w = 1 is the
 * plain kernel matrix).  c = w/(4 pi) (Laplace, Yukawa) or w (EXPR). */
static double kconst(int kernel, double weight)
{ return kernel == OR_KERNEL_EXPR ? weight : weight * ONE_OVER_4PI; }

static double kval(int kernel, double kappa, double c, const double* x, const double* y)
{
  const double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
  const double r = sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
  if (kernel == OR_KERNEL_LAPLACE) return c / r;
  return (exp(-kappa * r) * c) / r;
}

void or_kernel_block(int kernel, double kappa, double weight, const double* pts,
                     const unsigned* op, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                     double* out)
{
  const double c = kconst(kernel, weight);
  for (unsigned j = 0; j < n2; ++j) {
    const unsigned oj = op[b2 + j];
    for (unsigned i = 0; i < n1; ++i) {
      const unsigned oi = op[b1 + i];
      out[i + (size_t)j * n1] =
          oi == oj ? 1.0 : kval(kernel, kappa, c, pts + 3 * (size_t)oi, pts + 3 * (size_t)oj);
    }
  }
}

double or_kernel_scale(int kernel, double kappa, double weight, const double* pts,
                       const unsigned* op, unsigned b1, unsigned n1, unsigned b2, unsigned n2)
{
  double v;
  or_kernel_block(kernel, kappa, weight, pts, op, b1 + n1 / 2, 1, b2 + n2 / 2, 1, &v);
  return fabs(v);
}

typedef struct {
  void* ctx;
  void (*cmpbl)(void*, unsigned, unsigned, unsigned, unsigned, double*);
  double (*scale)(void*, unsigned, unsigned, unsigned, unsigned);
} matgen;

typedef struct { int kernel; double kappa, weight; const double* pts; const unsigned* op; } kgen_ctx;
static void kgen_cmpbl(void* c, unsigned b1, unsigned n1, unsigned b2, unsigned n2, double* o)
{ kgen_ctx* k = (kgen_ctx*)c; or_kernel_block(k->kernel, k->kappa, k->weight, k->pts, k->op, b1, n1, b2, n2, o); }
static double kgen_scale(void* c, unsigned b1, unsigned n1, unsigned b2, unsigned n2)
{ kgen_ctx* k = (kgen_ctx*)c; return or_kernel_scale(k->kernel, k->kappa, k->weight, k->pts, k->op, b1, n1, b2, n2); }

typedef struct { const double* A; unsigned lda; } dgen_ctx;
static void dgen_cmpbl(void* c, unsigned b1, unsigned n1, unsigned b2, unsigned n2, double* o)
{
  dgen_ctx* d = (dgen_ctx*)c;
  for (unsigned j = 0; j < n2; ++j)
    for (unsigned i = 0; i < n1; ++i) o[i + (size_t)j * n1] = d->A[(b1 + i) + (size_t)(b2 + j) * d->lda];
}

/* ======================================================================== */
/* ACA, Include/ACA.h:29-350                                                 */
/* ======================================================================== */

static unsigned nexti(unsigned n, const int* Z, unsigned j)
{
  for (;;) {
    if (Z[j] != -1) return j;
    j = (j + 1 < n) ? j + 1 : 0;
  }
}

/* ACA_row_step, Include/ACA.h:46-130 */
static int aca_row_step(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                        unsigned* klast, unsigned* i0, unsigned k, unsigned* no, int* Z,
                        int* S, double* U, double* V, double* nrmlsk2, double scale)
{
  unsigned l, j0 = 0;
  double nrmu2, nrmv2, absmax;
  double* const pu = U + (size_t)k * n1;
  double* const pv = V + (size_t)k * n2;
  do {
    G->cmpbl(G->ctx, b1 + *i0, 1, b2, n2, pv);
    for (l = 0; l < k; ++l) axpy(n2, -U[(size_t)l * n1 + *i0], V + (size_t)l * n2, pv);
    absmax = 0.0;
    for (l = 0; l < n2; ++l)
      if (S[l] >= 0) {
        const double eabs = fabs(pv[l]);
        if (eabs > absmax) { absmax = eabs; j0 = l; }
      }
    if (absmax < 1e-14 * scale) {
      Z[*i0] = -1;
      if (++(*no) < n1) *i0 = nexti(n1, Z, *i0);
    }
  } while (absmax < 1e-14 * scale && *no < n1);
  if (absmax < 1e-14 * scale) return 0;

  const double sqrtpiv = sqrt(absmax);
  nrmv2 = 0.0;
  const double sca = sqrtpiv / pv[j0];
  for (l = 0; l < n2; ++l)
    if (S[l] >= 0) {
      const double e = (pv[l] *= sca);
      nrmv2 += e * e;
      if (fabs(e) > 1e-8 * scale * sqrtpiv) ++S[l];
    }
  G->cmpbl(G->ctx, b1, n1, b2 + j0, 1, pu);
  for (l = 0; l < k; ++l) axpy(n1, -V[(size_t)l * n2 + j0], U + (size_t)l * n1, pu);

  *klast = (unsigned)Z[*i0];
  Z[*i0] = S[j0] = -1;
  ++(*no);
  const double dsca = 1.0 / sqrtpiv;
  absmax = nrmu2 = 0.0;
  for (l = 0; l < n1; ++l) {
    const double eabs = fabs(pu[l] *= dsca);
    if (Z[l] >= 0 && eabs > 1e-8 * scale * sqrtpiv) ++Z[l];
    nrmu2 += eabs * eabs;
    if (Z[l] >= 0 && eabs > absmax) { *i0 = l; absmax = eabs; }
  }
  *nrmlsk2 = nrmu2 * nrmv2;
  return 1;
}

/* ACA_col_step, Include/ACA.h:134-213 */
static int aca_col_step(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                        unsigned* j0, unsigned k, unsigned* no, int* Z, int* S, double* U,
                        double* V, double* nrmlsk2, double scale)
{
  double nrmu2, nrmv2, absmax;
  double* const pu = U + (size_t)k * n1;
  double* const pv = V + (size_t)k * n2;
  unsigned l, i0 = 0;
  do {
    G->cmpbl(G->ctx, b1, n1, b2 + *j0, 1, pu);
    for (l = 0; l < k; ++l) axpy(n1, -V[(size_t)l * n2 + *j0], U + (size_t)l * n1, pu);
    absmax = 0.0;
    for (l = 0; l < n1; ++l)
      if (Z[l] >= 0) {
        const double eabs = fabs(pu[l]);
        if (eabs > absmax) { absmax = eabs; i0 = l; }
      }
    if (absmax < 1e-14 * scale) {
      S[*j0] = -1;
      if (++(*no) < n2) *j0 = nexti(n2, S, *j0);
    }
  } while (absmax < 1e-14 * scale && *no < n2);
  if (absmax < 1e-14 * scale) return 0;

  const double sqrtpiv = sqrt(absmax);
  nrmu2 = 0.0;
  const double sca = sqrtpiv / pu[i0];
  for (l = 0; l < n1; ++l)
    if (Z[l] >= 0) {
      const double e = (pu[l] *= sca);
      nrmu2 += e * e;
      if (fabs(e) > 1e-8 * scale * sqrtpiv) ++Z[l];
    }
  G->cmpbl(G->ctx, b1 + i0, 1, b2, n2, pv);
  for (l = 0; l < k; ++l) axpy(n2, -U[(size_t)l * n1 + i0], V + (size_t)l * n2, pv);
  Z[i0] = S[*j0] = -1;
  ++(*no);
  const double dsca = 1.0 / sqrtpiv;
  absmax = nrmv2 = 0.0;
  for (l = 0; l < n2; ++l) {
    const double eabs = fabs(pv[l] *= dsca);
    if (S[l] >= 0 && eabs > 1e-8 * scale * sqrtpiv) ++S[l];
    nrmv2 += eabs * eabs;
    if (eabs > absmax) { *j0 = l; absmax = eabs; }
  }
  *nrmlsk2 = nrmu2 * nrmv2;
  return 1;
}

/* ACA, Include/ACA.h:217-350.  U,V preallocated (kmax+1)*n1, (kmax+1)*n2. */
static int aca(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
               double eps, unsigned kmax, unsigned i0, unsigned* kout, double* U, double* V)
{
  unsigned l, no = 0, j0, inext, klast = 0, k;
  int kmin = -2, keps = -1, ok, result;
  double nrmlsk2 = 0.0, nrms2 = 0.0, sum;
  double scale = G->scale(G->ctx, b1, n1, b2, n2);
  int* Z = (int*)xmalloc(sizeof(int) * n1);
  int* S = (int*)xmalloc(sizeof(int) * n2);
  for (l = 0; l < n1; ++l) Z[l] = 0;
  for (l = 0; l < n2; ++l) S[l] = 0;
  k = 0;

  do {
    inext = i0;
    ok = aca_row_step(G, b1, n1, b2, n2, &klast, &inext, k, &no, Z, S, U, V, &nrmlsk2, scale);
    if (!ok) { nrmlsk2 = 0.0; goto SUCC; }
    sum = 0.0;
    for (l = 0; l < k; ++l)
      sum += scpr(n1, U + (size_t)l * n1, U + (size_t)k * n1) *
             scpr(n2, V + (size_t)l * n2, V + (size_t)k * n2);
    nrms2 += 2.0 * sum + nrmlsk2;
    if (nrmlsk2 < eps * eps * nrms2 && keps == -1) keps = (int)(klast + 1u);
    ++k;
    scale = sqrt(nrmlsk2 / ((double)n1 * n2));
    if (keps > 0) {
      for (i0 = 0; i0 < n1 && Z[i0] < 0; ++i0) ;
      if (i0 == n1) { nrmlsk2 = 0.0; goto SUCC; }
      kmin = Z[i0];
      for (unsigned i = i0 + 1; i < n1; ++i)
        if (Z[i] >= 0 && Z[i] < kmin) kmin = Z[i0 = i];
      if ((unsigned)kmin == k) i0 = inext;
    } else i0 = inext;
  } while (no < n1 && k < kmax && kmin < keps);

  if ((no == n1 || k == kmax) && kmin < keps) goto NSUCC;

  do {
    for (j0 = 0; j0 < n2 && S[j0] < 0; ++j0) ;
    if (j0 == n2) { nrmlsk2 = 0.0; goto SUCC; }
    kmin = S[j0];
    for (unsigned j = j0 + 1; j < n2; ++j)
      if (S[j] >= 0 && S[j] < kmin) kmin = S[j0 = j];
    if (kmin >= keps) goto SUCC;
    ok = aca_col_step(G, b1, n1, b2, n2, &j0, k, &no, Z, S, U, V, &nrmlsk2, scale);
    if (!ok) { nrmlsk2 = 0.0; goto SUCC; }
    ++k;
  } while (no < n2 && k < kmax);

NSUCC:
  result = 0;
  goto DONE;
SUCC:
  {
    const double epsa = sqrt(nrmlsk2 / nrms2);
    if (epsa >= eps) goto NSUCC;   /* NaN (0/0) compares false: success */
  }
  result = 1;
DONE:
  free(S);
  free(Z);
  *kout = k;
  return result;
}

/* ACAr_step, Include/ACA.h:578-655: row pivoting only (no column
 * counters; every column is eligible), the cross scaled by sqrt|pivot| */
static int acar_step(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                     unsigned* i0, unsigned k, unsigned* no, int* Z, double* U, double* V,
                     double* nrmlsk2, double scale)
{
  unsigned l, j0 = 0;
  double nrmu2, nrmv2, absmax;
  double* const pu = U + (size_t)k * n1;
  double* const pv = V + (size_t)k * n2;
  do {
    G->cmpbl(G->ctx, b1 + *i0, 1, b2, n2, pv);
    for (l = 0; l < k; ++l) axpy(n2, -U[(size_t)l * n1 + *i0], V + (size_t)l * n2, pv);
    absmax = 0.0;
    for (l = 0; l < n2; ++l) {
      const double eabs = fabs(pv[l]);
      if (eabs > absmax) { absmax = eabs; j0 = l; }
    }
    if (absmax < 1e-14 * scale) {
      Z[*i0] = -1;
      if (++(*no) < n1) *i0 = nexti(n1, Z, *i0);
    }
  } while (absmax < 1e-14 * scale && *no < n1);
  if (absmax < 1e-14 * scale) return 0;

  const double sqrtpiv = sqrt(absmax);
  nrmv2 = 0.0;
  const double sca = sqrtpiv / pv[j0];
  for (l = 0; l < n2; ++l) {
    const double e = (pv[l] *= sca);
    nrmv2 += e * e;
  }
  G->cmpbl(G->ctx, b1, n1, b2 + j0, 1, pu);
  for (l = 0; l < k; ++l) axpy(n1, -V[(size_t)l * n2 + j0], U + (size_t)l * n1, pu);
  Z[*i0] = -1;
  ++(*no);
  const double dsca = 1.0 / sqrtpiv;
  absmax = nrmu2 = 0.0;
  for (l = 0; l < n1; ++l) {
    const double eabs = fabs(pu[l] *= dsca);
    nrmu2 += eabs * eabs;
    if (Z[l] >= 0 && eabs > absmax) { *i0 = l; absmax = eabs; }
  }
  *nrmlsk2 = nrmu2 * nrmv2;
  return 1;
}

/* ACAr, Include/ACA.h:657-720.  The loop test compares the number of rows
 * tried with n2, as the reference does. */
static int acar(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                double eps, unsigned kmax, unsigned i0, unsigned* kout, double* U, double* V)
{
  unsigned l, no = 0, k = 0;
  double nrmlsk2 = 0.0, nrms2 = 0.0, sum;
  double scale = G->scale(G->ctx, b1, n1, b2, n2);
  int* Z = (int*)xmalloc(sizeof(int) * n1);
  int result;
  for (l = 0; l < n1; ++l) Z[l] = 0;
  do {
    if (!acar_step(G, b1, n1, b2, n2, &i0, k, &no, Z, U, V, &nrmlsk2, scale)) {
      nrmlsk2 = 0.0;
      break;
    }
    sum = 0.0;
    for (l = 0; l < k; ++l)
      sum += scpr(n1, U + (size_t)l * n1, U + (size_t)k * n1) *
             scpr(n2, V + (size_t)l * n2, V + (size_t)k * n2);
    nrms2 += 2.0 * sum + nrmlsk2;
    ++k;
    scale = sqrt(nrmlsk2 / ((double)n1 * n2));
  } while (no < n2 && k < kmax && nrmlsk2 >= eps * eps * nrms2);
  {
    const double epsa = sqrt(nrmlsk2 / nrms2);
    result = !(epsa >= eps);   /* NaN (0/0) compares false: success */
  }
  free(Z);
  *kout = k;
  return result;
}

int or_acar_dense(const double* A, unsigned n1, unsigned n2, double eps, unsigned kmax,
                  unsigned i0, double scale, unsigned* k, double* U, double* V);

/* scale callback returning a fixed value (used by or_aca_dense) */
static __thread double tl_scale;
static double fixed_scale(void* c, unsigned a, unsigned b, unsigned d, unsigned e)
{
  (void)c; (void)a; (void)b; (void)d; (void)e;
  return tl_scale;
}

int or_aca_dense(const double* A, unsigned n1, unsigned n2, double eps, unsigned kmax,
                 unsigned i0, double scale, unsigned* k, double* U, double* V)
{
  blas_init();
  dgen_ctx c = {A, n1};
  matgen G = {&c, dgen_cmpbl, fixed_scale};
  tl_scale = scale;
  return aca(&G, 0, n1, 0, n2, eps, kmax, i0, k, U, V);
}

int or_acar_dense(const double* A, unsigned n1, unsigned n2, double eps, unsigned kmax,
                  unsigned i0, double scale, unsigned* k, double* U, double* V)
{
  blas_init();
  dgen_ctx c = {A, n1};
  matgen G = {&c, dgen_cmpbl, fixed_scale};
  tl_scale = scale;
  return acar(&G, 0, n1, 0, n2, eps, kmax, i0, k, U, V);
}

/* ======================================================================== */
/* Recompression: cpyLrM_cmpr -> addtrll, H/mblock_D.cpp:357-465,            */
/* Include/mblock.h:822-826 (existing rank 0, haarInfo NULL)                */
/* ======================================================================== */

#define EPS0 1e-64

/* returns new rank; *out = malloc'd [U (n1 x kt) | V (n2 x kt)] or NULL */
static unsigned addtrll(unsigned n1, unsigned n2, unsigned k, const double* U, unsigned ldU,
                        const double* V, unsigned ldV, double delta, unsigned kgoal, double** out)
{
  *out = NULL;
  if (k == 0) return 0;
  const unsigned ksum = k, LWORK = 10 * ksum;
  const unsigned mmin = n1 < ksum ? n1 : ksum, nmin = n2 < ksum ? n2 : ksum;
  const unsigned amin = mmin < nmin ? mmin : nmin;
  const size_t size = (size_t)ksum * (n1 + n2) + LWORK + mmin + (size_t)(mmin + amin + 1) * nmin + amin;
  double* tmp1 = (double*)xmalloc(sizeof(double) * size);
  double* tmp2 = tmp1 + (size_t)ksum * n1;
  for (unsigned l = 0; l < k; ++l) copyv(n1, U + (size_t)l * ldU, tmp1 + (size_t)l * n1);
  for (unsigned l = 0; l < k; ++l) copyv(n2, V + (size_t)l * ldV, tmp2 + (size_t)l * n2);
  double* const WORK = tmp2 + (size_t)n2 * ksum;
  double* const tau1 = WORK + LWORK;
  double* const tau2 = tau1 + mmin;
  int INFO, m_ = (int)n1, n_ = (int)ksum, lw = (int)LWORK;
  dgeqrf_(&m_, &n_, tmp1, &m_, tau1, WORK, &lw, &INFO);
  m_ = (int)n2;
  dgeqrf_(&m_, &n_, tmp2, &m_, tau2, WORK, &lw, &INFO);
  double* const R = tau2 + nmin;
  /* utrmmh, Include/blas.h:2767-2781 */
  for (unsigned j = 0; j < nmin; ++j)
    for (unsigned i = 0; i < mmin; ++i) {
      double d = 0.0;
      const unsigned ij = i > j ? i : j;
      for (unsigned l = ij; l < ksum; ++l) d += tmp1[i + (size_t)l * n1] * tmp2[j + (size_t)l * n2];
      R[i + (size_t)j * mmin] = d;
    }
  const double thresh = delta * nrm2(mmin * nmin, R) / sqrtf((float)mmin * nmin);
  for (unsigned i = 0; i < mmin; ++i)
    for (unsigned j = 0; j < nmin; ++j)
      if (fabs(R[i + (size_t)j * mmin]) < thresh) R[i + (size_t)j * mmin] = 0.0;
  double* const S = R + (size_t)mmin * nmin;
  double* const VT = S + amin;
  int mm = (int)mmin, nn = (int)nmin, ldvt = (int)amin;
  dgesvd_("O", "S", &mm, &nn, R, &mm, S, R, &mm, VT, &ldvt, WORK, &lw, &INFO);
  unsigned kt = amin < kgoal ? amin : kgoal;
  while (kt > 0 && (S[kt - 1] <= delta * S[0] || S[kt - 1] < EPS0)) --kt;
  if (kt > 0) {
    double* data = (double*)xmalloc(sizeof(double) * (size_t)kt * (n1 + n2));
    for (unsigned j = 0; j < kt; ++j) {
      unsigned i;
      for (i = 0; i < mmin; ++i) data[i + (size_t)j * n1] = S[j] * R[i + (size_t)j * mmin];
      for (; i < n1; ++i) data[i + (size_t)j * n1] = 0.0;
    }
    double* dataV = data + (size_t)kt * n1;
    for (unsigned j = 0; j < kt; ++j) {
      unsigned i;
      for (i = 0; i < nmin; ++i) dataV[i + (size_t)j * n2] = VT[(size_t)i * amin + j];
      for (; i < n2; ++i) dataV[i + (size_t)j * n2] = 0.0;
    }
    int kk = (int)kt;
    m_ = (int)n1; int p_ = (int)mmin;
    dormqr_("L", "N", &m_, &kk, &p_, tmp1, &m_, tau1, data, &m_, WORK, &lw, &INFO);
    m_ = (int)n2; p_ = (int)nmin;
    dormqr_("L", "N", &m_, &kk, &p_, tmp2, &m_, tau2, dataV, &m_, WORK, &lw, &INFO);
    *out = data;
  }
  free(tmp1);
  return kt;
}

void or_addtrll(unsigned n1, unsigned n2, unsigned k, const double* U, const double* V,
                double delta, unsigned kgoal, unsigned* kt, double* Uo, double* Vo)
{
  blas_init();
  double* d = NULL;
  *kt = addtrll(n1, n2, k, U, n1, V, n2, delta, kgoal, &d);
  if (*kt) {
    memcpy(Uo, d, sizeof(double) * (size_t)n1 * *kt);
    memcpy(Vo, d + (size_t)n1 * *kt, sizeof(double) * (size_t)n2 * *kt);
  }
  free(d);
}

/* ======================================================================== */
/* H-matrix container                                                        */
/* ======================================================================== */

struct or_hmat {
  unsigned N, nl;
  unsigned *b1, *n1, *b2, *n2;
  int* rank;           /* -1 dense, OR_HEM packed Hermitian diagonal leaf */
  size_t* off;
  size_t nvals;
  double* data;
  int sym;             /* symmetric: leaves on and above the diagonal only */
  unsigned nsched;     /* mltaHeHVec_ visiting order (or_tree.sched), or 0 */
  unsigned* sched;
};

static or_hmat* hmat_alloc(unsigned N, unsigned nl)
{
  or_hmat* h = (or_hmat*)calloc(1, sizeof(or_hmat));
  h->N = N; h->nl = nl;
  h->b1 = (unsigned*)xmalloc(sizeof(unsigned) * nl);
  h->n1 = (unsigned*)xmalloc(sizeof(unsigned) * nl);
  h->b2 = (unsigned*)xmalloc(sizeof(unsigned) * nl);
  h->n2 = (unsigned*)xmalloc(sizeof(unsigned) * nl);
  h->rank = (int*)xmalloc(sizeof(int) * nl);
  h->off = (size_t*)xmalloc(sizeof(size_t) * (nl + 1));
  return h;
}

static size_t leaf_nvals(unsigned n1, unsigned n2, int rank)
{
  if (rank == OR_HEM) return (size_t)n1 * (n1 + 1) / 2;
  return rank < 0 ? (size_t)n1 * n2 : (size_t)rank * (n1 + n2);
}

/* apprx_unsym, Include/apprx.h:136-172; method 1 uses ACAr as apprx_sym
 * does (Include/apprx.h:87-131) */
static double* apprx_unsym(const matgen* G, unsigned b1, unsigned n1, unsigned b2, unsigned n2,
                           int adm, unsigned i0, double eps, unsigned rankmax, int method, int* rank)
{
  /* 64-bit product: the reference's 32-bit n1*n2 wraps for n1*n2 >= 2^32 */
  unsigned maxk = (unsigned)((unsigned long long)n1 * n2 / (n1 + n2) + 4);
  if (maxk > rankmax) maxk = rankmax;
  if (adm && maxk) {
    unsigned k;
    double* U = (double*)xmalloc(sizeof(double) * (size_t)(maxk + 1) * n1);
    double* V = (double*)xmalloc(sizeof(double) * (size_t)(maxk + 1) * n2);
    int succ = method == 1 ? acar(G, b1, n1, b2, n2, eps, maxk, i0, &k, U, V)
                           : aca(G, b1, n1, b2, n2, eps, maxk, i0, &k, U, V);
    double* out = NULL;
    if (succ) *rank = (int)addtrll(n1, n2, k, U, n1, V, n2, eps, k, &out);
    free(U);
    free(V);
    if (succ) return out;
  }
  double* D = (double*)xmalloc(sizeof(double) * (size_t)n1 * n2);
  G->cmpbl(G->ctx, b1, n1, b2, n2, D);
  *rank = -1;
  return D;
}

or_hmat* or_hmat_assemble(const or_tree* t, const double* pts, int kernel, double kappa,
                          double weight, double eps, unsigned rankmax, int nthreads)
{
  return or_hmat_assemble_method(t, pts, kernel, kappa, weight, eps, rankmax, nthreads, 0);
}

or_hmat* or_hmat_assemble_method(const or_tree* t, const double* pts, int kernel, double kappa,
                                 double weight, double eps, unsigned rankmax, int nthreads,
                                 int method)
{
  blas_init();
  (void)pts;
  or_hmat* h = hmat_alloc(t->N, t->nl);
  double** parts = (double**)xmalloc(sizeof(double*) * t->nl);
  kgen_ctx kc = {kernel, kappa, weight, t->pts, t->op};
  matgen G = {&kc, kgen_cmpbl, kgen_scale};
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
  for (int l = 0; l < (int)t->nl; ++l) {
    const unsigned i0 = t->cl[t->rcl[l]].icom - t->b1[l];
    /* apprx_sym (Include/apprx.h:87-131): ACAr, and the diagonal blocks
       that stay dense are stored packed Hermitian (setHeM + cmpblsym) */
    parts[l] = apprx_unsym(&G, t->b1[l], t->n1[l], t->b2[l], t->n2[l], t->adm[l], i0, eps,
                           rankmax, t->sym ? 1 : method, &h->rank[l]);
    if (t->sym && h->rank[l] < 0 && t->b1[l] == t->b2[l]) {
      const unsigned n = t->n1[l];
      double* Pk = (double*)xmalloc(sizeof(double) * ((size_t)n * (n + 1) / 2 + 1));
      for (unsigned j = 0; j < n; ++j)
        for (unsigned i = 0; i <= j; ++i) Pk[i + (size_t)j * (j + 1) / 2] = parts[l][i + (size_t)j * n];
      free(parts[l]);
      parts[l] = Pk;
      h->rank[l] = OR_HEM;
    }
  }
  if (t->sym) {
    h->sym = 1;
    h->nsched = t->nsched;
    h->sched = (unsigned*)xmalloc(sizeof(unsigned) * 3 * (size_t)(t->nsched + 1));
    memcpy(h->sched, t->sched, sizeof(unsigned) * 3 * (size_t)t->nsched);
  }
  size_t off = 0;
  for (unsigned l = 0; l < t->nl; ++l) {
    h->b1[l] = t->b1[l]; h->n1[l] = t->n1[l]; h->b2[l] = t->b2[l]; h->n2[l] = t->n2[l];
    h->off[l] = off;
    off += leaf_nvals(h->n1[l], h->n2[l], h->rank[l]);
  }
  h->off[t->nl] = off;
  h->nvals = off;
  h->data = (double*)xmalloc(sizeof(double) * off);
  for (unsigned l = 0; l < t->nl; ++l) {
    const size_t nv = h->off[l + 1] - h->off[l];
    if (nv) memcpy(h->data + h->off[l], parts[l], sizeof(double) * nv);
    free(parts[l]);
  }
  free(parts);
  return h;
}

or_hmat* or_hmat_from_leaves(unsigned N, unsigned nl, const unsigned* b1, const unsigned* n1,
                             const unsigned* b2, const unsigned* n2, const int* rank,
                             const size_t* offset, const double* data)
{
  blas_init();
  or_hmat* h = hmat_alloc(N, nl);
  size_t off = 0;
  for (unsigned l = 0; l < nl; ++l) {
    h->b1[l] = b1[l]; h->n1[l] = n1[l]; h->b2[l] = b2[l]; h->n2[l] = n2[l];
    h->rank[l] = rank[l];
    h->off[l] = off;
    off += leaf_nvals(n1[l], n2[l], rank[l]);
    if (rank[l] == OR_HEM) h->sym = 1;
  }
  h->off[nl] = off;
  h->nvals = off;
  h->data = (double*)xmalloc(sizeof(double) * off);
  for (unsigned l = 0; l < nl; ++l) {
    const size_t nv = h->off[l + 1] - h->off[l];
    if (nv) memcpy(h->data + h->off[l], data + offset[l], sizeof(double) * nv);
  }
  return h;
}

void or_hmat_free(or_hmat* h)
{
  if (!h) return;
  free(h->b1); free(h->n1); free(h->b2); free(h->n2); free(h->rank); free(h->off);
  free(h->data);
  free(h->sched);
  free(h);
}

unsigned or_hmat_nleaves(const or_hmat* h) { return h->nl; }
size_t or_hmat_nvals(const or_hmat* h) { return h->nvals; }
const double* or_hmat_data(const or_hmat* h) { return h->data; }
void or_hmat_leaves(const or_hmat* h, unsigned* b1, unsigned* n1, unsigned* b2, unsigned* n2,
                    int* rank, size_t* offset)
{
  for (unsigned l = 0; l < h->nl; ++l) {
    b1[l] = h->b1[l]; n1[l] = h->n1[l]; b2[l] = h->b2[l]; n2[l] = h->n2[l];
    rank[l] = h->rank[l]; offset[l] = h->off[l];
  }
}

/* ======================================================================== */
/* Matvec                                                                   */
/* ======================================================================== */

/* mblock::mltaVec_ -> mltaLrMVec_ / mltaGeMVec (H/mblock.cpp:125-138,218-230;
 * Include/mblock.h:98-102) */
static int leaf_vec(const or_hmat* h, unsigned l, double d, const double* x, double* y)
{
  const unsigned n1 = h->n1[l], n2 = h->n2[l];
  const double* data = h->data + h->off[l];
  const double* xp = x + h->b2[l];
  double* yp = y + h->b1[l];
  if (h->rank[l] >= 0) {
    const unsigned k = (unsigned)h->rank[l];
    if (!k) return 0;
    for (unsigned q = 0; q < k; ++q) {
      const double e = d * scpr(n2, data + (size_t)k * n1 + (size_t)q * n2, xp);
      axpy(n1, e, data + (size_t)q * n1, yp);
    }
    return 1;
  }
  if (h->rank[l] == OR_HEM) {
    /* mltaHeMVec: blas::hemva = dspmv 'U' (Include/mblock.h:137-141,
       Include/blas.h:1311-1315) */
    int nn = (int)n1;
    dspmv_("U", &nn, &d, data, xp, &I1, &DONE, yp, &I1);
    return 1;
  }
  gemva(n1, n2, d, data, xp, yp);
  return 1;
}

static void leaf_vec_t(const or_hmat* h, unsigned l, double d, const double* x, double* y);

/* y += d A x for a symmetric H-matrix, mltaHeHVec_ (H/mltaGeHVec.cpp:
   420-446): every stored leaf forward, and the leaves off the diagonal also
   transposed.  With the tree's schedule the leaf ranges come in the
   recursion's order (forward over a son block's leaves, then transposed over
   the same leaves); without it (matrices built from leaf lists) all leaves
   forward, then the off-diagonal ones transposed. */
static void matvec_sym(const or_hmat* h, double d, const double* x, double* y)
{
  if (h->nsched) {
    for (unsigned s = 0; s < h->nsched; ++s) {
      const unsigned* e = h->sched + 3 * (size_t)s;
      for (unsigned l = e[1]; l < e[2]; ++l) {
        if (e[0] == 0) leaf_vec(h, l, d, x, y);
        else leaf_vec_t(h, l, d, x, y);
      }
    }
    return;
  }
  for (unsigned l = 0; l < h->nl; ++l) leaf_vec(h, l, d, x, y);
  for (unsigned l = 0; l < h->nl; ++l)
    if (h->b1[l] != h->b2[l]) leaf_vec_t(h, l, d, x, y);
}

/* mltaGeHVec_ (H/mltaGeHVec.cpp:16-35): the DFS recursion visits the leaves
 * in idx order, which is the order they are stored in here. */
int or_hmat_matvec(const or_hmat* h, double d, const double* x, double* y)
{
  blas_init();
  if (h->sym) {
    matvec_sym(h, d, x, y);
    return 1;
  }
  int changed = 0;
  for (unsigned l = 0; l < h->nl; ++l) changed |= leaf_vec(h, l, d, x, y);
  return changed;
}

/* y += d A^T x, the real case of mltaGeHhVec_ (H/mltaGeHVec.cpp:61-81):
 * leaves in idx order; low rank: e = d <U_l, x> then y += e V_l
 * (mltaLrMhVec_, H/mblock.cpp:159-170); dense: dgemv 'T' (mltaGeMhVec,
 * Include/mblock.h:112-115).  x is indexed by rows, y by columns. */
static void leaf_vec_t(const or_hmat* h, unsigned l, double d, const double* x, double* y)
{
  const unsigned n1 = h->n1[l], n2 = h->n2[l];
  const double* data = h->data + h->off[l];
  const double* xp = x + h->b1[l];
  double* yp = y + h->b2[l];
  if (h->rank[l] >= 0) {
    const unsigned k = (unsigned)h->rank[l];
    for (unsigned q = 0; q < k; ++q) {
      const double e = d * scpr(n1, data + (size_t)q * n1, xp);
      axpy(n2, e, data + (size_t)k * n1 + (size_t)q * n2, yp);
    }
  } else if (h->rank[l] == OR_HEM) {
    int nn = (int)n1;   /* symmetric block: A^T = A */
    dspmv_("U", &nn, &d, data, xp, &I1, &DONE, yp, &I1);
  } else {
    int mm = (int)n1, nn = (int)n2;
    dgemv_("T", &mm, &nn, &d, data, &mm, xp, &I1, &DONE, yp, &I1);
  }
}

int or_hmat_matvec_t(const or_hmat* h, double d, const double* x, double* y)
{
  blas_init();
  if (h->sym) {   /* a symmetric matrix is its own transpose */
    matvec_sym(h, d, x, y);
    return 1;
  }
  int changed = 0;
  for (unsigned l = 0; l < h->nl; ++l) {
    if (h->rank[l] == 0) continue;
    leaf_vec_t(h, l, d, x, y);
    changed = 1;
  }
  return changed;
}

/* mltaGeHGeM_ (H/mltaGeHVec.cpp:38-59), mltaLrMGeM_ (H/mblock.cpp:140-154),
 * mltaGeMGeM (Include/mblock.h:104-109) */
int or_hmat_matmat(const or_hmat* h, double d, unsigned p, const double* X, unsigned ldX,
                   double* Y, unsigned ldY)
{
  blas_init();
  if (h->sym) {   /* mltaHeHGeM_ (H/mltaGeHVec.cpp:448-472), column by column */
    for (unsigned c = 0; c < p; ++c) matvec_sym(h, d, X + (size_t)c * ldX, Y + (size_t)c * ldY);
    return 1;
  }
  int changed = 0;
  const int pp = (int)p, lx = (int)ldX, ly = (int)ldY;
  for (unsigned l = 0; l < h->nl; ++l) {
    const int n1 = (int)h->n1[l], n2 = (int)h->n2[l];
    const double* data = h->data + h->off[l];
    const double* Xp = X + h->b2[l];
    double* Yp = Y + h->b1[l];
    if (h->rank[l] >= 0) {
      const int k = h->rank[l];
      if (!k) continue;
      double* tmp = (double*)xmalloc(sizeof(double) * (size_t)k * p);
      dgemm_("T", "N", &k, &pp, &n2, &DONE, data + (size_t)k * n1, &n2, Xp, &lx, &DZERO, tmp, &k);
      dgemm_("N", "N", &n1, &pp, &k, &d, data, &n1, tmp, &k, &DONE, Yp, &ly);
      free(tmp);
    } else {
      dgemm_("N", "N", &n1, &pp, &n2, &d, data, &n1, Xp, &lx, &DONE, Yp, &ly);
    }
    changed = 1;
  }
  return changed;
}

int or_hmat_matvec_omp(const or_hmat* h, double d, const double* x, double* y, int nthreads)
{
  blas_init();
  if (h->sym) {
    matvec_sym(h, d, x, y);
    return 1;
  }
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const unsigned N = h->N;
  double* ybuf = (double*)calloc((size_t)nthreads * N, sizeof(double));
  int changed = 0;
  /* contiguous leaf intervals of equal stored-value count per thread */
  unsigned* cut = (unsigned*)xmalloc(sizeof(unsigned) * (nthreads + 1));
  cut[0] = 0;
  {
    unsigned l = 0;
    for (int t = 1; t < nthreads; ++t) {
      const size_t target = h->nvals / nthreads * (size_t)t;
      while (l < h->nl && h->off[l] < target) ++l;
      cut[t] = l;
    }
    cut[nthreads] = h->nl;
  }
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(| : changed)
#endif
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double* yt = ybuf + (size_t)t * N;
    for (unsigned l = cut[t]; l < cut[t + 1]; ++l) changed |= leaf_vec(h, l, d, x, yt);
  }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
  for (long i = 0; i < (long)N; ++i) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += ybuf[(size_t)t * N + i];
    y[i] += s;
  }
  free(cut);
  free(ybuf);
  return changed;
}

/* ======================================================================== */
/* Krylov solvers                                                            */
/* ======================================================================== */

typedef struct {
  unsigned n;
  void* ctx;
  void (*amux)(void*, double, const double*, double*);
} op_t;

static void hmat_amux(void* c, double d, const double* x, double* y)
{ or_hmat_matvec((const or_hmat*)c, d, x, y); }

typedef struct { unsigned n; const double* A; } dense_op;
static void dense_amux(void* c, double d, const double* x, double* y)
{ dense_op* o = (dense_op*)c; gemva(o->n, o->n, d, o->A, x, y); }

/* genPlRot / applPlRot, solvers/GMRES.cpp:33-55 */
static void genPlRot(double dx, double dy, double* cs, double* sn)
{
  if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) {
    const double tmp = dx / dy;
    *sn = 1.0 / sqrt(1.0 + tmp * tmp);
    *cs = tmp * *sn;
  } else {
    const double tmp = dy / dx;
    *cs = 1.0 / sqrt(1.0 + tmp * tmp);
    *sn = tmp * *cs;
  }
}
static void applPlRot(double* dx, double* dy, double cs, double sn)
{
  const double tmp = cs * *dx + sn * *dy;
  *dy = cs * *dy - sn * *dx;
  *dx = tmp;
}

/* update, solvers/GMRES.cpp:58-78 (no preconditioner) */
static void gmres_update(unsigned n, unsigned k, const double* H, unsigned ldH, const double* s,
                         const double* V, double* x)
{
  double* y = (double*)xmalloc(sizeof(double) * k);
  double* xh = (double*)calloc(n, sizeof(double));
  copyv(k, s, y);
  int kk = (int)k, one = 1, ld = (int)ldH, inf;
  dtrtrs_("U", "N", "N", &kk, &one, H, &ld, y, &kk, &inf);
  gemva(n, k, 1.0, V, y, xh);
  for (unsigned i = 0; i < n; ++i) x[i] += xh[i];
  free(xh);
  free(y);
}

/* GMRes, solvers/GMRES.cpp:80-181.  When nsteps ends a cycle early, the
 * update uses the i computed columns (the reference passes m and reads
 * unset columns). */
static unsigned gmres(const op_t* A, const double* b, double* x, double* eps, unsigned m,
                      unsigned* nsteps, double* hist, unsigned nhist)
{
  const unsigned n = A->n;
  double resid;
  unsigned j = 1, nh = 0;
  double* r = (double*)xmalloc(sizeof(double) * (2 * (size_t)n + (size_t)(n + m + 4) * (m + 1)));
  double* V = r + n;
  double* H = V + (size_t)n * (m + 1);
  double* cs = H + (size_t)(m + 1) * m;
  double* sn = cs + m + 1;
  double* s = sn + m + 1;
  double* xh = s + m + 1;
  const double normb = nrm2(n, b);
  if (normb == 0.0) {
    memset(x, 0, sizeof(double) * n);
    *eps = 0.0; *nsteps = 0; free(r); return 0;
  }
  copyv(n, b, r);
  A->amux(A->ctx, -1.0, x, r);
  double beta = nrm2(n, r);
  if ((resid = beta / normb) <= *eps) { *eps = resid; *nsteps = 0; free(r); return 0; }
  while (j <= *nsteps) {
    copyv(n, r, V);
    scal(n, 1.0 / beta, V);
    s[0] = beta;
    memset(s + 1, 0, sizeof(double) * m);
    unsigned i;
    for (i = 0; i < m && j <= *nsteps; i++, j++) {
      copyv(n, V + (size_t)i * n, xh);
      memset(V + (size_t)(i + 1) * n, 0, sizeof(double) * n);
      A->amux(A->ctx, 1.0, xh, V + (size_t)(i + 1) * n);
      for (unsigned k = 0; k <= i; k++) {
        H[k + i * (m + 1)] = scpr(n, V + (size_t)(i + 1) * n, V + (size_t)k * n);
        axpy(n, -H[k + i * (m + 1)], V + (size_t)k * n, V + (size_t)(i + 1) * n);
      }
      H[i * (m + 2) + 1] = nrm2(n, V + (size_t)(i + 1) * n);
      scal(n, 1.0 / H[i * (m + 2) + 1], V + (size_t)(i + 1) * n);
      for (unsigned k = 0; k < i; k++) applPlRot(&H[k + i * (m + 1)], &H[k + 1 + i * (m + 1)], cs[k], sn[k]);
      genPlRot(H[i * (m + 2)], H[i * (m + 2) + 1], &cs[i], &sn[i]);
      applPlRot(&H[i * (m + 2)], &H[i * (m + 2) + 1], cs[i], sn[i]);
      applPlRot(&s[i], &s[i + 1], cs[i], sn[i]);
      resid = fabs(s[i + 1] / normb);
      if (hist && nh < nhist) hist[nh++] = resid;
      if (resid < *eps) {
        gmres_update(n, i + 1, H, m + 1, s, V, x);
        *eps = resid; *nsteps = j; free(r); return 0;
      }
    }
    gmres_update(n, i, H, m + 1, s, V, x);
    copyv(n, b, r);
    A->amux(A->ctx, -1.0, x, r);
    beta = nrm2(n, r);
    if ((resid = beta / normb) < *eps) { *eps = resid; *nsteps = j; free(r); return 0; }
  }
  *eps = resid;
  free(r);
  return 1;
}

/* CG, solvers/CG.cpp:31-113 (no preconditioner) */
static unsigned cg(const op_t* A, const double* b, double* x, double* eps, unsigned* nsteps)
{
  const unsigned N = A->n;
  double rho, rho1 = 0.0;
  double* p = (double*)xmalloc(sizeof(double) * 4 * (size_t)N);
  double *q = p + N, *r = q + N, *rhat = r + N;
  const double nrmb = nrm2(N, b);
  if (nrmb < 1e-16) {
    memset(x, 0, sizeof(double) * N);
    *eps = 0.0; *nsteps = 0; free(p); return 0;
  }
  copyv(N, b, r);
  A->amux(A->ctx, -1.0, x, r);
  double resid = nrm2(N, r);
  if (resid <= *eps * nrmb) { *eps = resid / nrmb; *nsteps = 0; free(p); return 0; }
  for (unsigned l = 1; l <= *nsteps; ++l) {
    copyv(N, r, rhat);
    rho = scpr(N, r, rhat);
    if (l > 1) {
      const double beta = rho / rho1;
      scal(N, beta, p);
      axpy(N, 1.0, rhat, p);
    } else copyv(N, rhat, p);
    memset(q, 0, sizeof(double) * N);
    A->amux(A->ctx, 1.0, p, q);
    const double alpha = rho / scpr(N, p, q);
    axpy(N, alpha, p, x);
    axpy(N, -alpha, q, r);
    resid = nrm2(N, r);
    if (resid <= *eps * nrmb) { *eps = resid / nrmb; *nsteps = l; free(p); return 0; }
    rho1 = rho;
  }
  *eps = resid / nrmb;
  free(p);
  return 1;
}

unsigned or_gmres(const or_hmat* h, const double* b, double* x, double* eps, unsigned m,
                  unsigned* nsteps, double* hist, unsigned nhist)
{
  blas_init();
  op_t A = {h->N, (void*)h, hmat_amux};
  return gmres(&A, b, x, eps, m, nsteps, hist, nhist);
}

unsigned or_cg(const or_hmat* h, const double* b, double* x, double* eps, unsigned* nsteps)
{
  blas_init();
  op_t A = {h->N, (void*)h, hmat_amux};
  return cg(&A, b, x, eps, nsteps);
}

unsigned or_gmres_dense(unsigned n, const double* Am, const double* b, double* x, double* eps,
                        unsigned m, unsigned* nsteps)
{
  blas_init();
  dense_op o = {n, Am};
  op_t A = {n, &o, dense_amux};
  return gmres(&A, b, x, eps, m, nsteps, NULL, 0);
}

unsigned or_cg_dense(unsigned n, const double* Am, const double* b, double* x, double* eps,
                     unsigned* nsteps)
{
  blas_init();
  dense_op o = {n, Am};
  op_t A = {n, &o, dense_amux};
  return cg(&A, b, x, eps, nsteps);
}

/* ------------------------------------------------------------------------ */
/* Triangular H-solves (H/TH_solve.cpp:18-91, leaf solves H/mblock_D.cpp:
 * 1183-1205 and Include/mblock.h:694-697).                                  */

typedef struct { unsigned key, leaf; } or_key;

static int cmp_key(const void* a, const void* b)
{
  const or_key *x = (const or_key*)a, *y = (const or_key*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->leaf < y->leaf ? -1 : (x->leaf > y->leaf);
}

/* x[rows] -= leaf * x[cols] for one off-diagonal leaf (dense: dgemv as
 * Include/mblock.h:98-102; low rank: k dots then axpys, H/mblock.cpp:125-138) */
static void tri_apply(unsigned n1, unsigned n2, int kind, const double* d,
                      const double* xc, double* xr)
{
  if (kind >= 0) {
    const unsigned k = (unsigned)kind;
    for (unsigned q = 0; q < k; ++q) {
      const double e = -scpr(n2, d + (size_t)k * n1 + (size_t)q * n2, xc);
      axpy(n1, e, d + (size_t)q * n1, xr);
    }
  } else {
    const int mm = (int)n1, nn = (int)n2;
    const double mone = -1.0;
    dgemv_("N", &mm, &nn, &mone, d, &mm, xc, &I1, &DONE, xr, &I1);
  }
}

/* PL z = b on one diagonal leaf: z_i = b[ip_i], then unit lower forward
 * substitution on the packed columns (column j: entries j..n-1, entry j is
 * ip_j as a value) */
static void tri_diag_lower(unsigned n, const double* d, double* x, double* z)
{
  size_t o = 0;
  for (unsigned j = 0; j < n; ++j) {
    z[j] = x[(unsigned)d[o]];
    o += n - j;
  }
  o = 0;
  for (unsigned j = 0; j < n; ++j) {
    const double xj = z[j];
    for (unsigned i = j + 1; i < n; ++i) z[i] -= d[o + (i - j)] * xj;
    o += n - j;
  }
  memcpy(x, z, sizeof(double) * n);
}

/* U z = b on one diagonal leaf, packed upper columns (column j: rows 0..j
 * at offset j(j+1)/2), backward substitution */
static void tri_diag_upper(unsigned n, const double* d, double* x)
{
  for (unsigned j = n; j-- > 0;) {
    const size_t o = (size_t)j * (j + 1) / 2;
    x[j] /= d[o + j];
    const double xj = x[j];
    for (unsigned i = 0; i < j; ++i) x[i] -= d[o + i] * xj;
  }
}

/* x[cols] -= leaf^T x[rows] (the leaf products of mltaGeHhGeM,
 * H/mltaGeHVec.cpp:61-104, real case) */
static void tri_apply_t(unsigned n1, unsigned n2, int kind, const double* d,
                        const double* xr, double* xc)
{
  if (kind >= 0) {
    const unsigned k = (unsigned)kind;
    for (unsigned q = 0; q < k; ++q) {
      const double e = -scpr(n1, d + (size_t)q * n1, xr);
      axpy(n2, e, d + (size_t)k * n1 + (size_t)q * n2, xc);
    }
  } else {
    const int mm = (int)n1, nn = (int)n2;
    const double mone = -1.0;
    dgemv_("T", &mm, &nn, &mone, d, &mm, xr, &I1, &DONE, xc, &I1);
  }
}

/* U^T z = b on one diagonal leaf (dtpsv 'U','T','N', Include/blas.h:
 * 3962-3968): forward, column j of the packed upper triangle holds rows 0..j */
static void tri_diag_upper_t(unsigned n, const double* d, double* x)
{
  for (unsigned j = 0; j < n; ++j) {
    const size_t o = (size_t)j * (j + 1) / 2;
    double t = x[j];
    for (unsigned i = 0; i < j; ++i) t -= d[o + i] * x[i];
    x[j] = t / d[o + j];
  }
}

/* (PL)^T z = b on one diagonal leaf (mblock::ltrh_solve, H/mblock_D.cpp:
 * 1208-1230; dtpsv 'L','T','U'): unit upper backward substitution on a copy,
 * then scattered back through the permutation */
static void tri_diag_lower_t(unsigned n, const double* d, double* x, double* z)
{
  memcpy(z, x, sizeof(double) * n);
  size_t* oc = (size_t*)xmalloc(sizeof(size_t) * n);
  size_t o = 0;
  for (unsigned j = 0; j < n; ++j) {
    oc[j] = o;
    o += n - j;
  }
  for (unsigned j = n; j-- > 0;) {
    double t = z[j];
    for (unsigned i = j + 1; i < n; ++i) t -= d[oc[j] + (i - j)] * z[i];
    z[j] = t;
  }
  for (unsigned j = 0; j < n; ++j) x[(unsigned)d[oc[j]]] = z[j];
  free(oc);
}

int or_trih_solve(unsigned n, char uplo, unsigned nleaves, const unsigned* b1,
                  const unsigned* n1, const unsigned* b2, const unsigned* n2,
                  const int* kind, const size_t* offset, const double* data,
                  unsigned p, double* B, unsigned ldB)
{
  blas_init();
  /* factor stored lower ('L', 'l') or upper ('U', 'u'); lowercase solves
   * with its transpose; the sweep runs forward for L and U^T */
  const int stored_lower = uplo == 'L' || uplo == 'l';
  const int trans = uplo == 'l' || uplo == 'u';
  const int forward = stored_lower != trans;
  if (uplo != 'L' && uplo != 'U' && uplo != 'l' && uplo != 'u') return -1;
  or_key* dg = (or_key*)xmalloc(sizeof(or_key) * (nleaves + 1));
  or_key* od = (or_key*)xmalloc(sizeof(or_key) * (nleaves + 1));
  unsigned nd = 0, no = 0, maxn = 1;
  uint64_t covered = 0;
  for (unsigned l = 0; l < nleaves; ++l) {
    if (kind[l] == -2) continue;
    if (kind[l] == (stored_lower ? -3 : -4)) {
      if (b1[l] != b2[l] || n1[l] != n2[l]) goto bad;
      dg[nd].key = b1[l];
      dg[nd++].leaf = l;
      covered += n1[l];
      if (n1[l] > maxn) maxn = n1[l];
    } else if (kind[l] >= -1) {
      /* lower: columns end before the rows start; upper: the reverse */
      if (stored_lower ? (b2[l] + n2[l] > b1[l]) : (b1[l] + n1[l] > b2[l])) goto bad;
      /* the position whose solve makes the leaf's input final: forward
       * sweeps apply a leaf once its input range has ended, backward
       * sweeps once they reach its start */
      const unsigned in0 = trans ? b1[l] : b2[l], in1 = trans ? b1[l] + n1[l] : b2[l] + n2[l];
      od[no].key = forward ? in1 : in0;
      od[no++].leaf = l;
    } else {
      goto bad;
    }
  }
  if (covered != n) goto bad;
  qsort(dg, nd, sizeof(or_key), cmp_key);
  qsort(od, no, sizeof(or_key), cmp_key);
  double* z = (double*)xmalloc(sizeof(double) * maxn);
  for (unsigned c = 0; c < p; ++c) {
    double* x = B + (size_t)c * ldB;
    if (forward) {
      unsigned q = 0;
      for (unsigned i = 0; i < nd; ++i) {
        const unsigned l = dg[i].leaf, e = b1[l] + n1[l];
        if (trans) tri_diag_upper_t(n1[l], data + offset[l], x + b1[l]);
        else tri_diag_lower(n1[l], data + offset[l], x + b1[l], z);
        for (; q < no && od[q].key <= e; ++q) {
          const unsigned m = od[q].leaf;
          if (trans) tri_apply_t(n1[m], n2[m], kind[m], data + offset[m], x + b1[m], x + b2[m]);
          else tri_apply(n1[m], n2[m], kind[m], data + offset[m], x + b2[m], x + b1[m]);
        }
      }
    } else {
      unsigned q = no;
      for (unsigned i = nd; i-- > 0;) {
        const unsigned l = dg[i].leaf, s = b1[l];
        if (trans) tri_diag_lower_t(n1[l], data + offset[l], x + s, z);
        else tri_diag_upper(n1[l], data + offset[l], x + s);
        for (; q > 0 && od[q - 1].key >= s; --q) {
          const unsigned m = od[q - 1].leaf;
          if (trans) tri_apply_t(n1[m], n2[m], kind[m], data + offset[m], x + b1[m], x + b2[m]);
          else tri_apply(n1[m], n2[m], kind[m], data + offset[m], x + b2[m], x + b1[m]);
        }
      }
    }
  }
  free(z);
  free(dg);
  free(od);
  return 0;
bad:
  free(dg);
  free(od);
  return -1;
}
