// ref_harness.cpp -- drives the COMPILED REFERENCE (AHMED sources under
// /root/reference, built by oracle/Makefile into oracle/_ref/) to write golden
// vectors and to cross-check the oracle restatement in this container.
//
// TEST INFRASTRUCTURE ONLY.  This file is our own driver code: it includes the
// reference's public headers and calls the reference's own functions
// (cluster3d_bbx, bemcluster, blcluster::subdivide, matgenGeH_omp/apprx_unsym,
// ACA, mblock::cpyLrM_cmpr, mltaGeHVec, mltaGeHGeM, GMRes, CG, HLU and the
// triangular H-solves).  The MATGEN_T
// kernel class below is ours (the reference ships none).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include <omp.h>

#include "H.h"
#include "apprx.h"
#include "bemblcluster.h"
#include "bemcluster.h"
#include "blcluster.h"
#include "cluster_bbx.h"
#include "matgen_omp.h"
#include "matrix.h"
#include "solvers.h"

extern "C" int MKL_CBWR_Set(int);
extern void gen_BlSequence(blcluster*, blcluster**&);

namespace {

struct Point {
  double x[3];
  double getcenter(unsigned i) const { return x[i]; }
  double getradius2() const { return 0.0; }
};

const double ONE_OVER_4PI = 0.079577471545947667884;

// MATGEN_T plugin (contract Include/matgen_sqntl.h:21-29): same kernel
// definitions as the oracle's, entries in permuted ordering.
struct KernelGen {
  int kernel;
  double kappa;
  double weight;
  const Point* pts;
  const unsigned* op;
  double val(unsigned i, unsigned j) const {
    const unsigned oi = op[i], oj = op[j];
    if (oi == oj) return 1.0;
    const double dx = pts[oi].x[0] - pts[oj].x[0], dy = pts[oi].x[1] - pts[oj].x[1],
                 dz = pts[oi].x[2] - pts[oj].x[2];
    const double r = std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
    const double c = kernel == 2 ? weight : weight * ONE_OVER_4PI;
    if (kernel == 0) return c / r;
    return (std::exp(-kappa * r) * c) / r;
  }
  void cmpbl(unsigned b1, unsigned n1, unsigned b2, unsigned n2, double* d) const {
    for (unsigned j = 0; j < n2; ++j)
      for (unsigned i = 0; i < n1; ++i) d[i + (size_t)j * n1] = val(b1 + i, b2 + j);
  }
  double scale(unsigned b1, unsigned n1, unsigned b2, unsigned n2) const {
    return std::fabs(val(b1 + n1 / 2, b2 + n2 / 2));
  }
  // upper triangle of a diagonal block, packed column by column (the cmpblsym
  // contract, Include/matgen_omp.h:30-32)
  void cmpblsym(unsigned b1, unsigned n1, double* d) const {
    for (unsigned j = 0; j < n1; ++j)
      for (unsigned i = 0; i <= j; ++i) d[i + (size_t)j * (j + 1) / 2] = val(b1 + i, b1 + j);
  }
};

struct DenseGen {
  const double* A;
  unsigned lda;
  double sc;
  void cmpbl(unsigned b1, unsigned n1, unsigned b2, unsigned n2, double* d) const {
    for (unsigned j = 0; j < n2; ++j)
      for (unsigned i = 0; i < n1; ++i) d[i + (size_t)j * n1] = A[(b1 + i) + (size_t)(b2 + j) * lda];
  }
  double scale(unsigned, unsigned, unsigned, unsigned) const { return sc; }
};

struct Tree {
  int kind;
  bool sym;  // symmetric block tree (blcluster::subdivide_sym)
  unsigned N;
  std::vector<Point> pts;
  std::vector<unsigned> op, po;
  cluster* root;
  blcluster* bl;
  unsigned nblcks;
  std::vector<blcluster*> leaves;  // by idx
  std::vector<unsigned> i0;
};

// nl: leaf count, kept here so that freeing never reads the tree (the
// Python wrappers may free a tree before a matrix built on it, e.g. when the
// interpreter finalises both together)
struct HMat {
  Tree* t;
  mblock<double>** A;
  unsigned nl;
};

// H-LU factors on the block tree of t (H/HLU.cpp:20-66): L holds the lower
// and diagonal leaves, U the upper and diagonal leaves, the rest are NULL
struct Factors {
  Tree* t;
  mblock<double>** L;
  mblock<double>** U;
  unsigned nl;
};

// right-preconditioned operator: amux = A, precond_apply = HLU_solve
// (Include/matrix.h:30, Include/H.h:1383-1388)
struct PrecGeH : public GeHMatrix<double> {
  const Factors* f;
  PrecGeH(unsigned n, blcluster* bl, const Factors* ff) : GeHMatrix<double>(n, n, bl), f(ff) {}
  void precond_apply(double* x) const { HLU_solve(f->t->bl, f->L, f->U, x); }
};

// leaf kind codes of the factor layout: -2 absent, -1 dense, >= 0 low rank
// (the rank), -3 unit lower triangle with the permutation on its diagonal
// (mblock::isLtM, packed), -4 upper triangle (mblock::isUtM, packed)
int factor_kind(mblock<double>* m) {
  if (!m) return -2;
  if (m->isLrM()) return (int)m->rank();
  if (m->isLtM()) return -3;
  if (m->isUtM()) return -4;
  return -1;
}

// ACA start row for bbx trees: position in the row cluster closest to its
// centre of mass (first minimum), same rule as the oracle.
unsigned bbx_icom(const Tree* t, unsigned b1, unsigned n1) {
  double com[3] = {0, 0, 0};
  for (unsigned i = b1; i < b1 + n1; ++i)
    for (int j = 0; j < 3; ++j) com[j] += t->pts[t->op[i]].x[j];
  const double d = 1.0 / n1;
  for (int j = 0; j < 3; ++j) com[j] *= d;
  double smin = INFINITY;
  unsigned ic = b1;
  for (unsigned i = b1; i < b1 + n1; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = t->pts[t->op[i]].x[j] - com[j];
      s += e * e;
    }
    if (s < smin) { smin = s; ic = i; }
  }
  return ic;
}

// Include/apprx.h:136-172 restated for a plain blcluster with an explicit
// start row (apprx_unsym itself needs a bemblcluster).
void apprx_bbx(KernelGen& G, mblock<double>*& mbl, blcluster* bl, unsigned i0, double eps,
               unsigned rankmax, int method = 0) {
  unsigned n1 = bl->getn1(), n2 = bl->getn2(), b1 = bl->getb1(), b2 = bl->getb2();
  unsigned maxk = n1 * n2 / (n1 + n2) + 4;
  if (maxk > rankmax) maxk = rankmax;
  mbl = new mblock<double>(n1, n2);
  bool succ = false;
  if (bl->isadm() && maxk) {
    unsigned k;
    double *U, *V;
    // method 1: ACAr, the approximation apprx_sym uses (Include/apprx.h:109)
    succ = method == 1 ? ::ACAr(G, b1, n1, b2, n2, eps, maxk, i0, k, U, V)
                       : ::ACA(G, b1, n1, b2, n2, eps, maxk, i0, k, U, V);
    if (succ) mbl->cpyLrM_cmpr(k, U, n1, V, n2, eps, k);
    delete[] V;
    delete[] U;
  }
  if (!succ) {
    mbl->setGeM();
    G.cmpbl(b1, n1, b2, n2, mbl->getdata());
  }
}

// Include/apprx.h:87-131 (apprx_sym) restated for a plain blcluster with an
// explicit start row: ACAr on admissible blocks, diagonal blocks packed
// Hermitian (mblock::setHeM + cmpblsym), other blocks dense.
void apprx_sym_bbx(KernelGen& G, mblock<double>*& mbl, blcluster* bl, unsigned i0, double eps,
                   unsigned rankmax) {
  unsigned n1 = bl->getn1(), n2 = bl->getn2(), b1 = bl->getb1(), b2 = bl->getb2();
  unsigned maxk = n1 * n2 / (n1 + n2) + 4;
  if (maxk > rankmax) maxk = rankmax;
  mbl = new mblock<double>(n1, n2);
  bool succ = false;
  if (bl->isadm() && maxk) {
    unsigned k;
    double *U, *V;
    succ = ::ACAr(G, b1, n1, b2, n2, eps, maxk, i0, k, U, V);
    if (succ) mbl->cpyLrM_cmpr(k, U, n1, V, n2, eps, k);
    delete[] V;
    delete[] U;
  }
  if (!succ) {
    if (bl->isdbl()) {
      mbl->setHeM();
      G.cmpblsym(b1, n1, mbl->getdata());
    } else {
      mbl->setGeM();
      G.cmpbl(b1, n1, b2, n2, mbl->getdata());
    }
  }
}

// Leaf-parallel operator on all host cores for the CPU baseline: the leaves
// in idx order are cut into nthreads contiguous intervals of equal stored-
// value count; each thread applies the reference's own leaf product
// (mblock::mltaVec, H/mblock.cpp:321-323) into a private y, and the private
// vectors are summed.  The reference's mltaGeHVec_omp writes one shared y
// from every thread (H/mltaGeHVec_omp.cpp:23-36), which is a race.
struct OmpGeH : public Matrix<double> {
  const Tree* t;
  mblock<double>** A;
  int nthreads;
  std::vector<unsigned> cut;
  mutable std::vector<double> ybuf;
  OmpGeH(const Tree* tt, mblock<double>** AA, int nt)
      : Matrix<double>(tt->N, tt->N), t(tt), A(AA), nthreads(nt > 0 ? nt : 1) {
    size_t total = 0;
    for (unsigned l = 0; l < t->nblcks; ++l) total += A[l]->nvals();
    cut.assign(nthreads + 1, t->nblcks);
    cut[0] = 0;
    size_t acc = 0;
    unsigned l = 0;
    for (int k = 1; k < nthreads; ++k) {
      const size_t target = total / nthreads * (size_t)k;
      while (l < t->nblcks && acc < target) acc += A[l++]->nvals();
      cut[k] = l;
    }
    ybuf.resize((size_t)nthreads * t->N);
  }
  void amux(double d, double* x, double* y) const {
    const unsigned N = t->N;
#pragma omp parallel num_threads(nthreads)
    {
      const int k = omp_get_thread_num();
      double* yk = ybuf.data() + (size_t)k * N;
      std::memset(yk, 0, sizeof(double) * N);
      for (unsigned l = cut[k]; l < cut[k + 1]; ++l) {
        const blcluster* b = t->leaves[l];
        A[l]->mltaVec(d, x + b->getb2(), yk + b->getb1());
      }
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long i = 0; i < (long)N; ++i) {
      double s = 0.0;
      for (int k = 0; k < nthreads; ++k) s += ybuf[(size_t)k * N + i];
      y[i] += s;
    }
  }
};

void init() {
  static bool done = false;
  if (!done) { MKL_CBWR_Set(2); done = true; }
}

struct Silence {
  std::stringstream sink;
  std::streambuf* old;
  Silence() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~Silence() { std::cout.rdbuf(old); }
};

}  // namespace

extern "C" {

// kind: 0 bbx, 1 bemcluster (PCA); | 0x100: symmetric block tree
// (blcluster::subdivide_sym, cluster/blcluster.cpp:90-113)
void* ref_tree_build(const double* pts, unsigned N, unsigned bmin, double eta2, int kind) {
  init();
  Tree* t = new Tree;
  t->sym = (kind & 0x100) != 0;
  kind &= 0xff;
  t->kind = kind;
  t->N = N;
  t->pts.resize(N);
  for (unsigned i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) t->pts[i].x[j] = pts[3 * (size_t)i + j];
  t->op.resize(N);
  t->po.resize(N);
  for (unsigned i = 0; i < N; ++i) t->op[i] = t->po[i] = i;
  if (kind == 0) {
    cluster3d_bbx<Point>* r = new cluster3d_bbx<Point>(t->pts.data(), t->op.data(), 0, N);
    r->createClusterTree(bmin, t->op.data(), t->po.data());
    t->root = r;
    t->bl = new blcluster(0, 0, N, N);
  } else {
    bemcluster<Point>* r = new bemcluster<Point>(t->pts.data(), t->op.data(), 0, N);
    r->createClusterTree(bmin, t->op.data(), t->po.data());
    t->root = r;
    t->bl = new bemblcluster<Point, Point>(r, r);
  }
  if (t->sym)
    t->bl->subdivide_sym(t->root, eta2, t->nblcks);
  else
    t->bl->subdivide(t->root, t->root, eta2, t->nblcks);
  blcluster** list;
  gen_BlSequence(t->bl, list);
  t->leaves.assign(t->nblcks, nullptr);
  for (unsigned i = 0; i < t->nblcks; ++i) t->leaves[list[i]->getidx()] = list[i];
  delete[] list;
  t->i0.resize(t->nblcks);
  for (unsigned i = 0; i < t->nblcks; ++i) {
    blcluster* b = t->leaves[i];
    if (kind == 0) t->i0[i] = bbx_icom(t, b->getb1(), b->getn1()) - b->getb1();
    else t->i0[i] = ((bemblcluster<Point, Point>*)b)->getcl1()->geticom() - b->getb1();
  }
  return t;
}

void ref_tree_free(void* p) {
  Tree* t = (Tree*)p;
  delete t->bl;
  delete t->root;
  delete t;
}

unsigned ref_tree_nleaves(void* p) { return ((Tree*)p)->nblcks; }

void ref_tree_perm(void* p, unsigned* op, unsigned* po) {
  Tree* t = (Tree*)p;
  std::memcpy(op, t->op.data(), sizeof(unsigned) * t->N);
  std::memcpy(po, t->po.data(), sizeof(unsigned) * t->N);
}

void ref_tree_leaves(void* p, unsigned* b1, unsigned* n1, unsigned* b2, unsigned* n2,
                     unsigned char* adm, unsigned* i0) {
  Tree* t = (Tree*)p;
  for (unsigned i = 0; i < t->nblcks; ++i) {
    blcluster* b = t->leaves[i];
    b1[i] = b->getb1(); n1[i] = b->getn1(); b2[i] = b->getb2(); n2[i] = b->getn2();
    adm[i] = b->isadm() ? 1 : 0;
    i0[i] = t->i0[i];
  }
}

void* ref_hmat_assemble(void* tp, int kernel, double kappa, double weight, double eps,
                        unsigned rankmax) {
  init();
  Tree* t = (Tree*)tp;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  allocmbls(t->bl, h->A);
  KernelGen G = {kernel, kappa, weight, t->pts.data(), t->op.data()};
  Silence quiet;
  if (t->kind == 1) {
    matgenGeH_omp(G, t->nblcks, (bemblcluster<Point, Point>*)t->bl, eps, rankmax, h->A);
  } else {
    for (unsigned i = 0; i < t->nblcks; ++i)
      apprx_bbx(G, h->A[i], t->leaves[i], t->i0[i], eps, rankmax);
  }
  return h;
}

// matgenGeH_omp's leaf loop (Include/matgen_omp.h:59-75: an OpenMP dynamic
// schedule over the gen_BlSequence leaf list, one apprx call per leaf, the
// MATGEN_T object shared by all threads) run on every stride-th entry of that
// list with nthreads threads.  bbx trees take apprx_bbx (the start row their
// clusters lack); PCA trees apprx_unsym itself.  Leaves outside the sample
// stay rank 0 and are marked 0 in sampled[] (by idx).  *seconds: wall time of
// the loop.  stride 1 is the whole assembly.
void* ref_hmat_assemble_omp(void* tp, int kernel, double kappa, double weight, double eps,
                            unsigned rankmax, unsigned stride, int nthreads, double* seconds,
                            unsigned char* sampled) {
  init();
  Tree* t = (Tree*)tp;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  allocmbls(t->bl, h->A);
  KernelGen G = {kernel, kappa, weight, t->pts.data(), t->op.data()};
  blcluster** list;
  gen_BlSequence(t->bl, list);
  if (stride == 0) stride = 1;
  const int n = (int)t->nblcks;
  for (int i = 0; i < n; ++i) sampled[i] = 0;
  for (int i = 0; i < n; i += (int)stride) sampled[list[i]->getidx()] = 1;
  Silence quiet;
  const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
  for (int i = 0; i < n; ++i) {
    blcluster* b = list[i];
    const unsigned idx = b->getidx();
    if (i % (int)stride) continue;
    if (t->kind == 1)
      apprx_unsym(G, h->A[idx], (bemblcluster<Point, Point>*)b, eps, rankmax);
    else
      apprx_bbx(G, h->A[idx], b, t->i0[idx], eps, rankmax);
  }
  *seconds = omp_get_wtime() - t0;
  for (int i = 0; i < n; ++i) {
    const unsigned idx = list[i]->getidx();
    if (!h->A[idx]) {
      h->A[idx] = new mblock<double>(list[i]->getn1(), list[i]->getn2());
      h->A[idx]->setrank(0);
    }
  }
  delete[] list;
  return h;
}

// every leaf with ACAr on the admissible ones (both tree kinds, start row as
// in ref_hmat_assemble), unsymmetric storage
void* ref_hmat_assemble_acar(void* tp, int kernel, double kappa, double weight, double eps,
                             unsigned rankmax) {
  init();
  Tree* t = (Tree*)tp;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  allocmbls(t->bl, h->A);
  KernelGen G = {kernel, kappa, weight, t->pts.data(), t->op.data()};
  Silence quiet;
  for (unsigned i = 0; i < t->nblcks; ++i) apprx_bbx(G, h->A[i], t->leaves[i], t->i0[i], eps, rankmax, 1);
  return h;
}

// the symmetric assembly: matgenHeH_omp (Include/matgen_omp.h:77-100,
// apprx_sym per leaf) on bemcluster trees; on bbx trees the same leaf rule
// with the bbx start row (apprx_sym_bbx)
void* ref_hmat_assemble_sym(void* tp, int kernel, double kappa, double weight, double eps,
                            unsigned rankmax) {
  init();
  Tree* t = (Tree*)tp;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  allocmbls(t->bl, h->A);
  KernelGen G = {kernel, kappa, weight, t->pts.data(), t->op.data()};
  Silence quiet;
  if (t->kind == 1)
    matgenHeH_omp(G, t->nblcks, (bemblcluster<Point, Point>*)t->bl, eps, rankmax, h->A);
  else
    for (unsigned i = 0; i < t->nblcks; ++i) apprx_sym_bbx(G, h->A[i], t->leaves[i], t->i0[i], eps, rankmax);
  return h;
}

// an mblock array holding given leaf values (mblock layout, leaves in idx
// order of this tree): lets the reference's own mltaGeHVec run on a matrix
// assembled elsewhere
void* ref_hmat_from_data(void* tp, const int* rank, const size_t* offset, const double* data) {
  init();
  Tree* t = (Tree*)tp;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  allocmbls(t->bl, h->A);
  for (unsigned i = 0; i < t->nblcks; ++i) {
    blcluster* b = t->leaves[i];
    const unsigned n1 = b->getn1(), n2 = b->getn2();
    mblock<double>* m = new mblock<double>(n1, n2);
    if (rank[i] == -5) {
      m->setHeM();
      std::memcpy(m->getdata(), data + offset[i], sizeof(double) * (size_t)n1 * (n1 + 1) / 2);
    } else if (rank[i] < 0) {
      m->setGeM();
      std::memcpy(m->getdata(), data + offset[i], sizeof(double) * (size_t)n1 * n2);
    } else {
      m->setrank((unsigned)rank[i]);
      if (rank[i] > 0)
        std::memcpy(m->getdata(), data + offset[i], sizeof(double) * (size_t)rank[i] * (n1 + n2));
    }
    h->A[i] = m;
  }
  return h;
}

void ref_hmat_free(void* p) {
  HMat* h = (HMat*)p;
  freembls(h->nl, h->A);
  delete h;
}

// rank[i] = -1 for dense leaves, -5 for packed Hermitian (mblock::isHeM)
// diagonal leaves; returns total stored values
size_t ref_hmat_layout(void* p, int* rank, size_t* offset) {
  HMat* h = (HMat*)p;
  size_t off = 0;
  for (unsigned i = 0; i < h->t->nblcks; ++i) {
    mblock<double>* m = h->A[i];
    rank[i] = m->isLrM() ? (int)m->rank() : m->isHeM() ? -5 : -1;
    offset[i] = off;
    off += m->nvals();
  }
  return off;
}

void ref_hmat_data(void* p, double* out) {
  HMat* h = (HMat*)p;
  size_t off = 0;
  for (unsigned i = 0; i < h->t->nblcks; ++i) {
    mblock<double>* m = h->A[i];
    const size_t nv = m->nvals();
    if (nv) std::memcpy(out + off, m->getdata(), sizeof(double) * nv);
    off += nv;
  }
}

int ref_hmat_matvec(void* p, double d, double* x, double* y) {
  HMat* h = (HMat*)p;
  return mltaGeHVec(d, h->t->bl, h->A, x, y) ? 1 : 0;
}

// y += d A^H x (H/mltaGeHVec.cpp:61-81); x indexed by rows, y by columns
int ref_hmat_matvec_t(void* p, double d, double* x, double* y) {
  HMat* h = (HMat*)p;
  return mltaGeHhVec(d, h->t->bl, h->A, x, y) ? 1 : 0;
}

// y += d A x for a symmetric H-matrix (mltaHeHVec, H/mltaGeHVec.cpp:420-446)
int ref_hmat_matvec_sym(void* p, double d, double* x, double* y) {
  HMat* h = (HMat*)p;
  mltaHeHVec(d, h->t->bl, h->A, x, y);
  return 1;
}

// Y += d A X for a symmetric H-matrix (mltaHeHGeM, H/mltaGeHVec.cpp:448-472)
int ref_hmat_matmat_sym(void* p, double d, unsigned pp, double* X, unsigned ldX, double* Y,
                        unsigned ldY) {
  HMat* h = (HMat*)p;
  mltaHeHGeM(d, h->t->bl, h->A, pp, X, ldX, Y, ldY);
  return 1;
}

// the reference's CG (solvers/CG.cpp:31-113) and GMRes on HeHMatrix
// (Include/matrix.h:307-323), the operator of its own BEM example
unsigned ref_cg_sym(void* p, double* b, double* x, double* eps, unsigned* nsteps) {
  HMat* h = (HMat*)p;
  HeHMatrix<double> M(h->t->N, h->t->bl);
  M.blcks = h->A;
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = CG(M, b, x, e, ns);
  M.blcks = NULL;
  *eps = e;
  *nsteps = ns;
  return r;
}

unsigned ref_gmres_sym(void* p, double* b, double* x, double* eps, unsigned m, unsigned* nsteps) {
  HMat* h = (HMat*)p;
  HeHMatrix<double> M(h->t->N, h->t->bl);
  M.blcks = h->A;
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = GMRes(M, b, x, e, m, ns);
  M.blcks = NULL;
  *eps = e;
  *nsteps = ns;
  return r;
}

int ref_hmat_matmat(void* p, double d, unsigned pp, double* X, unsigned ldX, double* Y,
                    unsigned ldY) {
  HMat* h = (HMat*)p;
  return mltaGeHGeM(d, h->t->bl, h->A, pp, X, ldX, Y, ldY) ? 1 : 0;
}

unsigned ref_gmres(void* p, double* b, double* x, double* eps, unsigned m, unsigned* nsteps) {
  HMat* h = (HMat*)p;
  GeHMatrix<double> M(h->t->N, h->t->N, h->t->bl);
  M.blcks = h->A;
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = GMRes(M, b, x, e, m, ns);
  M.blcks = NULL;  // the H-matrix stays owned by h
  *eps = e;
  *nsteps = ns;
  return r;
}

// y += d A x on nthreads host cores (OmpGeH)
int ref_hmat_matvec_omp(void* p, double d, double* x, double* y, int nthreads) {
  HMat* h = (HMat*)p;
  OmpGeH M(h->t, h->A, nthreads);
  M.amux(d, x, y);
  return 1;
}

// the reference's GMRes (solvers/GMRES.cpp:80-181) on the leaf-parallel
// operator: products on nthreads cores, the solver's vector work serial
unsigned ref_gmres_omp(void* p, double* b, double* x, double* eps, unsigned m, unsigned* nsteps,
                       int nthreads) {
  HMat* h = (HMat*)p;
  OmpGeH M(h->t, h->A, nthreads);
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = GMRes(M, b, x, e, m, ns);
  *eps = e;
  *nsteps = ns;
  return r;
}

unsigned ref_cg(void* p, double* b, double* x, double* eps, unsigned* nsteps) {
  HMat* h = (HMat*)p;
  GeHMatrix<double> M(h->t->N, h->t->N, h->t->bl);
  M.blcks = h->A;
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = CG(M, b, x, e, ns);
  M.blcks = NULL;
  *eps = e;
  *nsteps = ns;
  return r;
}

// ACA on an explicit dense block; U,V hold (kmax+1)*n1, (kmax+1)*n2
int ref_aca_dense(const double* A, unsigned n1, unsigned n2, double eps, unsigned kmax,
                  unsigned i0, double scale, unsigned* k, double* U, double* V) {
  init();
  DenseGen G = {A, n1, scale};
  double *Up, *Vp;
  bool ok = ::ACA(G, 0, n1, 0, n2, eps, kmax, i0, *k, Up, Vp);
  std::memcpy(U, Up, sizeof(double) * (size_t)(kmax + 1) * n1);
  std::memcpy(V, Vp, sizeof(double) * (size_t)(kmax + 1) * n2);
  delete[] Up;
  delete[] Vp;
  return ok ? 1 : 0;
}

// ACAr on an explicit dense block (Include/ACA.h:657-720)
int ref_acar_dense(const double* A, unsigned n1, unsigned n2, double eps, unsigned kmax,
                   unsigned i0, double scale, unsigned* k, double* U, double* V) {
  init();
  DenseGen G = {A, n1, scale};
  double *Up, *Vp;
  bool ok = ::ACAr(G, 0, n1, 0, n2, eps, kmax, i0, *k, Up, Vp);
  std::memcpy(U, Up, sizeof(double) * (size_t)(kmax + 1) * n1);
  std::memcpy(V, Vp, sizeof(double) * (size_t)(kmax + 1) * n2);
  delete[] Up;
  delete[] Vp;
  return ok ? 1 : 0;
}

void ref_addtrll(unsigned n1, unsigned n2, unsigned k, double* U, double* V, double delta,
                 unsigned kgoal, unsigned* kt, double* Uo, double* Vo) {
  init();
  mblock<double> m(n1, n2);
  m.cpyLrM_cmpr(k, U, n1, V, n2, delta, kgoal);
  *kt = m.rank();
  if (*kt) {
    std::memcpy(Uo, m.getdata(), sizeof(double) * (size_t)n1 * *kt);
    std::memcpy(Vo, m.getdata() + (size_t)n1 * *kt, sizeof(double) * (size_t)n2 * *kt);
  }
}

// the reference's own on-disk format: savembls / loadmbls (Include/H.h:47-103,
// mblock::save/load Include/mblock.h:737-804)
int ref_hmat_save(void* hp, const char* path, double eta, unsigned bmin) {
  init();
  HMat* h = (HMat*)hp;
  std::ofstream os(path, std::ios::binary);
  if (!os) return 0;
  savembls(eta, bmin, h->t->bl, h->A, os);
  return os.good() ? 1 : 0;
}

// loadmbls into the leaves of a tree rebuilt from the points (the file holds
// no geometry); returns NULL when the leaf count or a leaf's size differs
void* ref_hmat_load(void* tp, const char* path, double* eta, unsigned* bmin) {
  init();
  Tree* t = (Tree*)tp;
  std::ifstream is(path, std::ios::binary);
  if (!is) return nullptr;
  mblock<double>** A = nullptr;
  loadmbls(*eta, *bmin, A, is);
  for (unsigned i = 0; i < t->nblcks; ++i)
    if (!A[i] || A[i]->getn1() != t->leaves[i]->getn1() || A[i]->getn2() != t->leaves[i]->getn2())
      return nullptr;
  HMat* h = new HMat;
  h->t = t;
  h->nl = t->nblcks;
  h->A = A;
  return h;
}


// ---- H-LU (H/HLU.cpp:389-395) and the triangular H-solves (H/TH_solve.cpp) ----

// LU factors of a copy of h's matrix (HLU destroys its input); NULL if HLU
// reports failure
void* ref_hlu(void* hp, double eps, unsigned rankmax) {
  init();
  HMat* h = (HMat*)hp;
  Tree* t = h->t;
  mblock<double>** A;
  allocmbls(t->bl, A);
  for (unsigned i = 0; i < t->nblcks; ++i) {
    mblock<double>* s = h->A[i];
    const unsigned n1 = s->getn1(), n2 = s->getn2();
    mblock<double>* m = new mblock<double>(n1, n2);
    if (s->isLrM()) {
      m->setrank(s->rank());
      if (s->rank()) std::memcpy(m->getdata(), s->getdata(), sizeof(double) * s->nvals());
    } else {
      m->setGeM();
      std::memcpy(m->getdata(), s->getdata(), sizeof(double) * (size_t)n1 * n2);
    }
    A[i] = m;
  }
  Factors* f = new Factors;
  f->t = t;
  f->nl = t->nblcks;
  initLtH_0(t->bl, f->L);
  initUtH_0(t->bl, f->U);
  Silence quiet;
  const bool ok = HLU(t->bl, A, f->L, f->U, eps, rankmax, NULL);
  freembls(t->bl, A);
  if (!ok) {
    freembls(t->bl, f->L);
    freembls(t->bl, f->U);
    delete f;
    return nullptr;
  }
  return f;
}

void ref_factors_free(void* p) {
  Factors* f = (Factors*)p;
  freembls(f->nl, f->L);
  freembls(f->nl, f->U);
  delete f;
}

// which = 0: L, 1: U.  kind[i] per factor_kind, offset[i] into the packed
// values; returns the number of values
size_t ref_factor_layout(void* p, int which, int* kind, size_t* offset) {
  Factors* f = (Factors*)p;
  mblock<double>** M = which ? f->U : f->L;
  size_t off = 0;
  for (unsigned i = 0; i < f->t->nblcks; ++i) {
    kind[i] = factor_kind(M[i]);
    offset[i] = off;
    if (M[i]) off += M[i]->nvals();
  }
  return off;
}

void ref_factor_data(void* p, int which, double* out) {
  Factors* f = (Factors*)p;
  mblock<double>** M = which ? f->U : f->L;
  size_t off = 0;
  for (unsigned i = 0; i < f->t->nblcks; ++i) {
    if (!M[i]) continue;
    const size_t nv = M[i]->nvals();
    if (nv) std::memcpy(out + off, M[i]->getdata(), sizeof(double) * nv);
    off += nv;
  }
}

// factors rebuilt from a layout (kind/offset per leaf as ref_factor_layout)
void* ref_factors_from_data(void* tp, const int* kind_l, const size_t* off_l, const double* dl,
                            const int* kind_u, const size_t* off_u, const double* du) {
  init();
  Tree* t = (Tree*)tp;
  Factors* f = new Factors;
  f->t = t;
  f->nl = t->nblcks;
  for (int w = 0; w < 2; ++w) {
    const int* kind = w ? kind_u : kind_l;
    const size_t* off = w ? off_u : off_l;
    const double* d = w ? du : dl;
    mblock<double>**& M = w ? f->U : f->L;
    allocmbls(t->bl, M);
    for (unsigned i = 0; i < t->nblcks; ++i) {
      if (kind[i] == -2) { M[i] = NULL; continue; }
      const unsigned n1 = t->leaves[i]->getn1(), n2 = t->leaves[i]->getn2();
      mblock<double>* m = new mblock<double>(n1, n2);
      if (kind[i] >= 0) m->setrank((unsigned)kind[i]);
      else if (kind[i] == -3) m->setLtM();
      else if (kind[i] == -4) m->setUtM();
      else m->setGeM();
      if (m->nvals()) std::memcpy(m->getdata(), d + off[i], sizeof(double) * m->nvals());
      M[i] = m;
    }
  }
  return f;
}

// which: 0 L X = B (LtHGeM_solve), 1 U X = B (UtHGeM_solve), 2 L U X = B
// (HLU_solve per column), 3 L^T X = B (LtHhGeM_solve), 4 U^T X = B
// (UtHhGeM_solve), 5 U^T U X = B (HCholesky_solve per column), 6 (LU)^T X = B
// (HLUh_solve per column); B is n x p, overwritten with X
void ref_trisolve(void* p, int which, unsigned pp, double* B, unsigned ldB) {
  Factors* f = (Factors*)p;
  blcluster* bl = f->t->bl;
  if (which == 0) LtHGeM_solve(bl, f->L, pp, B, ldB);
  else if (which == 1) UtHGeM_solve(bl, f->U, pp, B, ldB);
  else if (which == 3) LtHhGeM_solve(bl, f->L, pp, B, ldB);
  else if (which == 4) UtHhGeM_solve(bl, f->U, pp, B, ldB);
  else
    for (unsigned j = 0; j < pp; ++j) {
      double* b = B + (size_t)j * ldB;
      if (which == 2) HLU_solve(bl, f->L, f->U, b);
      else if (which == 5) HCholesky_solve(bl, f->U, b);
      else HLUh_solve(bl, f->L, f->U, b);
    }
}

// GMRes (solvers/GMRES.cpp:80-181) with HLU_solve as the right preconditioner
unsigned ref_gmres_hlu(void* hp, void* fp, double* b, double* x, double* eps, unsigned m,
                       unsigned* nsteps) {
  HMat* h = (HMat*)hp;
  PrecGeH M(h->t->N, h->t->bl, (const Factors*)fp);
  M.blcks = h->A;
  double e = *eps;
  unsigned ns = *nsteps;
  unsigned r = GMRes(M, b, x, e, m, ns);
  M.blcks = NULL;
  *eps = e;
  *nsteps = ns;
  return r;
}

}  // extern "C"
