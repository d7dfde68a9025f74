// surface.cpp -- the AHMED-compatible C++ entry points (include/ahmed/*.h)
// on top of the C ABI: H-matrix products, GPU assembly and Krylov solvers for
// trees and mblock arrays owned by the caller.
//
// A (tree, mblock array) pair is uploaded to the device on first use and kept
// in a cache; any mblock mutation bumps an epoch that forces a re-upload.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "ahm.h"
#include "ahm_internal.h"
#include "ahmed/H.h"
#include "ahmed/matgen_gpu.h"
#include "ahmed/matrix.h"
#include "ahmed/solvers.h"

namespace {

std::mutex g_mu;
unsigned long g_epoch = 1;
ahm_ctx g_ctx = nullptr;

struct Entry {
  ahm_hmat h = nullptr;
  unsigned long epoch = 0;
  bool changed = false;
};
std::map<std::pair<const void*, const void*>, Entry> g_cache;

ahm_ctx context() {
  if (!g_ctx) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (ahm_ctx_create(dev, &g_ctx) != AHM_OK) {
      std::fprintf(stderr, "ahmed: no GPU context: %s\n", ahm_last_error());
      return nullptr;
    }
  }
  return g_ctx;
}

template <class T>
void flatten(blcluster* bl, mblock<T>** A, std::vector<ahm_leaf_desc>& L, std::vector<T>& data,
             bool& changed) {
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  L.resize(lv.size());
  size_t total = 0;
  for (size_t i = 0; i < lv.size(); ++i) total += A[lv[i]->getidx()]->nvals();
  data.resize(total);
  size_t off = 0;
  changed = false;
  for (size_t i = 0; i < lv.size(); ++i) {
    const blcluster* b = lv[i];
    mblock<T>* m = A[b->getidx()];
    ahm_leaf_desc& d = L[i];
    d.b1 = b->getb1() - bl->getb1();
    d.n1 = b->getn1();
    d.b2 = b->getb2() - bl->getb2();
    d.n2 = b->getn2();
    d.rank = m->isLrM() ? (int32_t)m->rank() : -1;
    d.flags = m->isHeM() ? AHM_LEAF_HEM : 0u;
    d.offset = off;
    const size_t nv = m->nvals();
    if (nv) std::memcpy(data.data() + off, m->getdata(), nv * sizeof(T));
    off += nv;
    changed |= (d.rank != 0);
  }
}

// device copy of (bl, A); nullptr on failure (message printed)
template <class T>
ahm_hmat device_matrix(blcluster* bl, mblock<T>** A, bool& changed) {
  std::lock_guard<std::mutex> lock(g_mu);
  ahm_ctx c = context();
  if (!c) return nullptr;
  Entry& e = g_cache[{bl, A}];
  if (e.h && e.epoch == g_epoch) {
    changed = e.changed;
    return e.h;
  }
  if (e.h) ahm_hmat_destroy(e.h);
  e.h = nullptr;
  std::vector<ahm_leaf_desc> L;
  std::vector<T> data;
  flatten(bl, A, L, data, e.changed);
  const ahm_dtype dt = sizeof(T) == 8 ? AHM_F64 : AHM_F32;
  if (ahm_hmat_upload(c, bl->getn1(), bl->getn2(), (uint32_t)L.size(), L.data(), data.data(), dt,
                      dt, 0, &e.h) != AHM_OK) {
    std::fprintf(stderr, "ahmed: H-matrix upload failed: %s\n", ahm_last_error());
    e.h = nullptr;
    return nullptr;
  }
  e.epoch = g_epoch;
  changed = e.changed;
  return e.h;
}

// triangular factors: one cache entry per (tree, array, solve: L U l u)
struct TriEntry {
  ahm_trih h = nullptr;
  unsigned long epoch = 0;
};
std::map<std::pair<const void*, const void*>, TriEntry> g_tri[4];

ahm_trih device_factor(blcluster* bl, mblock<double>** M, char uplo) {
  std::lock_guard<std::mutex> lock(g_mu);
  ahm_ctx c = context();
  if (!c) return nullptr;
  const int slot = uplo == 'L' ? 0 : uplo == 'U' ? 1 : uplo == 'l' ? 2 : 3;
  TriEntry& e = g_tri[slot][{bl, M}];
  if (e.h && e.epoch == g_epoch) return e.h;
  if (e.h) ahm_trih_destroy(e.h);
  e.h = nullptr;
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  std::vector<ahm_leaf_desc> L;
  size_t total = 0;
  for (blcluster* b : lv)
    if (M[b->getidx()]) total += M[b->getidx()]->nvals();
  std::vector<double> data(total);
  size_t off = 0;
  for (blcluster* b : lv) {
    mblock<double>* m = M[b->getidx()];
    if (!m) continue;
    ahm_leaf_desc d;
    d.b1 = b->getb1() - bl->getb1();
    d.n1 = b->getn1();
    d.b2 = b->getb2() - bl->getb2();
    d.n2 = b->getn2();
    d.rank = m->isLrM() ? (int32_t)m->rank() : -1;
    d.flags = m->isLtM() ? AHM_LEAF_LTM : m->isUtM() ? AHM_LEAF_UTM : 0u;
    d.offset = off;
    const size_t nv = m->nvals();
    if (nv) std::memcpy(data.data() + off, m->getdata(), nv * sizeof(double));
    off += nv;
    L.push_back(d);
  }
  if (ahm_trih_upload(c, bl->getn1(), uplo, (uint32_t)L.size(), L.data(), data.data(), AHM_F64, &e.h) !=
      AHM_OK) {
    std::fprintf(stderr, "ahmed: triangular factor upload failed: %s\n", ahm_last_error());
    e.h = nullptr;
    return nullptr;
  }
  e.epoch = g_epoch;
  return e.h;
}

void tri_solve(blcluster* bl, mblock<double>** M, char uplo, unsigned p, double* B, unsigned ldB,
               const char* who) {
  ahm_trih t = device_factor(bl, M, uplo);
  if (!t) return;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (ahm_ctx_set_stream(g_ctx, nullptr) != AHM_OK) return;
  }
  if (ahm_trih_solve_host(t, B, ldB, p) != AHM_OK)
    std::fprintf(stderr, "ahmed: %s failed: %s\n", who, ahm_last_error());
}

unsigned start_row(const blcluster* b, const GpuKernel& G) {
  unsigned icom;
  if (b->aca_start_row(icom)) return icom - b->getb1();
  // row nearest the row cluster's centre of mass (first minimum)
  const unsigned b1 = b->getb1(), n1 = b->getn1();
  double c[3] = {0, 0, 0};
  for (unsigned i = b1; i < b1 + n1; ++i)
    for (int j = 0; j < 3; ++j) c[j] += G.pts[3 * (size_t)G.op_perm[i] + j];
  const double s = 1.0 / n1;
  for (int j = 0; j < 3; ++j) c[j] *= s;
  double best = INFINITY;
  unsigned arg = b1;
  for (unsigned i = b1; i < b1 + n1; ++i) {
    double d = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = G.pts[3 * (size_t)G.op_perm[i] + j] - c[j];
      d += e * e;
    }
    if (d < best) {
      best = d;
      arg = i;
    }
  }
  return arg - b1;
}

}  // namespace

extern "C" void ahm_surface_invalidate(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  ++g_epoch;
}

// Drops every cached device copy made from the leaf array A (any tree): the
// array is about to be freed, and a later array at the same address must not
// find its handles.  Called by freembls.
extern "C" void ahm_surface_forget(const void* A) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto it = g_cache.begin(); it != g_cache.end();) {
    if (it->first.second == A) {
      if (it->second.h) ahm_hmat_destroy(it->second.h);
      it = g_cache.erase(it);
    } else {
      ++it;
    }
  }
  for (auto& m : g_tri)
    for (auto it = m.begin(); it != m.end();) {
      if (it->first.second == A) {
        if (it->second.h) ahm_trih_destroy(it->second.h);
        it = m.erase(it);
      } else {
        ++it;
      }
    }
}

void LtHGeM_solve(blcluster* blL, mblock<double>** L, unsigned p, double* B, unsigned ldB) {
  tri_solve(blL, L, 'L', p, B, ldB, "LtHGeM_solve");
}

void UtHGeM_solve(blcluster* blU, mblock<double>** U, unsigned p, double* B, unsigned ldB) {
  tri_solve(blU, U, 'U', p, B, ldB, "UtHGeM_solve");
}

void LtHhGeM_solve(blcluster* blL, mblock<double>** L, unsigned p, double* B, unsigned ldB) {
  tri_solve(blL, L, 'l', p, B, ldB, "LtHhGeM_solve");
}

void UtHhGeM_solve(blcluster* blU, mblock<double>** U, unsigned p, double* B, unsigned ldB) {
  tri_solve(blU, U, 'u', p, B, ldB, "UtHhGeM_solve");
}

bool mltaGeHVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return false;
  if (ahm_hmat_matvec_host(h, d, x, bl->getn2(), y, bl->getn1(), 1) != AHM_OK) {
    std::fprintf(stderr, "ahmed: mltaGeHVec failed: %s\n", ahm_last_error());
    return false;
  }
  return changed;
}

bool mltaGeHGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X, unsigned ldX,
                double* Y, unsigned ldY) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return false;
  if (ahm_hmat_matvec_host(h, d, X, ldX, Y, ldY, p) != AHM_OK) {
    std::fprintf(stderr, "ahmed: mltaGeHGeM failed: %s\n", ahm_last_error());
    return false;
  }
  return changed;
}

// the device matrix of a symmetric leaf array is symmetric (its packed
// diagonal leaves mark it), so the plain product applies both triangles
void mltaHeHVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return;
  if (ahm_hmat_matvec_host(h, d, x, bl->getn2(), y, bl->getn1(), 1) != AHM_OK)
    std::fprintf(stderr, "ahmed: mltaHeHVec failed: %s\n", ahm_last_error());
}

void mltaHeHGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X, unsigned ldX,
                double* Y, unsigned ldY) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return;
  if (ahm_hmat_matvec_host(h, d, X, ldX, Y, ldY, p) != AHM_OK)
    std::fprintf(stderr, "ahmed: mltaHeHGeM failed: %s\n", ahm_last_error());
}

bool mltaGeHhVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return false;
  if (ahm_hmat_matvec_t_host(h, d, x, bl->getn1(), y, bl->getn2(), 1) != AHM_OK) {
    std::fprintf(stderr, "ahmed: mltaGeHhVec failed: %s\n", ahm_last_error());
    return false;
  }
  return changed;
}

bool mltaGeHhGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X, unsigned ldX,
                 double* Y, unsigned ldY) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return false;
  if (ahm_hmat_matvec_t_host(h, d, X, ldX, Y, ldY, p) != AHM_OK) {
    std::fprintf(stderr, "ahmed: mltaGeHhGeM failed: %s\n", ahm_last_error());
    return false;
  }
  return changed;
}

bool mltaGeHVec(float d, blcluster* bl, mblock<float>** A, float* x, float* y) {
  bool changed = false;
  ahm_hmat h = device_matrix(bl, A, changed);
  if (!h) return false;
  const size_t n1 = bl->getn1(), n2 = bl->getn2();
  float *dx = nullptr, *dy = nullptr;
  bool ok = hipMalloc(&dx, n2 * 4) == hipSuccess && hipMalloc(&dy, n1 * 4) == hipSuccess &&
            hipMemcpy(dx, x, n2 * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dy, y, n1 * 4, hipMemcpyHostToDevice) == hipSuccess;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (ok) ok = ahm_ctx_set_stream(g_ctx, nullptr) == AHM_OK;
  }
  if (ok) ok = ahm_hmat_matvec(h, d, dx, (uint32_t)n2, dy, (uint32_t)n1, 1, AHM_F32) == AHM_OK;
  if (ok) ok = hipMemcpy(y, dy, n1 * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (!ok) std::fprintf(stderr, "ahmed: mltaGeHVec(float) failed: %s\n", ahm_last_error());
  return ok && changed;
}

unsigned long sizeH(blcluster* bl, mblock<double>** A) {
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  unsigned long s = 0;
  for (blcluster* b : lv) s += A[b->getidx()]->size();
  return s;
}

unsigned Hmax_rank(blcluster* bl, mblock<double>** A) {
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  unsigned r = 0;
  for (blcluster* b : lv) {
    mblock<double>* m = A[b->getidx()];
    if (m->isLrM() && m->rank() > r) r = m->rank();
  }
  return r;
}

void copyH(blcluster* bl, mblock<double>** A, mblock<float>** B) {
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  for (blcluster* b : lv) {
    const unsigned i = b->getidx();
    mblock<double>* a = A[i];
    delete B[i];
    B[i] = new mblock<float>(b->getn1(), b->getn2());
    if (a->isLrM())
      B[i]->setrank(a->rank());
    else if (a->isHeM())
      B[i]->setHeM();
    else
      B[i]->setGeM();
    const unsigned long nv = a->nvals();
    for (unsigned long q = 0; q < nv; ++q) B[i]->getdata()[q] = (float)a->getdata()[q];
  }
}

namespace {

void matgen_gpu(const GpuKernel& G, unsigned nblcks, blcluster* bl, double eps, unsigned rankmax,
                mblock<double>** A, uint32_t flags, const char* who) {
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  if (lv.size() != nblcks) std::fprintf(stderr, "ahmed: %s: %zu leaves, nblcks=%u\n", who, lv.size(), nblcks);
  const size_t nl = lv.size();
  std::vector<uint32_t> b1(nl), n1(nl), b2(nl), n2(nl), i0(nl);
  std::vector<uint8_t> adm(nl);
  for (size_t l = 0; l < nl; ++l) {
    const blcluster* b = lv[l];
    b1[l] = b->getb1();
    n1[l] = b->getn1();
    b2[l] = b->getb2();
    n2[l] = b->getn2();
    adm[l] = b->isadm() ? 1 : 0;
    i0[l] = start_row(b, G);
  }
  ahm_kernel k{};
  k.kind = G.kind;
  k.kappa = G.kappa;
  k.weight = G.weight;
  k.flags = flags;
  ahm_hmat h = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!context()) return;
    if (ahm_assemble(g_ctx, G.pts, G.op_perm, G.N, (uint32_t)nl, b1.data(), n1.data(), b2.data(),
                     n2.data(), adm.data(), i0.data(), &k, eps, rankmax, nullptr, AHM_F64, 1,
                     nullptr, nullptr, &h) != AHM_OK) {
      std::fprintf(stderr, "ahmed: %s failed: %s\n", who, ahm_last_error());
      return;
    }
  }
  std::vector<ahm_leaf_desc> L(nl);
  ahm_hmat_download(h, L.data(), nullptr);
  size_t total = 0;
  for (const ahm_leaf_desc& d : L) total = std::max(total, (size_t)(d.offset + ahm::leaf_values(d)));
  std::vector<double> data(total);
  ahm_hmat_download(h, nullptr, data.data());
  for (size_t l = 0; l < nl; ++l) {
    const ahm_leaf_desc& d = L[l];
    mblock<double>*& m = A[lv[l]->getidx()];
    delete m;
    m = new mblock<double>(d.n1, d.n2);
    if (d.rank >= 0)
      m->setrank((unsigned)d.rank);
    else if (d.flags & AHM_LEAF_HEM)
      m->setHeM();
    else
      m->setGeM();
    if (m->nvals()) std::memcpy(m->getdata(), data.data() + d.offset, m->nvals() * sizeof(double));
  }
  // the device copy of exactly these leaves serves the following products
  std::lock_guard<std::mutex> lock(g_mu);
  Entry& e = g_cache[{bl, A}];
  if (e.h) ahm_hmat_destroy(e.h);
  e.h = h;
  e.epoch = g_epoch;
  e.changed = true;
}

}  // namespace

void matgenGeH_gpu(const GpuKernel& G, unsigned nblcks, blcluster* bl, double eps, unsigned rankmax,
                   mblock<double>** A) {
  matgen_gpu(G, nblcks, bl, eps, rankmax, A, 0u, "matgenGeH_gpu");
}

void matgenHeH_gpu(const GpuKernel& G, unsigned nblcks, blcluster* bl, double eps, unsigned rankmax,
                   mblock<double>** A) {
  matgen_gpu(G, nblcks, bl, eps, rankmax, A, AHM_ASSEMBLE_SYM, "matgenHeH_gpu");
}

namespace {

template <class F>
unsigned krylov(const Matrix<double>& A, double* b, double* x, F solve) {
  blcluster* tree = nullptr;
  mblock<double>** blcks = nullptr;
  if (const GeHMatrix<double>* G = dynamic_cast<const GeHMatrix<double>*>(&A)) {
    tree = G->blclTree;
    blcks = G->blcks;
  } else if (const HeHMatrix<double>* S = dynamic_cast<const HeHMatrix<double>*>(&A)) {
    tree = S->blclTree;
    blcks = S->blcks;
  } else {
    std::fprintf(stderr, "ahmed: the GPU solvers need a GeHMatrix<double> or HeHMatrix<double> operator\n");
    return 2;
  }
  bool changed;
  ahm_hmat h = device_matrix(tree, blcks, changed);
  if (!h) return 2;
  const size_t n = tree->getn1();
  double *db = nullptr, *dx = nullptr;
  if (hipMalloc(&db, n * 8) != hipSuccess || hipMalloc(&dx, n * 8) != hipSuccess) {
    std::fprintf(stderr, "ahmed: out of device memory\n");
    return 2;
  }
  unsigned rc = 2;
  if (hipMemcpy(db, b, n * 8, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dx, x, n * 8, hipMemcpyHostToDevice) == hipSuccess) {
    {
      std::lock_guard<std::mutex> lock(g_mu);
      ahm_ctx_set_stream(g_ctx, nullptr);
    }
    const int st = solve(h, db, dx);
    if (st == AHM_OK || st == AHM_ERR_NOT_CONVERGED) {
      rc = st == AHM_OK ? 0 : 1;
      if (hipMemcpy(x, dx, n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = 2;
    } else {
      std::fprintf(stderr, "ahmed: solver failed: %s\n", ahm_last_error());
    }
  }
  (void)hipFree(db);
  (void)hipFree(dx);
  return rc;
}

}  // namespace

unsigned GMRes(const Matrix<double>& A, double* const b, double* const x, double& eps,
               const unsigned m, unsigned& nsteps) {
  return krylov(A, b, x, [&](ahm_hmat h, double* db, double* dx) {
    uint32_t ns = nsteps;
    const int st = ahm_gmres(h, db, dx, &eps, m, &ns, nullptr, 0, nullptr);
    nsteps = ns;
    return st;
  });
}

unsigned CG(const Matrix<double>& A, double* const b, double* const x, double& eps,
            unsigned& nsteps) {
  return krylov(A, b, x, [&](ahm_hmat h, double* db, double* dx) {
    uint32_t ns = nsteps;
    const int st = ahm_cg(h, db, dx, &eps, &ns);
    nsteps = ns;
    return st;
  });
}
