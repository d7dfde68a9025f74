// tree_capi.cpp -- C ABI over the C++ cluster / block cluster trees
// (include/ahmed): builds the trees the reference builds and exports the
// permutation and the leaves in idx order for the device assembly.
#include <cmath>
#include <cstring>
#include <vector>

#include "ahm.h"
#include "ahm_internal.h"
#include "ahm_aca.h"
#include "ahmed/bemblcluster.h"
#include "ahmed/blcluster.h"
#include "ahmed/cluster_bbx.h"

namespace {

struct Point {
  double x[3];
  double getcenter(unsigned i) const { return x[i]; }
  double getradius2() const { return 0.0; }
};

// ACA start row of a bbx row cluster (the bbx tree has no icom): the position
// in [b1, b1+n1) whose point is closest to the points' centre of mass, centre
// summed in position order, first minimum wins.
unsigned bbx_start_row(const std::vector<Point>& pts, const std::vector<unsigned>& op, unsigned b1,
                       unsigned n1) {
  double c[3] = {0, 0, 0};
  for (unsigned i = b1; i < b1 + n1; ++i)
    for (int j = 0; j < 3; ++j) c[j] += pts[op[i]].x[j];
  const double s = 1.0 / n1;
  for (int j = 0; j < 3; ++j) c[j] *= s;
  double best = INFINITY;
  unsigned arg = b1;
  for (unsigned i = b1; i < b1 + n1; ++i) {
    double d = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = pts[op[i]].x[j] - c[j];
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

struct ahm_tree_s {
  int kind = 0;
  bool sym = false;   // symmetric block tree (subdivide_sym)
  uint32_t N = 0;
  std::vector<Point> pts;
  std::vector<unsigned> op, po;
  cluster* root = nullptr;
  blcluster* bl = nullptr;
  std::vector<blcluster*> leaves;
  std::vector<uint32_t> i0;
  uint32_t ncl = 0;
  ~ahm_tree_s() {
    delete bl;
    delete root;
  }
};

extern "C" {

int ahm_tree_build(const double* pts, uint32_t N, uint32_t bmin, double eta2, int kind,
                   ahm_tree* out) {
  if (!pts || !out || N == 0) return ahm::fail(AHM_ERR_INVALID, "ahm_tree_build: bad arguments");
  const bool sym = (kind & AHM_TREE_SYM) != 0;
  kind &= ~AHM_TREE_SYM;
  if (kind != AHM_TREE_BBX && kind != AHM_TREE_PCA)
    return ahm::fail(AHM_ERR_INVALID, "ahm_tree_build: unknown tree kind %d", kind);
  ahm_tree t = new ahm_tree_s;
  t->kind = kind;
  t->sym = sym;
  t->N = N;
  t->pts.resize(N);
  for (uint32_t i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) t->pts[i].x[j] = pts[3 * (size_t)i + j];
  t->op.resize(N);
  t->po.resize(N);
  for (uint32_t i = 0; i < N; ++i) t->op[i] = t->po[i] = i;
  unsigned nbl = 0;
  if (kind == AHM_TREE_BBX) {
    auto* r = new cluster3d_bbx<Point>(t->pts.data(), t->op.data(), 0, N);
    r->createClusterTree(bmin, t->op.data(), t->po.data());
    t->root = r;
    t->bl = new blcluster(0, 0, N, N);
  } else {
    auto* r = new bemcluster<Point>(t->pts.data(), t->op.data(), 0, N);
    r->createClusterTree(bmin, t->op.data(), t->po.data());
    t->root = r;
    t->bl = new bemblcluster<Point, Point>(r, r);
  }
  if (sym)
    t->bl->subdivide_sym(t->root, eta2, nbl);
  else
    t->bl->subdivide(t->root, t->root, eta2, nbl);
  t->bl->leaves(t->leaves);
  t->ncl = 1 + t->root->getncl();
  t->i0.resize(t->leaves.size());
  for (size_t l = 0; l < t->leaves.size(); ++l) {
    blcluster* b = t->leaves[l];
    if (kind == AHM_TREE_BBX)
      t->i0[l] = bbx_start_row(t->pts, t->op, b->getb1(), b->getn1());
    else
      t->i0[l] = static_cast<bemblcluster<Point, Point>*>(b)->getcl1()->geticom() - b->getb1();
  }
  *out = t;
  return AHM_OK;
}

int ahm_tree_destroy(ahm_tree t) {
  delete t;
  return AHM_OK;
}

int ahm_tree_info(ahm_tree t, uint32_t* n_leaves, uint32_t* n_clusters) {
  if (!t) return ahm::fail(AHM_ERR_INVALID, "ahm_tree_info: NULL tree");
  if (n_leaves) *n_leaves = (uint32_t)t->leaves.size();
  if (n_clusters) *n_clusters = t->ncl;
  return AHM_OK;
}

int ahm_tree_perm(ahm_tree t, uint32_t* op_perm, uint32_t* po_perm) {
  if (!t) return ahm::fail(AHM_ERR_INVALID, "ahm_tree_perm: NULL tree");
  if (op_perm) std::memcpy(op_perm, t->op.data(), sizeof(uint32_t) * t->N);
  if (po_perm) std::memcpy(po_perm, t->po.data(), sizeof(uint32_t) * t->N);
  return AHM_OK;
}

int ahm_tree_leaves(ahm_tree t, uint32_t* b1, uint32_t* n1, uint32_t* b2, uint32_t* n2,
                    uint8_t* adm, uint32_t* i0) {
  if (!t) return ahm::fail(AHM_ERR_INVALID, "ahm_tree_leaves: NULL tree");
  for (size_t l = 0; l < t->leaves.size(); ++l) {
    const blcluster* b = t->leaves[l];
    if (b1) b1[l] = b->getb1();
    if (n1) n1[l] = b->getn1();
    if (b2) b2[l] = b->getb2();
    if (n2) n2[l] = b->getn2();
    if (adm) adm[l] = b->isadm() ? 1 : 0;
    if (i0) i0[l] = t->i0[l];
  }
  return AHM_OK;
}

}  // extern "C"

extern "C" int ahm_assemble_tree(ahm_ctx ctx, ahm_tree t, const ahm_kernel* kernel, double eps,
                                 uint32_t rankmax, const uint32_t* rank_hint, ahm_dtype storage,
                                 int keep_leaves, ahm_assembly_stats* stats, uint32_t* aca_rank,
                                 ahm_hmat* out) {
  if (!t) return ahm::fail(AHM_ERR_INVALID, "ahm_assemble_tree: NULL tree");
  const size_t nl = t->leaves.size();
  std::vector<uint32_t> b1(nl), n1(nl), b2(nl), n2(nl);
  std::vector<uint8_t> adm(nl);
  ahm_tree_leaves(t, b1.data(), n1.data(), b2.data(), n2.data(), adm.data(), nullptr);
  std::vector<double> pts(3 * (size_t)t->N);
  for (uint32_t i = 0; i < t->N; ++i)
    for (int j = 0; j < 3; ++j) pts[3 * (size_t)i + j] = t->pts[i].x[j];
  if (!kernel) return ahm::fail(AHM_ERR_INVALID, "ahm_assemble_tree: NULL kernel");
  ahm_kernel k = *kernel;
  if (t->sym) k.flags |= AHM_ASSEMBLE_SYM;   // apprx_sym on a symmetric tree
  return ahm_assemble(ctx, pts.data(), t->op.data(), t->N, (uint32_t)nl, b1.data(), n1.data(),
                      b2.data(), n2.data(), adm.data(), t->i0.data(), &k, eps, rankmax,
                      rank_hint, storage, keep_leaves, stats, aca_rank, out);
}

// n_mat independent H-matrices (own points, tree and kernel) stacked into one
// block-diagonal H-matrix: one batched ACA / recompression launch for all
// leaves, one launch per phase for every product afterwards.
extern "C" int ahm_assemble_batch(ahm_ctx ctx, uint32_t n_mat, const ahm_tree* trees,
                                  const ahm_kernel* kernels, double eps, uint32_t rankmax,
                                  ahm_dtype storage, int keep_leaves, ahm_assembly_stats* stats,
                                  ahm_hmat* out) {
  if (!ctx || !trees || !kernels || !out || n_mat == 0)
    return ahm::fail(AHM_ERR_INVALID, "ahm_assemble_batch: bad arguments");
  std::vector<uint32_t> off(n_mat + 1, 0);
  for (uint32_t s = 0; s < n_mat; ++s) {
    if (!trees[s]) return ahm::fail(AHM_ERR_INVALID, "ahm_assemble_batch: tree %u is NULL", s);
    if (trees[s]->sym)
      return ahm::fail(AHM_ERR_UNSUPPORTED, "ahm_assemble_batch: tree %u is symmetric", s);
    if ((uint64_t)off[s] + trees[s]->N >= 0x7fffffffu)
      return ahm::fail(AHM_ERR_UNSUPPORTED, "ahm_assemble_batch: too many points");
    off[s + 1] = off[s] + trees[s]->N;
  }
  const uint32_t N = off[n_mat];
  std::vector<double> pts(3 * (size_t)N);
  std::vector<uint32_t> op(N), b1, n1, b2, n2, i0, lk;
  std::vector<uint8_t> adm;
  std::vector<ahm::KernelSpec> K(n_mat);
  for (uint32_t s = 0; s < n_mat; ++s) {
    const ahm_tree t = trees[s];
    for (uint32_t i = 0; i < t->N; ++i) {
      for (int j = 0; j < 3; ++j) pts[3 * ((size_t)off[s] + i) + j] = t->pts[i].x[j];
      op[off[s] + i] = off[s] + t->op[i];
    }
    for (size_t l = 0; l < t->leaves.size(); ++l) {
      const blcluster* b = t->leaves[l];
      b1.push_back(off[s] + b->getb1());
      n1.push_back(b->getn1());
      b2.push_back(off[s] + b->getb2());
      n2.push_back(b->getn2());
      adm.push_back(b->isadm() ? 1 : 0);
      i0.push_back(t->i0[l]);
      lk.push_back(s);
    }
    K[s] = ahm::make_kernel(kernels[s]);
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return ahm::fail(AHM_ERR_HIP, "hipSetDevice");
  std::vector<ahm_leaf_desc> leaves;
  ahm::DeviceBuf data;
  std::shared_ptr<ahm::PlanJob> plan;   // after leaves: destroyed (its thread joined) first
  AHM_TRY(ahm::assemble_leaves(ctx, pts.data(), op.data(), N, (uint32_t)b1.size(), b1.data(),
                               n1.data(), b2.data(), n2.data(), adm.data(), i0.data(), K.data(),
                               n_mat, lk.data(), eps, rankmax, nullptr, leaves, data, stats, nullptr, &plan));
  AHM_TRY(ahm::hmat_from_device(ctx, N, N, std::move(leaves), data, storage, keep_leaves, out, plan));
  (*out)->seg_off = off;
  return AHM_OK;
}
