// drop_in.cpp -- a reference-style user program on the MI355X surface.
//
// Follows the call sequence of the reference's BEM example
// (Examples/PrecondBEM.cpp:263-305,440-456): points -> cluster tree ->
// blcluster::subdivide -> allocmbls -> matrix generation -> GeHMatrix ->
// Krylov solve, here with the GPU assembly, products and solvers.  Checks the
// H-matvec against direct kernel sums and the solvers against the known
// solution x = 1.  Prints one "drop_in OK" line on success, exits 1 otherwise.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "ahmed/bemblcluster.h"
#include "ahmed/cluster_bbx.h"
#include "ahmed/matgen_gpu.h"
#include "ahmed/solvers.h"

struct Point {
  double x[3];
  double getcenter(unsigned i) const { return x[i]; }
  double getradius2() const { return 0.0; }
};

static double kernel(const Point& a, const Point& b, double w) {
  const double dx = a.x[0] - b.x[0], dy = a.x[1] - b.x[1], dz = a.x[2] - b.x[2];
  return w / (4.0 * M_PI * std::sqrt(dx * dx + dy * dy + dz * dz));
}

static int run(bool pca, unsigned N) {
  std::mt19937_64 rng(11);
  std::normal_distribution<double> nd;
  std::vector<Point> pts(N);
  std::vector<double> xyz(3 * (size_t)N);
  for (unsigned i = 0; i < N; ++i) {
    double v[3] = {nd(rng), nd(rng), nd(rng)};
    const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int j = 0; j < 3; ++j) xyz[3 * (size_t)i + j] = pts[i].x[j] = v[j] / r;
  }
  std::vector<unsigned> op(N), po(N);
  for (unsigned i = 0; i < N; ++i) op[i] = po[i] = i;
  cluster* root;
  blcluster* bl;
  if (pca) {
    auto* r = new bemcluster<Point>(pts.data(), op.data(), 0, N);
    r->createClusterTree(48, op.data(), po.data());
    root = r;
    bl = new bemblcluster<Point, Point>(r, r);
  } else {
    auto* r = new cluster3d_bbx<Point>(pts.data(), op.data(), 0, N);
    r->createClusterTree(48, op.data(), po.data());
    root = r;
    bl = new blcluster(0, 0, N, N);
  }
  unsigned nblcks = 0;
  bl->subdivide(root, root, 0.8 * 0.8, nblcks);
  std::fprintf(stderr, "[drop_in] %s tree: %u leaves\n", pca ? "bemcluster" : "cluster3d_bbx", nblcks);
  mblock<double>** A;
  allocmbls(bl, A);
  const double w = 4.0 * M_PI / N;
  GpuKernel G{AHM_KERNEL_LAPLACE, 0.0, w, xyz.data(), op.data(), N};
  matgenGeH_gpu(G, nblcks, bl, 1e-8, ~0u, A);
  std::fprintf(stderr, "[drop_in] assembled\n");
  for (unsigned i = 0; i < nblcks; ++i)
    if (!A[i]) {
      std::printf("leaf %u missing\n", i);
      return 1;
    }
  // H x against direct sums over sampled rows (permuted ordering)
  std::vector<double> x(N), y(N, 0.0);
  for (unsigned i = 0; i < N; ++i) x[i] = std::sin(0.37 * i);
  if (!mltaGeHVec(1.0, bl, A, x.data(), y.data())) return 1;
  double err = 0.0, ref = 0.0;
  for (unsigned s = 0; s < 40; ++s) {
    const unsigned i = (s * 7919u) % N;
    double d = 0.0;
    for (unsigned j = 0; j < N; ++j)
      d += (i == j ? 1.0 : kernel(pts[op[i]], pts[op[j]], w)) * x[j];
    err = std::fmax(err, std::fabs(d - y[i]));
    ref = std::fmax(ref, std::fabs(d));
  }
  if (!(err <= 1e-6 * ref)) {
    std::printf("matvec error %.3e (ref %.3e)\n", err, ref);
    return 1;
  }
  // GMRes and CG through the operator interface
  std::fprintf(stderr, "[drop_in] matvec checked\n");
  // GeHMatrix frees its leaves through the tree (freembls walks bl), so it
  // must be destroyed before the tree: the reference's order rule
  // (Include/matrix.h:351).  The inner scope ends M's life first.
  {
    GeHMatrix<double> M(N, N, bl);
    M.blcks = A;
    std::vector<double> ones(N, 1.0), b(N, 0.0), sol(N, 0.0);
    M.amux(1.0, ones.data(), b.data());
    double eps = 1e-10;
    unsigned steps = 500;
    unsigned rc = GMRes(M, b.data(), sol.data(), eps, 50, steps);
    double xerr = 0.0;
    for (unsigned i = 0; i < N; ++i) xerr = std::fmax(xerr, std::fabs(sol[i] - 1.0));
    if (rc != 0 || xerr > 1e-8) {
      std::printf("GMRes rc=%u eps=%.3e steps=%u err=%.3e\n", rc, eps, steps, xerr);
      return 1;
    }
    std::vector<double> sol2(N, 0.0);
    double eps2 = 1e-10;
    unsigned steps2 = 500;
    unsigned rc2 = CG(M, b.data(), sol2.data(), eps2, steps2);
    std::printf("drop_in %s N=%u leaves=%u size=%.1f MB max_rank=%u matvec_err=%.2e "
                "GMRes(%u steps, %.1e) CG(rc=%u, %u steps, %.1e)\n",
                pca ? "bemcluster" : "cluster3d_bbx", N, nblcks, sizeH(bl, A) / 1e6, Hmax_rank(bl, A),
                err / ref, steps, eps, rc2, steps2, eps2);
  }
  delete bl;
  delete root;
  return 0;
}

// The symmetric matrix of the reference's BEM example
// (Examples/PrecondBEM.cpp:360-375,440-456): subdivide_sym, the symmetric
// assembly, mltaHeHVec and CG / GMRes through HeHMatrix.
static int run_sym(unsigned N) {
  std::mt19937_64 rng(13);
  std::normal_distribution<double> nd;
  std::vector<Point> pts(N);
  std::vector<double> xyz(3 * (size_t)N);
  for (unsigned i = 0; i < N; ++i) {
    double v[3] = {nd(rng), nd(rng), nd(rng)};
    const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int j = 0; j < 3; ++j) xyz[3 * (size_t)i + j] = pts[i].x[j] = v[j] / r;
  }
  std::vector<unsigned> op(N), po(N);
  for (unsigned i = 0; i < N; ++i) op[i] = po[i] = i;
  auto* root = new cluster3d_bbx<Point>(pts.data(), op.data(), 0, N);
  root->createClusterTree(48, op.data(), po.data());
  blcluster* bl = new blcluster(0, 0, N, N);
  unsigned nblcks = 0;
  bl->subdivide_sym(root, 0.8 * 0.8, nblcks);
  std::fprintf(stderr, "[drop_in] symmetric tree: %u leaves\n", nblcks);
  const double w = 4.0 * M_PI / N;
  GpuKernel G{AHM_KERNEL_LAPLACE, 0.0, w, xyz.data(), op.data(), N};
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  {
    HeHMatrix<double> M(N, bl);
    allocmbls(bl, M.blcks);
    matgenHeH_gpu(G, nblcks, bl, 1e-8, ~0u, M.blcks);
    for (blcluster* b : lv) {
      mblock<double>* m = M.blcks[b->getidx()];
      if (!m || b->getb1() > b->getb2() || m->isHeM() != (b->getb1() == b->getb2())) {
        std::printf("leaf %u: missing, below the diagonal or wrong storage\n", b->getidx());
        return 1;
      }
    }
    // H x against direct sums over sampled rows (permuted ordering)
    std::vector<double> x(N), y(N, 0.0);
    for (unsigned i = 0; i < N; ++i) x[i] = std::sin(0.37 * i);
    mltaHeHVec(1.0, bl, M.blcks, x.data(), y.data());
    double err = 0.0, ref = 0.0;
    for (unsigned s = 0; s < 40; ++s) {
      const unsigned i = (s * 7919u) % N;
      double d = 0.0;
      for (unsigned j = 0; j < N; ++j) d += (i == j ? 1.0 : kernel(pts[op[i]], pts[op[j]], w)) * x[j];
      err = std::fmax(err, std::fabs(d - y[i]));
      ref = std::fmax(ref, std::fabs(d));
    }
    if (!(err <= 1e-6 * ref)) {
      std::printf("symmetric matvec error %.3e (ref %.3e)\n", err, ref);
      return 1;
    }
    std::vector<double> ones(N, 1.0), b(N, 0.0), sol(N, 0.0), sol2(N, 0.0);
    M.amux(1.0, ones.data(), b.data());
    double eps = 1e-10;
    unsigned steps = 500;
    const unsigned rc = CG(M, b.data(), sol.data(), eps, steps);
    double xerr = 0.0;
    for (unsigned i = 0; i < N; ++i) xerr = std::fmax(xerr, std::fabs(sol[i] - 1.0));
    double eps2 = 1e-10;
    unsigned steps2 = 500;
    const unsigned rc2 = GMRes(M, b.data(), sol2.data(), eps2, 30, steps2);
    if (rc != 0 || rc2 != 0 || xerr > 1e-8) {
      std::printf("CG rc=%u steps=%u err=%.3e GMRes rc=%u\n", rc, steps, xerr, rc2);
      return 1;
    }
    std::printf("drop_in symmetric N=%u leaves=%u size=%.1f MB matvec_err=%.2e CG(%u steps, %.1e) "
                "GMRes(%u steps, %.1e)\n",
                N, nblcks, sizeH(bl, M.blcks) / 1e6, err / ref, steps, eps, steps2, eps2);
  }   // M (and its leaves) go before the tree, as Include/matrix.h:305 requires
  delete bl;
  delete root;
  return 0;
}

int main(int argc, char** argv) {
  // unbuffered: every line reaches the caller even if the process dies
  std::setvbuf(stdout, nullptr, _IONBF, 0);
  if (argc > 1 && std::string(argv[1]) == "sym") {
    if (run_sym(8000)) return 1;
    std::printf("drop_in sym OK\n");
    return 0;
  }
  if (run(false, 12000) || run(true, 9000)) return 1;
  std::printf("drop_in OK\n");
  return 0;
}
