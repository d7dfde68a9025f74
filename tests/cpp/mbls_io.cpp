// mbls_io.cpp -- host-only use of the AHMED-style surface's file format, the
// way a reference program reloads a saved H-matrix: points -> cluster tree ->
// block cluster tree (bmin and eta from the file) -> loadmbls -> savembls.
// Usage: mbls_io POINTS IN OUT  (POINTS: N x 3 doubles).  No GPU call is made.
#include <cstdio>
#include <fstream>
#include <vector>

#include "ahmed/H.h"
#include "ahmed/cluster_bbx.h"

struct Point {
  double x[3];
  double getcenter(unsigned i) const { return x[i]; }
  double getradius2() const { return 0.0; }
};

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: mbls_io POINTS IN OUT\n");
    return 2;
  }
  std::ifstream ps(argv[1], std::ios::binary | std::ios::ate);
  const size_t bytes = (size_t)ps.tellg();
  ps.seekg(0);
  const unsigned N = (unsigned)(bytes / sizeof(Point));
  std::vector<Point> pts(N);
  ps.read((char*)pts.data(), (std::streamsize)(N * sizeof(Point)));

  std::ifstream is(argv[2], std::ios::binary);
  double eta = 0.0;
  unsigned bmin = 0;
  mblock<double>** A = nullptr;
  loadmbls(eta, bmin, A, is);

  std::vector<unsigned> op(N), po(N);
  for (unsigned i = 0; i < N; ++i) op[i] = po[i] = i;
  auto* root = new cluster3d_bbx<Point>(pts.data(), op.data(), 0, N);
  root->createClusterTree(bmin, op.data(), po.data());
  blcluster* bl = new blcluster(0, 0, N, N);
  unsigned nblcks = 0;
  bl->subdivide(root, root, eta * eta, nblcks);
  std::vector<blcluster*> lv;
  bl->leaves(lv);
  unsigned long nv = 0;
  for (blcluster* b : lv) {
    mblock<double>* m = A[b->getidx()];
    if (!m || m->getn1() != b->getn1() || m->getn2() != b->getn2()) {
      std::printf("leaf %u does not match the tree\n", b->getidx());
      return 1;
    }
    nv += m->nvals();
  }
  std::ofstream os(argv[3], std::ios::binary);
  savembls(eta, bmin, bl, A, os);
  os.close();
  std::printf("mbls_io eta=%.17g bmin=%u leaves=%u values=%lu\n", eta, bmin, nblcks, nv);
  freembls(bl, A);
  delete bl;
  delete root;
  return 0;
}
