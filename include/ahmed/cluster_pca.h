// cluster_pca.h -- principal-component clusters (drop-in for the reference's
// cluster_pca / cluster3d_pca, Include/cluster_pca.h:32-260).
//
// init (:82-115): centre of mass and diam2 = max |2 (x - com)|^2.
// Split (:118-202): main direction by 10 power iterations on the packed
// covariance, points with <x,u> >= <com,u> moved to the back by swaps; a
// one-sided split falls back to the middle.  Sons when both halves hold at
// least bmin points (:204-217).  Admissible iff min(diam2) < eta2 |com1-com2|^2.
#ifndef AHMED_CLUSTER_PCA_H
#define AHMED_CLUSTER_PCA_H

#include <cmath>
#include <utility>

#include "cluster.h"

template <class T>
class cluster3d_pca : public cluster_geo {
 protected:
  T* dofs;
  double com[3];

  void init(const unsigned* op_perm) {
    for (int j = 0; j < 3; ++j) com[j] = 0.0;
    for (unsigned i = nbeg; i < nend; ++i)
      for (int j = 0; j < 3; ++j) com[j] += dofs[op_perm[i]].getcenter(j);
    const double s = 1.0 / (nend - nbeg);
    for (int j = 0; j < 3; ++j) com[j] *= s;
    diam2 = 0.0;
    for (unsigned i = nbeg; i < nend; ++i) {
      double d2 = 0.0;
      for (int j = 0; j < 3; ++j) {
        const double d1 = 2 * (dofs[op_perm[i]].getcenter(j) - com[j]);
        d2 += d1 * d1;
      }
      if (d2 > diam2) diam2 = d2;
    }
  }

  virtual cluster3d_pca* clone(unsigned* op_perm, unsigned k, unsigned l) const {
    return new cluster3d_pca(dofs, op_perm, k, l);
  }

  unsigned split(unsigned* op_perm, unsigned* po_perm) {
    // packed upper covariance c[(j,k)], j <= k
    double c[6] = {0, 0, 0, 0, 0, 0}, d[3], u[3];
    for (unsigned i = nbeg; i < nend; ++i) {
      for (int j = 0; j < 3; ++j) d[j] = dofs[op_perm[i]].getcenter(j) - com[j];
      int l = 0;
      for (int j = 0; j < 3; ++j)
        for (int k = j; k < 3; ++k) c[l++] += d[j] * d[k];
    }
    for (int j = 0; j < 3; ++j) u[j] = 1.0 / std::sqrt(3.0);
    for (int it = 0; it < 10; ++it) {
      for (int j = 0; j < 3; ++j) d[j] = u[j];
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        int k = 0;
        for (; k < j; ++k) s += c[k * 3 - k * (k + 1) / 2 + j] * d[k];
        for (; k < 3; ++k) s += c[j * 3 - j * (j + 1) / 2 + k] * d[k];
        u[j] = s;
      }
      double nu = 0.0;
      for (int j = 0; j < 3; ++j) nu += u[j] * u[j];
      nu = 1.0 / std::sqrt(nu);
      for (int j = 0; j < 3; ++j) u[j] *= nu;
    }
    double mid = 0.0;
    for (int j = 0; j < 3; ++j) mid += u[j] * com[j];
    unsigned back = nend;
    for (unsigned i = nbeg; i < back;) {
      const T& v = dofs[op_perm[i]];
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += v.getcenter(j) * u[j];
      if (s < mid) {
        ++i;
        continue;
      }
      --back;
      std::swap(op_perm[i], op_perm[back]);
      po_perm[op_perm[i]] = i;
      po_perm[op_perm[back]] = back;
    }
    if (back == nbeg || back == nend) back = (nend + nbeg) / 2;
    return back;
  }

 public:
  cluster3d_pca(T* d, unsigned* op_perm, unsigned k, unsigned l) : cluster_geo(k, l), dofs(d) {
    init(op_perm);
  }
  unsigned dim() const override { return 3; }
  double getcom(int i) const { return com[i]; }

  bool isadm(double eta2, cluster* cl, bl_info& info) override {
    const cluster3d_pca* p = static_cast<const cluster3d_pca*>(cl);
    info.is_sep = false;
    double dc = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double e = com[j] - p->com[j];
      dc += e * e;
    }
    const double d2 = diam2 < p->diam2 ? diam2 : p->diam2;
    return (info.is_adm = d2 < eta2 * dc);
  }

  void createClusterTree(unsigned bmin, unsigned* op_perm, unsigned* po_perm) override {
    const unsigned sep = split(op_perm, po_perm);
    if (sep >= nbeg + bmin && nend >= sep + bmin) {
      cluster* s[2];
      s[0] = clone(op_perm, nbeg, sep);
      s[0]->createClusterTree(bmin, op_perm, po_perm);
      s[1] = clone(op_perm, sep, nend);
      s[1]->createClusterTree(bmin, op_perm, po_perm);
      setsons(2, s);
    }
  }
};

#endif
