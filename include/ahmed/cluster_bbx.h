// cluster_bbx.h -- bounding-box bisection clusters (drop-in for the
// reference's cluster_bbx / cluster3d_bbx, Include/cluster_bbx.h:24-302).
//
// T is the degree-of-freedom type: getcenter(i) and getradius2().
// Subdivision (Include/cluster_bbx.h:100-128): split the bounding box at its
// centre along the main direction (x when its extent exceeds 0.8 of the
// largest extent, else the largest), moving points with coordinate >= centre
// to the back by swaps; keep the sons only when both halves hold more than
// bmin points.  Admissibility (:245-297): min(diam1^2, diam2^2) < eta2 *
// dist^2 of the boxes, or boxes touching exactly at a corner pattern.
#ifndef AHMED_CLUSTER_BBX_H
#define AHMED_CLUSTER_BBX_H

#include <cmath>
#include <utility>

#include "cluster.h"

template <class T>
class cluster3d_bbx : public cluster_geo {
 protected:
  T* dofs;
  double lo[3], hi[3];
  unsigned maindir;
  double cntrdir;

  double dist2(const cluster3d_bbx* c) const {
    double d = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double e = c->lo[i] - hi[i];
      if (e > 0.0) d += e * e;
      else {
        const double e1 = lo[i] - c->hi[i];
        if (e1 > 0.0) d += e1 * e1;
      }
    }
    return d;
  }

  void init(const unsigned* op_perm) {
    for (int i = 0; i < 3; ++i) {
      const T& v = dofs[op_perm[nbeg]];
      const double r = std::sqrt(v.getradius2());
      lo[i] = v.getcenter(i) - r;
      hi[i] = v.getcenter(i) + r;
    }
    for (unsigned j = nbeg + 1; j < nend; ++j) {
      const T& v = dofs[op_perm[j]];
      for (int i = 0; i < 3; ++i) {
        const double x = v.getcenter(i), r = std::sqrt(v.getradius2());
        if (x - r < lo[i]) lo[i] = x - r;
        else if (x + r > hi[i]) hi[i] = x + r;
      }
    }
    diam2 = 0.0;
    unsigned imax = 0;
    double emax = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double e = hi[i] - lo[i];
      diam2 += e * e;
      if (e > emax) {
        emax = e;
        imax = (unsigned)i;
      }
    }
    maindir = (hi[0] - lo[0] > 0.8 * emax) ? 0u : imax;
    cntrdir = 0.5 * (lo[maindir] + hi[maindir]);
  }

  virtual cluster3d_bbx* clone(unsigned* op_perm, unsigned k, unsigned l) const {
    return new cluster3d_bbx(dofs, op_perm, k, l);
  }

 public:
  cluster3d_bbx(T* d, unsigned* op_perm, unsigned k, unsigned l) : cluster_geo(k, l), dofs(d) {
    init(op_perm);
  }
  unsigned dim() const override { return 3; }
  double getlo(int i) const { return lo[i]; }
  double gethi(int i) const { return hi[i]; }

  bool isadm(double eta2, cluster* cl, bl_info& info) override {
    const cluster3d_bbx* p = static_cast<const cluster3d_bbx*>(cl);
    info.is_sep = false;
    const double d2 = diam2 < p->diam2 ? diam2 : p->diam2;
    if (d2 < eta2 * dist2(p)) return (info.is_adm = true);
    const bool yc = hi[1] == p->lo[1] || p->hi[1] == lo[1];
    const bool zc = hi[2] == p->lo[2] || p->hi[2] == lo[2];
    if ((hi[0] == p->lo[0] || p->hi[0] == lo[0]) && yc && zc) return (info.is_adm = true);
    return (info.is_adm = false);
  }

  void createClusterTree(unsigned bmin, unsigned* op_perm, unsigned* po_perm) override {
    unsigned back = nend;
    for (unsigned i = nbeg; i < back;) {
      if (dofs[op_perm[i]].getcenter(maindir) < cntrdir) {
        ++i;
        continue;
      }
      --back;
      std::swap(op_perm[i], op_perm[back]);
      po_perm[op_perm[i]] = i;
      po_perm[op_perm[back]] = back;
    }
    if (back - nbeg > bmin && nend - back > bmin) {
      cluster* s[2];
      s[0] = clone(op_perm, nbeg, back);
      s[0]->createClusterTree(bmin, op_perm, po_perm);
      s[1] = clone(op_perm, back, nend);
      s[1]->createClusterTree(bmin, op_perm, po_perm);
      setsons(2, s);
    }
  }
};

#endif
