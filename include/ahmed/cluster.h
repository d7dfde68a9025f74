// cluster.h -- index-range clusters (drop-in for the reference's cluster,
// Include/cluster.h:34-173).  A cluster owns the permuted positions
// [nbeg, nend); sons split that range.  Host-side only.
#ifndef AHMED_CLUSTER_H
#define AHMED_CLUSTER_H

#include <cassert>
#include <vector>

struct bl_info {
  unsigned is_adm : 1;  // is the block admissible?
  unsigned is_sep : 1;  // (sparse matrices only)
};

class cluster {
 protected:
  unsigned nbeg, nend;
  unsigned nsons;
  cluster** sons;

 public:
  cluster(unsigned k, unsigned l) : nbeg(k), nend(l), nsons(0), sons(nullptr) {}
  virtual ~cluster() { setsons(0, nullptr); }
  cluster(const cluster&) = delete;
  cluster& operator=(const cluster&) = delete;

  unsigned getnbeg() const { return nbeg; }
  unsigned getnend() const { return nend; }
  void setnbeg(unsigned b) { nbeg = b; }
  void setnend(unsigned e) { nend = e; }
  unsigned size() const { return nend - nbeg; }
  bool isnleaf() const { return nsons != 0; }
  bool isleaf() const { return nsons == 0; }
  unsigned getns() const { return nsons; }
  cluster* getson(unsigned i) const {
    assert(i < nsons);
    return sons[i];
  }
  // takes ownership of the n sons p[0..n)
  void setsons(unsigned n, cluster** p) {
    for (unsigned i = 0; i < nsons; ++i) delete sons[i];
    delete[] sons;
    sons = nullptr;
    nsons = n;
    if (n) {
      sons = new cluster*[n];
      for (unsigned i = 0; i < n; ++i) sons[i] = p[i];
    }
  }

  // is the pair (this, cl) admissible for eta2 = eta^2?  Sets info.is_adm.
  virtual bool isadm(double eta2, cluster* cl, bl_info& info) = 0;
  // recursive subdivision; op_perm[pos] = original index, po_perm its inverse
  virtual void createClusterTree(unsigned bmin, unsigned* op_perm, unsigned* po_perm) = 0;

  unsigned getncl() const {
    unsigned n = nsons;
    for (unsigned i = 0; i < nsons; ++i) n += sons[i]->getncl();
    return n;
  }
  unsigned getMaxDepth() const {
    unsigned d = 0;
    for (unsigned i = 0; i < nsons; ++i) {
      const unsigned s = 1 + sons[i]->getMaxDepth();
      if (s > d) d = s;
    }
    return d;
  }
  // pre-order list of the leaves of this subtree
  void leaves(std::vector<const cluster*>& out) const {
    if (isleaf()) out.push_back(this);
    for (unsigned i = 0; i < nsons; ++i) sons[i]->leaves(out);
  }
};

// geometric cluster: squared diameter of its points
class cluster_geo : public cluster {
 protected:
  double diam2;

 public:
  cluster_geo(unsigned k, unsigned l) : cluster(k, l), diam2(0.0) {}
  virtual unsigned dim() const = 0;
  double getdiam2() const { return diam2; }
};

#endif
