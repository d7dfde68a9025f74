// blcluster.h -- block cluster tree (drop-in for the reference's blcluster,
// Include/blcluster.h:24-394, cluster/blcluster.cpp:40-66).
//
// A block (b1, n1, b2, n2) is a leaf when its cluster pair is admissible or
// when either cluster is a leaf; otherwise it gets ns1 x ns2 sons stored
// row-major.  Leaves are numbered (idx) in depth-first, row-major son order,
// which is also the order the H-matrix products visit them.  The symmetric
// tree (subdivide_sym) leaves the sons below the diagonal NULL.
#ifndef AHMED_BLCLUSTER_H
#define AHMED_BLCLUSTER_H

#include <vector>

#include "cluster.h"

class blcluster {
 protected:
  unsigned n1, n2, b1, b2;
  unsigned ns1, ns2;
  blcluster** sons;
  unsigned idx;
  bl_info info;

  void subdivide_(cluster* cl1, cluster* cl2, double eta2, unsigned& nblcks, unsigned lvl,
                  unsigned maxdepth) {
    if (cl1->isadm(eta2, cl2, info)) {
      setsons(0, 0, nullptr);
      idx = nblcks++;
      return;
    }
    if ((!maxdepth || lvl < maxdepth) && cl1->isnleaf() && cl2->isnleaf()) {
      ns1 = cl1->getns();
      ns2 = cl2->getns();
      sons = new blcluster*[ns1 * ns2];
      for (unsigned i = 0; i < ns1; ++i)
        for (unsigned j = 0; j < ns2; ++j) {
          blcluster* s = clone(cl1->getson(i), cl2->getson(j));
          sons[i * ns2 + j] = s;
          s->subdivide_(cl1->getson(i), cl2->getson(j), eta2, nblcks, lvl + 1, maxdepth);
        }
    } else {
      setsons(0, 0, nullptr);
      idx = nblcks++;
    }
  }

 public:
  blcluster(unsigned i1, unsigned i2, unsigned m1, unsigned m2)
      : n1(m1), n2(m2), b1(i1), b2(i2), ns1(0), ns2(0), sons(nullptr), idx(0) {
    info.is_adm = 0;
    info.is_sep = 0;
  }
  blcluster(cluster* c1, cluster* c2) : blcluster(c1->getnbeg(), c2->getnbeg(), c1->size(), c2->size()) {}
  virtual ~blcluster() { setsons(0, 0, nullptr); }
  blcluster(const blcluster&) = delete;
  blcluster& operator=(const blcluster&) = delete;

  void setsons(unsigned m1, unsigned m2, blcluster** p) {
    const unsigned ns = ns1 * ns2;
    for (unsigned i = 0; i < ns; ++i) delete sons[i];
    delete[] sons;
    sons = nullptr;
    ns1 = m1;
    ns2 = m2;
    if (m1 * m2) {
      sons = new blcluster*[m1 * m2];
      for (unsigned i = 0; i < m1 * m2; ++i) sons[i] = p[i];
    }
  }

  unsigned getb1() const { return b1; }
  unsigned getb2() const { return b2; }
  unsigned getn1() const { return n1; }
  unsigned getn2() const { return n2; }
  bool isadm() const { return info.is_adm; }
  bool isnadm() const { return !info.is_adm; }
  void setadm(bool b) { info.is_adm = b; }
  bool isleaf() const { return !(ns1 && ns2); }
  bool isnleaf() const { return ns1 && ns2; }
  unsigned getnrs() const { return ns1; }
  unsigned getncs() const { return ns2; }
  unsigned getns() const { return ns1 * ns2; }
  blcluster* getson(unsigned i, unsigned j) const { return sons[i * ns2 + j]; }
  bool isdbl() const { return b1 == b2; }
  unsigned getidx() const { return idx; }
  void setidx(unsigned i) { idx = i; }

  unsigned long nleaves() const {
    if (isleaf()) return 1;
    unsigned long n = 0;
    for (unsigned i = 0; i < getns(); ++i)
      if (sons[i]) n += sons[i]->nleaves();
    return n;
  }
  unsigned long nadmleaves() const {
    if (isleaf()) return isadm();
    unsigned long n = 0;
    for (unsigned i = 0; i < getns(); ++i)
      if (sons[i]) n += sons[i]->nadmleaves();
    return n;
  }
  // leaves in idx order (gen_BlSequence, parallel/gen_blseq.cpp:15-27,68-75)
  void leaves(std::vector<blcluster*>& out) {
    if (isleaf()) {
      out.push_back(this);
      return;
    }
    for (unsigned i = 0; i < getns(); ++i)
      if (sons[i]) sons[i]->leaves(out);
  }

  virtual blcluster* clone(cluster* c1, cluster* c2) const { return new blcluster(c1, c2); }

  // ACA start row of this leaf's row cluster when the tree carries one
  // (bemblcluster: bemcluster::icom, Include/apprx.h:155); false otherwise.
  virtual bool aca_start_row(unsigned& icom) const {
    (void)icom;
    return false;
  }

  virtual void subdivide(cluster* cl1, cluster* cl2, double eta2, unsigned& nblcks,
                         unsigned maxdepth = 0) {
    nblcks = 0;
    subdivide_(cl1, cl2, eta2, nblcks, 0, maxdepth);
  }

  // symmetric block cluster tree of (cl, cl) (cluster/blcluster.cpp:90-113):
  // a diagonal block is never admissible; its sons below the diagonal stay
  // NULL, the diagonal sons recurse symmetrically and the sons above the
  // diagonal are subdivided as in the unsymmetric tree, in row-major order
  virtual void subdivide_sym(cluster* cl, double eta2, unsigned& nblcks, unsigned maxdepth = 0) {
    nblcks = 0;
    subdivide_sym_(cl, eta2, nblcks, 0, maxdepth);
  }

 protected:
  void subdivide_sym_(cluster* cl, double eta2, unsigned& nblcks, unsigned lvl, unsigned maxdepth) {
    setadm(false);
    if ((!maxdepth || lvl < maxdepth) && cl->isnleaf()) {
      const unsigned ns = cl->getns();
      blcluster** s = new blcluster*[ns * ns]();
      ns1 = ns2 = ns;
      sons = s;
      for (unsigned i = 0; i < ns; ++i) {
        cluster* ci = cl->getson(i);
        blcluster* d = clone(ci, ci);
        s[i * ns + i] = d;
        d->subdivide_sym_(ci, eta2, nblcks, lvl + 1, maxdepth);
        for (unsigned j = i + 1; j < ns; ++j) {
          blcluster* u = clone(ci, cl->getson(j));
          s[i * ns + j] = u;
          u->subdivide_(ci, cl->getson(j), eta2, nblcks, lvl + 1, maxdepth);
        }
      }
    } else {
      setsons(0, 0, nullptr);
      idx = nblcks++;
    }
  }
};

// gen_BlSequence: array of the leaves in idx order (caller deletes[])
inline void gen_BlSequence(blcluster* bl, blcluster**& list) {
  std::vector<blcluster*> v;
  bl->leaves(v);
  list = new blcluster*[v.size()];
  for (size_t i = 0; i < v.size(); ++i) list[i] = v[i];
}

#endif
