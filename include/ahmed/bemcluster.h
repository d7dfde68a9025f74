// bemcluster.h -- PCA cluster carrying the ACA start row icom (drop-in for the
// reference's bemcluster, Include/bemcluster.h:17-59).
//
// icom is chosen exactly as the reference does it: among positions
// i in [nbeg, nend) the one whose dofs[i] -- indexed WITHOUT the permutation,
// as the reference does -- lies closest to the cluster's centre of mass.
#ifndef AHMED_BEMCLUSTER_H
#define AHMED_BEMCLUSTER_H

#include "cluster_pca.h"

template <class T>
class bemcluster : public cluster3d_pca<T> {
  unsigned icom;

  bemcluster* clone(unsigned* op_perm, unsigned k, unsigned l) const override {
    return new bemcluster(this->dofs, op_perm, k, l);
  }

 public:
  bemcluster(T* d, unsigned* op_perm, unsigned k, unsigned l) : cluster3d_pca<T>(d, op_perm, k, l) {
    const unsigned b = this->nbeg, e = this->nend;
    icom = b;
    double smin = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double x = this->com[j] - d[b].getcenter(j);
      smin += x * x;
    }
    for (unsigned i = b + 1; i < e; ++i) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) {
        const double x = this->com[j] - d[i].getcenter(j);
        s += x * x;
      }
      if (s < smin) {
        icom = i;
        smin = s;
      }
    }
  }
  unsigned geticom() const { return icom; }
};

#endif
