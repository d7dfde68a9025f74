// bemblcluster.h -- block cluster holding its two bemclusters (drop-in for
// the reference's bemblcluster, Include/bemblcluster.h:17-40).
#ifndef AHMED_BEMBLCLUSTER_H
#define AHMED_BEMBLCLUSTER_H

#include "bemcluster.h"
#include "blcluster.h"

template <class T1, class T2>
class bemblcluster : public blcluster {
  bemcluster<T1>* cl1;
  bemcluster<T2>* cl2;

 public:
  bemblcluster(unsigned i1, unsigned i2, unsigned m1, unsigned m2)
      : blcluster(i1, i2, m1, m2), cl1(nullptr), cl2(nullptr) {}
  bemblcluster(bemcluster<T1>* p1, bemcluster<T2>* p2) : blcluster(p1, p2), cl1(p1), cl2(p2) {}
  blcluster* clone(cluster* p1, cluster* p2) const override {
    return new bemblcluster(static_cast<bemcluster<T1>*>(p1), static_cast<bemcluster<T2>*>(p2));
  }
  bool aca_start_row(unsigned& icom) const override {
    if (!cl1) return false;
    icom = cl1->geticom();
    return true;
  }
  bemcluster<T1>* getcl1() const { return cl1; }
  bemcluster<T2>* getcl2() const { return cl2; }
};

#endif
