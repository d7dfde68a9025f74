// mblock.h -- leaf storage of an H-matrix (drop-in for the reference's
// mblock<T>, Include/mblock.h:29-400; storage only).
//
// A low-rank leaf U V^T is stored column-wise as (U | V): U is n1 x k, V is
// n2 x k (Include/mblock.h:48); a dense leaf is n1 x n2 column-major.  The
// products run on the GPU through the H-matrix functions of H.h; every
// mutator bumps a library-wide epoch so cached device copies of an H-matrix
// are refreshed.  Writing through getdata() must be followed by
// ahm_surface_invalidate().
#ifndef AHMED_MBLOCK_H
#define AHMED_MBLOCK_H

#include <cassert>
#include <cstring>

extern "C" void ahm_surface_invalidate(void);

template <class T>
class mblock {
 protected:
  T* data;
  unsigned n1, n2;
  unsigned bl_rank;
  struct st_info {
    unsigned is_LrM : 1;
  } info;

 public:
  mblock(unsigned m, unsigned n) : data(nullptr), n1(m), n2(n), bl_rank(0) { info.is_LrM = 1; }
  ~mblock() { delete[] data; }
  mblock(const mblock&) = delete;
  mblock& operator=(const mblock&) = delete;

  // number of stored values (Include/mblock.h:286-291)
  unsigned long nvals() const {
    return isGeM() ? (unsigned long)n1 * n2 : (unsigned long)(n1 + n2) * bl_rank;
  }
  unsigned long size() const { return sizeof(T) * nvals() + sizeof(mblock<T>); }
  unsigned getn1() const { return n1; }
  unsigned getn2() const { return n2; }
  bool isLrM() const { return info.is_LrM; }
  bool isGeM() const { return !info.is_LrM; }
  unsigned rank() const {
    assert(isLrM());
    return bl_rank;
  }
  T* getdata() const { return data; }

  void freedata() {
    delete[] data;
    data = nullptr;
    info.is_LrM = 1;
    bl_rank = 0;
    ahm_surface_invalidate();
  }
  void setrank(unsigned k) {
    delete[] data;
    bl_rank = k;
    info.is_LrM = 1;
    data = k ? new T[(size_t)k * (n1 + n2)] : nullptr;
    ahm_surface_invalidate();
  }
  void setGeM() {
    delete[] data;
    info.is_LrM = 0;
    data = new T[(size_t)n1 * n2];
    ahm_surface_invalidate();
  }
  // copy a low-rank matrix U V^T without recompression (mblock::cpyLrM)
  void cpyLrM(unsigned k, const T* U, unsigned ldU, const T* V, unsigned ldV) {
    setrank(k);
    for (unsigned l = 0; l < k; ++l) {
      std::memcpy(data + (size_t)l * n1, U + (size_t)l * ldU, sizeof(T) * n1);
      std::memcpy(data + (size_t)k * n1 + (size_t)l * n2, V + (size_t)l * ldV, sizeof(T) * n2);
    }
  }
  void cpyGeM(const T* A) {
    setGeM();
    std::memcpy(data, A, sizeof(T) * (size_t)n1 * n2);
  }
};

#endif
