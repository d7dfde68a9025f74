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
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

extern "C" void ahm_surface_invalidate(void);
// releases the device copies the surface keeps of the leaf array A
extern "C" void ahm_surface_forget(const void* A);

template <class T>
class mblock {
 protected:
  T* data;
  unsigned n1, n2;
  unsigned bl_rank;
  struct st_info {
    unsigned is_LrM : 1;
    unsigned is_LtM : 1;   // packed unit lower triangle, permutation on the diagonal
    unsigned is_UtM : 1;   // packed upper triangle
    unsigned is_HeM : 1;   // symmetric, upper triangle packed (diagonal leaves of HeH matrices)
  } info;

 public:
  mblock(unsigned m, unsigned n) : data(nullptr), n1(m), n2(n), bl_rank(0) {
    info.is_LrM = 1;
    info.is_LtM = info.is_UtM = info.is_HeM = 0;
  }
  ~mblock() { delete[] data; }
  mblock(const mblock&) = delete;
  mblock& operator=(const mblock&) = delete;

  // number of stored values (Include/mblock.h:286-291)
  unsigned long nvals() const {
    if (isGeM())
      return (info.is_LtM || info.is_UtM || info.is_HeM) ? (unsigned long)n1 * (n1 + 1) / 2
                                                         : (unsigned long)n1 * n2;
    return (unsigned long)(n1 + n2) * bl_rank;
  }
  unsigned long size() const { return sizeof(T) * nvals() + sizeof(mblock<T>); }
  unsigned getn1() const { return n1; }
  unsigned getn2() const { return n2; }
  bool isLrM() const { return info.is_LrM; }
  bool isGeM() const { return !info.is_LrM; }
  // triangular factor leaves on the diagonal of an H-LU (Include/mblock.h:
  // 322-328): n(n+1)/2 packed column-major values; an LtM leaf is unit lower
  // with column j's diagonal entry holding the row permutation index
  // (mblock::decomp_LU, H/mblock_D.cpp:1078-1107), a UtM leaf is upper
  bool isLtM() const { return !info.is_LrM && info.is_LtM; }
  bool isUtM() const { return !info.is_LrM && info.is_UtM; }
  // Hermitian (real: symmetric) dense leaf on the diagonal of a symmetric
  // H-matrix (Include/mblock.h:137-141,364-370): the upper triangle packed
  // column by column, (i, j) with i <= j at i + j(j+1)/2
  bool isHeM() const { return !info.is_LrM && info.is_HeM; }
  unsigned rank() const {
    assert(isLrM());
    return bl_rank;
  }
  T* getdata() const { return data; }

  void freedata() {
    delete[] data;
    data = nullptr;
    info.is_LrM = 1;
    info.is_LtM = info.is_UtM = info.is_HeM = 0;
    bl_rank = 0;
    ahm_surface_invalidate();
  }
  void setrank(unsigned k) {
    delete[] data;
    bl_rank = k;
    info.is_LrM = 1;
    info.is_LtM = info.is_UtM = info.is_HeM = 0;
    data = k ? new T[(size_t)k * (n1 + n2)] : nullptr;
    ahm_surface_invalidate();
  }
  void setGeM() {
    delete[] data;
    info.is_LrM = info.is_LtM = info.is_UtM = info.is_HeM = 0;
    data = new T[(size_t)n1 * n2];
    ahm_surface_invalidate();
  }
  void setLtM() {
    assert(n1 == n2);
    delete[] data;
    info.is_LrM = info.is_UtM = info.is_HeM = 0;
    info.is_LtM = 1;
    data = new T[(size_t)n1 * (n1 + 1) / 2];
    ahm_surface_invalidate();
  }
  void setUtM() {
    assert(n1 == n2);
    delete[] data;
    info.is_LrM = info.is_LtM = info.is_HeM = 0;
    info.is_UtM = 1;
    data = new T[(size_t)n1 * (n1 + 1) / 2];
    ahm_surface_invalidate();
  }
  void setHeM() {
    assert(n1 == n2);
    delete[] data;
    info.is_LrM = info.is_LtM = info.is_UtM = 0;
    info.is_HeM = 1;
    data = new T[(size_t)n1 * (n1 + 1) / 2];
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

  // The reference's on-disk leaf record (Include/mblock.h:737-804): n1, n2,
  // status (0 low rank, 4 dense), for low rank the rank, then the values.
  // Only the two storage kinds of the GeH path exist here; a record of the
  // reference's Hermitian / triangular / symmetric kinds stops the program
  // with a message, as the reference does for an unknown status.
  void save(std::ofstream& os) const {
    if (isLtM() || isUtM() || isHeM()) {
      std::cerr << "mblock::save: packed (triangular / Hermitian) leaves are not supported" << std::endl;
      std::exit(1);
    }
    const unsigned status = isLrM() ? 0u : 4u;
    os.write((const char*)&n1, sizeof(unsigned));
    os.write((const char*)&n2, sizeof(unsigned));
    os.write((const char*)&status, sizeof(unsigned));
    if (isLrM()) os.write((const char*)&bl_rank, sizeof(unsigned));
    if (nvals()) os.write((const char*)data, (std::streamsize)(sizeof(T) * nvals()));
  }
  void load(std::ifstream& is) {
    unsigned status = 0, k = 0;
    is.read((char*)&n1, sizeof(unsigned));
    is.read((char*)&n2, sizeof(unsigned));
    is.read((char*)&status, sizeof(unsigned));
    if (status == 0) {
      is.read((char*)&k, sizeof(unsigned));
      setrank(k);
    } else if (status == 4) {
      setGeM();
    } else {
      std::cerr << "mblock::load: leaf storage kind " << status << " is not supported" << std::endl;
      std::exit(1);
    }
    if (nvals()) is.read((char*)data, (std::streamsize)(sizeof(T) * nvals()));
  }
};

#endif
