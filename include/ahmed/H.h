// H.h -- H-matrix functions of the drop-in surface (the subset of the
// reference's Include/H.h on the matvec / assembly / Krylov path).
//
// The mblock<T>* array indexed by leaf idx is the reference's H-matrix
// (Include/H.h:21-45).  The products below run on the GPU: the first call on a
// (tree, array) pair uploads the leaves into a device arena that is reused
// until an mblock changes (see mblock.h).
#ifndef AHMED_H_H
#define AHMED_H_H

#include "blcluster.h"
#include "mblock.h"

template <class T>
void allocmbls(unsigned n, mblock<T>**& A) {
  A = new mblock<T>*[n];
  for (unsigned i = 0; i < n; ++i) A[i] = nullptr;
}
template <class T>
void allocmbls(blcluster* bl, mblock<T>**& A) {
  allocmbls((unsigned)bl->nleaves(), A);
}
template <class T>
void freembls(unsigned n, mblock<T>**& A) {
  if (A) {
    ahm_surface_forget(A);
    for (unsigned i = 0; i < n; ++i) delete A[i];
    delete[] A;
    A = nullptr;
  }
}
template <class T>
void freembls(blcluster* bl, mblock<T>**& A) {
  freembls((unsigned)bl->nleaves(), A);
}

// The reference's H-matrix file (Include/H.h:47-103): eta, bmin, the leaf
// count, then per leaf a presence byte and the mblock record (mblock::save).
// loadmbls allocates the array itself (no allocmbls needed), as there.
template <class T>
void savembls(const double eta, const unsigned bmin, blcluster* bl, mblock<T>**& A,
              std::ofstream& os) {
  if (!os) {
    std::cerr << "savembls: output stream not open" << std::endl;
    std::exit(1);
  }
  const unsigned n = (unsigned)bl->nleaves();
  os.write((const char*)&eta, sizeof(double));
  os.write((const char*)&bmin, sizeof(unsigned));
  os.write((const char*)&n, sizeof(unsigned));
  for (unsigned i = 0; i < n; ++i) {
    const bool present = A[i] != nullptr;
    os.write((const char*)&present, sizeof(bool));
    if (present) A[i]->save(os);
  }
}
template <class T>
void loadmbls(double& eta, unsigned& bmin, mblock<T>**& A, std::ifstream& is) {
  if (!is) {
    std::cerr << "loadmbls: input stream not open" << std::endl;
    std::exit(1);
  }
  unsigned n = 0;
  is.read((char*)&eta, sizeof(double));
  is.read((char*)&bmin, sizeof(unsigned));
  is.read((char*)&n, sizeof(unsigned));
  allocmbls(n, A);
  for (unsigned i = 0; i < n; ++i) {
    bool present = false;
    is.read((char*)&present, sizeof(bool));
    if (present) {
      A[i] = new mblock<T>(0, 0);
      A[i]->load(is);
    }
  }
}

// y += d A x (H/mltaGeHVec.cpp:567-571).  x, y are host vectors offset so
// that x[0] is column bl->getb2() and y[0] row bl->getb1().  Returns true
// if some leaf contributed.
bool mltaGeHVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y);
bool mltaGeHVec(float d, blcluster* bl, mblock<float>** A, float* x, float* y);
// Y += d A X for p columns (H/mltaGeHVec.cpp:573-577)
bool mltaGeHGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X,
                unsigned ldX, double* Y, unsigned ldY);

// y += d A x and Y += d A X for a symmetric H-matrix (mltaHeHVec /
// mltaHeHGeM, H/mltaGeHVec.cpp:420-472): bl is a symmetric block tree
// (subdivide_sym), A holds its leaves on and above the diagonal with packed
// Hermitian diagonal leaves (isHeM); every leaf is applied, the ones off the
// diagonal also transposed.  x, y offset to bl->getb1() (== getb2()).
void mltaHeHVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y);
void mltaHeHGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X, unsigned ldX,
                double* Y, unsigned ldY);

// y += d A^T x and Y += d A^T X (the real case of mltaGeHhVec / mltaGeHhGeM,
// H/mltaGeHVec.cpp:61-104): x, X offset to row bl->getb1(), y, Y to column
// bl->getb2().  Returns true if some leaf contributed.
bool mltaGeHhVec(double d, blcluster* bl, mblock<double>** A, double* x, double* y);
bool mltaGeHhGeM(double d, blcluster* bl, mblock<double>** A, unsigned p, double* X,
                 unsigned ldX, double* Y, unsigned ldY);

// Triangular H-solves (H/TH_solve.cpp:18-91, 475-535): B (n x p, column
// stride ldB, host memory) is overwritten with X.  blL / blU is the block
// cluster tree the factor lives on; L holds the leaves on and below the
// diagonal (isLtM diagonal leaves), U those on and above it (isUtM), the
// other entries may be NULL, as initLtH_0 / initUtH_0 leave them
// (H/initH.cpp:262-290).  The factor is uploaded on first use and cached
// like the matrices of mltaGeHVec.
void LtHGeM_solve(blcluster* blL, mblock<double>** L, unsigned p, double* B, unsigned ldB);
void UtHGeM_solve(blcluster* blU, mblock<double>** U, unsigned p, double* B, unsigned ldB);
// L x = b, U x = b and L U x = b in place (Include/H.h:1325-1388)
inline void LtHVec_solve(blcluster* blL, mblock<double>** L, double* b) {
  LtHGeM_solve(blL, L, 1, b, blL->getn1());
}
inline void UtHVec_solve(blcluster* blU, mblock<double>** U, double* b) {
  UtHGeM_solve(blU, U, 1, b, blU->getn1());
}
inline void HLU_solve(blcluster* bl, mblock<double>** L, mblock<double>** U, double* b) {
  LtHVec_solve(bl, L, b);
  UtHVec_solve(bl, U, b);
}
// solves with the transposes (real case of H/TH_solve.cpp:41-115): L^T X = B
// (mblock::ltrh_solve on the diagonal: the permutation is applied on the way
// out) and U^T X = B; each direction is uploaded and cached separately
void LtHhGeM_solve(blcluster* blL, mblock<double>** L, unsigned p, double* B, unsigned ldB);
void UtHhGeM_solve(blcluster* blU, mblock<double>** U, unsigned p, double* B, unsigned ldB);
inline void LtHhVec_solve(blcluster* blL, mblock<double>** L, double* b) {
  LtHhGeM_solve(blL, L, 1, b, blL->getn1());
}
inline void UtHhVec_solve(blcluster* blU, mblock<double>** U, double* b) {
  UtHhGeM_solve(blU, U, 1, b, blU->getn2());
}
// U^T U x = b (Include/H.h:1353-1358) and (L U)^T x = b (Include/H.h:1390-1396)
inline void HCholesky_solve(blcluster* bl, mblock<double>** U, double* b) {
  UtHhVec_solve(bl, U, b);
  UtHVec_solve(bl, U, b);
}
inline void HLUh_solve(blcluster* bl, mblock<double>** L, mblock<double>** U, double* b) {
  UtHhVec_solve(bl, U, b);
  LtHhVec_solve(bl, L, b);
}

// storage in bytes and largest rank (H/initH.cpp:16-78)
unsigned long sizeH(blcluster* bl, mblock<double>** A);
unsigned Hmax_rank(blcluster* bl, mblock<double>** A);
// FP64 -> FP32 copy of every leaf (H/initH.cpp:152-168)
void copyH(blcluster* bl, mblock<double>** A, mblock<float>** B);

#endif
