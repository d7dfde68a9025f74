// matrix.h -- operator interface of the Krylov solvers (drop-in for the
// reference's Matrix<T> / GeHMatrix<T> / HeHMatrix<T>, Include/matrix.h:
// 20-32,307-323,352-370).
#ifndef AHMED_MATRIX_H
#define AHMED_MATRIX_H

#include "H.h"

template <class T>
struct Matrix {
  unsigned m, n;
  Matrix(unsigned n1, unsigned n2) : m(n1), n(n2) {}
  Matrix() : m(0), n(0) {}
  virtual void amux(T d, T* x, T* y) const = 0;  // y += d A x
  virtual void precond_apply(T*) const {}
  virtual ~Matrix() {}
};

// General H-matrix; the tree must outlive it, the leaves are freed with it.
template <class T>
struct GeHMatrix : public Matrix<T> {
  mblock<T>** blcks;
  blcluster* blclTree;
  GeHMatrix(unsigned n1, unsigned n2, blcluster* tree = nullptr)
      : Matrix<T>(n1, n2), blcks(nullptr), blclTree(tree) {}
  ~GeHMatrix() {
    if (blclTree) freembls(blclTree, blcks);
  }
  void amux(T d, T* x, T* y) const override { mltaGeHVec(d, blclTree, blcks, x, y); }
};

// Symmetric H-matrix on a symmetric block tree (subdivide_sym), leaves on and
// above the diagonal (matgenHeH_gpu); the tree must outlive it, the leaves
// are freed with it.
template <class T>
struct HeHMatrix : public Matrix<T> {
  mblock<T>** blcks;
  blcluster* blclTree;
  HeHMatrix(unsigned n, blcluster* tree = nullptr) : Matrix<T>(n, n), blcks(nullptr), blclTree(tree) {}
  ~HeHMatrix() {
    if (blclTree) freembls(blclTree, blcks);
  }
  void amux(T d, T* x, T* y) const override { mltaHeHVec(d, blclTree, blcks, x, y); }
};

#endif
