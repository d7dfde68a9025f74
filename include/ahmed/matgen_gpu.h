// matgen_gpu.h -- H-matrix assembly on the GPU (drop-in for the reference's
// matgenGeH_omp, Include/matgen_omp.h:59-75, with apprx_unsym / ACA /
// cpyLrM_cmpr per leaf, Include/apprx.h:136-172).
//
// The reference's MATGEN_T plugin is user host code; kernels the device can
// evaluate are described by GpuKernel instead: entry (i,j) in permuted order
// is 1.0 when op_perm[i] == op_perm[j], else weight * G(|x_i - x_j|).
#ifndef AHMED_MATGEN_GPU_H
#define AHMED_MATGEN_GPU_H

#include "H.h"
#include "ahm.h"

struct GpuKernel {
  int kind;              // AHM_KERNEL_LAPLACE / YUKAWA / EXPR
  double kappa;
  double weight;
  const double* pts;     // N x 3, original order
  const unsigned* op_perm;
  unsigned N;
};

// Fills A[idx] for every leaf of bl (ACA start rows from bemblcluster's
// icom, or the row nearest the row cluster's centre of mass for other
// trees).  The device copy stays cached for the products of H.h.
void matgenGeH_gpu(const GpuKernel& G, unsigned nblcks, blcluster* bl, double eps,
                   unsigned rankmax, mblock<double>** A);

// The symmetric assembly (matgenHeH_omp, Include/matgen_omp.h:95-100, with
// apprx_sym per leaf, Include/apprx.h:87-131) on a symmetric block tree
// (subdivide_sym): ACAr on the admissible leaves, the diagonal leaves dense
// packed Hermitian (isHeM), the other dense leaves general.
void matgenHeH_gpu(const GpuKernel& G, unsigned nblcks, blcluster* bl, double eps,
                   unsigned rankmax, mblock<double>** A);

#endif
