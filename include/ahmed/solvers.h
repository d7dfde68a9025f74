// solvers.h -- Krylov solvers (drop-in for the reference's GMRes / CG,
// Include/solvers.h:14-34).  For a GeHMatrix<double> or HeHMatrix<double> the
// whole iteration runs on the GPU (ahm_gmres / ahm_cg); eps and nsteps are
// in/out exactly as in the reference (solvers/GMRES.cpp:80-181,
// solvers/CG.cpp:31-113).  Returns 0 on convergence, 1 otherwise, 2 if the
// operator is neither (a user-defined host operator cannot be applied on the
// device).
#ifndef AHMED_SOLVERS_H
#define AHMED_SOLVERS_H

#include "matrix.h"

unsigned GMRes(const Matrix<double>& A, double* const b, double* const x, double& eps,
               const unsigned m, unsigned& nsteps);
unsigned CG(const Matrix<double>& A, double* const b, double* const x, double& eps,
            unsigned& nsteps);

#endif
