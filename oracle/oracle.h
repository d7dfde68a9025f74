/*
 * oracle.h -- CPU restatement of AHMED's H-matrix hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (ahmed_amd/) never links, loads or calls anything here.
 *
 * Every function follows the reference file:line cited beside it.  Third-party
 * numerics (BLAS-1/2/3, LAPACK dgeqrf/dgesvd/dormqr/dtrtrs) are called through
 * MKL exactly where the reference calls its BLAS/LAPACK wrappers
 * (Include/blas.h), with MKL_CBWR_COMPATIBLE so results do not depend on the
 * host CPU's instruction set.
 *
 * Parity pins (what checks each function), stated per function below:
 *   golden  = bit-level or 1e-13 agreement with vectors written by the compiled
 *             reference (tests/golden/make_golden.py, oracle/_ref)
 *   closed  = closed form / invariant / brute force in tests/test_oracle.py
 */
#ifndef AHM_ORACLE_H
#define AHM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cluster tree kinds */
#define OR_TREE_BBX 0   /* cluster3d_bbx + centre-of-mass start row (golden) */
#define OR_TREE_PCA 1   /* bemcluster (cluster3d_pca + icom quirk)   (golden) */
#define OR_TREE_SYM 0x100 /* or-ed into the kind: symmetric block tree (subdivide_sym) */
#define OR_HEM (-5)     /* rank code of a packed Hermitian diagonal leaf (mblock::isHeM) */

/* kernel generators (MATGEN_T plugins; the reference ships none) */
#define OR_KERNEL_LAPLACE 0  /* 1/(4 pi r), diagonal 1.0                    */
#define OR_KERNEL_YUKAWA  1  /* exp(-kappa r)/(4 pi r), diagonal 1.0        */
#define OR_KERNEL_EXPR    2  /* exp(-kappa r)/r, diagonal 1.0               */

typedef struct or_tree or_tree;
typedef struct or_hmat or_hmat;

/* ---- trees: Include/cluster_bbx.h:52-128, Include/cluster_pca.h:82-217,
 *      Include/bemcluster.h:25-56, cluster/blcluster.cpp:40-66       [golden] */
or_tree* or_tree_build(const double* pts, unsigned N, unsigned bmin,
                       double eta2, int kind);
void     or_tree_free(or_tree* t);
unsigned or_tree_nleaves(const or_tree* t);
unsigned or_tree_nclusters(const or_tree* t);
void     or_tree_perm(const or_tree* t, unsigned* op_perm, unsigned* po_perm);
/* leaves in DFS (idx) order; i0 = ACA start row = icom(row cluster) - b1 */
void     or_tree_leaves(const or_tree* t, unsigned* b1, unsigned* n1,
                        unsigned* b2, unsigned* n2, unsigned char* adm,
                        unsigned* i0);
void     or_tree_clusters(const or_tree* t, unsigned* nbeg, unsigned* nend,
                          int* son0, int* son1, unsigned* icom);

/* ---- kernel entries (MATGEN_T::cmpbl / scale)                  [closed] */
/* entry (i,j) = 1.0 if op_perm[i]==op_perm[j], else weight*G(x_i, x_j);
 * scale = |entry at (b1+n1/2, b2+n2/2)| */
void   or_kernel_block(int kernel, double kappa, double weight, const double* pts,
                       const unsigned* op_perm, unsigned b1, unsigned n1,
                       unsigned b2, unsigned n2, double* out);
double or_kernel_scale(int kernel, double kappa, double weight, const double* pts,
                       const unsigned* op_perm, unsigned b1, unsigned n1,
                       unsigned b2, unsigned n2);

/* ---- ACA on an explicit dense block (Include/ACA.h:217-350)  [golden+closed]
 * A is n1 x n2 col-major; U,V must hold (kmax+1)*n1 and (kmax+1)*n2.
 * Returns 1 on success (reference returns true). */
int or_aca_dense(const double* A, unsigned n1, unsigned n2, double eps,
                 unsigned kmax, unsigned i0, double scale, unsigned* k,
                 double* U, double* V);

/* ---- ACAr on an explicit dense block (Include/ACA.h:578-720): the
 * simplified row-pivoted ACA of apprx_sym; same buffers.      [golden+closed] */
int or_acar_dense(const double* A, unsigned n1, unsigned n2, double eps,
                  unsigned kmax, unsigned i0, double scale, unsigned* k,
                  double* U, double* V);

/* ---- recompression addtrll via cpyLrM_cmpr (H/mblock_D.cpp:357-465)
 * Input U (n1 x k), V (n2 x k); output rank *kt and Uo (n1 x kt), Vo (n2 x kt)
 * (buffers of n1*k, n2*k).                                    [golden+closed] */
void or_addtrll(unsigned n1, unsigned n2, unsigned k, const double* U,
                const double* V, double delta, unsigned kgoal, unsigned* kt,
                double* Uo, double* Vo);

/* ---- assembly: apprx_unsym per leaf (Include/apprx.h:136-172) driven like
 * matgenGeH_omp (Include/matgen_omp.h:59-75)                       [golden] */
or_hmat* or_hmat_assemble(const or_tree* t, const double* pts, int kernel,
                          double kappa, double weight, double eps,
                          unsigned rankmax, int nthreads);
/* method 0: ACA (apprx_unsym), 1: ACAr per admissible leaf (the
 * approximation apprx_sym uses, Include/apprx.h:87-131)          [golden] */
or_hmat* or_hmat_assemble_method(const or_tree* t, const double* pts, int kernel,
                                 double kappa, double weight, double eps,
                                 unsigned rankmax, int nthreads, int method);
/* H-matrix from explicit leaves (mblock layout: dense n1*n2 col-major;
 * low-rank U (n1 x k) then V (n2 x k)); rank[i] < 0 marks a dense leaf. */
or_hmat* or_hmat_from_leaves(unsigned N, unsigned nleaves, const unsigned* b1,
                             const unsigned* n1, const unsigned* b2,
                             const unsigned* n2, const int* rank,
                             const size_t* offset, const double* data);
void     or_hmat_free(or_hmat* h);
unsigned or_hmat_nleaves(const or_hmat* h);
size_t   or_hmat_nvals(const or_hmat* h);
void     or_hmat_leaves(const or_hmat* h, unsigned* b1, unsigned* n1,
                        unsigned* b2, unsigned* n2, int* rank, size_t* offset);
const double* or_hmat_data(const or_hmat* h);

/* ---- matvec: mltaGeHVec / mltaGeHGeM (H/mltaGeHVec.cpp:16-59,567-577),
 * leaf products H/mblock.cpp:125-154, Include/mblock.h:98-109     [golden] */
int or_hmat_matvec(const or_hmat* h, double d, const double* x, double* y);
int or_hmat_matvec_t(const or_hmat* h, double d, const double* x, double* y);
int or_hmat_matmat(const or_hmat* h, double d, unsigned p, const double* X,
                   unsigned ldX, double* Y, unsigned ldY);
/* leaf-parallel all-core variant with per-thread y (CPU baseline only) */
int or_hmat_matvec_omp(const or_hmat* h, double d, const double* x, double* y,
                       int nthreads);

/* ---- Krylov: GMRes (solvers/GMRES.cpp:80-181), CG (solvers/CG.cpp:31-113)
 * hist (may be NULL) receives the residual estimate per iteration. [golden] */
unsigned or_gmres(const or_hmat* h, const double* b, double* x, double* eps,
                  unsigned m, unsigned* nsteps, double* hist, unsigned nhist);
unsigned or_cg(const or_hmat* h, const double* b, double* x, double* eps,
               unsigned* nsteps);
unsigned or_gmres_dense(unsigned n, const double* A, const double* b,
                        double* x, double* eps, unsigned m, unsigned* nsteps);
unsigned or_cg_dense(unsigned n, const double* A, const double* b, double* x,
                     double* eps, unsigned* nsteps);

/* ---- triangular H-solves: LtHGeM_solve_ / UtHGeM_solve_ (H/TH_solve.cpp:18-91)
 * with the leaf solves mblock::ltr_solve (H/mblock_D.cpp:1183-1205, dtpsv
 * 'L','N','U' Include/blas.h:3749-3755) and mblock::utr_solve
 * (Include/mblock.h:694-697, dtpsv 'U','N','N' Include/blas.h:3930-3936).
 * B (n x p, ld ldB) is overwritten with op^{-1} B.  uplo 'L': the factor's
 * leaves are the diagonal ones (kind -3: packed unit lower triangle, column
 * j holding its permutation entry on the diagonal) and lower ones; 'U':
 * diagonal (kind -4: packed upper triangle) and upper ones.  Off-diagonal
 * leaves: kind -1 dense, k >= 0 low rank (U then V); kind -2 is skipped.
 * uplo 'l' / 'u' solve with the transpose of the stored factor:
 * LtHhGeM_solve_ / UtHhGeM_solve_ (H/TH_solve.cpp:41-115, real case), leaf
 * solves mblock::ltrh_solve (H/mblock_D.cpp:1208-1230, dtpsv 'L','T','U'
 * then the permutation scatter) and utrh_solve (dtpsv 'U','T','N').
 * Restated as a sweep over the diagonal leaves (ascending for L and U^T,
 * descending for U and L^T); every off-diagonal leaf is applied whole as
 * soon as the diagonal leaves covering its input range are solved, which
 * the reference's recursion also guarantees, in another summation order.
 * Returns 0, or -1 for a malformed factor.                       [golden] */
int or_trih_solve(unsigned n, char uplo, unsigned nleaves, const unsigned* b1,
                  const unsigned* n1, const unsigned* b2, const unsigned* n2,
                  const int* kind, const size_t* offset, const double* data,
                  unsigned p, double* B, unsigned ldB);

/* MKL identification for fixtures */
void or_blas_version(char* buf, int len);

#ifdef __cplusplus
}
#endif
#endif
