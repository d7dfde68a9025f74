/*
 * ahm.h -- C ABI of the MI355X H-matrix library (libahmed_amd.so).
 *
 * This is the drop-in boundary the AHMED-compatible C++ surface (include/ahmed)
 * calls into, and what any FFI (ctypes, cgo, JNI) binds.  Plain pointers and
 * sizes only; every call returns an ahm_status; ahm_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Data conventions (same as the reference):
 *   - vectors are in cluster-permuted order (op_perm[pos] = original index,
 *     Include/cluster_bbx.h:106-112);
 *   - a leaf is (b1, n1, b2, n2): rows [b1, b1+n1), columns [b2, b2+n2);
 *   - leaf values use mblock<T>'s storage (Include/mblock.h:36-54): a dense
 *     leaf is n1*n2 column-major; a low-rank leaf U V^T is U (n1 x k,
 *     column-major) followed by V (n2 x k, column-major);
 *   - products ACCUMULATE: y += alpha * H x (H/mltaGeHVec.cpp:16-35).
 * Device pointers are HIP device memory on the context's device; every
 * device-side call is ordered on the context's stream.  A handle is not
 * thread-safe: calls on one handle must be serialised by the caller.
 */
#ifndef AHM_H
#define AHM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AHM_OK = 0,
  AHM_ERR_INVALID = 1,       /* bad argument / shape                         */
  AHM_ERR_HIP = 2,           /* HIP runtime error                            */
  AHM_ERR_NOMEM = 3,         /* device or host allocation failed             */
  AHM_ERR_UNSUPPORTED = 4,   /* combination not implemented                  */
  AHM_ERR_NOT_CONVERGED = 5  /* Krylov solver hit nsteps (results still set) */
} ahm_status;

typedef enum { AHM_F64 = 0, AHM_F32 = 1, AHM_BF16 = 2 } ahm_dtype;

typedef struct ahm_ctx_s* ahm_ctx;
typedef struct ahm_hmat_s* ahm_hmat;
typedef struct ahm_tree_s* ahm_tree;
typedef struct ahm_trih_s* ahm_trih;

/* cluster tree kinds */
#define AHM_TREE_BBX 0   /* cluster3d_bbx (Include/cluster_bbx.h:236-302)      */
#define AHM_TREE_PCA 1   /* bemcluster  (Include/bemcluster.h:17-59)          */
/* or'ed into the kind: symmetric block cluster tree (blcluster::subdivide_sym,
 * cluster/blcluster.cpp:90-113): only the blocks on and above the diagonal
 * are generated, diagonal blocks are never admissible */
#define AHM_TREE_SYM 0x100

/* One leaf of the block cluster tree with its storage.  rank < 0: dense leaf
 * (mblock::isGeM), rank >= 0: low-rank leaf (mblock::isLrM, rank 0 = empty).
 * offset: element offset of the leaf's values in the data array. */
typedef struct {
  uint32_t b1, n1, b2, n2;
  int32_t rank;
  uint32_t flags;            /* 0, or AHM_LEAF_LTM / AHM_LEAF_UTM (triangular factors) */
  uint64_t offset;
} ahm_leaf_desc;

/* Diagonal leaves of triangular factors (ahm_trih_upload), n1 == n2 == n,
 * n(n+1)/2 packed values as the reference's mblock stores them:
 * AHM_LEAF_LTM  unit lower triangle (mblock::isLtM), column j holds rows
 *               j..n-1 and its entry for row j is the row permutation
 *               index ip_j as a value (mblock::decomp_LU, H/mblock_D.cpp:
 *               1078-1107): the leaf represents P^T L with (P b)_i = b[ip_i];
 * AHM_LEAF_UTM  upper triangle (mblock::isUtM), column j holds rows 0..j. */
#define AHM_LEAF_LTM 1u
#define AHM_LEAF_UTM 2u
/* Dense diagonal leaf of a symmetric H-matrix (mblock::isHeM,
 * Include/mblock.h:137-141): b1 == b2, n1 == n2 = n, rank < 0, the upper
 * triangle packed column by column, n(n+1)/2 values ((i, j), i <= j, at
 * i + j(j+1)/2).  A matrix holding such leaves is symmetric: its leaves are
 * those of a symmetric block tree (on and above the diagonal), and every
 * product applies each stored leaf and the transpose of each leaf off the
 * diagonal (mltaHeHVec, H/mltaGeHVec.cpp:420-446). */
#define AHM_LEAF_HEM 4u

typedef struct {
  uint64_t n_leaves, n_dense, n_lowrank;
  uint64_t nvals;            /* sum of stored values (mblock::nvals)          */
  uint64_t nvals_dense;      /* values in dense leaves (sum n1*n2)            */
  uint64_t nvals_u;          /* values in U factors (sum k*n1)                */
  uint64_t nvals_v;          /* values in V factors (sum k*n2)                */
  uint64_t t_len;            /* sum of low-rank ranks = len(V^T x)            */
  uint64_t bytes_per_value;  /* storage precision                             */
  uint64_t n_tasks_a, n_tasks_b;
  uint32_t max_rank;
  uint32_t n_rows, n_cols;
  uint32_t symmetric;        /* 1: symmetric storage (AHM_LEAF_HEM leaves)    */
} ahm_hmat_stats;

const char* ahm_last_error(void);

/* -------- context: one device, one HIP stream ---------------------------- */
int ahm_ctx_create(int device, ahm_ctx* out);
int ahm_ctx_destroy(ahm_ctx ctx);
/* run on an existing hipStream_t from now on (NULL = HIP's default stream);
 * until this is called the context uses a stream of its own */
int ahm_ctx_set_stream(ahm_ctx ctx, void* hip_stream);
int ahm_ctx_sync(ahm_ctx ctx);
/* device blocks of >= 1 GiB are not freed when released but cached (up to
 * three quarters of the device memory) for the next large allocation: a
 * 50 GB hipMalloc right after large frees took 2.3 s, against under 1 ms on
 * unused memory.
 * The cache is emptied when an allocation fails, when a context of the
 * device is destroyed, and by this call (e.g. before handing the memory to
 * another allocator).  No counterpart in the reference (host memory). */
int ahm_ctx_release_cached(ahm_ctx ctx);
/* bytes the allocation cache holds on the context's device */
int ahm_ctx_cached_bytes(ahm_ctx ctx, uint64_t* bytes);

/* -------- trees (host): the C++ surface's cluster tree + block cluster tree
 * pts: N x 3 (original order).  Replaces the caller-side sequence
 * createClusterTree + blcluster::subdivide (Examples/PrecondBEM.cpp:263-277,
 * Include/cluster_bbx.h:100-128, cluster/blcluster.cpp:40-66).
 * Leaves come out in idx order; i0[l] is the ACA start row of leaf l relative
 * to its b1 (bemcluster::icom - b1 as in Include/apprx.h:155; for bbx trees
 * the row closest to the row cluster's centre of mass). */
int ahm_tree_build(const double* pts, uint32_t N, uint32_t bmin, double eta2, int kind,
                   ahm_tree* out);
int ahm_tree_destroy(ahm_tree t);
int ahm_tree_info(ahm_tree t, uint32_t* n_leaves, uint32_t* n_clusters);
int ahm_tree_perm(ahm_tree t, uint32_t* op_perm, uint32_t* po_perm);
int ahm_tree_leaves(ahm_tree t, uint32_t* b1, uint32_t* n1, uint32_t* b2, uint32_t* n2,
                    uint8_t* adm, uint32_t* i0);

/* -------- H-matrix upload / download / info -------------------------------
 * Replaces the per-leaf mblock<T>* array of the reference (Include/H.h:21-45,
 * allocmbls/freembls) with one device arena in the matvec layout.
 * data_dtype: element type of `data` (AHM_F64 or AHM_F32, host memory);
 * storage: AHM_F64 or AHM_F32 (copyH double->float, H/initH.cpp:152-168).
 * keep_leaves != 0 keeps a device copy in mblock layout for download. */
int ahm_hmat_upload(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols, uint32_t n_leaves,
                    const ahm_leaf_desc* leaves, const void* data, ahm_dtype data_dtype,
                    ahm_dtype storage, int keep_leaves, ahm_hmat* out);
/* Re-layout an existing matrix in another storage precision: FP64 -> FP32
 * (copyH, H/initH.cpp:152-168), or FP64 -> AHM_BF16: the FP32 matrix plus a
 * BF16 copy of every V factor, so that products with 9 or more right-hand
 * sides form V^T X on BF16 MFMA (V and X rounded to BF16 by round-to-nearest-
 * even, FP32 accumulation) and U (V^T X) and the dense blocks in FP32.  An
 * AHM_BF16 matrix takes AHM_F32 vectors; with fewer right-hand sides it
 * multiplies exactly as its FP32 copy. */
int ahm_hmat_convert(ahm_hmat src, ahm_dtype storage, ahm_hmat* out);
int ahm_hmat_destroy(ahm_hmat h);
int ahm_hmat_stats_get(ahm_hmat h, ahm_hmat_stats* st);
/* leaves (n_leaves descs, offsets into data) and values in mblock layout, in
 * the element type they were uploaded in (data_dtype; FP64 for matrices
 * assembled on the GPU); needs keep_leaves.  data may be NULL to query
 * descriptors only (any matrix). */
int ahm_hmat_download(ahm_hmat h, ahm_leaf_desc* leaves, void* data);

/* -------- on-disk H-matrix: the reference's savembls / loadmbls format -----
 * (Include/H.h:47-103; mblock::save/load Include/mblock.h:737-804), so files
 * written by either side load in the other.  Layout, native byte order:
 *   double eta, uint32 bmin, uint32 n_leaves, then per leaf in idx order a
 *   one-byte presence flag (1) followed by uint32 n1, n2, status; status 0
 *   (low rank): uint32 rank, rank*(n1+n2) values (U then V, column-major);
 *   status 4 (dense): n1*n2 values column-major.  Values are double for
 *   AHM_F64 files (mblock<double>), float for AHM_F32.
 * The file holds no geometry: b1/b2 of the leaves come from the block cluster
 * tree rebuilt from the points with the stored bmin and eta.  The reference's
 * Hermitian / triangular / symmetric dense leaves (status 1, 2, 3, 5) are
 * outside the GeH path and reported as AHM_ERR_UNSUPPORTED. */
int ahm_mbls_save(const char* path, double eta, uint32_t bmin, uint32_t n_leaves,
                  const ahm_leaf_desc* leaves, const void* data, ahm_dtype dtype);
/* header of a file and the number of values it holds (to size ahm_mbls_read) */
int ahm_mbls_info(const char* path, ahm_dtype dtype, double* eta, uint32_t* bmin,
                  uint32_t* n_leaves, uint64_t* n_values);
/* leaves[i]: n1, n2, rank (-1 dense), offset (values packed in leaf order);
 * b1 = b2 = 0.  n_leaves must equal the file's count. */
int ahm_mbls_read(const char* path, ahm_dtype dtype, uint32_t n_leaves, ahm_leaf_desc* leaves,
                  void* data);

/* -------- assembly on the GPU ---------------------------------------------
 * The reference's matgenGeH_omp / apprx_unsym / ACA / cpyLrM_cmpr pipeline
 * (Include/matgen_omp.h:59-75, Include/apprx.h:136-172, Include/ACA.h:217-350,
 * H/mblock_D.cpp:357-465) for kernels the device can evaluate (the MATGEN_T
 * plugin, Include/matgen_sqntl.h:21-29, becomes an ahm_kernel).  Entry (i,j)
 * in permuted order is 1.0 on the diagonal, otherwise weight*G(|x_i - x_j|). */
#define AHM_KERNEL_LAPLACE 0   /* G = 1/(4 pi r)            */
#define AHM_KERNEL_YUKAWA  1   /* G = exp(-kappa r)/(4 pi r) */
#define AHM_KERNEL_EXPR    2   /* G = exp(-kappa r)/r        */
typedef struct {
  int32_t kind;
  uint32_t flags;    /* 0, or AHM_ASSEMBLE_ACAR */
  double kappa;
  double weight;
} ahm_kernel;
/* approximate the admissible leaves with ACAr, the row-pivoted variant
 * apprx_sym uses (Include/ACA.h:578-720, Include/apprx.h:87-131), instead of
 * ACA (Include/ACA.h:217-350); storage and recompression stay as in
 * apprx_unsym */
#define AHM_ASSEMBLE_ACAR 1u
/* symmetric storage, the reference's matgenHeH_omp / apprx_sym
 * (Include/matgen_omp.h:77-100, Include/apprx.h:87-131): the leaves are those
 * of a symmetric block tree; admissible leaves take ACAr, dense leaves on the
 * diagonal (b1 == b2) are stored packed (AHM_LEAF_HEM) and the matrix is
 * symmetric.  ahm_assemble_tree sets it for trees built with AHM_TREE_SYM. */
#define AHM_ASSEMBLE_SYM 2u

typedef struct {
  uint64_t n_candidates;    /* admissible leaves that ran ACA              */
  uint64_t n_retries;       /* ACA reruns after a rank slab was too small   */
  uint64_t n_lowrank, n_dense;
  uint64_t sum_rank;        /* sum of ranks after recompression            */
  uint64_t sum_aca_rank;    /* sum of ACA ranks before recompression       */
  uint64_t slab_bytes;      /* ACA scratch allocated over all batches/reruns */
  uint64_t peak_slab_bytes; /* largest single batch                          */
  double aca_ms;            /* wall time of the ACA batches (host clock)     */
} ahm_assembly_stats;

/* pts: N x 3 in original order, op_perm from the tree; leaves in idx order
 * with admissibility adm[] and ACA start rows i0[] (relative to b1).
 * rank_hint (may be NULL): per-leaf rank-slab size; ACA keeps going past an
 * under-sized slab (the leaf is rerun with a larger one), so hints change
 * memory and time, never the result.  stats may be NULL; aca_rank (may be
 * NULL) receives each leaf's ACA rank before recompression (0 for leaves
 * stored dense). */
int ahm_assemble(ahm_ctx ctx, const double* pts, const uint32_t* op_perm, uint32_t N,
                 uint32_t n_leaves, const uint32_t* b1, const uint32_t* n1, const uint32_t* b2,
                 const uint32_t* n2, const uint8_t* adm, const uint32_t* i0,
                 const ahm_kernel* kernel, double eps, uint32_t rankmax,
                 const uint32_t* rank_hint, ahm_dtype storage, int keep_leaves,
                 ahm_assembly_stats* stats, uint32_t* aca_rank, ahm_hmat* out);
/* same, taking the points, permutation and leaves from a tree */
int ahm_assemble_tree(ahm_ctx ctx, ahm_tree tree, const ahm_kernel* kernel, double eps,
                      uint32_t rankmax, const uint32_t* rank_hint, ahm_dtype storage,
                      int keep_leaves, ahm_assembly_stats* stats, uint32_t* aca_rank,
                      ahm_hmat* out);

/* n_mat independent H-matrices (each with its own points, tree and kernel)
 * assembled in one batched launch and stacked block-diagonally: system s
 * owns rows/columns [offsets[s], offsets[s+1]) of the result, with offsets
 * the running sums of the trees' point counts.  Every product of the result
 * is one launch per phase for all systems. */
int ahm_assemble_batch(ahm_ctx ctx, uint32_t n_mat, const ahm_tree* trees,
                       const ahm_kernel* kernels, double eps, uint32_t rankmax,
                       ahm_dtype storage, int keep_leaves, ahm_assembly_stats* stats,
                       ahm_hmat* out);
/* number of independent systems and their offsets (n_seg + 1 values);
 * a matrix not built by ahm_assemble_batch is one system. */
int ahm_hmat_segments(ahm_hmat h, uint32_t* n_seg, uint32_t* offsets);

/* Rank-slab sizes for the ACA of each leaf from an 8-64-64-1 MLP (BF16
 * inference, FP32 accumulation).  feats: n x 8 host floats (tree level,
 * log n1, log n2, diam_row/diam_col, distance/max diam, log10 eps, kappa,
 * log10 max diam); params: packed FP32 W1(8x64) b1 W2(64x64) b2 W3(64) b3
 * mean(8) std(8); hint[l] = clamp(ceil(pad*(exp(pred)-1)) + margin, 1,
 * maxk[l]) (maxk may be NULL); pred (may be NULL) receives log(1+rank).
 * Feed hint to ahm_assemble's rank_hint.  No counterpart in the reference. */
int ahm_rank_predict(ahm_ctx ctx, const float* feats, uint32_t n, const float* params, float pad,
                     uint32_t margin, const uint32_t* maxk, uint32_t* hint, float* pred);

/* -------- products: mltaGeHVec / mltaGeHGeM (H/mltaGeHVec.cpp:567-577) ----
 * Y += alpha * H * X for p right-hand sides; X is n_cols x p column-major
 * with leading dimension ldx, Y n_rows x p with ldy (device pointers).
 * vec: AHM_F64 or AHM_F32 (must match storage unless storage=F32, vec=F64
 * is asked for, which is refused).  p = 1 is the single-vector path. */
int ahm_hmat_matvec(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                    uint32_t ldy, uint32_t p, ahm_dtype vec);
/* The same product for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Its workspace (V^T x, the
 * partial sums of chunked panels, the interleaved right-hand sides) is one
 * static 256 MiB device array per GPU, held from the first call on that GPU
 * until the process ends; the call allocates nothing.  Products that need
 * more workspace than that (t_len + partial slots, times the right-hand-side
 * slots) are refused with AHM_ERR_INVALID; ahm_hmat_matvec serves them.  All
 * calls on one GPU share the array: calls that overlap in time on one GPU
 * (other threads, other streams) race and must be serialised by the caller.
 * Results are bit-identical to ahm_hmat_matvec's. */
int ahm_hmat_matvec_noalloc(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                            uint32_t ldy, uint32_t p, ahm_dtype vec);
/* Runs the product `reps` times with HIP events around each of the two
 * launches (V^T x, then the row panels) and returns their mean durations in
 * milliseconds (instrumentation; same arguments as ahm_hmat_matvec). */
int ahm_hmat_matvec_profile(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                            uint32_t ldy, uint32_t p, ahm_dtype vec, uint32_t reps,
                            double* ms_phase_a, double* ms_phase_b);
/* Autotuner: times `reps` products with p right-hand sides for each tile
 * size (column-chunk size of the row panels and V^T panels: 4K .. 192K
 * values, or none) and, for p <= 8, the single-launch forms
 * (ahm_hmat_set_fused), and keeps the fastest.  fused (may be NULL) receives
 * the launch form (0: three kernels, 1: row-panel single launch, 2:
 * persistent single launch).  Products agree to the rounding of the chunk
 * sums (the order of the summands changes). */
int ahm_hmat_autotune(ahm_hmat h, uint32_t p, ahm_dtype vec, uint32_t reps,
                      uint64_t* chunk_values, double* best_ms, int* fused);
/* Launch form of products of 1 to 8 right-hand sides (ahm_hmat_matvec and
 * the calls built on it).  on == 1: ONE launch with one workgroup per row
 * panel, which forms the V^T x values of its own low-rank leaves (each
 * leaf's t serves its own row panel only), so no workgroup waits for
 * another, and X / Y are used in place; the first call builds the plan (row
 * panels, U-column groups); not for the solvers' gated products.  on == 2:
 * ONE persistent cooperative launch of as many workgroups as stay resident,
 * running the tasks of the three-kernel form phase after phase with grid
 * barriers between them (same results bit for bit; the solvers' single-system
 * products too; products recorded into a hipGraph keep the separate
 * launches).  on == 0: the three-kernel form.  Symmetric matrices: form 0
 * only.  No counterpart in the reference (whose mltaGeHVec loops over the
 * leaves, H/mltaGeHVec.cpp:38-59). */
int ahm_hmat_set_fused(ahm_hmat h, int on);
/* aborted: always 0 (the single launch has no bounded waits to give up). */
int ahm_hmat_fused_status(ahm_hmat h, int* aborted);
/* Same with host vectors (the drop-in for mltaGeHVec's host pointers):
 * copies X in and Y in/out over PCIe around the device product. */
int ahm_hmat_matvec_host(ahm_hmat h, double alpha, const double* X, uint32_t ldx,
                         double* Y, uint32_t ldy, uint32_t p);

/* Y += alpha A^T X: the real case of mltaGeHhVec / mltaGeHhGeM
 * (H/mltaGeHVec.cpp:61-104, leaf products H/mblock.cpp:159-170 and
 * Include/mblock.h:112-115).  X: p columns of n_rows values (ldx apart),
 * Y: p columns of n_cols values (ldy apart), cluster-permuted as in
 * ahm_hmat_matvec; runs on the forward layout (no second copy of the leaves),
 * deterministic.  The first call on a matrix builds its column index. */
int ahm_hmat_matvec_t(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                      uint32_t ldy, uint32_t p, ahm_dtype vec);
/* the same with host pointers (FP64 matrices), as ahm_hmat_matvec_host */
int ahm_hmat_matvec_t_host(ahm_hmat h, double alpha, const double* X, uint32_t ldx, double* Y,
                           uint32_t ldy, uint32_t p);

/* -------- Krylov solvers on the device --------------------------------------
 * b, x: FP64 device vectors of length n_rows (x is the initial guess and the
 * result).  eps: in = tolerance, out = achieved relative residual; nsteps:
 * in = step limit, out = steps taken on convergence (unchanged otherwise).
 * Returns AHM_OK on convergence and AHM_ERR_NOT_CONVERGED otherwise (the
 * reference's 0 / 1).  With an FP32 matrix only the H-matvec runs in FP32;
 * Krylov vectors, orthogonalisation and rotations stay FP64.
 * hist (host, may be NULL) receives up to nhist residual estimates;
 * nmatvec (may be NULL) the number of H-matvecs. */
/* GMRes, solvers/GMRES.cpp:80-181 (MGS + Givens, restart length m).  The
 * Hessenberg matrix, the rotations and the convergence test run on the
 * device (ahm_hmat_set_arnoldi); the host queues the Arnoldi steps ahead and
 * reads the convergence flags two steps behind.  The Krylov basis
 * (n x (m+1) FP64) stays with the handle between calls. */
int ahm_gmres(ahm_hmat h, const double* b, double* x, double* eps, uint32_t m,
              uint32_t* nsteps, double* hist, uint32_t nhist, uint32_t* nmatvec);
/* on != 0: the GMRES entry points of this handle capture each Arnoldi step
 * into a hipGraph on its second use and replay it from then on (the graphs
 * stay with the handle); 0 (the default): they launch the same kernels one
 * by one, which measured as fast or faster on MI355X.  Both give
 * bit-identical iterates.  No counterpart in the reference. */
int ahm_hmat_set_graphs(ahm_hmat h, int on);
/* Form of the GMRES orthogonalisation and rotations after each product:
 * 1: chunked launches over all systems (dot, one launch per MGS pass,
 * normalisation, rotations); 2: one workgroup per system does all of it in
 * one launch, for systems of at most 16384 rows (larger ones keep the
 * chunked launches); 0 (the default): the second form where the systems
 * are small or many.  Both give bit-identical iterates.  No counterpart in
 * the reference. */
int ahm_hmat_set_arnoldi(ahm_hmat h, int mode);
/* CG,solvers/CG.cpp:31-113 */
int ahm_cg(ahm_hmat h, const double* b, double* x, double* eps, uint32_t* nsteps);
/* FP64 iterative refinement: r = b - A64 x (FP64 H-matvec), inner GMRES(m)
 * on `lo` (typically the FP32 copy of the same matrix) to inner_eps relative,
 * x += d, until |r|/|b| < eps or max_outer refinements.  No counterpart in
 * the reference; its FEMMatrix (Include/matrix.h:489-507) is the closest
 * mixed-precision precedent. */
int ahm_gmres_ir(ahm_hmat hi, ahm_hmat lo, const double* b, double* x, double* eps, uint32_t m,
                 double inner_eps, uint32_t inner_steps, uint32_t max_outer, uint32_t* outer,
                 uint32_t* total_inner, double* hist, uint32_t nhist);

/* GMRes(m) on every independent system of a batched matrix
 * (ahm_assemble_batch) at once: per-system Krylov spaces, rotations and
 * convergence tests (solvers/GMRES.cpp:80-181 each) on the device, Arnoldi
 * steps in lockstep so one H-matvec launch per phase serves all systems; a
 * converged system is frozen on the device and its product tasks skipped.
 * With use_graph != 0 each Arnoldi step is captured into a hipGraph on its
 * second use (kept with the handle) and replayed.  b, x: FP64 device
 * vectors over all systems.  Per system (may be NULL): achieved residual, steps (the
 * reference's nsteps output), converged flag.  AHM_OK if every system
 * converged. */
int ahm_gmres_batch(ahm_hmat h, const double* b, double* x, double eps, uint32_t m,
                    uint32_t nsteps, int use_graph, double* eps_out, uint32_t* steps_out,
                    int* converged);

/* -------- triangular H-matrix factors: forward / backward substitution ----
 * The reference's LtHGeM_solve / UtHGeM_solve (H/TH_solve.cpp:18-91,
 * 475-535) with the leaf solves mblock::ltr_solve (H/mblock_D.cpp:
 * 1183-1205) and mblock::utr_solve (Include/mblock.h:694-697); HLU_solve
 * (Include/H.h:1383-1388) is an 'L' solve followed by a 'U' solve.
 * uplo 'L': the unit lower factor (diagonal leaves AHM_LEAF_LTM, the other
 * leaves strictly below the diagonal); 'U': the upper factor (AHM_LEAF_UTM,
 * other leaves strictly above).  'l' / 'u': the same factor set up for
 * solves with its transpose, LtHhGeM_solve / UtHhGeM_solve (H/TH_solve.cpp:
 * 41-115, real case; mblock::ltrh_solve H/mblock_D.cpp:1208-1230 applies the
 * permutation after the unit upper solve); HCholesky_solve (Include/H.h:
 * 1353-1358) is a 'u' solve then a 'U' solve with the same leaves.  Off-diagonal leaves are dense (rank < 0) or
 * low rank (mblock layout, U then V); the diagonal leaves must tile [0, n).
 * data: host values of data_dtype (AHM_F64); the factor is stored in FP64.
 * One persistent launch per right-hand side: the diagonal leaves are solved
 * in order while the off-diagonal leaf products run as soon as the columns
 * they read are solved (dependency counters in device memory, a work queue
 * in an order the host has checked to be free of cyclic waits). */
int ahm_trih_upload(ahm_ctx ctx, uint32_t n, char uplo, uint32_t n_leaves,
                    const ahm_leaf_desc* leaves, const void* data, ahm_dtype data_dtype,
                    ahm_trih* out);
int ahm_trih_destroy(ahm_trih t);
/* B (n x p, column stride ldb, FP64 device memory) <- op^{-1} B, in place,
 * on the context stream (the reference's "B is destroyed, X stored in B"). */
int ahm_trih_solve(ahm_trih t, void* B, uint32_t ldb, uint32_t p);
/* the same on host memory (copies in and out) */
int ahm_trih_solve_host(ahm_trih t, double* B, uint32_t ldb, uint32_t p);
/* diagonal leaves, off-diagonal items and stored values of a factor */
int ahm_trih_info(ahm_trih t, uint32_t* n_diag, uint64_t* n_items, uint64_t* nvals);

/* GMRes / CG with precond_apply (Include/matrix.h:30) = HLU_solve
 * (Include/H.h:1383-1388): v <- U^{-1} L^{-1} v; GMRes preconditions from
 * the right (w = A M v_i, x += M V y: solvers/GMRES.cpp:70-73,125-130), CG
 * uses r^ = M r (solvers/CG.cpp:69-71).  L or U may be NULL (that factor is
 * skipped); both must live on h's context.  Arguments and returns otherwise
 * as ahm_gmres / ahm_cg. */
int ahm_gmres_prec(ahm_hmat h, ahm_trih L, ahm_trih U, const double* b, double* x, double* eps,
                   uint32_t m, uint32_t* nsteps, double* hist, uint32_t nhist, uint32_t* nmatvec);
int ahm_cg_prec(ahm_hmat h, ahm_trih L, ahm_trih U, const double* b, double* x, double* eps,
                uint32_t* nsteps);

#ifdef __cplusplus
}
#endif
#endif
