// ahm_internal.h -- shared declarations of libahmed_amd (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <memory>
#include <vector>

#include "ahm.h"

namespace ahm {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define AHM_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::ahm::fail(AHM_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                         \
  } while (0)

#define AHM_TRY(expr)          \
  do {                         \
    int s_ = (expr);           \
    if (s_ != AHM_OK) return s_; \
  } while (0)

// Block size of the panel kernels: a panel tile has at most kPanelThreads rows.
constexpr int kPanelThreads = 256;

// One GEMV task: out(m) (+)= P (m x W, column-major, ld m) * z, where
// z[c] = src(zsrc[zoff + c]).  Phase A tasks write V^T x into the t buffer,
// phase B tasks accumulate alpha * (D x | U t) into y.
struct GemvTask {
  uint64_t data;   // element offset of P in the arena
  uint64_t zoff;   // offset of the task's column-source list
  uint32_t m, W;
  uint32_t out;    // first output row: t index (phase A), y row (phase B), or
                   // partial-sum slot (column chunk of either phase)
  uint32_t partial;  // 1: write the chunk's row sums to the partial buffer
};

// Row panels wider than one chunk are cut into column chunks of about
// kChunkValues values; their row sums meet in a fixed-order reduction.
constexpr uint64_t kChunkValues = 48 * 1024;

// One low-rank leaf's U columns in a row panel (single-launch products):
// panel columns col .. col+k-1 take V^T x from rows row0 .. row0+k-1 of the
// V^T panel apanel.
struct FusedGroup {
  uint32_t col, k, apanel, row0;
};

struct ReduceTask {
  uint32_t row;      // first output row (t index in phase A, y row in phase B)
  uint32_t m;        // rows
  uint32_t part;     // first partial slot (chunk c, row r at part + c*m + r)
  uint32_t nchunks;
};

// Zeroed values after the last panel of an arena: the MFMA products load
// whole 16-row tiles of 16 columns, so a panel (at most kPanelThreads rows)
// may be read up to 15 columns and 15 values past its end.
constexpr uint64_t kArenaSlack = 16 * (uint64_t)kPanelThreads + 64;

// std::allocator whose value-less construct leaves trivial elements
// uninitialised: resize() of a large host array that is filled right after
// skips a serial zeroing pass over it.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
template <typename T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

// zsrc entry: bit 31 set -> index into t, else index into x
constexpr uint32_t kFromT = 0x80000000u;

// Device-side switch of the product kernels used inside solvers: a
// workgroup whose task belongs to a frozen system (frozen[seg[task]] != 0,
// set on the device by the solver's convergence test) returns at once, so
// steps queued ahead of the host's convergence check cost launches only.
// frozen == nullptr: no gate; seg == nullptr: every task is system 0's.
struct Gate {
  const uint32_t* frozen;
  const uint32_t* seg;
};
__device__ __forceinline__ bool gated(const Gate g) {
  return g.frozen && g.frozen[g.seg ? g.seg[blockIdx.x] : 0];
}

// stored values of a leaf in mblock layout: n1*n2 dense, n(n+1)/2 packed
// Hermitian (AHM_LEAF_HEM), k*(n1+n2) low rank
inline uint64_t leaf_values(const ahm_leaf_desc& d) {
  if (d.rank >= 0) return (uint64_t)d.rank * ((uint64_t)d.n1 + d.n2);
  if (d.flags & AHM_LEAF_HEM) return (uint64_t)d.n1 * (d.n1 + 1) / 2;
  return (uint64_t)d.n1 * d.n2;
}

// Copy of a rows x cols block from the mblock-layout buffer into the arena.
// transpose = 0: dst[r + c*dst_ld] = src[r + c*src_ld]
// transpose = 1: dst[c + r*dst_ld] = src[r + c*src_ld]   (V -> V^T tile)
// transpose = 2: rows r0 = pad .. pad+rows-1 of a symmetric n x n block
//                packed upper (n = src_ld): dst[r + c*dst_ld] = A(pad+r, c),
//                A(i, j) = src[min + max(max+1)/2] with min/max of (i, j)
struct CopyJob {
  uint64_t src, dst;
  uint32_t rows, cols, src_ld, dst_ld;
  uint32_t transpose, pad;
};

// Device allocation.  Blocks of at least 1 GiB are not freed on release but
// kept in a process-wide cache (per device, at most 3/4 of its memory) and handed to
// the next large request of at least half their size (best fit): a 50 GB
// hipMalloc right after large frees took 2.3 s on MI355X, against under
// 1 ms on memory not used before, so a second assembly in one process would
// pay that for each large buffer.  A failed hipMalloc empties the device's
// cache and retries; cache_flush(device) (context destruction,
// ahm_ctx_release_cached) returns the blocks.
struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;   // requested size
  size_t cap = 0;     // size of the block (>= bytes when it came from the cache)
  int dev = 0;
  int alloc(size_t b);
  void release();
  ~DeviceBuf() { release(); }
};
size_t cache_bytes(int device);   // bytes held by the cache on the device
void cache_flush(int device);

size_t dtype_size(ahm_dtype t);

// Host-side phase timing of the set-up paths (assembly, layout), printed to
// stderr when the environment variable AHM_TIMING is set: one line per
// mark with the milliseconds since the previous mark.
struct PhaseTimer {
  explicit PhaseTimer(const char* what);
  void mark(const char* phase);
  bool on;
  const char* what;
  double t_last;
};

// y += alpha H x on a caller-provided workspace (hmat.hip); no allocation
int matvec_on_workspace(ahm_hmat h, void* base, size_t bytes, double alpha, const void* X, uint32_t ldx,
                        void* Y, uint32_t ldy, uint32_t p, ahm_dtype vec, const char* who);

// Single-RHS products of the solvers (hmat.hip).  matvec_prepare sizes the
// handle's workspace for them (call it before a graph capture: the gated
// product allocates nothing and refuses to run on a workspace too small);
// seg[0..3] give the system of each phase-A, phase-B, phase-A reduction and
// phase-B reduction task (nullptr: all system 0).  matvec_signature lists
// every value a p = 1 launch bakes in (task lists, counts, arena and
// workspace addresses): a captured product stays valid while it is unchanged.
int matvec_prepare(ahm_hmat h, ahm_dtype vec);
int matvec_gated(ahm_hmat h, double alpha, const void* x, void* y, ahm_dtype vec, const uint32_t* frozen,
                 const uint32_t* const seg[4]);
void matvec_signature(ahm_hmat h, std::vector<uint64_t>& sig);

// The transposed part of symmetric products (hmat_t.hip): sym_prepare builds
// its column index and sizes its buffers for vectors of type vec and p
// right-hand sides (allocates; call before a graph capture); sym_transposed
// adds alpha S^T X for the leaves off the diagonal (32 right-hand sides at a
// time when p >= 2) and allocates nothing.  Both do nothing for unsymmetric
// matrices.
int sym_prepare(ahm_hmat h, ahm_dtype vec, uint32_t p = 1);
int sym_transposed(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y, uint32_t ldy, uint32_t p,
                   ahm_dtype vec, Gate g);
// Single right-hand side, symmetric: the phase-B launch that reads each row
// panel value once for both directions (y rows += alpha [D | U] z, or the
// chunk's partial sums, and the column dots of the panel with x_rows into the
// slots), then, after the phase-B chunk reduction, sym_rest: t' = U^T x,
// V t', and y += alpha (slot sums) (hmat_t.hip).  After sym_prepare.
int sym_phase_b(ahm_hmat h, double alpha, const void* x, const void* t, void* y, void* part, ahm_dtype vec,
                Gate g);
int sym_rest(ahm_hmat h, double alpha, void* y, ahm_dtype vec, Gate g);

// solver state kept with a handle between calls (gmres.hip)
struct GmresCache;
void gmres_cache_free(ahm_hmat h);

// precond_apply (Include/matrix.h:30) as HLU_solve (Include/H.h:1383-1388):
// v <- U^{-1} L^{-1} v with the triangular H-solves (trih.hip); either factor
// may be absent.  Runs on the factors' context stream, which must be the
// operator's.
struct Prec {
  ahm_trih L = nullptr, U = nullptr;
  uint32_t n = 0;
  bool on() const { return L || U; }
  int apply(double* v) const {
    if (L) AHM_TRY(ahm_trih_solve(L, v, n, 1));
    if (U) AHM_TRY(ahm_trih_solve(U, v, n, 1));
    return AHM_OK;
  }
};

// Restarted MGS-GMRES(m) (solvers/GMRES.cpp:80-181) on S independent
// systems (gmres.hip); see the file for the device state and the step
// protocol.  S = 1 unless batch (the handle's diagonal blocks).
struct GmresArgs {
  const double* b = nullptr;
  double* x = nullptr;
  double eps = 0.0;
  uint32_t m = 0, maxsteps = 0;
  int use_graph = 1;
  bool batch = false;
  const Prec* M = nullptr;
  double* hist = nullptr;     // single system: implicit residual per step
  uint32_t nhist = 0;
  // outputs, S entries each (S = 1 unless batch)
  double* eps_out = nullptr;
  uint32_t* steps_out = nullptr;
  int* conv_out = nullptr;
  uint32_t* nmatvec = nullptr;   // single system: H-matvecs done
};
int gmres_run(ahm_hmat h, const GmresArgs& a);

struct PlanJob;
int hmat_build(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols, std::vector<ahm_leaf_desc>&& leaves,
               struct DeviceBuf& raw, ahm_dtype data_dtype, ahm_dtype storage, int keep_leaves,
               ahm_hmat* out, std::shared_ptr<PlanJob> pre = nullptr);

}  // namespace ahm

struct ahm_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own = nullptr;
};

struct ahm_hmat_s {
  ahm_ctx ctx = nullptr;
  uint32_t n_rows = 0, n_cols = 0;
  ahm_dtype storage = AHM_F64;
  ahm_dtype leaf_dtype = AHM_F64;
  uint64_t nvals = 0;
  uint32_t t_len = 0;
  uint32_t max_rank = 0;
  uint64_t n_dense = 0, n_lowrank = 0;
  uint64_t nv_dense = 0, nv_u = 0, nv_v = 0;
  ahm::DeviceBuf arena;       // nvals values in panel layout
  ahm::DeviceBuf tasks_a, tasks_b, zsrc, tasks_r, tasks_ra;
  uint32_t n_a = 0, n_b = 0, n_r = 0, n_ra = 0;
  // multi-RHS MFMA launches: their own chunking of the panels (tasks,
  // reductions) and workgroup maps (task << 2 | row piece)
  ahm::DeviceBuf mm_tasks_a, mm_tasks_b, mm_tasks_ra, mm_tasks_r, mm_map_a, mm_map_b;
  uint32_t n_mm_a = 0, n_mm_b = 0, n_mm_ra = 0, n_mm_r = 0, n_mm_ta = 0;
  std::vector<ahm::GemvTask> h_mm_tasks_a;   // host copy (the BF16 V^T copy follows it)
  uint32_t part_len = 0;      // partial-sum slots per right-hand side (shared by
                              // the phase-A and phase-B chunk sums, used in turn)
  ahm::DeviceBuf pbuf;        // partial sums: part_len * p values
  size_t pbuf_cap = 0;
  std::vector<ahm::ReduceTask> h_tasks_r, h_tasks_ra;
  // independent systems stacked block-diagonally (batched assembly):
  // system s owns rows/columns [seg_off[s], seg_off[s+1])
  std::vector<uint32_t> seg_off;
  // V^T panels and row panels before chunking, and the chunk size in use
  std::vector<ahm::GemvTask> h_panels, h_panels_a;
  uint64_t chunk_values = ahm::kChunkValues;
  uint64_t n_zsrc = 0;
  ahm::DeviceBuf tbuf;        // V^T X workspace: t_len values per right-hand side slot
  size_t tbuf_cap = 0;        // bytes
  ahm::DeviceBuf xbuf;        // interleaved copy of X for multi-RHS products
  size_t xbuf_cap = 0;        // bytes
  // AHM_BF16 storage (FP32 values, V^T X of multi-RHS products on BF16
  // MFMA): BF16 copy of the phase-A tasks in k-groups of 8 columns, the
  // offset of each task's copy (parallel to mm_tasks_a), the BF16 interleaved X
  ahm::DeviceBuf bfarena, bfoff;
  uint64_t bf_vals = 0;
  ahm::DeviceBuf xbbuf;
  size_t xbbuf_cap = 0;       // bytes
  // single-launch products of 1..8 right-hand sides (ahm_hmat_set_fused):
  // row panels in launch order, V^T panels, U-column groups and their
  // offsets per row panel
  int fused = 0, fz_built = 0;   // launch form: 0 three kernels, 1 row-panel single launch, 2 persistent
  int pgrid[4] = {0, 0, 0, 0};    // resident workgroups of the persistent kernel (f64/f32 x 1/several RHS)
  ahm::DeviceBuf fz_panels, fz_apanels, fz_groups, fz_goff;
  uint32_t fz_n = 0;
  int sym = 0;                // symmetric storage (AHM_LEAF_HEM diagonal leaves)
  ahm::DeviceBuf leafdata;    // optional mblock-layout copy (keep_leaves)
  std::vector<ahm_leaf_desc> leaves;
  // host copies of the plan (used by ahm_hmat_convert)
  std::vector<ahm::GemvTask> h_tasks_a, h_tasks_b;
  ahm::HostVec<uint32_t> h_zsrc;
  // transposed product (built on first use, hmat_t.hip): column-dot tasks of
  // both phases and, per destination (t index / column), the panel columns
  // that contribute to it (CSR, fixed order)
  struct Transposed {
    bool built = false;
    ahm::DeviceBuf tasks_a, tasks_b, off_t, pos_t, off_y, pos_y;
    uint32_t n_a = 0, n_b = 0;
    ahm::DeviceBuf tasks_a_mm, tasks_b_mm;   // the same columns in longer pieces (several RHS)
    uint32_t n_a_mm = 0, n_b_mm = 0;
    // symmetric products: the phase-B tasks (same chunks and partial-sum
    // slots as tasks_b of the handle) with the first row of their panel,
    // for the single pass that forms row sums and column dots together
    bool sym_built = false;
    ahm::DeviceBuf sym_tasks;
    uint32_t n_sym = 0;
    ahm::DeviceBuf contrib, tprime;   // one value per panel column; t' = U^T x
    size_t contrib_cap = 0, tprime_cap = 0;
    ahm::DeviceBuf xi;                // several right-hand sides: X rows, 32 values each
    size_t xi_cap = 0;
  } tr;
  ahm::GmresCache* gm = nullptr;   // GMRES buffers and captured steps (gmres.hip)
  int solver_graphs = 0;           // GMRES steps replayed as hipGraphs (ahm_hmat_set_graphs)
  int arnoldi = 0;                 // GMRES orthogonalisation form (ahm_hmat_set_arnoldi)
};

// triangular H-matrix factor (trih.hip): work queue, arena, counters
struct ahm_trih_s {
  ahm_ctx ctx = nullptr;
  uint32_t n = 0;
  int mode = 0;   // 0 L, 1 U, 2 U^T, 3 L^T (trih_kernel's modes)
  uint32_t n_diag = 0, n_lr = 0;
  uint64_t n_items = 0, nvals = 0, arena_vals = 0;
  ahm::DeviceBuf arena, items, perm, rinv, pend0, counters, tpart;
  uint64_t t_len = 0;
  uint32_t grid = 0;         // resident workgroups of the one-chain trih_kernel
  uint32_t grid_multi = 0;   // resident workgroups of the several-chain trih_kernel
  uint32_t chains = 0;   // right-hand sides the counters and slots are sized for
};
