// trih.hip -- forward / backward substitution with triangular H-matrix factors.
//
// Replaces the reference's LtHGeM_solve_ / UtHGeM_solve_ (H/TH_solve.cpp:
// 18-91) with the leaf solves mblock::ltr_solve (H/mblock_D.cpp:1183-1205:
// gather by the row permutation, then dtpsv 'L','N','U', Include/blas.h:
// 3749-3755) and mblock::utr_solve (Include/mblock.h:694-697: dtpsv 'U','N',
// 'N', Include/blas.h:3930-3936).
//
// The reference recurses over the block tree: solve the first diagonal son,
// subtract the off-diagonal son times the solved part from the rest, solve
// the second diagonal son.  Only the diagonal leaves form a chain; an
// off-diagonal leaf can run as soon as the diagonal leaves covering its
// columns are solved, and must finish before the diagonal leaves covering
// its rows start.  Here that is one persistent launch over a work queue:
//   S  solve one diagonal leaf (one workgroup; a wavefront substitutes 64
//      rows at a time from a 64x64 triangle staged in LDS, the workgroup
//      applies the rectangle below (above, for 'U') the triangle);
//   D  a tile (<= 256 rows) of a dense off-diagonal leaf: acc[rows] -= D x;
//   A  a column chunk of V^T x of a low-rank leaf, into its own slot;
//   B  a row tile of a low-rank leaf: acc[rows] -= U t, t = fixed-order sum
//      of the leaf's A slots.
// Workgroups take items in queue order (one atomic counter).  An item only
// waits for items earlier in the queue (the host builds the order and checks
// it by running the queue sequentially), so every wait ends: the items it
// waits for were taken by workgroups that are already running.  Waits are
// bounded anyway; a wait that runs out sets an abort flag that ends every
// workgroup and the call returns an error.
//
// acc is the right-hand side vector itself (solved in place).  D and B items
// subtract with FP64 atomics, so the order of the contributions to one row
// (not their values) depends on timing; S items, A slots and t sums are in a
// fixed order.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

namespace {

constexpr uint32_t kItemS = 0, kItemA = 1, kItemB = 2, kItemD = 3;
constexpr int kTri = 64;             // rows substituted per wavefront pass
constexpr uint32_t kMaxDiag = 2048;  // largest diagonal leaf (LDS vector)
constexpr uint32_t kMaxW = 2048;     // columns per gemv item (LDS vector)
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMaxChains = 16;  // right-hand sides solved by one launch
constexpr uint32_t kPushVals = 12288;  // LDS values of a dense block an S item applies itself
constexpr uint32_t kSpinLimit = 1u << 22;

struct TriItem {
  uint64_t data;   // arena offset: panel (A/B/D) or the diagonal block (S)
  uint32_t type;
  uint32_t m;      // rows (A/B/D), or the diagonal leaf's size (S)
  uint32_t W;      // columns (A/B/D)
  uint32_t ld;     // column stride of the panel (A/B/D)
  uint32_t in;     // A/D: first column in acc; B: slot of chunk 0, rank 0; S: pushed block or kNone
  uint32_t out;    // A: first slot; B/D: first row in acc; S: first row
  uint32_t wait;   // S: diagonal index; A/D: diagonal that must be solved; B: low-rank leaf
  uint32_t wait2;  // S: previous diagonal in solve order (kNone for the first); B: A-item count
  uint32_t sig0;   // B/D: first diagonal whose pending count drops; A: low-rank leaf
  uint32_t sig1;   // B/D: last such diagonal
  uint32_t aux;    // B: rank k (slot stride); S: permutation (unit) or reciprocal-pivot (non-unit) offset
  uint32_t parts;  // B: slots summed per rank entry (the leaf's column chunks)
};

// counters layout per chain: [0] queue head, [1] abort flag (chain 0's is
// the one all chains use), then solved[n_diag], tdone[n_lr], pending[n_diag]
struct Counters {
  uint32_t* head;
  uint32_t* abort;
  uint32_t* solved;
  uint32_t* tdone;
  uint32_t* pending;
};

__device__ __forceinline__ uint32_t load_rlx(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spin until *p == v (eq) or *p >= v with relaxed (agent-coherent) loads;
// an acquire load would invalidate the caches on every poll.  False if the
// abort flag is (or gets) set.
__device__ bool spin_until(const uint32_t* p, uint32_t v, bool eq, uint32_t* abort) {
  for (uint32_t it = 0;; ++it) {
    const uint32_t c = load_rlx(p);
    if (eq ? c == v : c >= v) return true;
    if ((it & 31) == 31 && load_rlx(abort)) return false;
    if (it > kSpinLimit) {
      __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Values produced by one workgroup and read by another (the vector, the
// low-rank slots) move through agent-coherent accesses (the sc1 loads and
// stores of relaxed agent-scope atomics), so no cache has to be written back
// or invalidated around a dependency: a producer waits for its own stores
// (s_waitcnt) before the workgroup barrier that precedes its flag, and a
// consumer reads them only after its spin on the flag has succeeded.
__device__ __forceinline__ double ld_c(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_c(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

struct TriSmem {
  double z[kMaxDiag];          // S: the leaf's vector; A/B/D: the input slice
  double tri[kTri * kTri];     // S: one triangle
  double red[kPanelThreads];   // cross-group partial sums
  uint32_t item;
  uint32_t ok;
};

constexpr int kBatch = 16;     // loads in flight per thread in tile_gemv

// out(r) = sum_c P[r + c*ld] * z[c] for r < m (m <= 256, W <= kMaxW), z in
// LDS; thread t < G*m owns row t % m and columns g, g+G, ... (G = 256/m),
// loaded kBatch at a time with clamped (always valid) addresses so that the
// loads of a batch are in flight together; returns the row sum on threads
// t < m (fixed-order reduction across the groups)
__device__ __forceinline__ double tile_gemv(const double* __restrict__ P, uint32_t m, uint32_t W,
                                            uint32_t ld, const double* z, double* red) {
  const int tid = threadIdx.x;
  const int G = kPanelThreads / (int)m;
  const int row = tid % (int)m, g = tid / (int)m;
  double a = 0;
  if (g < G) {
    const double* Pr = P + row;
    for (int c0 = g; c0 < (int)W; c0 += kBatch * G) {
      double v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int c = c0 + u * G;
        v[u] = Pr[(uint64_t)(c < (int)W ? c : c0) * ld];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int c = c0 + u * G;
        a = fma(v[u], c < (int)W ? z[c] : 0.0, a);
      }
    }
  }
  red[tid] = g < G ? a : 0.0;
  __syncthreads();
  double s = 0;
  if (tid < (int)m)
    for (int q = 0; q < G; ++q) s += red[q * (int)m + tid];
  return s;
}

// A/B/D items are cut so that a thread owns at most 2 * kBatch columns of
// its row (W <= kItemBatches * kBatch * G): their whole tile is loaded into
// registers before the item waits for its inputs.
constexpr int kItemBatches = 2;
constexpr int kItemCols = kItemBatches * kBatch;   // 32 columns per thread

__device__ __forceinline__ void gemv_prefetch(const double* __restrict__ P, uint32_t m, uint32_t W, uint32_t ld,
                                              double (&v)[kItemCols]) {
  const int tid = threadIdx.x;
  const int G = kPanelThreads / (int)m;
  const int row = tid % (int)m, g = tid / (int)m;
  const int gg = g < G ? g : 0;
#pragma unroll
  for (int u = 0; u < kItemCols; ++u) {
    const int c = gg + u * G;
    v[u] = P[row + (uint64_t)(c < (int)W ? c : 0) * ld];
  }
}

__device__ __forceinline__ double gemv_apply(const double (&v)[kItemCols], uint32_t m, uint32_t W, const double* z,
                                             double* red) {
  const int tid = threadIdx.x;
  const int G = kPanelThreads / (int)m;
  const int g = tid / (int)m;
  double a = 0;
  if (g < G) {
#pragma unroll
    for (int u = 0; u < kItemCols; ++u) {
      const int c = g + u * G;
      a = fma(v[u], c < (int)W ? z[c] : 0.0, a);
    }
  }
  red[tid] = g < G ? a : 0.0;
  __syncthreads();
  double s = 0;
  if (tid < (int)m)
    for (int q = 0; q < G; ++q) s += red[q * (int)m + tid];
  return s;
}

constexpr int kTriPerThread = kTri * kTri / kPanelThreads;   // 16

__device__ __forceinline__ void tri_load(const double* __restrict__ blk, double* t) {
#pragma unroll
  for (int q = 0; q < kTriPerThread; ++q) t[q] = blk[threadIdx.x + q * kPanelThreads];
}
__device__ __forceinline__ void tri_store(const double* t, double* tri) {
#pragma unroll
  for (int q = 0; q < kTriPerThread; ++q) tri[threadIdx.x + q * kPanelThreads] = t[q];
}

// Substitution of wave 0's 64 rows through a 64x64 unit triangle in LDS
// (zero on the diagonal and on the other side of it; a non-unit factor
// arrives row-scaled): lane r holds z_r; step j broadcasts z_j (final after
// the previous steps) with readlane and every lane subtracts its multiple,
// a no-op where the entry is zero.  The entries are read from LDS 16 steps
// ahead into registers, so a step costs only the readlane and the fma.
__device__ __forceinline__ double subst_fwd(double zr, const double* tri) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int jb = 0; jb < kTri; jb += 16) {
    double tv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) tv[u] = tri[(jb + u) * kTri + lane];
#pragma unroll
    for (int u = 0; u < 16; ++u) zr = fma(-tv[u], readlane_d(zr, jb + u), zr);
  }
  return zr;
}

__device__ __forceinline__ double subst_bwd(double zr, const double* tri) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int jb = kTri - 16; jb >= 0; jb -= 16) {
    double tv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) tv[u] = tri[(jb + u) * kTri + lane];
#pragma unroll
    for (int u = 15; u >= 0; --u) zr = fma(-tv[u], readlane_d(zr, jb + u), zr);
  }
  return zr;
}

// Data of a diagonal leaf that does not depend on the vector, loaded before
// the item waits: the first triangle (into LDS), the second triangle and the
// rectangle next to the first (registers, when it has at most 64 rows), and
// for 'L' the row permutation.
struct SPre {
  uint32_t ip[kMaxDiag / kPanelThreads];
  double rv[kMaxDiag / kPanelThreads];
};
// the register prefetch of an item: an A/B/D tile, or for an S item the
// second triangle (first kTriPerThread) and the first rectangle (last kBatch)
using Pf = double[kItemCols];
static_assert(kTriPerThread + kBatch == kItemCols, "S prefetch layout");

// thread mapping of a rows x 64 rectangle with rows <= 64 (tile_gemv's, one
// batch): row tid % rows, columns g, g+G, ... (G = 256 / rows >= 4, <= 16 each)
__device__ __forceinline__ void rect_load(const double* __restrict__ R, uint32_t rows, double* v) {
  const int tid = threadIdx.x;
  const int G = kPanelThreads / (int)rows;
  const int row = tid % (int)rows, g = tid / (int)rows;
  const int cg = g < G ? g : 0;
#pragma unroll
  for (int u = 0; u < kBatch; ++u) {
    const int c = cg + u * G;
    v[u] = R[(uint64_t)(c < kTri ? c : cg) * rows + row];
  }
}

__device__ __forceinline__ double rect_apply(const double* v, uint32_t rows, const double* z, double* red) {
  const int tid = threadIdx.x;
  const int G = kPanelThreads / (int)rows;
  const int g = tid / (int)rows;
  double a = 0;
  if (g < G) {
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int c = g + u * G;
      a = fma(v[u], c < kTri ? z[c] : 0.0, a);
    }
  }
  red[tid] = g < G ? a : 0.0;
  __syncthreads();
  double s = 0;
  if (tid < (int)rows)
    for (int q = 0; q < G; ++q) s += red[q * (int)rows + tid];
  return s;
}

__device__ __forceinline__ void solve_prefetch(const TriItem& it, const double* __restrict__ arena,
                                               const uint32_t* __restrict__ perm, const double* __restrict__ rinv,
                                               bool forward, bool unit, TriSmem& sm, SPre& pre, Pf& pf) {
  const int tid = threadIdx.x;
  const uint32_t n = it.m;
  const uint32_t np = (n + kTri - 1) / kTri;
  const double* blk = arena + it.data;
  double t[kTriPerThread];
  tri_load(blk, t);
  // first rectangle: below the first triangle (forward) or above the last
  const uint32_t rows0 = forward ? (n > kTri ? n - kTri : 0) : (np - 1) * kTri;
  if (rows0 > 0 && rows0 <= kTri) rect_load(blk + kTri * kTri, rows0, pf + kTriPerThread);
  if (np > 1) tri_load(blk + kTri * kTri + (uint64_t)rows0 * kTri, pf);
#pragma unroll
  for (int q = 0; q < kMaxDiag / kPanelThreads; ++q) {
    const uint32_t r = tid + q * kPanelThreads;
    if (unit) pre.ip[q] = r < n ? perm[it.aux + r] : 0u;
    else pre.rv[q] = r < n ? rinv[it.aux + r] : 0.0;
  }
  tri_store(t, sm.tri);
}

// one diagonal leaf, forward (L, U^T): for the unit lower factor the slice
// is first gathered through the row permutation (P^T L, mblock::ltr_solve);
// 64 rows at a time, then the rectangle below each triangle
template <bool UNIT>
__device__ void solve_fwd(const TriItem& it, const double* __restrict__ arena, double* acc, TriSmem& sm,
                          const SPre& pre, const Pf& pf) {
  const int tid = threadIdx.x;
  const uint32_t n = it.m, s0 = it.out;
  const uint32_t np = (n + kTri - 1) / kTri;
#pragma unroll
  for (int q = 0; q < kMaxDiag / kPanelThreads; ++q) {
    const uint32_t r = tid + q * kPanelThreads;
    if (r < np * kTri) sm.z[r] = r < n ? (UNIT ? ld_c(acc + s0 + pre.ip[q]) : ld_c(acc + s0 + r) * pre.rv[q]) : 0.0;
  }
  const double* blk = arena + it.data;
  for (uint32_t p = 0; p < np; ++p) {
    const uint32_t c0 = p * kTri;
    const uint32_t below = n > c0 + kTri ? n - c0 - kTri : 0;
    __syncthreads();
    if (tid < kTri) sm.z[c0 + tid] = subst_fwd(sm.z[c0 + tid], sm.tri);
    __syncthreads();
    const double* rect = blk + kTri * kTri;
    const double* next = rect + (uint64_t)below * kTri;
    double t[kTriPerThread];
    if (p > 0 && p + 1 < np) tri_load(next, t);
    if (p == 0 && below > 0 && below <= kTri) {
      const double s = rect_apply(pf + kTriPerThread, below, sm.z + c0, sm.red);
      if (tid < (int)below) sm.z[c0 + kTri + tid] -= s;
      __syncthreads();
    } else {
      for (uint32_t r0 = 0; r0 < below; r0 += kPanelThreads) {
        const uint32_t m = min(below - r0, (uint32_t)kPanelThreads);
        const double s = tile_gemv(rect + r0, m, kTri, below, sm.z + c0, sm.red);
        if (tid < (int)m) sm.z[c0 + kTri + r0 + tid] -= s;
        __syncthreads();
      }
    }
    if (p + 1 < np) tri_store(p == 0 ? pf : t, sm.tri);
    blk = next;
  }
  __syncthreads();
  for (uint32_t r = tid; r < n; r += kPanelThreads) st_c(acc + s0 + r, sm.z[r]);
}

// one diagonal leaf, backward (U, L^T): panels are stored last first, each
// a triangle followed by the rectangle above it; for the unit factor (L^T)
// the result is scattered back through the row permutation
// (mblock::ltrh_solve)
template <bool UNIT>
__device__ void solve_bwd(const TriItem& it, const double* __restrict__ arena, double* acc, TriSmem& sm,
                          const SPre& pre, const Pf& pf) {
  const int tid = threadIdx.x;
  const uint32_t n = it.m, s0 = it.out;
  const uint32_t np = (n + kTri - 1) / kTri;
#pragma unroll
  for (int q = 0; q < kMaxDiag / kPanelThreads; ++q) {
    const uint32_t r = tid + q * kPanelThreads;
    if (r < np * kTri) sm.z[r] = r < n ? (UNIT ? ld_c(acc + s0 + r) : ld_c(acc + s0 + r) * pre.rv[q]) : 0.0;
  }
  const double* blk = arena + it.data;
  for (uint32_t pp = np; pp-- > 0;) {
    const uint32_t c0 = pp * kTri;
    const bool first = pp + 1 == np;
    __syncthreads();
    if (tid < kTri) sm.z[c0 + tid] = subst_bwd(sm.z[c0 + tid], sm.tri);
    __syncthreads();
    const double* rect = blk + kTri * kTri;
    const double* next = rect + (uint64_t)c0 * kTri;
    double t[kTriPerThread];
    if (!first && pp > 0) tri_load(next, t);
    if (first && c0 > 0 && c0 <= kTri) {
      const double s = rect_apply(pf + kTriPerThread, c0, sm.z + c0, sm.red);
      if (tid < (int)c0) sm.z[tid] -= s;
      __syncthreads();
    } else {
      for (uint32_t r0 = 0; r0 < c0; r0 += kPanelThreads) {
        const uint32_t m = min(c0 - r0, (uint32_t)kPanelThreads);
        const double s = tile_gemv(rect + r0, m, kTri, c0, sm.z + c0, sm.red);
        if (tid < (int)m) sm.z[r0 + tid] -= s;
        __syncthreads();
      }
    }
    if (pp > 0) tri_store(first ? pf : t, sm.tri);
    blk = next;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kMaxDiag / kPanelThreads; ++q) {
    const uint32_t r = tid + q * kPanelThreads;
    if (r < n) st_c(acc + s0 + (UNIT ? pre.ip[q] : r), sm.z[r]);
  }
}

// One launch solves several right-hand sides as independent chains: chain
// blockIdx.y has its own vector, slots, queue head and counters (block of
// cstride words) and shares the read-only items, arena and the abort flag.
//
// PUSH_LDS: an S item stages the dense block it pushes in LDS while it waits
// (one workgroup per CU: the fastest single chain); without, it reads that
// block from the arena after its solve (two workgroups per CU, for launches
// of several chains, whose workers are shared out).
template <bool PUSH_LDS>
__global__ __launch_bounds__(kPanelThreads) void trih_kernel(
    const TriItem* __restrict__ items, uint32_t n_items, const double* __restrict__ arena,
    const uint32_t* __restrict__ perm, const double* __restrict__ rinv, double* acc, uint64_t ldb, double* tpart,
    uint64_t tstride,
    uint32_t* counters, uint32_t cstride, uint32_t n_diag, uint32_t n_lr, int mode) {
  {
    const uint32_t j = blockIdx.y;
    acc += j * ldb;
    tpart += j * tstride;
  }
  uint32_t* cb = counters + (uint64_t)blockIdx.y * cstride;
  const Counters cn{cb, counters + 1, cb + 2, cb + 2 + n_diag, cb + 2 + n_diag + n_lr};
  // mode: 0 L (forward, unit, gather), 1 U (backward), 2 U^T (forward),
  // 3 L^T (backward, unit, scatter)
  const bool forward = mode == 0 || mode == 2;
  const bool unit = mode == 0 || mode == 3;
  __shared__ TriSmem sm;
  __shared__ double push_lds[PUSH_LDS ? kPushVals : 1];   // S: the dense block it pushes
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) {
      const uint32_t i = __hip_atomic_fetch_add(cn.head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sm.item = (i < n_items && !load_rlx(cn.abort)) ? i : kNone;
    }
    __syncthreads();
    const uint32_t idx = sm.item;
    __syncthreads();
    if (idx == kNone) return;
    const TriItem it = items[idx];
    SPre pre;
    Pf pf;
    if (it.type == kItemS) {
      solve_prefetch(it, arena, perm, rinv, forward, unit, sm, pre, pf);
      if (PUSH_LDS && it.in != kNone) {
        // the block it pushes into the next leaf, staged in LDS while it waits
        const TriItem tp = items[n_items + it.in];
        const double* src = arena + tp.data;
        const uint32_t nv = tp.m * tp.W;
        for (uint32_t e0 = 0; e0 < nv; e0 += 16 * kPanelThreads) {
          double v[16];
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const uint32_t e = e0 + tid + q * kPanelThreads;
            v[q] = src[e < nv ? e : 0];
          }
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const uint32_t e = e0 + tid + q * kPanelThreads;
            if (e < nv) push_lds[e] = v[q];
          }
        }
      }
    } else {
      gemv_prefetch(arena + it.data, it.m, it.W, it.ld, pf);
    }
    if (tid == 0) {
      bool ok;
      if (it.type == kItemS) {
        ok = spin_until(cn.pending + it.wait, 0u, true, cn.abort);
        if (ok && it.wait2 != kNone) ok = spin_until(cn.solved + it.wait2, 1u, false, cn.abort);
      } else if (it.type == kItemB) {
        ok = spin_until(cn.tdone + it.wait, it.wait2, false, cn.abort);
      } else {
        ok = spin_until(cn.solved + it.wait, 1u, false, cn.abort);
      }
      sm.ok = ok ? 1u : 0u;
    }
    __syncthreads();
    if (!sm.ok) return;
    if (it.type == kItemS) {
      if (mode == 0)
        solve_fwd<true>(it, arena, acc, sm, pre, pf);
      else if (mode == 1)
        solve_bwd<false>(it, arena, acc, sm, pre, pf);
      else if (mode == 2)
        solve_fwd<false>(it, arena, acc, sm, pre, pf);
      else
        solve_bwd<true>(it, arena, acc, sm, pre, pf);
      TriItem tp{};
      if (it.in != kNone) {
        // the dense block that reads only this leaf's solution (still in
        // sm.z) and feeds the next leaf: applied here, saving a hand-over
        tp = items[n_items + it.in];
        const int m = (int)tp.m;
        const int G = kPanelThreads / m;
        const int row = tid % m, g = tid / m;
        const double* zc = sm.z + (tp.in - it.out);
        const double* pb = PUSH_LDS ? push_lds : arena + tp.data;
        double a = 0;
        if (g < G)
          for (int c = g; c < (int)tp.W; c += G) a = fma(pb[row + c * m], zc[c], a);
        sm.red[tid] = g < G ? a : 0.0;
        __syncthreads();
        if (tid < m) {
          double sum = 0;
          for (int q = 0; q < G; ++q) sum += sm.red[q * m + tid];
          unsafeAtomicAdd(acc + tp.out + tid, -sum);
        }
      }
      stores_done();
      __syncthreads();
      if (tid == 0) {
        __hip_atomic_store(cn.solved + it.wait, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (it.in != kNone)
          for (uint32_t d = tp.sig0; d <= tp.sig1; ++d)
            __hip_atomic_fetch_sub(cn.pending + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      continue;
    }
    // stage the input slice
    if (it.type == kItemB) {
      const uint32_t k = it.aux, nparts = it.parts;
      for (uint32_t l = tid; l < it.W; l += kPanelThreads) {
        double t = 0;
        for (uint32_t c = 0; c < nparts; ++c) t += ld_c(tpart + it.in + (uint64_t)c * k + l);
        sm.z[l] = t;
      }
    } else {
      for (uint32_t c = tid; c < it.W; c += kPanelThreads) sm.z[c] = ld_c(acc + it.in + c);
    }
    __syncthreads();
    const double s = gemv_apply(pf, it.m, it.W, sm.z, sm.red);
    if (tid < (int)it.m) {
      if (it.type == kItemA)
        st_c(tpart + it.out + tid, s);
      else
        unsafeAtomicAdd(acc + it.out + tid, -s);
    }
    stores_done();
    __syncthreads();
    if (tid == 0) {
      if (it.type == kItemA) {
        __hip_atomic_fetch_add(cn.tdone + it.sig0, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        for (uint32_t d = it.sig0; d <= it.sig1; ++d)
          __hip_atomic_fetch_sub(cn.pending + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// counters of chain blockIdx.y <- (head 0, abort 0, solved = 0, tdone = 0,
// pending = pend0)
__global__ void trih_reset_kernel(uint32_t* counters, uint32_t cstride, uint32_t n_zero, const uint32_t* pend0,
                                  uint32_t n_diag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t* c = counters + (uint64_t)blockIdx.y * cstride;
  if (i < n_zero) c[i] = 0;
  else if (i < n_zero + n_diag) c[i] = pend0[i - n_zero];
}

// ---------------------------------------------------------------- host plan

struct Diag {
  uint32_t leaf, s, n;
};

}  // namespace

// The work queue of one factor, built and checked on the host.
struct TriPlan {
  std::vector<double> arena;
  std::vector<uint32_t> perm;
  std::vector<double> rinv;      // non-unit factors: reciprocal pivots, per diagonal leaf row
  std::vector<TriItem> queue;
  std::vector<TriItem> push;     // dense blocks applied by the S item that solves their columns
  std::vector<uint32_t> pend0;   // per diagonal: B/D tiles and pushed blocks touching its rows
  uint32_t n_diag = 0, n_lr = 0;
  uint64_t t_len = 0, nvals = 0;
};

// op: 'L' / 'U' solve with the stored lower / upper factor, 'l' / 'u' with
// its transpose.  L and U^T sweep forward, U and L^T backward; the unit
// factor (L) carries the row permutation.  A transposed factor is turned
// into the solve's orientation here: leaf (b1, n1, b2, n2) becomes
// (b2, n2, b1, n1), a dense block its transpose, U V^T becomes V U^T.
int build_plan(uint32_t n, char op, uint32_t n_leaves, const ahm_leaf_desc* leaves, const double* src,
               TriPlan& P) {
  const bool stored_lower = op == 'L' || op == 'l';
  const bool trans = op == 'l' || op == 'u';
  const bool forward = stored_lower != trans;
  const bool unit = stored_lower;
  const char uplo = stored_lower ? 'L' : 'U';
  const uint32_t dflag = stored_lower ? AHM_LEAF_LTM : AHM_LEAF_UTM;

  // diagonal leaves, sorted by first row; they must tile [0, n)
  std::vector<Diag> dg;
  for (uint32_t l = 0; l < n_leaves; ++l) {
    const ahm_leaf_desc& d = leaves[l];
    if (d.flags == dflag) {
      if (d.b1 != d.b2 || d.n1 != d.n2 || d.n1 == 0)
        return fail(AHM_ERR_INVALID, "ahm_trih_upload: diagonal leaf %u is not square on the diagonal", l);
      if (d.n1 > kMaxDiag)
        return fail(AHM_ERR_UNSUPPORTED, "ahm_trih_upload: diagonal leaf of %u rows (at most %u)", d.n1,
                    kMaxDiag);
      dg.push_back({l, d.b1, d.n1});
    } else if (d.flags != 0) {
      return fail(AHM_ERR_INVALID, "ahm_trih_upload: leaf %u has flags %u (expected 0 or %u for '%c')", l,
                  d.flags, dflag, uplo);
    }
  }
  std::sort(dg.begin(), dg.end(), [](const Diag& a, const Diag& b) { return a.s < b.s; });
  {
    uint32_t pos = 0;
    for (const Diag& d : dg) {
      if (d.s != pos) return fail(AHM_ERR_INVALID, "ahm_trih_upload: diagonal leaves do not tile [0, %u)", n);
      pos += d.n;
    }
    if (pos != n) return fail(AHM_ERR_INVALID, "ahm_trih_upload: diagonal leaves do not tile [0, %u)", n);
  }
  const uint32_t D = (uint32_t)dg.size();
  std::vector<uint32_t> row_diag(n);
  for (uint32_t i = 0; i < D; ++i)
    for (uint32_t r = 0; r < dg[i].n; ++r) row_diag[dg[i].s + r] = i;
  // solve order: step(d) = d (forward) or D-1-d (backward)
  auto step = [&](uint32_t d) { return forward ? d : D - 1 - d; };

  std::vector<double>& ar = P.arena;
  std::vector<uint32_t>& perm = P.perm;
  std::vector<TriItem> its;
  std::vector<uint32_t> need;   // step that needs the item (non-S items)
  std::vector<uint32_t> cls;    // ordering class: A 0, D 1, B 2
  uint64_t nvals = 0;

  // diagonal leaves: 64-row panels, each a 64x64 triangle followed by the
  // rectangle below it (lower) or above it (upper, panels last first)
  std::vector<TriItem> sitems(D);
  for (uint32_t i = 0; i < D; ++i) {
    const ahm_leaf_desc& d = leaves[dg[i].leaf];
    const uint32_t nn = d.n1;
    const double* pk = src + d.offset;
    nvals += (uint64_t)nn * (nn + 1) / 2;
    TriItem it{};
    it.type = kItemS;
    it.m = nn;
    it.out = d.b1;
    it.data = ar.size();
    it.wait = i;
    const uint32_t np = (nn + kTri - 1) / kTri;
    // packed column j starts at: lower sum_{q<j} (n-q); upper j(j+1)/2
    auto lo_at = [&](uint32_t r, uint32_t c) -> double {  // r > c
      const uint64_t o = (uint64_t)c * nn - (uint64_t)c * (c - 1) / 2;
      return pk[o + (r - c)];
    };
    auto up_at = [&](uint32_t r, uint32_t c) -> double {  // r <= c
      return pk[(uint64_t)c * (c + 1) / 2 + r];
    };
    // entry (r, c) of the triangle in the solve's orientation
    auto elem = [&](uint32_t r, uint32_t c) -> double {
      if (!trans) return stored_lower ? lo_at(r, c) : up_at(r, c);
      return stored_lower ? lo_at(c, r) : up_at(c, r);
    };
    if (unit) {
      it.aux = (uint32_t)perm.size();
      std::vector<char> seen(nn, 0);
      for (uint32_t j = 0; j < nn; ++j) {
        const double v = pk[(uint64_t)j * nn - (uint64_t)j * (j - 1) / 2];
        const uint32_t ip = (uint32_t)v;
        if (!(v >= 0.0) || (double)ip != v || ip >= nn || seen[ip])
          return fail(AHM_ERR_INVALID, "ahm_trih_upload: diagonal leaf at row %u: entry %u is not a permutation index",
                      d.b1, j);
        seen[ip] = 1;
        perm.push_back(ip);
      }
    }
    // A non-unit triangle is solved as the unit one D^{-1} T (rows scaled by
    // the reciprocal pivots) on D^{-1} z: every stored entry of row r
    // (triangle and rectangles) is multiplied by 1/T_rr here and the slice
    // by the same reciprocals once it is loaded, so one substitution routine
    // serves both kinds and no step divides.
    std::vector<double> rs(nn, 1.0);
    if (!unit) {
      it.aux = (uint32_t)P.rinv.size();
      for (uint32_t r = 0; r < nn; ++r) {
        const double dr = elem(r, r);
        if (dr == 0.0)
          return fail(AHM_ERR_INVALID, "ahm_trih_upload: zero on the diagonal of the leaf at row %u", d.b1);
        rs[r] = 1.0 / dr;
        P.rinv.push_back(rs[r]);
      }
    }
    auto sel = [&](uint32_t r, uint32_t c) { return elem(r, c) * rs[r]; };
    if (forward) {
      for (uint32_t p = 0; p < np; ++p) {
        const uint32_t c0 = p * kTri;
        const size_t base = ar.size();
        ar.resize(base + kTri * kTri, 0.0);
        for (uint32_t j = 0; j < kTri && c0 + j < nn; ++j)
          for (uint32_t r = j + 1; r < kTri && c0 + r < nn; ++r) ar[base + j * kTri + r] = sel(c0 + r, c0 + j);
        const uint32_t below = nn > c0 + kTri ? nn - c0 - kTri : 0;
        const size_t rb = ar.size();
        ar.resize(rb + (size_t)below * kTri, 0.0);
        for (uint32_t j = 0; j < kTri && c0 + j < nn; ++j)
          for (uint32_t r = 0; r < below; ++r) ar[rb + (size_t)j * below + r] = sel(c0 + kTri + r, c0 + j);
      }
    } else {
      for (uint32_t pp = np; pp-- > 0;) {
        const uint32_t c0 = pp * kTri;
        const size_t base = ar.size();
        ar.resize(base + kTri * kTri, 0.0);
        for (uint32_t j = 0; j < kTri && c0 + j < nn; ++j)
          for (uint32_t r = 0; r < j; ++r) ar[base + j * kTri + r] = sel(c0 + r, c0 + j);
        const size_t rb = ar.size();
        ar.resize(rb + (size_t)c0 * kTri, 0.0);
        for (uint32_t j = 0; j < kTri && c0 + j < nn; ++j)
          for (uint32_t r = 0; r < c0; ++r) ar[rb + (size_t)j * c0 + r] = sel(r, c0 + j);
      }
    }
    it.wait2 = kNone;
    it.in = kNone;
    sitems[i] = it;
  }
  for (uint32_t i = 0; i < D; ++i) {
    // previous diagonal in solve order
    if (forward && i > 0) sitems[i].wait2 = i - 1;
    if (!forward && i + 1 < D) sitems[i].wait2 = i + 1;
  }

  // off-diagonal leaves
  std::vector<uint32_t> a_count;  // per low-rank leaf
  uint64_t t_len = 0;
  for (uint32_t l = 0; l < n_leaves; ++l) {
    const ahm_leaf_desc& ld = leaves[l];
    if (ld.flags) continue;
    if ((uint64_t)ld.b1 + ld.n1 > n || (uint64_t)ld.b2 + ld.n2 > n)
      return fail(AHM_ERR_INVALID, "ahm_trih_upload: leaf %u outside [0, %u)", l, n);
    if (ld.n1 == 0 || ld.n2 == 0) continue;
    const bool stored_ok = stored_lower ? (uint64_t)ld.b2 + ld.n2 <= ld.b1 : (uint64_t)ld.b1 + ld.n1 <= ld.b2;
    if (!stored_ok)
      return fail(AHM_ERR_INVALID, "ahm_trih_upload: leaf %u is not strictly %s the diagonal", l,
                  stored_lower ? "below" : "above");
    // the leaf in the solve's orientation
    ahm_leaf_desc d = ld;
    if (trans) {
      d.b1 = ld.b2;
      d.n1 = ld.n2;
      d.b2 = ld.b1;
      d.n2 = ld.n1;
    }
    // diagonal whose solve completes a column range [c0, c1)
    auto ready_of = [&](uint32_t c0, uint32_t c1) { return forward ? row_diag[c1 - 1] : row_diag[c0]; };
    // diagonals touched by rows [r0, r1), and the step needing them first
    auto rows_of = [&](uint32_t r0, uint32_t r1, uint32_t& d0, uint32_t& d1) {
      d0 = row_diag[r0];
      d1 = row_diag[r1 - 1];
      return forward ? step(d0) : step(d1);
    };
    if (d.rank < 0) {
      nvals += (uint64_t)d.n1 * d.n2;
      const size_t base = ar.size();
      const double* D0 = src + ld.offset;
      if (!trans) {
        ar.insert(ar.end(), D0, D0 + (size_t)d.n1 * d.n2);
      } else {  // transpose: entry (r, c) of the solve's block is (c, r) of the stored one
        ar.resize(base + (size_t)d.n1 * d.n2);
        for (uint32_t c = 0; c < d.n2; ++c)
          for (uint32_t r = 0; r < d.n1; ++r) ar[base + r + (size_t)c * d.n1] = D0[c + (size_t)r * ld.n1];
      }
      {
        // a block whose columns lie in one diagonal leaf and whose rows are
        // the next step's is applied by that leaf's S item (from LDS) right
        // after its solve; not for L^T, whose solution leaves the S item
        // permuted
        const uint32_t dc = row_diag[d.b2];
        uint32_t p0, p1;
        const uint32_t nd = rows_of(d.b1, d.b1 + d.n1, p0, p1);
        if (!(unit && !forward) && row_diag[d.b2 + d.n2 - 1] == dc && nd == step(dc) + 1 &&
            d.n1 <= (uint32_t)kPanelThreads && (uint64_t)d.n1 * d.n2 <= kPushVals && sitems[dc].in == kNone) {
          TriItem tp{};
          tp.type = kItemD;
          tp.data = base;
          tp.m = d.n1;
          tp.W = d.n2;
          tp.ld = d.n1;
          tp.in = d.b2;
          tp.out = d.b1;
          tp.wait = dc;
          tp.sig0 = p0;
          tp.sig1 = p1;
          sitems[dc].in = (uint32_t)P.push.size();
          P.push.push_back(tp);
          continue;
        }
      }
      const uint32_t mrow = std::min<uint32_t>(d.n1, kPanelThreads);
      const uint32_t wc = std::min<uint32_t>(kMaxW, kItemCols * (kPanelThreads / mrow));
      for (uint32_t r0 = 0; r0 < d.n1; r0 += kPanelThreads) {
        const uint32_t m = std::min<uint32_t>(kPanelThreads, d.n1 - r0);
        for (uint32_t c0 = 0; c0 < d.n2; c0 += wc) {
          const uint32_t W = std::min(wc, d.n2 - c0);
          TriItem it{};
          it.type = kItemD;
          it.data = base + r0 + (uint64_t)c0 * d.n1;
          it.m = m;
          it.W = W;
          it.ld = d.n1;
          it.in = d.b2 + c0;
          it.out = d.b1 + r0;
          it.wait = ready_of(d.b2 + c0, d.b2 + c0 + W);
          need.push_back(rows_of(d.b1 + r0, d.b1 + r0 + m, it.sig0, it.sig1));
          cls.push_back(1);
          its.push_back(it);
        }
      }
    } else if (d.rank > 0) {
      const uint32_t k = (uint32_t)d.rank;
      if (k > kMaxW) return fail(AHM_ERR_UNSUPPORTED, "ahm_trih_upload: rank %u above %u", k, kMaxW);
      nvals += (uint64_t)k * (d.n1 + d.n2);
      const uint32_t lr = (uint32_t)a_count.size();
      // stored U (ld.n1 x k) then V (ld.n2 x k); transposed, V U^T
      const double* Us = src + ld.offset;
      const double* Vs = Us + (size_t)k * ld.n1;
      const double* U = trans ? Vs : Us;
      const double* V = trans ? Us : Vs;
      const size_t ub = ar.size();
      ar.insert(ar.end(), U, U + (size_t)k * d.n1);
      const size_t vb = ar.size();
      ar.resize(vb + (size_t)k * d.n2);
      for (uint32_t c = 0; c < d.n2; ++c)
        for (uint32_t q = 0; q < k; ++q) ar[vb + (size_t)c * k + q] = V[(size_t)q * d.n2 + c];
      // the leaf's first need: its first (forward) / last (backward) row diagonal
      uint32_t e0, e1;
      const uint32_t leaf_need = rows_of(d.b1, d.b1 + d.n1, e0, e1);
      // A items: rank slices of <= 256 x column chunks
      const uint32_t mk = std::min<uint32_t>(k, kPanelThreads);
      const uint32_t wc = std::min<uint32_t>(kMaxW, kItemCols * (kPanelThreads / mk));
      const uint32_t nchunks = (d.n2 + wc - 1) / wc;
      uint32_t na = 0;
      for (uint32_t ci = 0; ci < nchunks; ++ci) {
        const uint32_t c0 = ci * wc, W = std::min(wc, d.n2 - c0);
        for (uint32_t q0 = 0; q0 < k; q0 += kPanelThreads) {
          TriItem it{};
          it.type = kItemA;
          it.data = vb + (uint64_t)c0 * k + q0;
          it.m = std::min<uint32_t>(kPanelThreads, k - q0);
          it.W = W;
          it.ld = k;
          it.in = d.b2 + c0;
          it.out = (uint32_t)(t_len + (uint64_t)ci * k + q0);
          it.wait = ready_of(d.b2 + c0, d.b2 + c0 + W);
          it.sig0 = lr;
          need.push_back(leaf_need);
          cls.push_back(0);
          its.push_back(it);
          ++na;
        }
      }
      a_count.push_back(na);
      // B items: row tiles of U times the summed t, at most 256 / ceil(k / 32)
      // rows so that a thread owns at most 32 of the k columns
      const uint32_t mb = kPanelThreads / ((k + kItemCols - 1) / kItemCols);
      for (uint32_t r0 = 0; r0 < d.n1; r0 += mb) {
        const uint32_t m = std::min<uint32_t>(mb, d.n1 - r0);
        TriItem it{};
        it.type = kItemB;
        it.data = ub + r0;
        it.m = m;
        it.W = k;
        it.ld = d.n1;
        it.in = (uint32_t)t_len;
        it.out = d.b1 + r0;
        it.wait = lr;
        it.wait2 = na;        // A items of the leaf
        it.aux = k;
        it.parts = nchunks;   // slots summed per rank entry
        need.push_back(rows_of(d.b1 + r0, d.b1 + r0 + m, it.sig0, it.sig1));
        cls.push_back(2);
        its.push_back(it);
      }
      t_len += (uint64_t)nchunks * k;
      if (t_len > 0xffffffffull) return fail(AHM_ERR_UNSUPPORTED, "ahm_trih_upload: too many low-rank slots");
    }
  }
  // queue: for each step, the items needed there (A, D, B), then its S item
  const uint32_t n_off = (uint32_t)its.size();
  std::vector<uint32_t> ord(n_off);
  for (uint32_t i = 0; i < n_off; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
    if (need[a] != need[b]) return need[a] < need[b];
    return cls[a] < cls[b];
  });
  std::vector<TriItem>& q = P.queue;
  q.reserve(n_off + D);
  std::vector<uint32_t>& pend0 = P.pend0;
  pend0.assign(D, 0);
  {
    size_t k = 0;
    for (uint32_t st = 0; st < D; ++st) {
      for (; k < n_off && need[ord[k]] == st; ++k) q.push_back(its[ord[k]]);
      q.push_back(sitems[forward ? st : D - 1 - st]);
    }
    if (k != n_off) return fail(AHM_ERR_INVALID, "ahm_trih_upload: internal: unplaced items");
  }
  for (const TriItem& it : q)
    if (it.type == kItemB || it.type == kItemD)
      for (uint32_t d = it.sig0; d <= it.sig1; ++d) ++pend0[d];
  for (const TriItem& it : P.push)
    for (uint32_t d = it.sig0; d <= it.sig1; ++d) ++pend0[d];

  // every A/B/D tile fits the per-thread register prefetch
  for (const TriItem& it : q)
    if (it.type != kItemS &&
        (it.m == 0 || it.m > (uint32_t)kPanelThreads || it.W > (uint32_t)kItemCols * (kPanelThreads / it.m)))
      return fail(AHM_ERR_INVALID, "ahm_trih_upload: internal: item of %u x %u exceeds the tile limits", it.m, it.W);
  // run the queue once on the host: every wait must hold when reached
  {
    std::vector<uint32_t> solved(D, 0), tdone(a_count.size(), 0), pend(pend0);
    for (const TriItem& it : q) {
      bool ok;
      if (it.type == kItemS)
        ok = pend[it.wait] == 0 && (it.wait2 == kNone || solved[it.wait2]);
      else if (it.type == kItemB)
        ok = tdone[it.wait] >= it.wait2;
      else
        ok = solved[it.wait] != 0;
      if (!ok) return fail(AHM_ERR_INVALID, "ahm_trih_upload: internal: queue order violates a dependency");
      if (it.type == kItemS) {
        solved[it.wait] = 1;
        if (it.in != kNone)
          for (uint32_t d = P.push[it.in].sig0; d <= P.push[it.in].sig1; ++d) --pend[d];
      } else if (it.type == kItemA) {
        ++tdone[it.sig0];
      }
      else
        for (uint32_t d = it.sig0; d <= it.sig1; ++d) --pend[d];
    }
  }
  P.n_diag = D;
  P.n_lr = (uint32_t)a_count.size();
  P.t_len = t_len;
  P.nvals = nvals;
  return AHM_OK;
}

}  // namespace ahm

extern "C" {

int ahm_trih_upload(ahm_ctx ctx, uint32_t n, char uplo, uint32_t n_leaves, const ahm_leaf_desc* leaves,
                    const void* data, ahm_dtype data_dtype, ahm_trih* out) {
  using namespace ahm;
  if (!ctx || !out || (n_leaves && (!leaves || !data)))
    return fail(AHM_ERR_INVALID, "ahm_trih_upload: NULL argument");
  if (uplo != 'L' && uplo != 'U' && uplo != 'l' && uplo != 'u')
    return fail(AHM_ERR_INVALID, "ahm_trih_upload: uplo must be 'L', 'U', 'l' or 'u'");
  if (data_dtype != AHM_F64) return fail(AHM_ERR_UNSUPPORTED, "ahm_trih_upload: leaf values must be AHM_F64");
  *out = nullptr;
  TriPlan P;
  AHM_TRY(build_plan(n, uplo, n_leaves, leaves, (const double*)data, P));
  const std::vector<double>& ar = P.arena;
  const std::vector<uint32_t>& perm = P.perm;
  const std::vector<TriItem>& q = P.queue;
  const std::vector<uint32_t>& pend0 = P.pend0;
  const uint32_t D = P.n_diag;

  std::unique_ptr<ahm_trih_s> h(new ahm_trih_s);
  h->ctx = ctx;
  h->n = n;
  h->mode = uplo == 'L' ? 0 : uplo == 'U' ? 1 : uplo == 'u' ? 2 : 3;
  h->n_diag = D;
  h->n_lr = P.n_lr;
  h->n_items = q.size();
  h->nvals = P.nvals;
  h->arena_vals = ar.size();
  h->t_len = P.t_len;
  AHM_HIP(hipSetDevice(ctx->device));
  AHM_TRY(h->arena.alloc(std::max<size_t>(ar.size(), 1) * sizeof(double)));
  AHM_TRY(h->items.alloc(std::max<size_t>(q.size() + P.push.size(), 1) * sizeof(TriItem)));
  AHM_TRY(h->perm.alloc(std::max<size_t>(perm.size(), 1) * sizeof(uint32_t)));
  AHM_TRY(h->rinv.alloc(std::max<size_t>(P.rinv.size(), 1) * sizeof(double)));
  if (!P.rinv.empty())
    AHM_HIP(hipMemcpy(h->rinv.p, P.rinv.data(), P.rinv.size() * sizeof(double), hipMemcpyHostToDevice));
  AHM_TRY(h->pend0.alloc(std::max<size_t>(D, 1) * sizeof(uint32_t)));

  if (!ar.empty()) AHM_HIP(hipMemcpy(h->arena.p, ar.data(), ar.size() * sizeof(double), hipMemcpyHostToDevice));
  if (!q.empty()) AHM_HIP(hipMemcpy(h->items.p, q.data(), q.size() * sizeof(TriItem), hipMemcpyHostToDevice));
  if (!P.push.empty())
    AHM_HIP(hipMemcpy((TriItem*)h->items.p + q.size(), P.push.data(), P.push.size() * sizeof(TriItem),
                      hipMemcpyHostToDevice));
  if (!perm.empty())
    AHM_HIP(hipMemcpy(h->perm.p, perm.data(), perm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (D) AHM_HIP(hipMemcpy(h->pend0.p, pend0.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice));
  int per_cu = 0, n_cu = 0;
  AHM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trih_kernel<true>, kPanelThreads, 0));
  AHM_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  h->grid = (uint32_t)std::max<uint64_t>(1, (uint64_t)std::max(per_cu, 1) * n_cu);   // resident workgroups
  AHM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trih_kernel<false>, kPanelThreads, 0));
  h->grid_multi = (uint32_t)std::max<uint64_t>(1, (uint64_t)std::max(per_cu, 1) * n_cu);
  *out = h.release();
  return AHM_OK;
}

int ahm_trih_destroy(ahm_trih t) {
  if (!t) return AHM_OK;
  (void)hipSetDevice(t->ctx->device);
  delete t;
  return AHM_OK;
}

int ahm_trih_info(ahm_trih t, uint32_t* n_diag, uint64_t* n_items, uint64_t* nvals) {
  using namespace ahm;
  if (!t) return fail(AHM_ERR_INVALID, "ahm_trih_info: NULL handle");
  if (n_diag) *n_diag = t->n_diag;
  if (n_items) *n_items = t->n_items;
  if (nvals) *nvals = t->nvals;
  return AHM_OK;
}

int ahm_trih_solve(ahm_trih t, void* B, uint32_t ldb, uint32_t p) {
  using namespace ahm;
  if (!t || (!B && p)) return fail(AHM_ERR_INVALID, "ahm_trih_solve: NULL argument");
  if (p > 1 && ldb < t->n) return fail(AHM_ERR_INVALID, "ahm_trih_solve: ldb %u < n %u", ldb, t->n);
  if (p == 0 || t->n == 0) return AHM_OK;
  AHM_HIP(hipSetDevice(t->ctx->device));
  hipStream_t s = t->ctx->stream;
  const uint32_t D = t->n_diag;
  const uint32_t n_zero = 2 + D + t->n_lr;
  const uint32_t cstride = n_zero + D;
  const uint32_t chains = std::min(p, kMaxChains);
  // per-chain counters and slots, grown to the widest call
  if (t->chains < chains) {
    AHM_TRY(t->counters.alloc((size_t)cstride * chains * sizeof(uint32_t)));
    AHM_TRY(t->tpart.alloc(std::max<uint64_t>(t->t_len, 1) * chains * sizeof(double)));
    t->chains = chains;
  }
  uint32_t* cnt = (uint32_t*)t->counters.p;
  const uint64_t step = p > 1 ? ldb : 0;
  for (uint32_t j0 = 0; j0 < p; j0 += chains) {
    const uint32_t c = std::min(chains, p - j0);
    double* x = (double*)B + (uint64_t)j0 * step;
    hipLaunchKernelGGL(trih_reset_kernel, dim3((cstride + 255) / 256, c), dim3(256), 0, s, cnt, cstride, n_zero,
                       (const uint32_t*)t->pend0.p, D);
    // one chain: LDS-staged pushes, one workgroup per CU; several chains:
    // two per CU without the staging, shared out among the chains
    const uint32_t gx =
        std::max<uint32_t>(1, std::min<uint64_t>(t->n_items, (c > 1 ? t->grid_multi : t->grid) / c));
    if (c > 1)
      hipLaunchKernelGGL(trih_kernel<false>, dim3(gx, c), dim3(kPanelThreads), 0, s, (const TriItem*)t->items.p,
                         (uint32_t)t->n_items, (const double*)t->arena.p, (const uint32_t*)t->perm.p,
                         (const double*)t->rinv.p, x, step, (double*)t->tpart.p, std::max<uint64_t>(t->t_len, 1),
                         cnt, cstride, D, t->n_lr, t->mode);
    else
      hipLaunchKernelGGL(trih_kernel<true>, dim3(gx, c), dim3(kPanelThreads), 0, s, (const TriItem*)t->items.p,
                         (uint32_t)t->n_items, (const double*)t->arena.p, (const uint32_t*)t->perm.p,
                         (const double*)t->rinv.p, x, step, (double*)t->tpart.p, std::max<uint64_t>(t->t_len, 1),
                         cnt, cstride, D, t->n_lr, t->mode);
    AHM_HIP(hipGetLastError());
  }
  uint32_t ab = 0;
  AHM_HIP(hipMemcpyAsync(&ab, cnt + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  if (ab) return fail(AHM_ERR_HIP, "ahm_trih_solve: a dependency wait ran out (internal error)");
  return AHM_OK;
}

int ahm_trih_solve_host(ahm_trih t, double* B, uint32_t ldb, uint32_t p) {
  using namespace ahm;
  if (!t || (!B && p)) return fail(AHM_ERR_INVALID, "ahm_trih_solve_host: NULL argument");
  if (p == 0) return AHM_OK;
  if (p > 1 && ldb < t->n) return fail(AHM_ERR_INVALID, "ahm_trih_solve_host: ldb %u < n %u", ldb, t->n);
  AHM_HIP(hipSetDevice(t->ctx->device));
  hipStream_t s = t->ctx->stream;
  DeviceBuf db;
  AHM_TRY(db.alloc((size_t)t->n * p * 8));
  AHM_HIP(hipMemcpy2DAsync(db.p, (size_t)t->n * 8, B, (size_t)(p > 1 ? ldb : t->n) * 8, (size_t)t->n * 8, p,
                           hipMemcpyHostToDevice, s));
  AHM_TRY(ahm_trih_solve(t, db.p, t->n, p));
  AHM_HIP(hipMemcpy2DAsync(B, (size_t)(p > 1 ? ldb : t->n) * 8, db.p, (size_t)t->n * 8, (size_t)t->n * 8, p,
                           hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  return AHM_OK;
}

}  // extern "C"
