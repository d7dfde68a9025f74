// hmat_t.hip -- the transposed H-matvec y += alpha A^T x on the same arena.
//
// Replaces the reference's mltaGeHhVec_ / mltaGeHhGeM_ (H/mltaGeHVec.cpp:
// 61-104) with its leaf products mltaLrMhVec_ (H/mblock.cpp:159-170) and
// mltaGeMhVec (Include/mblock.h:112-115), real case (A^H = A^T).
//
// No second copy of the matrix: every panel column of the forward layout
// (hmat.hip) becomes one dot product of that column with the vector slice of
// its rows, written to its own slot (the column's position in the source
// list).  Four launches per right-hand side:
//   1. dots of the row panels' columns with x  (dense columns -> y, U columns -> t')
//   2. t'[i] = sum of the U-column dots of low-rank block i      (CSR, fixed order)
//   3. dots of the V^T panels' columns with t'                   (-> y)
//   4. y[j] += alpha * sum of the dots that belong to column j   (CSR, fixed order)
// Each output is summed by one thread in a fixed order: deterministic, no
// atomics.  The panels are read once; the slots cost 8 bytes per panel
// column (one per 65-331 values at the benchmark sizes).
//
// Symmetric matrices (AHM_LEAF_HEM diagonal leaves; mltaHeHVec,
// H/mltaGeHVec.cpp:420-446) store the leaves on and above the block diagonal
// only: their product is the forward product plus these four launches with
// the slots of the diagonal leaves left out of step 4 (those blocks are
// stored whole in the arena, so the forward product already applied them).
#include <algorithm>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

namespace {

struct ColTask {
  uint64_t data;   // element offset of the panel piece in the arena
  uint64_t zoff;   // slot (source-list position) of its first column
  uint32_t m, W;   // rows, columns
  uint32_t base;   // index of row 0 in the vector (x row, or t index)
  uint32_t pad;
};

// slots[zoff + c] = sum_r P[r][c] v[base + r].  Lanes form groups of GS
// (32 for m > 128, 16 for m > 64, 8 for m > 4, else the power of two >= m);
// a group takes four columns at a time and its lanes hold rows gl, gl + GS,
// ... (at most 8 passes, m <= 256); the group sums
// are fixed butterflies of shuffles.  Out-of-range rows and columns read a
// clamped, valid address (row 0 / the group's first column) and are weighted
// by zero or not stored, so every load is unconditional and a group's loads
// are in flight together.
template <typename TS, typename TV, int GS, int NQ>
__device__ __forceinline__ void coldot_body(const ColTask& tk, const TS* __restrict__ arena,
                                            const TV* __restrict__ v, TV* __restrict__ slots) {
  constexpr int kPerWave = 64 / GS;
  constexpr int kGroups = 4 * kPerWave;
  const int m = (int)tk.m;
  const int lane = threadIdx.x & 63;
  const int gl = lane & (GS - 1);
  const int grp = (int)(threadIdx.x >> 6) * kPerWave + lane / GS;
  const int nq = (m + GS - 1) / GS;  // row passes, uniform over the workgroup
  TV vr[NQ];
  int roff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = gl + GS * q;
    const bool in = r < m;
    roff[q] = in ? r : 0;
    vr[q] = in ? v[tk.base + r] : (TV)0;
  }
  const TS* P = arena + tk.data;
  // the workgroup takes 4 * kGroups consecutive columns per pass; each
  // group's first lane stores its four sums (no workgroup barrier)
  for (uint32_t base = 0; base < tk.W; base += 4u * kGroups) {
    TV s[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t c = base + (uint32_t)(u * kGroups + grp);
      const TS* Pc = P + (uint64_t)(c < tk.W ? c : base) * m;
      s[u] = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q < nq) s[u] = fma((TV)Pc[roff[q]], vr[q], s[u]);
    }
#pragma unroll
    for (int o = GS / 2; o > 0; o >>= 1)
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += __shfl_xor(s[u], o, 64);
    if (gl == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t c = base + (uint32_t)(u * kGroups + grp);
        if (c < tk.W) slots[tk.zoff + c] = s[u];
      }
  }
}

template <typename TS, typename TV>
__global__ __launch_bounds__(256) void coldot_kernel(const ColTask* __restrict__ tasks,
                                                     const TS* __restrict__ arena,
                                                     const TV* __restrict__ v,
                                                     TV* __restrict__ slots, Gate g) {
  if (gated(g)) return;
  const ColTask tk = tasks[blockIdx.x];
  const uint32_t m = tk.m;
  if (m > 128)
    coldot_body<TS, TV, 32, 8>(tk, arena, v, slots);
  else if (m > 64)
    coldot_body<TS, TV, 16, 8>(tk, arena, v, slots);
  else if (m > 4)
    coldot_body<TS, TV, 8, 8>(tk, arena, v, slots);
  else if (m > 2)
    coldot_body<TS, TV, 4, 1>(tk, arena, v, slots);
  else if (m > 1)
    coldot_body<TS, TV, 2, 1>(tk, arena, v, slots);
  else
    coldot_body<TS, TV, 1, 1>(tk, arena, v, slots);
}

// Symmetric matrices, one right-hand side: a phase-B task (row panel chunk
// of m <= 256 rows, W columns) read once for both directions.  The lanes of
// a group of GS hold rows gl, gl + GS, ... (as coldot_body); a group takes
// four columns per pass; per value one fma into the column dot with x_rows
// (slots, summed over the group by shuffles) and one into the lane's row
// sums with z (the column's source value, staged in LDS per pass).  The row
// sums of the kGroups groups meet in LDS in group order.
struct SymTask {
  uint64_t data;   // arena offset of the chunk
  uint64_t zoff;   // source-list position of its first column (= its first slot)
  uint32_t m, W;
  uint32_t row0;   // first row of the chunk's panel (rows of x for the column dots)
  uint32_t out;    // y row (partial == 0) or partial-sum offset (partial == 1)
  uint32_t partial, pad;
};

template <typename TS, typename TV, int GS, int NQ>
__device__ __forceinline__ void sym_body(const SymTask& tk, const uint32_t* __restrict__ zsrc,
                                         const TS* __restrict__ arena, const TV* __restrict__ x,
                                         const TV* __restrict__ t, TV* __restrict__ y, TV* __restrict__ part,
                                         TV* __restrict__ slots, TV alpha, TV* zs, TV* red) {
  constexpr int kPerWave = 64 / GS;
  constexpr int kGroups = 4 * kPerWave;
  constexpr uint32_t kZ = 1024;   // columns of z staged at a time
  const int m = (int)tk.m;
  const int lane = threadIdx.x & 63;
  const int gl = lane & (GS - 1);
  const int grp = (int)(threadIdx.x >> 6) * kPerWave + lane / GS;
  const int nq = (m + GS - 1) / GS;
  TV xr[NQ], acc[NQ];
  int roff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = gl + GS * q;
    const bool in = r < m;
    roff[q] = in ? r : 0;
    xr[q] = in ? x[tk.row0 + r] : (TV)0;
    acc[q] = 0;
  }
  for (uint32_t c0 = 0; c0 < tk.W; c0 += kZ) {
    const uint32_t cw = min(kZ, tk.W - c0);
    const TS* P = arena + tk.data + (uint64_t)c0 * m;
    __syncthreads();   // the previous chunk's z is consumed
    {
      // source entries first, then their values, four of each in flight
      uint32_t src[kZ / 256];
#pragma unroll
      for (int k = 0; k < (int)(kZ / 256); ++k) {
        const uint32_t i = threadIdx.x + 256u * k;
        src[k] = zsrc[tk.zoff + c0 + (i < cw ? i : 0)];
      }
#pragma unroll
      for (int k = 0; k < (int)(kZ / 256); ++k) {
        const uint32_t i = threadIdx.x + 256u * k;
        const TV v = (src[k] & kFromT) ? t[src[k] & ~kFromT] : x[src[k]];
        zs[i] = i < cw ? v : (TV)0;
      }
    }
    __syncthreads();
    for (uint32_t base = 0; base < cw; base += 4u * kGroups) {
      TV sd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t c = base + (uint32_t)(u * kGroups + grp);
        const bool in = c < cw;
        const TS* Pc = P + (uint64_t)(in ? c : base) * m;
        const TV zc = zs[in ? c : 0];
        TV s = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (q < nq) {
            const TV v = (TV)Pc[roff[q]];
            s = fma(v, xr[q], s);
            acc[q] = fma(v, in ? zc : (TV)0, acc[q]);
          }
        sd[u] = s;
      }
#pragma unroll
      for (int o = GS / 2; o > 0; o >>= 1)
#pragma unroll
        for (int u = 0; u < 4; ++u) sd[u] += __shfl_xor(sd[u], o, 64);
      if (gl == 0)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t c = base + (uint32_t)(u * kGroups + grp);
          if (c < cw) slots[tk.zoff + c0 + c] = sd[u];
        }
    }
  }
  // row sums: group g's partial sums of row r at red[g * m + r], summed in group order
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = gl + GS * q;
    if (q < nq && r < m) red[grp * m + r] = acc[q];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < m; r += 256) {
    TV sum = 0;
    for (int g2 = 0; g2 < kGroups; ++g2) sum += red[g2 * m + r];
    if (tk.partial)
      part[tk.out + r] = sum;
    else
      y[tk.out + r] += alpha * sum;
  }
}

template <typename TS, typename TV>
__global__ __launch_bounds__(256) void sym_panel_kernel(const SymTask* __restrict__ tasks,
                                                        const uint32_t* __restrict__ zsrc,
                                                        const TS* __restrict__ arena, const TV* __restrict__ x,
                                                        const TV* __restrict__ t, TV* __restrict__ y,
                                                        TV* __restrict__ part, TV* __restrict__ slots, TV alpha,
                                                        Gate g) {
  // 1024 staged z values; kGroups * m <= 2048 row partials
  __shared__ TV zs[1024];
  __shared__ TV red[2048];
  if (gated(g)) return;
  const SymTask tk = tasks[blockIdx.x];
  const uint32_t m = tk.m;
  if (m > 128)
    sym_body<TS, TV, 32, 8>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
  else if (m > 64)
    sym_body<TS, TV, 16, 8>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
  else if (m > 4)
    sym_body<TS, TV, 8, 8>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
  else if (m > 2)
    sym_body<TS, TV, 4, 1>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
  else if (m > 1)
    sym_body<TS, TV, 2, 1>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
  else
    sym_body<TS, TV, 1, 1>(tk, zsrc, arena, x, t, y, part, slots, alpha, zs, red);
}

// out[i] (=|+= alpha *) sum_{j in [off[i], off[i+1])} slots[pos[j]]
template <typename TV, bool ASSIGN>
__global__ __launch_bounds__(256) void csr_sum_kernel(const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ pos,
                                                      const TV* __restrict__ slots, uint32_t n,
                                                      TV* __restrict__ out, TV alpha, Gate g) {
  if (gated(g)) return;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  TV s = 0;
  // eight positions, then their eight slots, in flight together; the sum
  // keeps the slot order
  uint32_t j = off[i];
  const uint32_t e = off[i + 1];
  for (; j + 8 <= e; j += 8) {
    uint32_t q[8];
    TV v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = pos[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = slots[q[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; j < e; ++j) s += slots[pos[j]];
  if (ASSIGN)
    out[i] = s;
  else
    out[i] += alpha * s;
}

// ---- several right-hand sides (p >= 2): 32 at a time on MFMA ----------
//
// The same four steps with 32-wide rows: x, t' and the slots hold the 32
// right-hand sides of one row / column contiguously (slot q at 32 q).
// A task's column dots are the small product P^T V (W x m times m x 32) on
// v_mfma_{f64,f32}_16x16x4: A = 16 panel columns x 4 rows, B = the same 4
// rows of v (staged in LDS for the task's m <= 256 rows), C = 16 columns x
// 16 right-hand sides.  In k-step s of a 16-row block, lane group g
// (lanes 16g..16g+15) takes row 4g + s, so a lane's four panel values of the
// block are four consecutive rows of its column; A and B use the same row
// order, so each dot is a sum over the task's rows (in a fixed order).
constexpr int kTmRhs = 32;
constexpr int kTmThreads = 512;
constexpr uint64_t kTmChunkValues = 4 * kChunkValues;

template <typename T> struct TmAcc;
template <> struct TmAcc<double> {
  typedef double __attribute__((ext_vector_type(4))) type;
  static __device__ __forceinline__ type mma(double a, double b, type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return lane / 16 + 4 * r; }
};
template <> struct TmAcc<float> {
  typedef float __attribute__((ext_vector_type(4))) type;
  static __device__ __forceinline__ type mma(float a, float b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane / 16) + r; }
};

// a lane's four consecutive panel values of a 16-row block, in two 16-byte
// loads (8-byte aligned; the target reads unaligned vector loads)
typedef double tm_d2 __attribute__((ext_vector_type(2), aligned(8)));
typedef float tm_f4 __attribute__((ext_vector_type(4), aligned(4)));
template <typename TS, typename TV>
__device__ __forceinline__ void tm_load4(TV (&A)[4], const TS* p) {
  if constexpr (sizeof(TS) == 8) {
    const tm_d2 lo = *(const tm_d2*)p, hi = *(const tm_d2*)(p + 2);
    A[0] = (TV)lo.x;
    A[1] = (TV)lo.y;
    A[2] = (TV)hi.x;
    A[3] = (TV)hi.y;
  } else {
    const tm_f4 v = *(const tm_f4*)p;
    A[0] = (TV)v.x;
    A[1] = (TV)v.y;
    A[2] = (TV)v.z;
    A[3] = (TV)v.w;
  }
}

// slots[(zoff + c) * 32 + j] = sum_r P[r][c] v[(base + r) * 32 + j] for the
// task's columns c < W; eight wavefronts take 16-column tiles in turn.  Rows
// past m (the next column's values or the arena's slack, at most 31 values
// past the task) are weighted by zero; columns past W read the last column
// and are not stored.
template <typename TS, typename TV>
__global__ __launch_bounds__(kTmThreads) void coldot_mm_kernel(const ColTask* __restrict__ tasks,
                                                               const TS* __restrict__ arena,
                                                               const TV* __restrict__ v,
                                                               TV* __restrict__ slots) {
  typedef typename TmAcc<TV>::type acc_t;
  __shared__ TV vs[256][kTmRhs + 1];
  const ColTask tk = tasks[blockIdx.x];
  const int m = (int)tk.m;
  const int nrb = (m + 15) / 16;
  for (int e = threadIdx.x; e < nrb * 16 * kTmRhs; e += kTmThreads) {
    const int r = e / kTmRhs, j = e % kTmRhs;
    vs[r][j] = r < m ? v[((uint64_t)tk.base + r) * kTmRhs + j] : (TV)0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int li = lane & 15, lg = lane >> 4;
  const uint32_t nct = (tk.W + 15) / 16;
  for (uint32_t ct = (uint32_t)w; ct < nct; ct += kTmThreads / 64) {
    const uint32_t c = ct * 16 + li;
    const TS* Pc = arena + tk.data + (uint64_t)(c < tk.W ? c : tk.W - 1) * (uint32_t)m + 4 * lg;
    acc_t a0 = acc_t{0, 0, 0, 0}, a1 = acc_t{0, 0, 0, 0};
    // two 16-row blocks per step, the next two in flight
    TV A[2][4], An[2][4];
    tm_load4<TS, TV>(A[0], Pc);
    tm_load4<TS, TV>(A[1], Pc + 16);
    for (int rb = 0; rb < nrb; rb += 2) {
      if (rb + 2 < nrb) {
        tm_load4<TS, TV>(An[0], Pc + (rb + 2) * 16);
        tm_load4<TS, TV>(An[1], Pc + (rb + 3) * 16);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (rb + h >= nrb) break;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = (rb + h) * 16 + 4 * lg + s;
          const TV a = r < m ? A[h][s] : (TV)0;
          a0 = TmAcc<TV>::mma(a, vs[r][li], a0);
          a1 = TmAcc<TV>::mma(a, vs[r][16 + li], a1);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int s = 0; s < 4; ++s) A[h][s] = An[h][s];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t cc = ct * 16 + (uint32_t)TmAcc<TV>::row(lane, q);
      if (cc < tk.W) {
        TV* o = slots + (tk.zoff + cc) * kTmRhs;
        o[li] = a0[q];
        o[16 + li] = a1[q];
      }
    }
  }
}

// sum of slots[pos[k] * 32 + j] over k in [k0, k1), in k order; eight
// positions, then their eight slot values, in flight together
template <typename TV>
__device__ __forceinline__ TV tm_slot_sum(uint32_t k0, uint32_t k1, const uint32_t* __restrict__ pos,
                                          const TV* __restrict__ slots, uint32_t j) {
  TV s = 0;
  uint32_t k = k0;
  for (; k + 8 <= k1; k += 8) {
    uint32_t q[8];
    TV v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = pos[k + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = slots[(uint64_t)q[u] * kTmRhs + j];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; k < k1; ++k) s += slots[(uint64_t)pos[k] * kTmRhs + j];
  return s;
}

// t'[i * 32 + j] = sum_{k in [off[i], off[i+1])} slots[pos[k] * 32 + j]
template <typename TV>
__global__ __launch_bounds__(256) void csr_mm_t_kernel(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ pos,
                                                       const TV* __restrict__ slots, uint32_t n,
                                                       TV* __restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint32_t i = (uint32_t)(g / kTmRhs), j = (uint32_t)(g % kTmRhs);
  if (i >= n) return;
  out[(uint64_t)i * kTmRhs + j] = tm_slot_sum(off[i], off[i + 1], pos, slots, j);
}

// Y[j * ldy + i] += alpha * sum_{k in [off[i], off[i+1])} slots[pos[k] * 32 + j]
// for the pj right-hand sides of the group: 64 destinations per workgroup,
// summed with the right-hand sides across lanes, stored through LDS with the
// destinations across lanes
template <typename TV>
__global__ __launch_bounds__(256) void csr_mm_y_kernel(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ pos,
                                                       const TV* __restrict__ slots, uint32_t n,
                                                       TV* __restrict__ Y, uint32_t ldy, uint32_t pj, TV alpha) {
  __shared__ TV sum[kTmRhs][64 + 1];
  const uint32_t i0 = blockIdx.x * 64u;
  const uint32_t j = threadIdx.x % kTmRhs;
  for (uint32_t il = threadIdx.x / kTmRhs; il < 64; il += 256 / kTmRhs) {
    const uint32_t i = i0 + il;
    sum[j][il] = i < n ? tm_slot_sum(off[i], off[i + 1], pos, slots, j) : (TV)0;
  }
  __syncthreads();
  const uint32_t il = threadIdx.x & 63;
  for (uint32_t jj = threadIdx.x >> 6; jj < pj; jj += 4)
    if (i0 + il < n) Y[(uint64_t)jj * ldy + i0 + il] += alpha * sum[jj][il];
}

// rows [0, n) of the pj <= 32 columns X[j * ldx + i] -> xi[i * 32 + j] (zero
// for j >= pj), 64 rows per workgroup through LDS
template <typename TV>
__global__ __launch_bounds__(256) void tm_interleave_kernel(const TV* __restrict__ X, uint32_t ldx, uint32_t n,
                                                            uint32_t pj, TV* __restrict__ xi) {
  __shared__ TV tile[kTmRhs][64 + 1];
  const uint32_t i0 = blockIdx.x * 64u;
  for (uint32_t e = threadIdx.x; e < kTmRhs * 64; e += 256) {
    const uint32_t jj = e >> 6, il = e & 63;
    tile[jj][il] = (jj < pj && i0 + il < n) ? X[(uint64_t)jj * ldx + i0 + il] : (TV)0;
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < kTmRhs * 64; e += 256) {
    const uint32_t il = e / kTmRhs, jj = e % kTmRhs;
    if (i0 + il < n) xi[((uint64_t)i0 + il) * kTmRhs + jj] = tile[jj][il];
  }
}

// column pieces of about kChunkValues values per task
void col_tasks(const std::vector<GemvTask>& panels, std::vector<ColTask>& out,
               uint64_t chunk_values = kChunkValues) {
  for (const GemvTask& p : panels) {
    const uint32_t cols = (uint32_t)std::max<uint64_t>(1, chunk_values / std::max(1u, p.m));
    for (uint32_t c0 = 0; c0 < p.W; c0 += cols) {
      ColTask t{};
      t.data = p.data + (uint64_t)c0 * p.m;
      t.zoff = p.zoff + c0;
      t.m = p.m;
      t.W = std::min(cols, p.W - c0);
      t.base = p.out;
      out.push_back(t);
    }
  }
  std::stable_sort(out.begin(), out.end(), [](const ColTask& a, const ColTask& b) {
    return (uint64_t)a.m * a.W > (uint64_t)b.m * b.W;
  });
}

// CSR of the slots by destination, slots in increasing position
void csr(const std::vector<uint32_t>& dest, const std::vector<uint32_t>& slot, uint32_t n,
         std::vector<uint32_t>& off, std::vector<uint32_t>& pos) {
  off.assign((size_t)n + 1, 0);
  for (uint32_t d : dest) off[d + 1]++;
  for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
  pos.resize(dest.size());
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  for (size_t k = 0; k < dest.size(); ++k) pos[fill[dest[k]]++] = slot[k];
}

template <typename T>
int upload(DeviceBuf& b, const std::vector<T>& v, hipStream_t s) {
  AHM_TRY(b.alloc(std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) AHM_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return AHM_OK;
}

int build_transposed(ahm_hmat h) {
  auto& T = h->tr;
  if (T.built) return AHM_OK;
  if (h->n_zsrc >= 0xFFFFFFFFull) return fail(AHM_ERR_UNSUPPORTED, "transposed product: too many panel columns");
  std::vector<ColTask> ta, tb, ta_mm, tb_mm;
  col_tasks(h->h_panels_a, ta);
  col_tasks(h->h_panels, tb);
  // multi-RHS column dots: longer column pieces (a task stages its rows of
  // the vector in LDS once, and does 32 times the arithmetic per value)
  col_tasks(h->h_panels_a, ta_mm, kTmChunkValues);
  col_tasks(h->h_panels, tb_mm, kTmChunkValues);
  // the multi-RHS column dots stage a task's rows in LDS (at most 256)
  for (const auto* v : {&ta, &tb, &ta_mm, &tb_mm})
    for (const ColTask& t : *v)
      if (t.m > 256) return fail(AHM_ERR_UNSUPPORTED, "transposed product: panel of %u rows", t.m);
  // symmetric storage: the rows of each diagonal leaf; a row panel lies in
  // one of them, and its dense columns inside that range are the diagonal
  // leaf's own (blocks of a block tree overlap the diagonal only there)
  std::vector<uint8_t> skip;
  if (h->sym) {
    std::vector<uint32_t> dlo(h->n_rows, 0), dhi(h->n_rows, 0);
    for (const ahm_leaf_desc& d : h->leaves)
      if (d.flags & AHM_LEAF_HEM)
        for (uint32_t r = d.b1; r < d.b1 + d.n1; ++r) {
          dlo[r] = d.b1;
          dhi[r] = d.b1 + d.n1;
        }
    skip.assign(h->h_zsrc.size(), 0);
    for (const GemvTask& p : h->h_panels) {
      const uint32_t lo = dlo[p.out], hi = dhi[p.out];
      for (uint32_t c = 0; c < p.W; ++c) {
        const uint32_t src = h->h_zsrc[p.zoff + c];
        if (!(src & kFromT) && src >= lo && src < hi) skip[p.zoff + c] = 1;
      }
    }
  }
  std::vector<uint32_t> dt, st, dy, sy;
  for (uint32_t q = 0; q < (uint32_t)h->h_zsrc.size(); ++q) {
    const uint32_t s = h->h_zsrc[q];
    if (s & kFromT) {
      dt.push_back(s & ~kFromT);
      st.push_back(q);
    } else if (skip.empty() || !skip[q]) {
      dy.push_back(s);
      sy.push_back(q);
    }
  }
  std::vector<uint32_t> off_t, pos_t, off_y, pos_y;
  csr(dt, st, h->t_len, off_t, pos_t);
  csr(dy, sy, h->n_cols, off_y, pos_y);
  hipStream_t s = h->ctx->stream;
  AHM_TRY(upload(T.tasks_a, ta, s));
  AHM_TRY(upload(T.tasks_b, tb, s));
  AHM_TRY(upload(T.tasks_a_mm, ta_mm, s));
  AHM_TRY(upload(T.tasks_b_mm, tb_mm, s));
  AHM_TRY(upload(T.off_t, off_t, s));
  AHM_TRY(upload(T.pos_t, pos_t, s));
  AHM_TRY(upload(T.off_y, off_y, s));
  AHM_TRY(upload(T.pos_y, pos_y, s));
  AHM_HIP(hipStreamSynchronize(s));
  T.n_a = (uint32_t)ta.size();
  T.n_b = (uint32_t)tb.size();
  T.n_a_mm = (uint32_t)ta_mm.size();
  T.n_b_mm = (uint32_t)tb_mm.size();
  T.built = true;
  return AHM_OK;
}

template <typename TS, typename TV>
int launch_t(ahm_hmat h, TV alpha, const TV* x, TV* y, Gate g = Gate{}) {
  auto& T = h->tr;
  hipStream_t s = h->ctx->stream;
  TV* slots = (TV*)T.contrib.p;
  TV* tp = (TV*)T.tprime.p;
  const TS* arena = (const TS*)h->arena.p;
  if (T.n_b)
    hipLaunchKernelGGL((coldot_kernel<TS, TV>), dim3(T.n_b), dim3(256), 0, s,
                       (const ColTask*)T.tasks_b.p, arena, x, slots, g);
  if (h->t_len)
    hipLaunchKernelGGL((csr_sum_kernel<TV, true>), dim3((h->t_len + 255) / 256), dim3(256), 0, s,
                       (const uint32_t*)T.off_t.p, (const uint32_t*)T.pos_t.p, (const TV*)slots,
                       h->t_len, tp, (TV)1, g);
  if (T.n_a)
    hipLaunchKernelGGL((coldot_kernel<TS, TV>), dim3(T.n_a), dim3(256), 0, s,
                       (const ColTask*)T.tasks_a.p, arena, (const TV*)tp, slots, g);
  if (h->n_cols)
    hipLaunchKernelGGL((csr_sum_kernel<TV, false>), dim3((h->n_cols + 255) / 256), dim3(256), 0, s,
                       (const uint32_t*)T.off_y.p, (const uint32_t*)T.pos_y.p, (const TV*)slots,
                       h->n_cols, y, alpha, g);
  AHM_HIP(hipGetLastError());
  return AHM_OK;
}

// y += alpha A^T X (or the transposed part of a symmetric product) for
// p >= 2 right-hand sides, 32 at a time; after ensure_t_buffers(h, es, 32)
template <typename TS, typename TV>
int launch_t_mm(ahm_hmat h, TV alpha, const TV* X, uint32_t ldx, TV* Y, uint32_t ldy, uint32_t p) {
  auto& T = h->tr;
  hipStream_t s = h->ctx->stream;
  TV* slots = (TV*)T.contrib.p;
  TV* tp = (TV*)T.tprime.p;
  TV* xi = (TV*)T.xi.p;
  const TS* arena = (const TS*)h->arena.p;
  for (uint32_t g0 = 0; g0 < p; g0 += kTmRhs) {
    const uint32_t pj = std::min<uint32_t>(kTmRhs, p - g0);
    if (h->n_rows)
      hipLaunchKernelGGL((tm_interleave_kernel<TV>), dim3((h->n_rows + 63) / 64), dim3(256), 0, s,
                         X + (size_t)g0 * ldx, ldx, h->n_rows, pj, xi);
    if (T.n_b_mm)
      hipLaunchKernelGGL((coldot_mm_kernel<TS, TV>), dim3(T.n_b_mm), dim3(kTmThreads), 0, s,
                         (const ColTask*)T.tasks_b_mm.p, arena, (const TV*)xi, slots);
    if (h->t_len)
      hipLaunchKernelGGL((csr_mm_t_kernel<TV>), dim3((uint32_t)(((uint64_t)h->t_len * kTmRhs + 255) / 256)),
                         dim3(256), 0, s, (const uint32_t*)T.off_t.p, (const uint32_t*)T.pos_t.p,
                         (const TV*)slots, h->t_len, tp);
    if (T.n_a_mm)
      hipLaunchKernelGGL((coldot_mm_kernel<TS, TV>), dim3(T.n_a_mm), dim3(kTmThreads), 0, s,
                         (const ColTask*)T.tasks_a_mm.p, arena, (const TV*)tp, slots);
    if (h->n_cols)
      hipLaunchKernelGGL((csr_mm_y_kernel<TV>), dim3((h->n_cols + 63) / 64), dim3(256), 0, s,
                         (const uint32_t*)T.off_y.p, (const uint32_t*)T.pos_y.p, (const TV*)slots, h->n_cols,
                         Y + (size_t)g0 * ldy, ldy, pj, alpha);
  }
  AHM_HIP(hipGetLastError());
  return AHM_OK;
}

// slots and t' sized for vectors of es bytes, width values per slot / row
// (1, or 32 for several right-hand sides, which also sizes the interleaved
// copy of X); allocates; before any capture
int ensure_t_buffers(ahm_hmat h, size_t es, size_t width = 1) {
  auto& T = h->tr;
  const size_t ew = es * width;
  if (h->n_zsrc * ew > T.contrib_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(T.contrib.alloc(std::max<size_t>(1, h->n_zsrc) * ew));
    T.contrib_cap = h->n_zsrc * ew;
  }
  if ((size_t)h->t_len * ew > T.tprime_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(T.tprime.alloc(std::max<size_t>(1, h->t_len) * ew));
    T.tprime_cap = (size_t)h->t_len * ew;
  }
  if (width > 1 && (size_t)h->n_rows * ew > T.xi_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(T.xi.alloc(std::max<size_t>(1, h->n_rows) * ew));
    T.xi_cap = (size_t)h->n_rows * ew;
  }
  return AHM_OK;
}

}  // namespace

// the handle's phase-B tasks as SymTasks: a chunk's panel is the row panel
// whose arena range holds the chunk's values
int build_sym_tasks(ahm_hmat h) {
  auto& T = h->tr;
  if (T.sym_built) return AHM_OK;
  std::vector<std::pair<uint64_t, uint32_t>> start;   // (arena offset, first row) of every panel
  start.reserve(h->h_panels.size());
  for (const GemvTask& p : h->h_panels) start.emplace_back(p.data, p.out);
  std::sort(start.begin(), start.end());
  std::vector<SymTask> st(h->h_tasks_b.size());
  for (size_t i = 0; i < st.size(); ++i) {
    const GemvTask& b = h->h_tasks_b[i];
    if (b.m > 256) return fail(AHM_ERR_UNSUPPORTED, "symmetric product: row panel of %u rows", b.m);
    auto it = std::upper_bound(start.begin(), start.end(), std::make_pair(b.data, 0xFFFFFFFFu));
    if (it == start.begin()) return fail(AHM_ERR_INVALID, "symmetric product: chunk outside the panels");
    --it;
    SymTask& s = st[i];
    s.data = b.data;
    s.zoff = b.zoff;
    s.m = b.m;
    s.W = b.W;
    s.row0 = it->second;
    s.out = b.out;
    s.partial = b.partial;
    s.pad = 0;
    if (!b.partial && b.out != s.row0) return fail(AHM_ERR_INVALID, "symmetric product: panel rows disagree");
  }
  AHM_TRY(upload(T.sym_tasks, st, h->ctx->stream));
  AHM_HIP(hipStreamSynchronize(h->ctx->stream));
  T.n_sym = (uint32_t)st.size();
  T.sym_built = true;
  return AHM_OK;
}

int sym_prepare(ahm_hmat h, ahm_dtype vec, uint32_t p) {
  if (!h->sym) return AHM_OK;
  AHM_HIP(hipSetDevice(h->ctx->device));
  AHM_TRY(build_transposed(h));
  AHM_TRY(build_sym_tasks(h));
  return ensure_t_buffers(h, dtype_size(vec), p > 1 ? kTmRhs : 1);
}

int sym_phase_b(ahm_hmat h, double alpha, const void* x, const void* t, void* y, void* part, ahm_dtype vec,
                Gate g) {
  auto& T = h->tr;
  const size_t es = dtype_size(vec);
  if (!T.built || !T.sym_built || h->n_zsrc * es > T.contrib_cap)
    return fail(AHM_ERR_INVALID, "symmetric product: not prepared");
  if (!T.n_sym) return AHM_OK;
  hipStream_t s = h->ctx->stream;
  if (vec == AHM_F64)
    hipLaunchKernelGGL((sym_panel_kernel<double, double>), dim3(T.n_sym), dim3(256), 0, s,
                       (const SymTask*)T.sym_tasks.p, (const uint32_t*)h->zsrc.p, (const double*)h->arena.p,
                       (const double*)x, (const double*)t, (double*)y, (double*)part, (double*)T.contrib.p,
                       alpha, g);
  else
    hipLaunchKernelGGL((sym_panel_kernel<float, float>), dim3(T.n_sym), dim3(256), 0, s,
                       (const SymTask*)T.sym_tasks.p, (const uint32_t*)h->zsrc.p, (const float*)h->arena.p,
                       (const float*)x, (const float*)t, (float*)y, (float*)part, (float*)T.contrib.p,
                       (float)alpha, g);
  AHM_HIP(hipGetLastError());
  return AHM_OK;
}

// the transposed part after the slots of the phase-B columns are filled:
// t' = sum of U-column dots, the V^T panels' column dots with t', and
// y += alpha * (dense-column and V-column slot sums)
int sym_rest(ahm_hmat h, double alpha, void* y, ahm_dtype vec, Gate g) {
  auto& T = h->tr;
  hipStream_t s = h->ctx->stream;
  auto run = [&](auto tag) {
    using TV = decltype(tag);
    TV* slots = (TV*)T.contrib.p;
    TV* tp = (TV*)T.tprime.p;
    if (h->t_len)
      hipLaunchKernelGGL((csr_sum_kernel<TV, true>), dim3((h->t_len + 255) / 256), dim3(256), 0, s,
                         (const uint32_t*)T.off_t.p, (const uint32_t*)T.pos_t.p, (const TV*)slots, h->t_len, tp,
                         (TV)1, g);
    if (T.n_a)
      hipLaunchKernelGGL((coldot_kernel<TV, TV>), dim3(T.n_a), dim3(256), 0, s, (const ColTask*)T.tasks_a.p,
                         (const TV*)h->arena.p, (const TV*)tp, slots, g);
    if (h->n_cols)
      hipLaunchKernelGGL((csr_sum_kernel<TV, false>), dim3((h->n_cols + 255) / 256), dim3(256), 0, s,
                         (const uint32_t*)T.off_y.p, (const uint32_t*)T.pos_y.p, (const TV*)slots, h->n_cols,
                         (TV*)y, (TV)alpha, g);
  };
  if (vec == AHM_F64)
    run(double{});
  else
    run(float{});
  AHM_HIP(hipGetLastError());
  return AHM_OK;
}

// the transposed part of a symmetric product: y += alpha S^T x over the
// leaves off the diagonal, column by column of X / Y; allocates nothing
// after sym_prepare
int sym_transposed(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y, uint32_t ldy, uint32_t p,
                   ahm_dtype vec, Gate g) {
  if (!h->sym) return AHM_OK;
  const size_t es = dtype_size(vec);
  const size_t ew = es * (p > 1 ? kTmRhs : 1);
  if (!h->tr.built || h->n_zsrc * ew > h->tr.contrib_cap || (size_t)h->t_len * ew > h->tr.tprime_cap ||
      (p > 1 && (size_t)h->n_rows * ew > h->tr.xi_cap))
    return fail(AHM_ERR_INVALID, "symmetric product: transposed part not prepared");
  if (p > 1) {
    if (vec == AHM_F64)
      return launch_t_mm<double, double>(h, alpha, (const double*)X, ldx, (double*)Y, ldy, p);
    return launch_t_mm<float, float>(h, (float)alpha, (const float*)X, ldx, (float*)Y, ldy, p);
  }
  if (vec == AHM_F64)
    return launch_t<double, double>(h, alpha, (const double*)X, (double*)Y, g);
  return launch_t<float, float>(h, (float)alpha, (const float*)X, (float*)Y, g);
}

}  // namespace ahm

extern "C" {

int ahm_hmat_matvec_t(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y, uint32_t ldy,
                      uint32_t p, ahm_dtype vec) {
  using namespace ahm;
  if (!h || !X || !Y) return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_t: NULL argument");
  if (p == 0) return AHM_OK;
  if (p > 1 && (ldx < h->n_rows || ldy < h->n_cols))
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_t: ldx=%u ldy=%u for %u x %u", ldx, ldy, h->n_rows,
                h->n_cols);
  const bool f64 = h->storage == AHM_F64 && vec == AHM_F64;
  const bool f32 = (h->storage == AHM_F32 || h->storage == AHM_BF16) && vec == AHM_F32;
  if (!f64 && !f32)
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_matvec_t: storage %d with vectors %d", h->storage, vec);
  // a symmetric matrix is its own transpose
  if (h->sym) return ahm_hmat_matvec(h, alpha, X, ldx, Y, ldy, p, vec);
  AHM_HIP(hipSetDevice(h->ctx->device));
  AHM_TRY(build_transposed(h));
  const size_t es = f64 ? 8 : 4;
  AHM_TRY(ensure_t_buffers(h, es, p > 1 ? kTmRhs : 1));
  if (p > 1) {
    if (f64) return launch_t_mm<double, double>(h, alpha, (const double*)X, ldx, (double*)Y, ldy, p);
    return launch_t_mm<float, float>(h, (float)alpha, (const float*)X, ldx, (float*)Y, ldy, p);
  }
  if (f64) return launch_t<double, double>(h, alpha, (const double*)X, (double*)Y);
  return launch_t<float, float>(h, (float)alpha, (const float*)X, (float*)Y);
}

int ahm_hmat_matvec_t_host(ahm_hmat h, double alpha, const double* X, uint32_t ldx, double* Y,
                           uint32_t ldy, uint32_t p) {
  using namespace ahm;
  if (!h || !X || !Y) return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_t_host: NULL argument");
  if (h->storage != AHM_F64) return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_matvec_t_host: needs an F64 matrix");
  if (p == 0) return AHM_OK;
  if (p > 1 && (ldx < h->n_rows || ldy < h->n_cols))
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_t_host: bad leading dimensions");
  AHM_HIP(hipSetDevice(h->ctx->device));
  hipStream_t s = h->ctx->stream;
  DeviceBuf dx, dy;
  AHM_TRY(dx.alloc((size_t)h->n_rows * p * 8));
  AHM_TRY(dy.alloc((size_t)h->n_cols * p * 8));
  AHM_HIP(hipMemcpy2DAsync(dx.p, (size_t)h->n_rows * 8, X, (size_t)ldx * 8, (size_t)h->n_rows * 8, p,
                           hipMemcpyHostToDevice, s));
  AHM_HIP(hipMemcpy2DAsync(dy.p, (size_t)h->n_cols * 8, Y, (size_t)ldy * 8, (size_t)h->n_cols * 8, p,
                           hipMemcpyHostToDevice, s));
  AHM_TRY(ahm_hmat_matvec_t(h, alpha, dx.p, h->n_rows, dy.p, h->n_cols, p, AHM_F64));
  AHM_HIP(hipMemcpy2DAsync(Y, (size_t)ldy * 8, dy.p, (size_t)h->n_cols * 8, (size_t)h->n_cols * 8, p,
                           hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  return AHM_OK;
}

}  // extern "C"
