// hmat.hip -- device layout of an H-matrix and the H-matvec kernels.
//
// Replaces the reference's recursive mltaGeHVec_ (H/mltaGeHVec.cpp:16-35) and
// its per-leaf BLAS-2 products (mltaLrMVec_ H/mblock.cpp:125-138, mltaGeMVec
// Include/mblock.h:98-102) by two batched launches over a re-laid-out arena:
//
//   phase A  t = V^T x for every low-rank leaf.  Leaves sharing a column
//            cluster (b2, n2) are stacked into one V^T panel (sum k rows x n2
//            columns, column-major), so x[b2:b2+n2) is staged once per panel.
//   phase B  y += alpha * [D pieces | U pieces] z per row segment.  Rows are
//            cut at every leaf boundary (the leaf clusters of the row tree);
//            one panel holds, for its m rows, the matching rows of every dense
//            block and every U that covers them, so each output row is
//            written by exactly one workgroup: deterministic, no atomics.
//
// Panels of either phase wider than about kChunkValues values are cut into
// column chunks (one workgroup each) whose row sums meet in a fixed-order
// reduction, so a 32K-column V^T panel does not leave one CU streaming it
// alone at the end of the launch.  Every stored value is read exactly once
// per product; x and t are gathered into LDS per panel through a per-column
// source list.
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>

#include <hip/hip_cooperative_groups.h>

#include "ahm_internal.h"
#include "ahm_aca.h"

namespace ahm {

// ------------------------------------------------------------------------
// kernels

// LDS tile of the transposed copy: rb rows x cols columns (cols <= 256, the
// panel height) with one value of padding per column
constexpr uint32_t kRelayoutTile = 4096;

template <typename TD, typename TS>
__global__ __launch_bounds__(256) void relayout_kernel(const CopyJob* __restrict__ jobs,
                                                       const TS* __restrict__ src,
                                                       TD* __restrict__ dst) {
  const CopyJob j = jobs[blockIdx.x];
  if (j.rows == 0) return;
  if (j.transpose == 2) {
    // rows pad .. pad+rows-1 of a packed symmetric block: A(i, k) with
    // i <= k at i + k(k+1)/2, else at k + i(i+1)/2
    const TS* s = src + j.src;
    for (uint32_t c = 0; c < j.cols; ++c) {
      const uint64_t cc = (uint64_t)c * (c + 1) / 2;
      TD* d = dst + j.dst + (uint64_t)c * j.dst_ld;
      for (uint32_t r = threadIdx.x; r < j.rows; r += blockDim.x) {
        const uint64_t i = (uint64_t)j.pad + r;
        d[r] = (TD)s[i <= c ? i + cc : c + i * (i + 1) / 2];
      }
    }
    return;
  }
  if (j.transpose) {
    // V (rows x cols, column-major) -> V^T rows of the panel: a tile of rb
    // rows is read column by column and written row by row, so both sides
    // move contiguous runs (rb values per column, cols values per row)
    __shared__ TD tile[kRelayoutTile + 256];
    const uint32_t cols = j.cols;
    uint32_t lg = 0;  // rb = 2^lg, the largest power of two with rb * cols <= kRelayoutTile
    while ((2u << lg) * cols <= kRelayoutTile) ++lg;
    const uint32_t rb = 1u << lg, ld = rb + 1;
    const uint32_t lgl = lg < 8 ? lg : 8;  // load lanes per column: min(rb, 256)
    const uint32_t lr = threadIdx.x & ((1u << lgl) - 1), lc = threadIdx.x >> lgl;
    const uint32_t lstep = 256u >> lgl;
    const uint32_t step = cols, per = 256 / step;  // store: per rows of cols lanes
    const uint32_t sub = threadIdx.x / step, c0 = threadIdx.x - sub * step;
    for (uint32_t rbase = 0; rbase < j.rows; rbase += rb) {
      const uint32_t rn = j.rows - rbase < rb ? j.rows - rbase : rb;
      for (uint32_t c = lc; c < cols; c += lstep) {
        const TS* s = src + j.src + (uint64_t)c * j.src_ld + rbase;
        for (uint32_t r = lr; r < rn; r += 1u << lgl) tile[c * ld + r] = (TD)s[r];
      }
      __syncthreads();
      if (sub < per) {
        TD* d = dst + j.dst + (uint64_t)rbase * j.dst_ld + c0;
        for (uint32_t r = sub; r < rn; r += per) d[(uint64_t)r * j.dst_ld] = tile[c0 * ld + r];
      }
      __syncthreads();
    }
    return;
  }
  // column by column, no per-element division: a column shorter than the
  // workgroup shares it with the next blockDim.x / rows - 1 columns; each
  // thread has four columns' loads in flight before it stores
  const uint32_t step = j.rows >= blockDim.x ? blockDim.x : j.rows;
  const uint32_t per = blockDim.x / step;
  const uint32_t sub = threadIdx.x / step, r0 = threadIdx.x - sub * step;
  if (sub >= per) return;
  const TS* s = src + j.src;
  TD* d = dst + j.dst;
  for (uint32_t r = r0; r < j.rows; r += step)
    for (uint32_t c = sub; c < j.cols; c += 4 * per) {
      TS v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t cc = c + u * per;
        v[u] = cc < j.cols ? s[(uint64_t)cc * j.src_ld + r] : (TS)0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t cc = c + u * per;
        if (cc < j.cols) d[(uint64_t)cc * j.dst_ld + r] = (TD)v[u];
      }
    }
}

template <typename TD, typename TS>
__global__ void cast_kernel(const TS* __restrict__ src, TD* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = (TD)src[i];
}

// z chunk staged in LDS per pass over a panel
constexpr int kZChunk = 2048;

// out(m) (+)= P z  for one task per workgroup, blockIdx.y = right-hand side.
// Thread t < G*m (G = 256/m column groups) owns row t%m and the columns
// c = t/m, t/m + G, ...; consecutive groups read consecutive columns, so a
// workgroup's loads cover G*m contiguous values.  Rows are reduced across
// groups through LDS in a fixed order.  Phase B tasks either accumulate
// alpha * sum into y or, for a column chunk of a wide panel, store the row
// sums in the partial buffer.
// COH selects agent-coherent accesses (the sc1 loads and stores of relaxed
// agent-scope atomics) for t and the chunk partial sums, for a form in which
// one workgroup reads what another wrote within the same launch.  The
// launches of this file pass COH = false: their kernel boundaries order
// those values.
template <typename T>
__device__ __forceinline__ T ld_coh(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_coh(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool COH, typename T>
__device__ __forceinline__ T ld_x(const T* p) {
  if constexpr (COH) return ld_coh(p);
  return *p;
}
template <bool COH, typename T>
__device__ __forceinline__ void st_x(T* p, T v) {
  if constexpr (COH) st_coh(p, v);
  else *p = v;
}

template <typename TS, typename TV, bool PHASE_B, bool COH>
__device__ __forceinline__ void gemv_task(const GemvTask& tk, uint32_t rhs, const uint32_t* __restrict__ zsrc,
                                          const TS* __restrict__ arena, const TV* __restrict__ x, uint32_t ldx,
                                          TV* __restrict__ t, uint32_t t_len, TV* __restrict__ y, uint32_t ldy,
                                          TV* __restrict__ part, uint32_t part_len, TV alpha, TV* zs, TV* red) {
  const TV* xr = x + (uint64_t)rhs * ldx;
  TV* tr = t + (uint64_t)rhs * t_len;
  const int m = (int)tk.m;
  const int G = kPanelThreads / m;
  const int tid = threadIdx.x;
  const int row = tid % m, g = tid / m;
  const bool active = g < G;
  const TS* P = arena + tk.data + row;
  TV a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (uint32_t c0 = 0; c0 < tk.W; c0 += kZChunk) {
    const int cw = (int)min((uint32_t)kZChunk, tk.W - c0);
    __syncthreads();
    for (int c = tid; c < cw; c += kPanelThreads) {
      const uint32_t s = zsrc[tk.zoff + c0 + c];
      zs[c] = (s & kFromT) ? ld_x<COH>(tr + (s & ~kFromT)) : xr[s];
    }
    __syncthreads();
    if (active) {
      const TS* Pc = P + (uint64_t)c0 * m;
      const uint32_t step = (uint32_t)(G * m);
      uint32_t off = (uint32_t)g * m;
      int c = g;
      for (; c + 7 * G < cw; c += 8 * G, off += 8 * step) {
        const TV p0 = (TV)Pc[off];
        const TV p1 = (TV)Pc[off + step];
        const TV p2 = (TV)Pc[off + 2 * step];
        const TV p3 = (TV)Pc[off + 3 * step];
        const TV p4 = (TV)Pc[off + 4 * step];
        const TV p5 = (TV)Pc[off + 5 * step];
        const TV p6 = (TV)Pc[off + 6 * step];
        const TV p7 = (TV)Pc[off + 7 * step];
        a0 = fma(p0, zs[c], a0);
        a1 = fma(p1, zs[c + G], a1);
        a2 = fma(p2, zs[c + 2 * G], a2);
        a3 = fma(p3, zs[c + 3 * G], a3);
        a0 = fma(p4, zs[c + 4 * G], a0);
        a1 = fma(p5, zs[c + 5 * G], a1);
        a2 = fma(p6, zs[c + 6 * G], a2);
        a3 = fma(p7, zs[c + 7 * G], a3);
      }
      for (; c < cw; c += G, off += step) a0 = fma((TV)Pc[off], zs[c], a0);
    }
  }
  red[tid] = active ? (a0 + a1) + (a2 + a3) : (TV)0;
  __syncthreads();
  if (tid < m) {
    TV s = 0;
    for (int q = 0; q < G; ++q) s += red[q * m + tid];
    if (PHASE_B) {
      if (tk.partial)
        st_x<COH>(part + (uint64_t)rhs * part_len + tk.out + tid, s);
      else
        y[(uint64_t)rhs * ldy + tk.out + tid] += alpha * s;
    } else if (tk.partial) {
      st_x<COH>(part + (uint64_t)rhs * part_len + tk.out + tid, s);
    } else {
      st_x<COH>(tr + tk.out + tid, s);
    }
  }
}

template <typename TS, typename TV, bool PHASE_B>
__global__ __launch_bounds__(kPanelThreads) void panel_gemv_kernel(
    const GemvTask* __restrict__ tasks, const uint32_t* __restrict__ zsrc,
    const TS* __restrict__ arena, const TV* __restrict__ x, uint32_t ldx, TV* __restrict__ t,
    uint32_t t_len, TV* __restrict__ y, uint32_t ldy, TV* __restrict__ part, uint32_t part_len,
    TV alpha, Gate g) {
  __shared__ TV zs[kZChunk];
  __shared__ TV red[kPanelThreads];
  if (gated(g)) return;
  const GemvTask tk = tasks[blockIdx.x];
  gemv_task<TS, TV, PHASE_B, false>(tk, blockIdx.y, zsrc, arena, x, ldx, t, t_len, y, ldy, part, part_len,
                                    alpha, zs, red);
}

// ------------------------------------------------------------------------
// 2 <= p <= 8 right-hand sides: the single-RHS kernel's thread mapping and
// streaming with NP = 8 accumulators per thread.  At these widths the
// products are bound by the panel stream, not by arithmetic, and the flat
// per-row loads of the single-RHS kernel stream at the full rate; Z comes
// from the NP-wide interleaved copies of x and t (one NP*8-byte row per
// column), staged in LDS and read as broadcasts.  From 9 right-hand sides on
// the MFMA kernel below is faster (one 16-column tile up to 16).
constexpr int kZChunkP = 256;   // columns of Z staged per pass

template <typename TS, typename TV, bool PHASE_B, int NP, bool COH>
__device__ __forceinline__ void gemvp_task(const GemvTask& tk, const uint32_t* __restrict__ zsrc,
                                           const TS* __restrict__ arena, const TV* __restrict__ xi,
                                           TV* __restrict__ ti, uint32_t t_len, TV* __restrict__ y, uint32_t ldy,
                                           TV* __restrict__ part, uint32_t part_len, TV alpha, uint32_t p,
                                           TV* zs) {
  constexpr int LNP = NP == 16 ? 4 : 3;
  const int m = (int)tk.m;
  const int G = kPanelThreads / m;
  const int tid = threadIdx.x;
  const int row = tid % m, g = tid / m;
  const bool active = g < G;
  const TS* P = arena + tk.data + row;
  TV a[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) a[j] = 0;
  for (uint32_t c0 = 0; c0 < tk.W; c0 += kZChunkP) {
    const int cw = (int)min((uint32_t)kZChunkP, tk.W - c0);
    __syncthreads();
    {
      // source-list entries first, then the rows (NP threads per row)
      constexpr int kG = kZChunkP * NP / kPanelThreads;
      uint32_t src[kG];
#pragma unroll
      for (int i = 0; i < kG; ++i) {
        const int c = (tid + i * kPanelThreads) >> LNP;
        src[i] = zsrc[tk.zoff + c0 + (c < cw ? c : 0)];
      }
#pragma unroll
      for (int i = 0; i < kG; ++i) {
        const int e = tid + i * kPanelThreads;
        const uint32_t sv = src[i];
        const TV* r = (sv & kFromT) ? ti + (uint64_t)(sv & ~kFromT) * NP : xi + (uint64_t)sv * NP;
        const TV v = ld_x<COH>(r + (e & (NP - 1)));
        zs[e] = (e >> LNP) < cw ? v : (TV)0;
      }
    }
    __syncthreads();
    if (active) {
      const TS* Pc = P + (uint64_t)c0 * m;
      const uint32_t step = (uint32_t)(G * m);
      uint32_t off = (uint32_t)g * m;
      int c = g;
      for (; c + 3 * G < cw; c += 4 * G, off += 4 * step) {
        const TV p0 = (TV)Pc[off];
        const TV p1 = (TV)Pc[off + step];
        const TV p2 = (TV)Pc[off + 2 * step];
        const TV p3 = (TV)Pc[off + 3 * step];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          a[j] = fma(p0, zs[NP * c + j], a[j]);
          a[j] = fma(p1, zs[NP * (c + G) + j], a[j]);
          a[j] = fma(p2, zs[NP * (c + 2 * G) + j], a[j]);
          a[j] = fma(p3, zs[NP * (c + 3 * G) + j], a[j]);
        }
      }
      for (; c < cw; c += G, off += step) {
        const TV p0 = (TV)Pc[off];
#pragma unroll
        for (int j = 0; j < NP; ++j) a[j] = fma(p0, zs[NP * c + j], a[j]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j) zs[j * kPanelThreads + tid] = active ? a[j] : (TV)0;
  __syncthreads();
  if (tid < m) {
    for (uint32_t j = 0; j < p; ++j) {
      TV s = 0;
      for (int q = 0; q < G; ++q) s += zs[j * kPanelThreads + q * m + tid];
      if (tk.partial)
        st_x<COH>(part + (uint64_t)j * part_len + tk.out + tid, s);
      else if (PHASE_B)
        y[(uint64_t)j * ldy + tk.out + tid] += alpha * s;
      else
        st_x<COH>(ti + (uint64_t)(tk.out + tid) * NP + j, s);
    }
  }
}

template <typename TS, typename TV, bool PHASE_B, int NP>
__global__ __launch_bounds__(kPanelThreads) void panel_gemvp_kernel(
    const GemvTask* __restrict__ tasks, const uint32_t* __restrict__ zsrc,
    const TS* __restrict__ arena, const TV* __restrict__ xi, TV* __restrict__ ti, uint32_t t_len,
    TV* __restrict__ y, uint32_t ldy, TV* __restrict__ part, uint32_t part_len, TV alpha,
    uint32_t p) {
  __shared__ TV zs[kZChunkP * NP];   // also the cross-group reduction buffer (256 x NP)
  const GemvTask tk = tasks[blockIdx.x];
  gemvp_task<TS, TV, PHASE_B, NP, false>(tk, zsrc, arena, xi, ti, t_len, y, ldy, part, part_len, alpha, p, zs);
}

// ------------------------------------------------------------------------
// 1 <= p <= 8 right-hand sides in ONE launch (ahm_hmat_set_fused).  One
// workgroup per row panel [D | U]; the V^T x values its U columns need are
// formed by the workgroup itself from the V^T panels (each low-rank leaf's
// t is used by its own row panel only), so no workgroup waits for another
// and no chunk sums are needed.  X and Y are read and written in place
// (p columns ldx / ldy apart), without the interleaved copies.  The price:
// x of a column cluster is read once per low-rank leaf instead of once per
// V^T panel, and the V^T values are read k at a time down a column of their
// panel.
//
// Per chunk of kZChunkP panel columns the workgroup stages Z (8 values per
// column, zero beyond p) in LDS: x columns directly, and each piece of a
// FusedGroup in the chunk by one wavefront: lane (r, cg) (r < kk rows of the
// piece, cg < 64/kk column groups) sums the V^T row r against x over the
// columns cg, cg + 64/kk, ... of the V^T panel, and the column groups are
// summed through LDS in order.  Then the panel stream runs as in the
// p <= 8 kernel above.
template <typename TS, typename TV>
__global__ __launch_bounds__(kPanelThreads) void fused_kernel(
    const GemvTask* __restrict__ panels, const GemvTask* __restrict__ apanels,
    const FusedGroup* __restrict__ groups, const uint32_t* __restrict__ goff, const uint32_t* __restrict__ zsrc,
    const TS* __restrict__ arena, const TV* __restrict__ X, uint32_t ldx, TV* __restrict__ Y, uint32_t ldy,
    TV alpha, uint32_t p) {
  constexpr int NP = 8;
  __shared__ TV zs[kZChunkP * NP];            // also the cross-group reduction buffer
  __shared__ TV wred[kPanelThreads / 64][64 * NP];
  const GemvTask tk = panels[blockIdx.x];
  const int m = (int)tk.m;
  const int G = kPanelThreads / m;
  const int tid = threadIdx.x;
  const int row = tid % m, g = tid / m;
  const bool active = g < G;
  const int w = tid >> 6, lane = tid & 63;
  const TS* P = arena + tk.data + row;
  const uint32_t g0 = goff[blockIdx.x], g1 = goff[blockIdx.x + 1];
  TV a[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) a[j] = 0;
  for (uint32_t c0 = 0; c0 < tk.W; c0 += kZChunkP) {
    const uint32_t cw = min((uint32_t)kZChunkP, tk.W - c0);
    __syncthreads();
    // x columns (the t columns are written by the pieces below)
    for (uint32_t e = tid; e < cw * NP; e += kPanelThreads) {
      const uint32_t c = e / NP, j = e % NP;
      const uint32_t src = zsrc[tk.zoff + c0 + c];
      if (!(src & kFromT)) zs[e] = j < p ? X[(uint64_t)j * ldx + src] : (TV)0;
    }
    // pieces of the groups in this chunk, round robin over the wavefronts
    uint32_t q = 0;
    for (uint32_t gi = g0; gi < g1; ++gi) {
      const FusedGroup fg = groups[gi];
      if (fg.col >= c0 + cw || fg.col + fg.k <= c0) continue;
      if ((int)(q++ % (kPanelThreads / 64)) != w) continue;
      const uint32_t cs = max(fg.col, c0), ce = min(fg.col + fg.k, c0 + cw);
      const GemvTask A = apanels[fg.apanel];
      const TS* Pa = arena + A.data + fg.row0 + (cs - fg.col);
      for (uint32_t rb = 0; rb < ce - cs; rb += 64) {
        const uint32_t nr = min(64u, ce - cs - rb);
        const uint32_t r = lane % nr, cg = lane / nr, ncg = 64 / nr;
        TV acc[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) acc[j] = 0;
        if (cg < ncg) {
          // the V^T panel's columns are the x entries xb, xb+1, ... (checked
          // when the plan is built); four columns' loads in flight
          const TV* Xc = X + zsrc[A.zoff];
          const TS* Pr = Pa + rb + r;
          uint32_t cc = cg;
          for (; cc + 3 * ncg < A.W; cc += 4 * ncg) {
            TV v[4], xv[4][NP];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              v[u] = (TV)Pr[(uint64_t)(cc + u * ncg) * A.m];
#pragma unroll
              for (int j = 0; j < NP; ++j) xv[u][j] = (uint32_t)j < p ? Xc[(uint64_t)j * ldx + cc + u * ncg] : (TV)0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int j = 0; j < NP; ++j) acc[j] = fma(v[u], xv[u][j], acc[j]);
          }
          for (; cc < A.W; cc += ncg) {
            const TV v = (TV)Pr[(uint64_t)cc * A.m];
#pragma unroll
            for (int j = 0; j < NP; ++j)
              if ((uint32_t)j < p) acc[j] = fma(v, Xc[(uint64_t)j * ldx + cc], acc[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) wred[w][lane * NP + j] = cg < ncg ? acc[j] : (TV)0;
        // the column groups are summed by other lanes of this wavefront
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < nr)
#pragma unroll
          for (int j = 0; j < NP; ++j) {
            TV t = 0;
            for (uint32_t qq = 0; qq < ncg; ++qq) t += wred[w][(lane + qq * nr) * NP + j];
            zs[(cs - c0 + rb + lane) * NP + j] = t;
          }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    __syncthreads();
    if (active) {
      const TS* Pc = P + (uint64_t)c0 * m;
      const uint32_t step = (uint32_t)(G * m);
      uint32_t off = (uint32_t)g * m;
      int c = g;
      for (; c + 3 * G < (int)cw; c += 4 * G, off += 4 * step) {
        const TV p0 = (TV)Pc[off];
        const TV p1 = (TV)Pc[off + step];
        const TV p2 = (TV)Pc[off + 2 * step];
        const TV p3 = (TV)Pc[off + 3 * step];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          a[j] = fma(p0, zs[NP * c + j], a[j]);
          a[j] = fma(p1, zs[NP * (c + G) + j], a[j]);
          a[j] = fma(p2, zs[NP * (c + 2 * G) + j], a[j]);
          a[j] = fma(p3, zs[NP * (c + 3 * G) + j], a[j]);
        }
      }
      for (; c < (int)cw; c += G, off += step) {
        const TV p0 = (TV)Pc[off];
#pragma unroll
        for (int j = 0; j < NP; ++j) a[j] = fma(p0, zs[NP * c + j], a[j]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j) zs[j * kPanelThreads + tid] = active ? a[j] : (TV)0;
  __syncthreads();
  if (tid < m)
    for (uint32_t j = 0; j < p; ++j) {
      TV sum = 0;
      for (int q2 = 0; q2 < G; ++q2) sum += zs[j * kPanelThreads + q2 * m + tid];
      Y[(uint64_t)j * ldy + tk.out + tid] += alpha * sum;
    }
}

// ------------------------------------------------------------------------
// p >= 9 right-hand sides: the same panels as a matrix product on the MFMA
// units.
//
// out(m x p) (+)= P (m x W) * Z (W x p), one task per workgroup and 32
// right-hand sides per blockIdx.y.  The output is tiled 16 x 16: nrt row
// tiles times two column tiles (right-hand sides 0-15, 16-31 of the group).
// The 2*nrt units are spread evenly over the four wavefronts: row tile
// w + 4q goes to wavefront w whole (both column tiles, one panel load for
// two MFMAs); when nrt is not a multiple of 4 the last row tiles are split by
// column tile so that no wavefront holds more than ceil(2*nrt/4) units.  Z is
// gathered into LDS in chunks of kMmChunk columns from interleaved copies of
// x and t (the 32 right-hand sides of one source row contiguous).  Per k-step
// of 4:
//   A (16 x 4 of P):  lane l holds P[row0 + l%16][k0 + l/16]
//   B (4 x 16 of Z):  lane l holds Z[k0 + l/16][col0 + l%16]
//   C (16 x 16):      f64: lane l, value r -> (l/16 + 4r, l%16)
//                     f32: lane l, value r -> (4(l/16) + r, l%16)
// (lane layouts of v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 on gfx950).
// A lane's panel element of a k-step sits at a fixed offset l/16*m + l%16
// from a base that is uniform over the wavefront, so the loads take a scalar
// base and a 32-bit lane offset.  A tile's rows beyond m are loaded too (the
// next column's values, or the zeroed slack after the last panel): they only
// reach output rows beyond m, which are never stored.
constexpr int kMmChunk = 128;
constexpr int kMmRhs = 32;
// LDS output tile of a task piece (at most 8 row tiles): [kMmRhs][128 + 1]
// over the Z chunk's kMmChunk x (kMmRhs + 1) values
constexpr int kMmStageLd = 129;
static_assert(kMmRhs * kMmStageLd <= kMmChunk * (kMmRhs + 1), "output tile exceeds the Z chunk's LDS");

typedef double mm_f64x4 __attribute__((ext_vector_type(4)));
typedef float mm_f32x4 __attribute__((ext_vector_type(4)));

template <typename T> struct MmAcc;
template <> struct MmAcc<double> {
  typedef mm_f64x4 type;
  static __device__ __forceinline__ type mma(double a, double b, type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return lane / 16 + 4 * r; }
};
template <> struct MmAcc<float> {
  typedef mm_f32x4 type;
  static __device__ __forceinline__ type mma(float a, float b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane / 16) + r; }
};

// Z chunk into LDS: zs[c][j] = value j of the source row of column c, zero
// for c >= cw.  In two passes of 8 columns per thread: each thread loads its
// source-list entries, then its values, so that the two rounds of loads are
// each in flight together; columns beyond cw read column 0's (valid) source.
template <typename T>
__device__ __forceinline__ void mm_gather(T (&zs)[kMmChunk][kMmRhs + 1], const uint32_t* __restrict__ zsrc,
                                          int cw, const T* __restrict__ xg, const T* __restrict__ tg) {
  constexpr int kCs = kPanelThreads / kMmRhs;           // columns per step
  constexpr int kG = kMmChunk / kCs / 2;                 // steps per pass
  const int j = threadIdx.x & (kMmRhs - 1);
  const int cb = threadIdx.x / kMmRhs;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t src[kG];
#pragma unroll
    for (int i = 0; i < kG; ++i) {
      const int c = cb + (h * kG + i) * kCs;
      src[i] = zsrc[c < cw ? c : 0];
    }
    T val[kG];
#pragma unroll
    for (int i = 0; i < kG; ++i) {
      const uint32_t sv = src[i];
      const T* row = (sv & kFromT) ? tg + (uint64_t)(sv & ~kFromT) * kMmRhs : xg + (uint64_t)sv * kMmRhs;
      val[i] = row[j];
    }
#pragma unroll
    for (int i = 0; i < kG; ++i) {
      const int c = cb + (h * kG + i) * kCs;
      zs[c][j] = c < cw ? val[i] : (T)0;
    }
  }
}

// Panel loads of round r (16 columns, four k-steps) for the wavefront's NW
// whole row tiles (w, w+4, ...) and, with HALF, row tile hrt.  lob: the
// lane's byte offset within a k-step.  The last round of a task may run past
// its W columns (into the next task's values or the arena's slack); those
// columns are zeroed in mm_mma.
template <typename T, int NW, bool HALF, int NA>
__device__ __forceinline__ void mm_load(T (&A)[4][NA], const char* __restrict__ P, uint32_t r, uint32_t m,
                                        uint32_t lob, int w, int hrt) {
  const char* Pr = P + (uint64_t)r * 16u * m * sizeof(T);
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const char* Ps = Pr + (uint64_t)(4u * st * m) * sizeof(T) + lob;
#pragma unroll
    for (int q = 0; q < NW; ++q) A[st][q] = *(const T*)(Ps + (uint64_t)(16u * (w + 4 * q)) * sizeof(T));
    if (HALF) A[st][NA - 1] = *(const T*)(Ps + (uint64_t)(16u * hrt) * sizeof(T));
  }
}

// MFMAs of round r on the loaded panel values and the Z rows of the round in
// LDS; whole tiles against both column tiles (TWO), the half tile against
// column tile hct.  LAST: columns at or beyond W use zero.
template <typename T, int NW, bool HALF, bool TWO, bool LAST, int NA>
__device__ __forceinline__ void mm_mma(const T (&A)[4][NA], uint32_t r, uint32_t W, int li, int lk, int hct,
                                       const T (&zs)[kMmChunk][kMmRhs + 1],
                                       typename MmAcc<T>::type (&acc)[4][2],
                                       typename MmAcc<T>::type& acch) {
  const int kb = (int)(r & (kMmChunk / 16 - 1)) * 16;
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int k = kb + 4 * st + lk;   // zero-filled beyond the chunk's width
    const T b0 = zs[k][li];
    const T b1 = zs[k][16 + li];
    const bool in = !LAST || r * 16u + 4u * st + lk < W;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const T a = in ? A[st][q] : (T)0;
      acc[q][0] = MmAcc<T>::mma(a, b0, acc[q][0]);
      if (TWO) acc[q][1] = MmAcc<T>::mma(a, b1, acc[q][1]);
    }
    if (HALF) acch = MmAcc<T>::mma(in ? A[st][NA - 1] : (T)0, hct ? b1 : b0, acch);
  }
}

template <typename T, bool PHASE_B>
__device__ __forceinline__ void mm_store(const GemvTask& tk, int rt, int ct, const typename MmAcc<T>::type& c,
                                         int lane, uint32_t rhs0, uint32_t p, T* __restrict__ tg,
                                         T* __restrict__ y, uint32_t ldy, T* __restrict__ part,
                                         uint32_t part_len, T alpha) {
  const int jj = ct * 16 + (lane & 15);
  const uint32_t j = rhs0 + jj;
  if (j >= p) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t row = rt * 16 + MmAcc<T>::row(lane, r);
    if (row >= tk.m) continue;
    const T v = c[r];
    if (tk.partial)
      part[(uint64_t)j * part_len + tk.out + row] = v;
    else if (PHASE_B)
      y[(uint64_t)j * ldy + tk.out + row] += alpha * v;
    else
      tg[(uint64_t)(tk.out + row) * kMmRhs + jj] = v;
  }
}

// The whole task for one wavefront shape (NW whole row tiles, HALF, TWO).
// The panel loads run one round (16 columns) ahead of the MFMAs, across the
// Z chunk boundaries too; every wavefront of the workgroup walks the same
// rounds, so the workgroup barriers of the Z gathers pair up across shapes.
template <typename T, bool PHASE_B, int NW, bool HALF, bool TWO>
__device__ __forceinline__ void mm_task(const GemvTask& tk, const uint32_t* __restrict__ zsrc,
                                        const T* __restrict__ arena, const T* __restrict__ xg, T* __restrict__ tg,
                                        T (&zs)[kMmChunk][kMmRhs + 1], int w, int hrt, int hct, int rt0,
                                        int cnt, uint32_t rhs0, uint32_t p, T* __restrict__ y, uint32_t ldy,
                                        T* __restrict__ part, uint32_t part_len, T alpha) {
  typedef typename MmAcc<T>::type acc_t;
  constexpr int NA = NW + (HALF ? 1 : 0) > 0 ? NW + (HALF ? 1 : 0) : 1;
  constexpr bool WORK = NW > 0 || HALF;
  const int lane = threadIdx.x & 63;
  const uint32_t m = tk.m, W = tk.W;
  acc_t acc[4][2], acch = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  const uint32_t lob = ((uint32_t)lk * m + (uint32_t)li) * (uint32_t)sizeof(T);
  const char* P = (const char*)(arena + tk.data);
  const uint32_t R = (W + 15) / 16;
  // FP32: each Z chunk's products are summed in accumulators of their own
  // and then added to the running sums, so a long row panel is not one FP32
  // chain of thousands of terms
  constexpr bool kSplitSum = sizeof(T) == 4;
  acc_t tot[4][2], toth = acc_t{0, 0, 0, 0};
  if constexpr (kSplitSum) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) tot[a][b] = acc_t{0, 0, 0, 0};
  }
  auto fold = [&]() {
    if constexpr (kSplitSum) {
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        tot[q][0] += acc[q][0];
        acc[q][0] = acc_t{0, 0, 0, 0};
        if (TWO) {
          tot[q][1] += acc[q][1];
          acc[q][1] = acc_t{0, 0, 0, 0};
        }
      }
      if (HALF) {
        toth += acch;
        acch = acc_t{0, 0, 0, 0};
      }
    }
  };
  auto gather = [&](uint32_t r) {
    if (r % (kMmChunk / 16)) return;
    const uint32_t c0 = r * 16;
    if (r) fold();
    __syncthreads();
    mm_gather<T>(zs, zsrc + tk.zoff + c0, (int)min((uint32_t)kMmChunk, W - c0), xg, tg);
    __syncthreads();
  };
  // up to two whole row tiles: the panel loads of round r + 1 are issued
  // before the MFMAs of round r; wider shapes load and multiply round by
  // round to stay within the register budget of three wavefronts per SIMD
  constexpr bool PIPE = NW <= 2;
  T A0[4][NA], A1[4][PIPE ? NA : 1];
  if (WORK && PIPE) mm_load<T, NW, HALF, NA>(A0, P, 0, m, lob, w, hrt);
  uint32_t r = 0;
  for (; r + 1 < R; ++r) {
    gather(r);
    if (!WORK) continue;
    if (PIPE) {
      mm_load<T, NW, HALF, PIPE ? NA : 1>(A1, P, r + 1, m, lob, w, hrt);
    } else {
      mm_load<T, NW, HALF, NA>(A0, P, r, m, lob, w, hrt);
    }
    mm_mma<T, NW, HALF, TWO, false, NA>(A0, r, W, li, lk, hct, zs, acc, acch);
    if (PIPE) {
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int q = 0; q < NA; ++q) A0[st][q] = A1[st][PIPE ? q : 0];
    }
  }
  gather(r);   // the last round (W >= 1, so R >= 1)
  if (WORK) {
    if (!PIPE) mm_load<T, NW, HALF, NA>(A0, P, r, m, lob, w, hrt);
    mm_mma<T, NW, HALF, TWO, true, NA>(A0, r, W, li, lk, hct, zs, acc, acch);
  }
  fold();
  if constexpr (kSplitSum) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      acc[q][0] = tot[q][0];
      if (TWO) acc[q][1] = tot[q][1];
    }
    if (HALF) acch = toth;
  }
  if (!PHASE_B && !tk.partial) {
    // t rows hold their 32 right-hand sides contiguously: direct stores
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      mm_store<T, PHASE_B>(tk, w + 4 * q, 0, acc[q][0], lane, rhs0, p, tg, y, ldy, part, part_len, alpha);
      if (TWO) mm_store<T, PHASE_B>(tk, w + 4 * q, 1, acc[q][1], lane, rhs0, p, tg, y, ldy, part, part_len, alpha);
    }
    if (HALF) mm_store<T, PHASE_B>(tk, hrt, hct, acch, lane, rhs0, p, tg, y, ldy, part, part_len, alpha);
    return;
  }
  // y and the partial sums are right-hand-side major: the piece's output
  // goes through LDS (ys[rhs][row], over the Z chunk's space) so that each
  // wavefront store covers 64 consecutive rows of one right-hand side
  T (*ys)[kMmStageLd] = reinterpret_cast<T (*)[kMmStageLd]>(&zs[0][0]);
  auto put = [&](int rt, int ct, const acc_t& c) {
    const int jj = ct * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) ys[jj][(rt - rt0) * 16 + MmAcc<T>::row(lane, r)] = c[r];
  };
  __syncthreads();   // every wavefront is done with the Z chunk
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    put(w + 4 * q, 0, acc[q][0]);
    if (TWO) put(w + 4 * q, 1, acc[q][1]);
  }
  if (HALF) put(hrt, hct, acch);
  __syncthreads();
  const uint32_t rbase = (uint32_t)rt0 * 16u;
  const uint32_t rows = min((uint32_t)cnt * 16u, m - rbase);
  for (int jj = (int)(threadIdx.x >> 6); jj < kMmRhs; jj += kPanelThreads / 64) {
    const uint32_t j = rhs0 + jj;
    if (j >= p) break;
    for (uint32_t i = (uint32_t)lane; i < rows; i += 64) {
      const T v = ys[jj][i];
      if (tk.partial)
        part[(uint64_t)j * part_len + tk.out + rbase + i] = v;
      else
        y[(uint64_t)j * ldy + tk.out + rbase + i] += alpha * v;
    }
  }
}

// Workgroup map of the MFMA launches: entry (task << 2) | piece, piece 0 the
// whole task, 1 / 2 the first ceil(nrt/2) / last floor(nrt/2) row tiles of a
// task cut in two (tasks of more than kMmSplitTiles row tiles), so that a
// wavefront holds at most two whole row tiles and its panel loads run one
// round ahead; both halves gather the same Z chunks.
constexpr uint32_t kMmSplitTiles = 8;

__device__ __forceinline__ void mm_piece(uint32_t e, int nrt, int& rt0, int& cnt) {
  const uint32_t piece = e & 3u;
  rt0 = piece == 2 ? (nrt + 1) / 2 : 0;
  cnt = piece == 0 ? nrt : piece == 1 ? (nrt + 1) / 2 : nrt / 2;
}

template <typename T, bool PHASE_B>
__global__ __launch_bounds__(kPanelThreads) __attribute__((amdgpu_waves_per_eu(4))) void panel_mm_kernel(
    const GemvTask* __restrict__ tasks, const uint32_t* __restrict__ map, const uint32_t* __restrict__ zsrc,
    const T* __restrict__ arena, const T* __restrict__ xi, uint32_t n_x, T* __restrict__ ti,
    uint32_t t_len, T* __restrict__ y, uint32_t ldy, T* __restrict__ part,
    uint32_t part_len, T alpha, uint32_t p) {
  __shared__ T zs[kMmChunk][kMmRhs + 1];
  const uint32_t e = map[blockIdx.x];
  const GemvTask tk = tasks[e >> 2];
  const uint32_t g = blockIdx.y, rhs0 = g * kMmRhs;
  const T* xg = xi + (uint64_t)g * n_x * kMmRhs;
  T* tg = ti + (uint64_t)g * t_len * kMmRhs;
  int rt0, nrt;
  mm_piece(e, (int)(tk.m + 15) / 16, rt0, nrt);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const bool two = p - rhs0 > 16;   // second column tile needed
  // this wavefront's units: nwhole row tiles w + 4q with both column tiles,
  // and at most one row tile hrt with column tile hct alone
  int nwhole, hrt = -1, hct = 0;
  if (two) {
    const int full = nrt >> 2, rem = nrt & 3;
    nwhole = full + (rem == 3 && w < 2 ? 1 : 0);
    if (rem == 1 && w < 2) {
      hrt = 4 * full;
      hct = w;
    } else if (rem == 2) {
      hrt = 4 * full + (w >> 1);
      hct = w & 1;
    } else if (rem == 3 && w >= 2) {
      hrt = 4 * full + 2;
      hct = w - 2;
    }
  } else {
    nwhole = (nrt - w + 3) >> 2;   // row tiles w, w+4, ... below nrt
  }
#define AHM_MM_TASK(NW, HALF, TWO)                                                                    \
  mm_task<T, PHASE_B, NW, HALF, TWO>(tk, zsrc, arena, xg, tg, zs, rt0 + w, hrt, hct, rt0, nrt, rhs0, p, y, \
                                     ldy, part, part_len, alpha)
  const int shape = nwhole * 4 + (hrt >= 0 ? 2 : 0) + (two ? 1 : 0);
  if (hrt >= 0) hrt += rt0;   // row tiles of the task, not of the piece
  switch (shape) {
    case 0: case 1: AHM_MM_TASK(0, false, false); break;
    case 2: AHM_MM_TASK(0, true, false); break;
    case 3: AHM_MM_TASK(0, true, true); break;
    case 4: AHM_MM_TASK(1, false, false); break;
    case 5: AHM_MM_TASK(1, false, true); break;
    case 6: AHM_MM_TASK(1, true, false); break;
    case 7: AHM_MM_TASK(1, true, true); break;
    case 8: AHM_MM_TASK(2, false, false); break;
    case 9: AHM_MM_TASK(2, false, true); break;
    case 10: AHM_MM_TASK(2, true, false); break;
    default: AHM_MM_TASK(2, true, true); break;
  }
#undef AHM_MM_TASK
}

// X (n x p, right-hand sides ldx apart) -> Xi: groups of 32 right-hand sides,
// row i of group g holding its pw values contiguously at (g*n + i)*pw
// (zero beyond p).  64 rows x 32 right-hand sides per block through LDS.
template <typename T, bool COH = false>
__device__ __forceinline__ void interleave_tile(const T* __restrict__ X, uint32_t ldx, uint32_t n, uint32_t p,
                                                int pw, T* __restrict__ Xi, uint32_t bx, uint32_t g,
                                                T (&tile)[32][65]) {
  const uint32_t i0 = bx * 64u;
  for (int e = threadIdx.x; e < 32 * 64; e += 256) {
    const int jj = e >> 6, ii = e & 63;
    const uint32_t rhs = g * 32u + jj, i = i0 + ii;
    tile[jj][ii] = (jj < pw && rhs < p && i < n) ? X[(uint64_t)rhs * ldx + i] : (T)0;
  }
  __syncthreads();
  const int lpw = pw > 16 ? 5 : pw > 8 ? 4 : 3;
  for (int e = threadIdx.x; e < (64 << lpw); e += 256) {
    const int ii = e >> lpw, jj = e & (pw - 1);
    const uint32_t i = i0 + ii;
    if (i < n) st_x<COH>(Xi + ((uint64_t)g * n + i) * pw + jj, tile[jj][ii]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void interleave_kernel(const T* __restrict__ X, uint32_t ldx,
                                                         uint32_t n, uint32_t p, int pw,
                                                         T* __restrict__ Xi) {
  __shared__ T tile[32][65];
  interleave_tile<T>(X, ldx, n, p, pw, Xi, blockIdx.x, blockIdx.y, tile);
}

// Sum over chunks (chunk order fixed) for chunked panels: phase B adds
// alpha * sum to y rows, phase A (ASSIGN) stores the sum into t.
// pw > 0: y is interleaved (element (row, rhs) at ((rhs/32)*ldy + row)*pw +
// rhs%32, the layout of t in multi-RHS products), else y[rhs*ldy + row].
template <typename TV, bool ASSIGN, bool COH = false>
__device__ __forceinline__ void reduce_task(const ReduceTask& r, uint32_t rhs, const TV* __restrict__ part,
                                            uint32_t part_len, TV* __restrict__ y, uint32_t ldy, TV alpha,
                                            int pw) {
  const TV* pr = part + (uint64_t)rhs * part_len + r.part;
  const uint64_t base = pw ? (uint64_t)(rhs >> 5) * ldy * pw + (rhs & 31) : (uint64_t)rhs * ldy;
  const uint64_t stride = pw ? (uint64_t)pw : 1;
  for (uint32_t i = threadIdx.x; i < r.m; i += kPanelThreads) {
    TV s = 0;
    uint32_t c = 0;
    for (; c + 8 <= r.nchunks; c += 8) {   // eight loads in flight, summed in order
      TV v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = ld_x<COH>(pr + (uint64_t)(c + u) * r.m + i);
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < r.nchunks; ++c) s += ld_x<COH>(pr + (uint64_t)c * r.m + i);
    TV* o = y + base + (r.row + i) * stride;
    if (ASSIGN)
      st_x<COH>(o, s);
    else
      *o += alpha * s;
  }
}

template <typename TV, bool ASSIGN>
__global__ __launch_bounds__(kPanelThreads) void chunk_reduce_kernel(
    const ReduceTask* __restrict__ tasks, const TV* __restrict__ part, uint32_t part_len,
    TV* __restrict__ y, uint32_t ldy, TV alpha, int pw, Gate g) {
  if (gated(g)) return;
  reduce_task<TV, ASSIGN>(tasks[blockIdx.x], blockIdx.y, part, part_len, y, ldy, alpha, pw);
}

// ------------------------------------------------------------------------
// One persistent launch per product of 1 <= p <= 8 right-hand sides
// (ahm_hmat_set_fused(h, 2)): the grid holds as many workgroups as the
// device keeps resident (a cooperative launch, which the runtime refuses
// otherwise) and walks the same task lists as the separate launches, phase
// after phase, with a grid barrier between them: [interleave X,] V^T x, its
// chunk sums, the row panels, their chunk sums.  Tasks go round-robin over
// the workgroups in list order (largest first).  Values one workgroup writes
// and another reads (t, partial sums, interleaved X) are stored and loaded
// with agent-coherent accesses across the grid barriers.  The arithmetic and
// its order are those of the separate launches, so the results are the same
// bit for bit.  A gated
// product (single system, frozen) returns at once in every workgroup.
template <typename TS, typename TV>
struct PersistArgs {
  const GemvTask* tasks_a;
  const GemvTask* tasks_b;
  const ReduceTask* tasks_ra;
  const ReduceTask* tasks_r;
  const uint32_t* zsrc;
  const TS* arena;
  const TV* X;
  TV* Y;
  TV* t;
  TV* part;
  TV* xi;
  uint32_t ldx, ldy, p, n_cols, t_len, part_len;
  uint32_t n_a, n_b, n_ra, n_r;
  TV alpha;
  const uint32_t* frozen;
};

// the grid barrier between phases
__device__ __forceinline__ void phase_barrier(cooperative_groups::grid_group& grid) {
  grid.sync();
}

template <typename TS, typename TV, bool MULTI>
__global__ __launch_bounds__(kPanelThreads) void persistent_kernel(PersistArgs<TS, TV> a) {
  // gemv: 2048 staged z + 256 row sums; gemvp: 2048; interleave: 32 x 65
  __shared__ TV lds[2304];
  if (a.frozen && a.frozen[0]) return;   // uniform over the grid
  auto grid = cooperative_groups::this_grid();
  const uint32_t G = gridDim.x, b = blockIdx.x;
  if (MULTI) {
    const uint32_t nt = (a.n_cols + 63) / 64;
    for (uint32_t q = b; q < nt; q += G)
      interleave_tile<TV, true>(a.X, a.ldx, a.n_cols, a.p, 8, a.xi, q, 0, *(TV(*)[32][65])lds);
    phase_barrier(grid);
  }
  for (uint32_t q = b; q < a.n_a; q += G) {
    if (MULTI)
      gemvp_task<TS, TV, false, 8, true>(a.tasks_a[q], a.zsrc, a.arena, a.xi, a.t, a.t_len, a.Y, a.ldy, a.part,
                                          a.part_len, a.alpha, a.p, lds);
    else
      gemv_task<TS, TV, false, true>(a.tasks_a[q], 0, a.zsrc, a.arena, a.X, a.ldx, a.t, a.t_len, a.Y, a.ldy,
                                      a.part, a.part_len, a.alpha, lds, lds + kZChunk);
  }
  phase_barrier(grid);
  for (uint32_t q = b; q < a.n_ra * a.p; q += G)
    reduce_task<TV, true, true>(a.tasks_ra[q % a.n_ra], q / a.n_ra, a.part, a.part_len, a.t, a.t_len, a.alpha,
                                MULTI ? 8 : 0);
  phase_barrier(grid);
  for (uint32_t q = b; q < a.n_b; q += G) {
    if (MULTI)
      gemvp_task<TS, TV, true, 8, true>(a.tasks_b[q], a.zsrc, a.arena, a.xi, a.t, a.t_len, a.Y, a.ldy, a.part,
                                         a.part_len, a.alpha, a.p, lds);
    else
      gemv_task<TS, TV, true, true>(a.tasks_b[q], 0, a.zsrc, a.arena, a.X, a.ldx, a.t, a.t_len, a.Y, a.ldy,
                                     a.part, a.part_len, a.alpha, lds, lds + kZChunk);
  }
  phase_barrier(grid);
  for (uint32_t q = b; q < a.n_r * a.p; q += G)
    reduce_task<TV, false, true>(a.tasks_r[q % a.n_r], q / a.n_r, a.part, a.part_len, a.Y, a.ldy, a.alpha, 0);
}

// ------------------------------------------------------------------------
// V^T X of multi-RHS products on BF16 MFMA (AHM_BF16 storage): V and X are
// rounded to BF16, the products accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16.  Lane layout (gfx950): lane l holds
//   A (16 x 32): A[l%16][8*(l/16) .. 8*(l/16)+7]
//   B (32 x 16): B[8*(l/16) .. 8*(l/16)+7][l%16]
//   C (16 x 16): value r -> (4*(l/16) + r, l%16).
// The BF16 copy of a phase-A task (m x W values, column-major in the FP32
// arena) holds ceil(W/8) k-groups, each m rows of 8 consecutive columns, so
// a lane's A operand is one 16-byte load: element (r, c) at
// ((c/8)*m + r)*8 + c%8, zero for c >= W.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// FP32 -> BF16, round to nearest even (finite values)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// one workgroup per phase-A task: its BF16 copy at bfoff[task]
__global__ __launch_bounds__(256) void bf16_panels_kernel(const GemvTask* __restrict__ tasks,
                                                          const uint64_t* __restrict__ bfoff,
                                                          const float* __restrict__ arena,
                                                          uint16_t* __restrict__ out) {
  const GemvTask tk = tasks[blockIdx.x];
  const uint64_t base = bfoff[blockIdx.x];
  const uint32_t m = tk.m, W = tk.W;
  const uint64_t n = (uint64_t)m * ((W + 7) & ~7u);
  for (uint64_t e = threadIdx.x; e < n; e += 256) {
    const uint32_t i = (uint32_t)(e & 7), q = (uint32_t)(e >> 3);
    const uint32_t r = q % m, c = (q / m) * 8 + i;
    out[base + e] = c < W ? bf16_rne(arena[tk.data + (uint64_t)c * m + r]) : (uint16_t)0;
  }
}

__global__ void bf16_copy_kernel(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = bf16_rne(src[i]);
}

// Z columns staged per pass (one per thread)
constexpr int kBfChunk = kPanelThreads;

// t(m x 32) = V^T X for one phase-A task and 32 right-hand sides
// (blockIdx.y): Z gathered into LDS transposed (zs[rhs][column], BF16 rows
// of the interleaved X), wavefront w owns row tiles w, w+4, w+8, w+12 with
// both 16-column tiles.  Output as panel_mm_kernel's phase A: t interleaved
// (32 values per row), or the chunk's partial sums.
__global__ __launch_bounds__(kPanelThreads) void panel_bf16_kernel(
    const GemvTask* __restrict__ tasks, const uint64_t* __restrict__ bfoff, const uint32_t* __restrict__ zsrc,
    const uint16_t* __restrict__ bfa, const uint16_t* __restrict__ xb, uint32_t n_x, float* __restrict__ ti,
    uint32_t t_len, float* __restrict__ part, uint32_t part_len, uint32_t p) {
  __shared__ __attribute__((aligned(16))) uint16_t zs[kMmRhs][kBfChunk + 8];
  const GemvTask tk = tasks[blockIdx.x];
  const uint64_t base = bfoff[blockIdx.x];
  const uint32_t g = blockIdx.y, rhs0 = g * kMmRhs;
  const uint16_t* xg = xb + (uint64_t)g * n_x * kMmRhs;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t m = tk.m, W = tk.W, ngr = (W + 7) / 8;
  const int nrt = (int)(m + 15) / 16;
  const bool two = p - rhs0 > 16;
  const int li = lane & 15, lk = lane >> 4;
  mm_f32x4 acc[4][2];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q][0] = acc[q][1] = mm_f32x4{0, 0, 0, 0};
  for (uint32_t c0 = 0; c0 < W; c0 += kBfChunk) {
    const uint32_t cw = min((uint32_t)kBfChunk, W - c0);
    __syncthreads();
    {
      const uint32_t c = threadIdx.x;
      uint4 v[4] = {};
      if (c < cw) {
        const uint4* row = (const uint4*)(xg + (uint64_t)zsrc[tk.zoff + c0 + c] * kMmRhs);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = row[q];
      }
      const uint16_t* vv = (const uint16_t*)v;
#pragma unroll
      for (int j = 0; j < kMmRhs; ++j) zs[j][c] = vv[j];
    }
    __syncthreads();
    const uint32_t nks = (cw + 31) / 32;
    for (uint32_t ks = 0; ks < nks; ++ks) {
      const uint32_t kc = ks * 32 + (uint32_t)lk * 8;   // chunk-local first column of the lane's group
      const uint32_t kg = (c0 + kc) / 8;                // the task's k-group (c0 is a multiple of 8)
      const bf16x8_t b0 = __builtin_bit_cast(bf16x8_t, *(const uint4*)&zs[li][kc]);
      const bf16x8_t b1 = __builtin_bit_cast(bf16x8_t, *(const uint4*)&zs[16 + li][kc]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rt = w + 4 * q;
        if (rt >= nrt) break;
        uint4 av = {0, 0, 0, 0};
        if (kg < ngr) av = *(const uint4*)(bfa + base + ((uint64_t)kg * m + (uint32_t)(rt * 16 + li)) * 8);
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, av);
        acc[q][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[q][0], 0, 0, 0);
        if (two) acc[q][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[q][1], 0, 0, 0);
      }
    }
  }
  float* tg = ti + (uint64_t)g * t_len * kMmRhs;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rt = w + 4 * q;
    if (rt >= nrt) break;
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      if (ct == 1 && !two) break;
      const int jj = ct * 16 + li;
      const uint32_t j = rhs0 + jj;
      if (j >= p) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t row = (uint32_t)(rt * 16 + 4 * lk + r);
        if (row >= m) continue;
        const float v = acc[q][ct][r];
        if (tk.partial)
          part[(uint64_t)j * part_len + tk.out + row] = v;
        else
          tg[(uint64_t)(tk.out + row) * kMmRhs + jj] = v;
      }
    }
  }
}

// ------------------------------------------------------------------------
// plan: host-side construction of the tasks and copy jobs

struct Plan {
  // a_panels / panels: V^T panels and row panels before chunking
  std::vector<GemvTask> a, b, a_panels, panels;
  std::vector<ReduceTask> r, ra;
  uint32_t part_len = 0;
  HostVec<uint32_t> zsrc;
  HostVec<CopyJob> jobs;
  uint64_t nvals = 0;
  uint32_t t_len = 0, max_rank = 0;
  uint64_t n_dense = 0, n_lowrank = 0;
  uint64_t nv_dense = 0, nv_u = 0, nv_v = 0;
};

// Tasks from the panels of one phase: panels wider than about chunk_values
// values are cut into equal column chunks whose row sums meet in a reduction
// task; tasks sorted largest first.
static void chunk_panels(const std::vector<GemvTask>& panels, uint64_t chunk_values,
                         std::vector<GemvTask>& b, std::vector<ReduceTask>& r, uint32_t& part_len) {
  b.clear();
  r.clear();
  part_len = 0;
  for (const GemvTask& tk : panels) {
    const uint64_t vals = (uint64_t)tk.m * tk.W;
    const uint32_t nch = (uint32_t)std::min<uint64_t>(tk.W, (vals + chunk_values - 1) / chunk_values);
    if (nch <= 1) {
      b.push_back(tk);
      continue;
    }
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t c0 = (uint32_t)((uint64_t)tk.W * c / nch);
      const uint32_t c1 = (uint32_t)((uint64_t)tk.W * (c + 1) / nch);
      GemvTask ck = tk;
      ck.data = tk.data + (uint64_t)c0 * tk.m;
      ck.zoff = tk.zoff + c0;
      ck.W = c1 - c0;
      ck.out = part_len + c * tk.m;
      ck.partial = 1;
      b.push_back(ck);
    }
    r.push_back(ReduceTask{tk.out, tk.m, part_len, nch});
    part_len += nch * tk.m;
  }
  std::stable_sort(b.begin(), b.end(), [](const GemvTask& x, const GemvTask& y) {
    return (uint64_t)x.m * x.W > (uint64_t)y.m * y.W;
  });
}

// Host threads of the plan construction (at most 16).
static unsigned plan_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, std::min(hw ? hw : 1u, 16u));
}

// Calls f(lo, hi) on host threads for contiguous ranges [lo, hi) of the n
// items whose weight prefix sums are wsum (wsum[0] = 0, wsum[n] = total), cut
// so that each range carries about the same weight.  f writes only to its own
// items' pre-sized output ranges.
template <typename F>
static void parallel_by_weight(const std::vector<uint64_t>& wsum, const F& f) {
  const size_t n = wsum.size() - 1;
  const unsigned T = plan_threads();
  if (T == 1 || n < 4 * (size_t)T) {
    f((size_t)0, n);
    return;
  }
  std::vector<size_t> cut(T + 1, n);
  cut[0] = 0;
  for (unsigned k = 1; k < T; ++k)
    cut[k] = (size_t)(std::lower_bound(wsum.begin(), wsum.end(), wsum[n] / T * k) - wsum.begin());
  std::vector<std::thread> th;
  th.reserve(T);
  for (unsigned k = 0; k < T; ++k)
    if (cut[k] < cut[k + 1]) th.emplace_back(f, cut[k], cut[k + 1]);
  for (auto& t : th) t.join();
}

// The plan of a leaf list: phase A V^T panels (low-rank leaves grouped by
// column cluster, <= 256 rows per panel) and phase B row panels (one per
// row segment between leaf row boundaries, <= 256 rows each, [D | U]
// columns of every leaf covering the segment, in leaf order), their column
// source lists and the copy jobs that lay the leaves out.  Sizes are counted
// first; the arrays are then filled in parallel, group by group and segment
// by segment, each at its counted offset.
static int build_plan(uint32_t n_rows, uint32_t n_cols, uint32_t nl, const ahm_leaf_desc* L,
                      Plan& P) {
  const uint32_t TM = kPanelThreads;
  // validate
  for (uint32_t l = 0; l < nl; ++l) {
    const ahm_leaf_desc& d = L[l];
    if ((uint64_t)d.b1 + d.n1 > n_rows || (uint64_t)d.b2 + d.n2 > n_cols || d.n1 == 0 ||
        d.n2 == 0)
      return fail(AHM_ERR_INVALID, "leaf %u (%u,%u,%u,%u) outside %u x %u", l, d.b1, d.n1,
                  d.b2, d.n2, n_rows, n_cols);
    if (d.rank >= 0 && (uint64_t)d.rank > std::min(d.n1, d.n2) + 64u)
      return fail(AHM_ERR_INVALID, "leaf %u rank %d exceeds its dimensions", l, d.rank);
    if ((d.flags & AHM_LEAF_HEM) && (d.rank >= 0 || d.b1 != d.b2 || d.n1 != d.n2))
      return fail(AHM_ERR_INVALID, "leaf %u: a packed Hermitian leaf must be a dense diagonal block", l);
    if (d.flags & ~AHM_LEAF_HEM)
      return fail(AHM_ERR_INVALID, "leaf %u: flags %u (triangular leaves belong to ahm_trih_upload)", l, d.flags);
  }
  // low-rank leaves ordered by (b2, n2), leaf order among equals: a counting
  // sort on b2 (stable), then an insertion sort on n2 inside each b2 bucket
  std::vector<uint32_t> lr;
  {
    std::vector<uint32_t> start((size_t)n_cols + 1, 0);
    size_t nlr = 0;
    for (uint32_t l = 0; l < nl; ++l)
      if (L[l].rank > 0) {
        start[L[l].b2 + 1]++;
        ++nlr;
      }
    for (uint32_t c = 0; c < n_cols; ++c) start[c + 1] += start[c];
    lr.resize(nlr);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t l = 0; l < nl; ++l)
      if (L[l].rank > 0) lr[fill[L[l].b2]++] = l;
    for (uint32_t c = 0; c < n_cols; ++c)
      for (uint32_t q = start[c] + 1; q < start[c + 1]; ++q) {  // insertion sort, stable
        const uint32_t v = lr[q];
        uint32_t w = q;
        for (; w > start[c] && L[lr[w - 1]].n2 > L[v].n2; --w) lr[w] = lr[w - 1];
        lr[w] = v;
      }
  }
  // phase A groups: runs of equal (b2, n2); K = sum of their ranks
  struct GroupA {
    size_t i, j;           // lr[i, j)
    uint64_t K, arena, zoff, job, task, t0;
  };
  std::vector<GroupA> ga;
  std::vector<uint64_t> wa(1, 0);
  uint64_t arena = 0, t_len = 0, nz = 0, nj = 0, nt = 0;
  for (size_t i = 0; i < lr.size();) {
    size_t j = i;
    const ahm_leaf_desc& d0 = L[lr[i]];
    uint64_t K = 0, jobs = 0;
    while (j < lr.size() && L[lr[j]].b2 == d0.b2 && L[lr[j]].n2 == d0.n2) {
      const uint64_t k = (uint64_t)L[lr[j]].rank;
      jobs += (K + k - 1) / TM - K / TM + 1;  // panels the leaf's rows fall into
      K += k;
      ++j;
    }
    const uint64_t ntile = (K + TM - 1) / TM;
    ga.push_back(GroupA{i, j, K, arena, nz, nj, nt, t_len});
    arena += K * d0.n2;
    nz += ntile * d0.n2;
    nj += jobs;
    nt += ntile;
    t_len += K;
    wa.push_back(wa.back() + ntile * d0.n2 + 8 * jobs);
    i = j;
  }
  if (t_len >= kFromT) return fail(AHM_ERR_UNSUPPORTED, "sum of ranks %llu too large", (unsigned long long)t_len);
  std::vector<uint32_t> toff(nl, 0);
  P.a_panels.resize(nt);
  // row segments: cut at every leaf row boundary
  std::vector<uint32_t> seg_at((size_t)n_rows + 1, 0);  // 1 at cuts, then the cut's index
  seg_at[0] = seg_at[n_rows] = 1;
  for (uint32_t l = 0; l < nl; ++l) seg_at[L[l].b1] = seg_at[L[l].b1 + L[l].n1] = 1;
  std::vector<uint32_t> cuts;
  for (uint32_t r = 0; r <= n_rows; ++r)
    if (seg_at[r]) {
      seg_at[r] = (uint32_t)cuts.size();
      cuts.push_back(r);
    }
  const size_t nseg = cuts.size() - 1;
  // leaves covering each segment (leaf order = reference traversal order)
  std::vector<uint32_t> cnt(nseg + 1, 0);
  for (uint32_t l = 0; l < nl; ++l) {
    if (L[l].rank == 0) continue;
    const uint32_t s0 = seg_at[L[l].b1], s1 = seg_at[L[l].b1 + L[l].n1];
    for (uint32_t q = s0; q < s1; ++q) cnt[q + 1]++;
  }
  for (size_t q = 0; q < nseg; ++q) cnt[q + 1] += cnt[q];
  std::vector<uint32_t> cover(cnt[nseg]);
  {
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (uint32_t l = 0; l < nl; ++l) {
      if (L[l].rank == 0) continue;
      const uint32_t s0 = seg_at[L[l].b1], s1 = seg_at[L[l].b1 + L[l].n1];
      for (uint32_t q = s0; q < s1; ++q) cover[fill[q]++] = l;
    }
  }
  // phase B segment sizes: W columns, one panel per 256 rows
  struct SegB {
    uint64_t arena, zoff, job, task;
    uint32_t W;
  };
  std::vector<SegB> sb(nseg);
  std::vector<uint64_t> wb(nseg + 1, 0);
  uint64_t ntb = 0;
  for (size_t q = 0; q < nseg; ++q) {
    uint64_t W = 0;
    for (uint32_t c = cnt[q]; c < cnt[q + 1]; ++c) {
      const ahm_leaf_desc& d = L[cover[c]];
      W += d.rank < 0 ? d.n2 : (uint32_t)d.rank;
    }
    if (W > 0xFFFFFFFFull) return fail(AHM_ERR_UNSUPPORTED, "row segment %zu has %llu columns", q, (unsigned long long)W);
    const uint64_t rows = cuts[q + 1] - cuts[q], ntile = W ? (rows + TM - 1) / TM : 0;
    const uint64_t jobs = ntile * (cnt[q + 1] - cnt[q]);
    sb[q] = SegB{arena, nz, nj, ntb, (uint32_t)W};
    arena += rows * W;
    nz += ntile * W;
    nj += jobs;
    ntb += ntile;
    wb[q + 1] = wb[q] + ntile * W + 8 * jobs;
  }
  P.zsrc.resize(nz);
  P.jobs.resize(nj);
  P.panels.resize(ntb);
  uint32_t* const Z = P.zsrc.data();
  CopyJob* const J = P.jobs.data();
  // phase A fill: V^T panel of K rows, tiled into <= TM rows
  parallel_by_weight(wa, [&](size_t g0, size_t g1) {
    for (size_t g = g0; g < g1; ++g) {
      const GroupA& G = ga[g];
      const uint32_t b2 = L[lr[G.i]].b2, n2 = L[lr[G.i]].n2;
      uint64_t K = 0;
      for (size_t q = G.i; q < G.j; ++q) {
        toff[lr[q]] = (uint32_t)(G.t0 + K);
        K += (uint64_t)L[lr[q]].rank;
      }
      uint64_t ar = G.arena, zo = G.zoff, jo = G.job, ta = G.task;
      for (uint64_t r0 = 0; r0 < K; r0 += TM) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(TM, K - r0);
        GemvTask tk{};
        tk.data = ar;
        tk.m = m;
        tk.W = n2;
        tk.zoff = zo;
        tk.out = (uint32_t)(G.t0 + r0);
        for (uint32_t c = 0; c < n2; ++c) Z[zo + c] = b2 + c;
        zo += n2;
        // copy the V columns of the leaves that fall into rows [r0, r0+m)
        uint64_t rowbase = 0;
        for (size_t q = G.i; q < G.j; ++q) {
          const ahm_leaf_desc& d = L[lr[q]];
          const uint64_t k = (uint64_t)d.rank;
          const uint64_t lo = std::max(rowbase, r0), hi = std::min(rowbase + k, r0 + m);
          if (lo < hi) {
            CopyJob cj{};
            cj.src = d.offset + (uint64_t)d.n1 * k + (lo - rowbase) * d.n2;
            cj.src_ld = d.n2;
            cj.rows = d.n2;
            cj.cols = (uint32_t)(hi - lo);
            cj.dst = ar + (lo - r0);
            cj.dst_ld = m;
            cj.transpose = 1;
            J[jo++] = cj;
          }
          rowbase += k;
        }
        P.a_panels[ta++] = tk;
        ar += (uint64_t)m * n2;
      }
    }
  });
  // phase B fill
  parallel_by_weight(wb, [&](size_t q0, size_t q1) {
    for (size_t q = q0; q < q1; ++q) {
      const SegB& S = sb[q];
      if (!S.W) continue;
      const uint32_t r0s = cuts[q], r1s = cuts[q + 1];
      uint64_t ar = S.arena, zo = S.zoff, jo = S.job, ta = S.task;
      for (uint32_t r0 = r0s; r0 < r1s; r0 += TM) {
        const uint32_t m = std::min(TM, r1s - r0);
        GemvTask tk{};
        tk.data = ar;
        tk.m = m;
        tk.zoff = zo;
        tk.out = r0;
        uint32_t W = 0;
        for (uint32_t c = cnt[q]; c < cnt[q + 1]; ++c) {
          const ahm_leaf_desc& d = L[cover[c]];
          const uint32_t cols = d.rank < 0 ? d.n2 : (uint32_t)d.rank;
          CopyJob cj{};
          cj.src = d.offset + (r0 - d.b1);
          cj.src_ld = d.n1;
          cj.rows = m;
          cj.cols = cols;
          cj.dst = ar + (uint64_t)W * m;
          cj.dst_ld = m;
          cj.transpose = 0;
          if (d.flags & AHM_LEAF_HEM) {   // expanded from the packed upper triangle
            cj.src = d.offset;
            cj.transpose = 2;
            cj.pad = r0 - d.b1;
          }
          J[jo++] = cj;
          if (d.rank < 0)
            for (uint32_t k = 0; k < cols; ++k) Z[zo + k] = d.b2 + k;
          else
            for (uint32_t k = 0; k < cols; ++k) Z[zo + k] = kFromT | (toff[cover[c]] + k);
          zo += cols;
          W += cols;
        }
        tk.W = W;
        tk.partial = 0;
        P.panels[ta++] = tk;
        ar += (uint64_t)m * W;
      }
    }
  });
  // wide panels of both phases cut into column chunks, biggest tasks first
  // (the hardware dispatches workgroups in order); the two phases use the
  // partial-sum slots in turn
  uint32_t part_a = 0, part_b = 0;
  chunk_panels(P.a_panels, kChunkValues, P.a, P.ra, part_a);
  chunk_panels(P.panels, kChunkValues, P.b, P.r, part_b);
  P.part_len = std::max(part_a, part_b);
  P.nvals = arena;
  P.t_len = (uint32_t)t_len;
  for (uint32_t l = 0; l < nl; ++l) {
    if (L[l].rank < 0) {
      P.n_dense++;
      P.nv_dense += (uint64_t)L[l].n1 * L[l].n2;
    } else {
      P.n_lowrank++;
      P.max_rank = std::max(P.max_rank, (uint32_t)L[l].rank);
      P.nv_u += (uint64_t)L[l].rank * L[l].n1;
      P.nv_v += (uint64_t)L[l].rank * L[l].n2;
    }
  }
  return AHM_OK;
}

template <typename T, typename A>
static int upload_vec(DeviceBuf& b, const std::vector<T, A>& v, hipStream_t s) {
  AHM_TRY(b.alloc(v.size() * sizeof(T)));
  if (!v.empty()) AHM_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return AHM_OK;
}

// The MFMA launches (multi-RHS products of 9 or more right-hand sides) run
// their own chunking of the panels: a task there carries 32 times the
// arithmetic and output per stored value of a single-RHS task, so row panels
// run whole (kMmChunkB is never reached) and V^T panels in chunks of about
// kMmChunkA values, which cuts the partial sums written and reduced.
// Each task of more than kMmSplitTiles row tiles runs as two workgroups
// (mm_piece); the map lists the workgroups as (task << 2) | piece.
constexpr uint64_t kMmChunkA = 512 * 1024;
constexpr uint64_t kMmChunkB = 1ull << 40;

static void mm_map(const std::vector<GemvTask>& T, std::vector<uint32_t>& map) {
  map.clear();
  map.reserve(T.size() * 2);
  for (size_t i = 0; i < T.size(); ++i) {
    const uint32_t e = (uint32_t)i << 2;
    if ((T[i].m + 15) / 16 > kMmSplitTiles) {
      map.push_back(e | 1u);
      map.push_back(e | 2u);
    } else {
      map.push_back(e);
    }
  }
}

// the MFMA task lists of h's panels; widens part_len to their partial sums
static int build_mm_tasks(ahm_hmat h, hipStream_t s) {
  std::vector<GemvTask> a, b;
  std::vector<ReduceTask> ra, r;
  uint32_t part_a = 0, part_b = 0;
  chunk_panels(h->h_panels_a, kMmChunkA, a, ra, part_a);
  chunk_panels(h->h_panels, kMmChunkB, b, r, part_b);
  if (a.size() >= (1u << 30) || b.size() >= (1u << 30))
    return fail(AHM_ERR_UNSUPPORTED, "too many panel tasks for the MFMA workgroup map");
  std::vector<uint32_t> ma, mb;
  mm_map(a, ma);
  mm_map(b, mb);
  AHM_TRY(upload_vec(h->mm_tasks_a, a, s));
  AHM_TRY(upload_vec(h->mm_tasks_b, b, s));
  AHM_TRY(upload_vec(h->mm_tasks_ra, ra, s));
  AHM_TRY(upload_vec(h->mm_tasks_r, r, s));
  AHM_TRY(upload_vec(h->mm_map_a, ma, s));
  AHM_TRY(upload_vec(h->mm_map_b, mb, s));
  AHM_HIP(hipStreamSynchronize(s));
  h->n_mm_a = (uint32_t)ma.size();
  h->n_mm_ta = (uint32_t)a.size();
  h->h_mm_tasks_a = a;
  h->n_mm_b = (uint32_t)mb.size();
  h->n_mm_ra = (uint32_t)ra.size();
  h->n_mm_r = (uint32_t)r.size();
  const uint32_t pl = std::max(part_a, part_b);
  if (pl > h->part_len) {
    h->part_len = pl;
    h->pbuf_cap = 0;
    h->pbuf.release();
  }
  return AHM_OK;
}

// multi-RHS products keep x and t interleaved in groups of 32 right-hand
// sides (8 when p <= 8): rhs_width values per row and group
static uint32_t rhs_width(uint32_t p) { return p <= 8 ? 8u : (uint32_t)kMmRhs; }
static uint32_t rhs_slots(uint32_t p) {
  return p == 1 ? 1u : ((p + kMmRhs - 1) / kMmRhs) * rhs_width(p);
}

static size_t ws_t_bytes(ahm_hmat h, uint32_t p, size_t esize) {
  return (size_t)h->t_len * rhs_slots(p) * esize;
}
static size_t ws_part_bytes(ahm_hmat h, uint32_t p, size_t esize) {
  return (size_t)h->part_len * p * esize;
}
static size_t ws_x_bytes(ahm_hmat h, uint32_t p, size_t esize) {
  return p == 1 ? 0 : (size_t)h->n_cols * rhs_slots(p) * esize;
}

static int ensure_tbuf(ahm_hmat h, uint32_t p, size_t esize) {
  const size_t need = ws_t_bytes(h, p, esize);
  if (need > h->tbuf_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(h->tbuf.alloc(need));
    h->tbuf_cap = need;
  }
  const size_t xbneed = h->storage == AHM_BF16 && p >= 9 ? ws_x_bytes(h, p, esize) / 2 : 0;
  if (xbneed > h->xbbuf_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(h->xbbuf.alloc(xbneed));
    h->xbbuf_cap = xbneed;
  }
  const size_t xneed = ws_x_bytes(h, p, esize);
  if (xneed > h->xbuf_cap) {
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
    AHM_TRY(h->xbuf.alloc(xneed));
    h->xbuf_cap = xneed;
  }
  return AHM_OK;
}

static int ensure_pbuf(ahm_hmat h, uint32_t p, size_t esize) {
  const size_t need = ws_part_bytes(h, p, esize);
  if (need <= h->pbuf_cap) return AHM_OK;
  AHM_HIP(hipStreamSynchronize(h->ctx->stream));
  AHM_TRY(h->pbuf.alloc(need));
  h->pbuf_cap = need;
  return AHM_OK;
}

// Device workspace of one product: V^T X (t_len values per right-hand-side
// slot), the partial sums of chunked panels (part_len per right-hand side)
// and, for p >= 2, the interleaved copy of X (n_cols per slot).
struct Workspace {
  void* t;
  void* part;
  void* xi;
  void* xb;   // BF16 copy of xi (AHM_BF16 storage, p >= 9)
  // single-RHS products of a solver: gates of the phase-A, phase-B,
  // phase-A reduction and phase-B reduction tasks (none by default)
  Gate gate[4];
};

static Workspace own_workspace(ahm_hmat h) { return Workspace{h->tbuf.p, h->pbuf.p, h->xbuf.p, h->xbbuf.p}; }

// one persistent cooperative launch (form 2 of ahm_hmat_set_fused); the grid
// is the device's resident capacity for the kernel, at most the largest
// phase's task count
template <typename TS, typename TV>
static int launch_persistent(ahm_hmat h, const Workspace& ws, TV alpha, const TV* X, uint32_t ldx, TV* Y,
                             uint32_t ldy, uint32_t p) {
  const bool multi = p > 1;
  const void* kern = multi ? reinterpret_cast<const void*>(&persistent_kernel<TS, TV, true>)
                           : reinterpret_cast<const void*>(&persistent_kernel<TS, TV, false>);
  int& cap = h->pgrid[(sizeof(TV) == 8 ? 0 : 2) + (multi ? 1 : 0)];
  if (!cap) {
    int nb = 0, ncu = 0;
    AHM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kPanelThreads, 0));
    AHM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->ctx->device));
    if (nb <= 0 || ncu <= 0) return fail(AHM_ERR_UNSUPPORTED, "persistent product: no resident workgroups");
    cap = nb * ncu;
  }
  uint64_t work = std::max<uint64_t>({h->n_a, h->n_b, (uint64_t)h->n_ra * p, (uint64_t)h->n_r * p,
                                      multi ? (h->n_cols + 63) / 64 : 0, 1});
  const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)cap, work);
  PersistArgs<TS, TV> a{(const GemvTask*)h->tasks_a.p, (const GemvTask*)h->tasks_b.p,
                        (const ReduceTask*)h->tasks_ra.p, (const ReduceTask*)h->tasks_r.p,
                        (const uint32_t*)h->zsrc.p, (const TS*)h->arena.p, X, Y, (TV*)ws.t, (TV*)ws.part,
                        (TV*)ws.xi, ldx, ldy, p, h->n_cols, h->t_len, h->part_len,
                        h->n_a, h->n_b, h->n_ra, h->n_r, alpha, ws.gate[0].frozen};
  void* args[] = {&a};
  AHM_HIP(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(kPanelThreads), args, 0, h->ctx->stream));
  return AHM_OK;
}

template <typename TS, typename TV>
static int launch_matvec(ahm_hmat h, const Workspace& ws, TV alpha, const TV* X, uint32_t ldx, TV* Y,
                         uint32_t ldy, uint32_t p, hipEvent_t ev_mid = nullptr) {
  hipStream_t s = h->ctx->stream;
  TV* t = (TV*)ws.t;
  TV* part = (TV*)ws.part;
  const ahm_dtype vec = std::is_same<TV, double>::value ? AHM_F64 : AHM_F32;
  // symmetric storage, one right-hand side: the transposed part rides on the
  // row-panel pass (after sym_prepare); wider products add it afterwards
  const bool sym1 = p == 1 && h->sym && h->tr.sym_built;
  // the phase-B reduction: the MFMA launches have their own
  const ReduceTask* reduce_b = (const ReduceTask*)h->tasks_r.p;
  uint32_t n_reduce_b = h->n_r;
  if (h->fused == 2 && p <= 8 && !ws.gate[0].seg && !h->sym && !ev_mid) {
    // a cooperative launch is not recorded into a graph: captured products
    // take the separate launches
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AHM_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) return launch_persistent<TS, TV>(h, ws, alpha, X, ldx, Y, ldy, p);
  }
  if (h->fused == 1 && p <= 8 && !ws.gate[0].frozen && !h->sym) {
    if (h->fz_n)
      hipLaunchKernelGGL((fused_kernel<TS, TV>), dim3(h->fz_n), dim3(kPanelThreads), 0, s,
                         (const GemvTask*)h->fz_panels.p, (const GemvTask*)h->fz_apanels.p,
                         (const FusedGroup*)h->fz_groups.p, (const uint32_t*)h->fz_goff.p,
                         (const uint32_t*)h->zsrc.p, (const TS*)h->arena.p, X, ldx, Y, ldy, alpha, p);
    if (ev_mid) AHM_HIP(hipEventRecord(ev_mid, s));
    AHM_HIP(hipGetLastError());
    return AHM_OK;
  }
  if (p == 1) {
    if (h->n_a)
      hipLaunchKernelGGL((panel_gemv_kernel<TS, TV, false>), dim3(h->n_a, 1), dim3(kPanelThreads), 0, s,
                         (const GemvTask*)h->tasks_a.p, (const uint32_t*)h->zsrc.p,
                         (const TS*)h->arena.p, X, ldx, t, h->t_len, Y, ldy, part, h->part_len, alpha,
                         ws.gate[0]);
    if (h->n_ra)
      hipLaunchKernelGGL((chunk_reduce_kernel<TV, true>), dim3(h->n_ra, 1), dim3(kPanelThreads), 0, s,
                         (const ReduceTask*)h->tasks_ra.p, (const TV*)part, h->part_len, t, h->t_len, alpha, 0,
                         ws.gate[2]);
    if (ev_mid) AHM_HIP(hipEventRecord(ev_mid, s));
    if (sym1)   // both directions of the row panels in one pass
      AHM_TRY(sym_phase_b(h, (double)alpha, X, t, Y, part, vec, ws.gate[1]));
    else if (h->n_b)
      hipLaunchKernelGGL((panel_gemv_kernel<TS, TV, true>), dim3(h->n_b, 1), dim3(kPanelThreads), 0, s,
                         (const GemvTask*)h->tasks_b.p, (const uint32_t*)h->zsrc.p,
                         (const TS*)h->arena.p, X, ldx, t, h->t_len, Y, ldy, part, h->part_len, alpha,
                         ws.gate[1]);
  } else {
    // p >= 2 on interleaved copies of x and t: p <= 8 with the streaming
    // kernel, wider with MFMA panel products (32 right-hand sides per
    // workgroup)
    const uint32_t ng = (p + kMmRhs - 1) / kMmRhs;
    const int pw = (int)rhs_width(p);
    TV* xi = (TV*)ws.xi;
    hipLaunchKernelGGL((interleave_kernel<TV>), dim3((h->n_cols + 63) / 64, ng), dim3(256), 0, s, X,
                       ldx, h->n_cols, p, pw, xi);
    if (pw == 8) {
      auto run = [&](auto np) {
        constexpr int NP = decltype(np)::value;
        if (h->n_a)
          hipLaunchKernelGGL((panel_gemvp_kernel<TS, TV, false, NP>), dim3(h->n_a), dim3(kPanelThreads), 0,
                             s, (const GemvTask*)h->tasks_a.p, (const uint32_t*)h->zsrc.p,
                             (const TS*)h->arena.p, (const TV*)xi, t, h->t_len, Y, ldy, part,
                             h->part_len, alpha, p);
        if (h->n_ra)
          hipLaunchKernelGGL((chunk_reduce_kernel<TV, true>), dim3(h->n_ra, p), dim3(kPanelThreads), 0, s,
                             (const ReduceTask*)h->tasks_ra.p, (const TV*)part, h->part_len, t,
                             h->t_len, alpha, pw, Gate{});
        if (ev_mid) (void)hipEventRecord(ev_mid, s);
        if (h->n_b)
          hipLaunchKernelGGL((panel_gemvp_kernel<TS, TV, true, NP>), dim3(h->n_b), dim3(kPanelThreads), 0,
                             s, (const GemvTask*)h->tasks_b.p, (const uint32_t*)h->zsrc.p,
                             (const TS*)h->arena.p, (const TV*)xi, t, h->t_len, Y, ldy, part,
                             h->part_len, alpha, p);
      };
      run(std::integral_constant<int, 8>());
    } else {
      bool bf16_a = false;
      if constexpr (std::is_same<TV, float>::value) bf16_a = h->storage == AHM_BF16;
      if (h->n_mm_ta && bf16_a) {
        // V^T X on BF16 MFMA: X rounded to BF16, FP32 accumulation
        hipLaunchKernelGGL(bf16_copy_kernel, dim3(2048), dim3(256), 0, s, (const float*)xi, (uint16_t*)ws.xb,
                           (uint64_t)h->n_cols * kMmRhs * ng);
        // one workgroup per MFMA phase-A task (the BF16 copy follows that list)
        hipLaunchKernelGGL(panel_bf16_kernel, dim3(h->n_mm_ta, ng), dim3(kPanelThreads), 0, s,
                           (const GemvTask*)h->mm_tasks_a.p, (const uint64_t*)h->bfoff.p,
                           (const uint32_t*)h->zsrc.p, (const uint16_t*)h->bfarena.p,
                           (const uint16_t*)ws.xb, h->n_cols, (float*)t, h->t_len, (float*)part,
                           h->part_len, p);
      } else if (h->n_a) {
        hipLaunchKernelGGL((panel_mm_kernel<TV, false>), dim3(h->n_mm_a, ng), dim3(kPanelThreads), 0, s,
                           (const GemvTask*)h->mm_tasks_a.p, (const uint32_t*)h->mm_map_a.p,
                           (const uint32_t*)h->zsrc.p,
                           (const TS*)h->arena.p, xi, h->n_cols, t, h->t_len, Y, ldy, part,
                           h->part_len, alpha, p);
      }
      // the BF16 V^T X tasks keep the single-RHS chunking (their copy is
      // laid out per task)
      const uint32_t nra = h->n_mm_ra;
      if (nra)
        hipLaunchKernelGGL((chunk_reduce_kernel<TV, true>), dim3(nra, p), dim3(kPanelThreads), 0, s,
                           (const ReduceTask*)h->mm_tasks_ra.p, (const TV*)part,
                           h->part_len, t, h->t_len, alpha, pw, Gate{});
      if (ev_mid) AHM_HIP(hipEventRecord(ev_mid, s));
      if (h->n_b)
        hipLaunchKernelGGL((panel_mm_kernel<TV, true>), dim3(h->n_mm_b, ng), dim3(kPanelThreads), 0, s,
                           (const GemvTask*)h->mm_tasks_b.p, (const uint32_t*)h->mm_map_b.p,
                           (const uint32_t*)h->zsrc.p,
                           (const TS*)h->arena.p, xi, h->n_cols, t, h->t_len, Y, ldy, part,
                           h->part_len, alpha, p);
      reduce_b = (const ReduceTask*)h->mm_tasks_r.p;
      n_reduce_b = h->n_mm_r;
    }
  }
  if (n_reduce_b)
    hipLaunchKernelGGL((chunk_reduce_kernel<TV, false>), dim3(n_reduce_b, p), dim3(kPanelThreads), 0, s,
                       reduce_b, (const TV*)part, h->part_len, Y, ldy, alpha, 0,
                       p == 1 ? ws.gate[3] : Gate{});
  AHM_HIP(hipGetLastError());
  if (sym1) AHM_TRY(sym_rest(h, (double)alpha, Y, vec, ws.gate[3]));
  return AHM_OK;
}

// One product on a workspace the caller provides (bytes at base, 256-byte
// aligned): t, then the partial sums, then the interleaved X, each rounded up
// to 256 bytes.  Allocates nothing; refuses workspaces that are too small.
int matvec_on_workspace(ahm_hmat h, void* base, size_t bytes, double alpha, const void* X, uint32_t ldx,
                        void* Y, uint32_t ldy, uint32_t p, ahm_dtype vec, const char* who) {
  if (!h || !X || !Y) return fail(AHM_ERR_INVALID, "%s: NULL argument", who);
  if (p == 0) return AHM_OK;
  if (p > 65535) return fail(AHM_ERR_INVALID, "%s: p=%u > 65535", who, p);
  if (p > 1 && (ldx < h->n_cols || ldy < h->n_rows))
    return fail(AHM_ERR_INVALID, "%s: ldx=%u ldy=%u for %u x %u", who, ldx, ldy, h->n_rows, h->n_cols);
  // the static workspace carries no BF16 copy of X: F64 and F32 storage only
  if (!((h->storage == AHM_F64 && vec == AHM_F64) || (h->storage == AHM_F32 && vec == AHM_F32)))
    return fail(AHM_ERR_UNSUPPORTED, "%s: storage %d with vectors %d", who, h->storage, vec);
  if (h->sym) return fail(AHM_ERR_UNSUPPORTED, "%s: symmetric matrices need the transposed part's buffers", who);
  const size_t es = dtype_size(vec);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t tb = up(ws_t_bytes(h, p, es)), pb = up(ws_part_bytes(h, p, es)), xb = up(ws_x_bytes(h, p, es));
  if (tb + pb + xb > bytes)
    return fail(AHM_ERR_INVALID,
                "%s: this product needs %zu bytes of workspace (V^T x %zu, partial sums %zu, "
                "interleaved x %zu), more than the %zu available",
                who, tb + pb + xb, tb, pb, xb, bytes);
  char* b = (char*)base;
  const Workspace ws{b, b + tb, b + tb + pb, nullptr};
  AHM_HIP(hipSetDevice(h->ctx->device));
  if (vec == AHM_F64)
    return launch_matvec<double, double>(h, ws, alpha, (const double*)X, ldx, (double*)Y, ldy, p);
  return launch_matvec<float, float>(h, ws, (float)alpha, (const float*)X, ldx, (float*)Y, ldy, p);
}

int matvec_prepare(ahm_hmat h, ahm_dtype vec) {
  AHM_HIP(hipSetDevice(h->ctx->device));
  AHM_TRY(ensure_tbuf(h, 1, dtype_size(vec)));
  AHM_TRY(ensure_pbuf(h, 1, dtype_size(vec)));
  return sym_prepare(h, vec);
}

int matvec_gated(ahm_hmat h, double alpha, const void* x, void* y, ahm_dtype vec, const uint32_t* frozen,
                 const uint32_t* const seg[4]) {
  const size_t es = dtype_size(vec);
  if (ws_t_bytes(h, 1, es) > h->tbuf_cap || ws_part_bytes(h, 1, es) > h->pbuf_cap)
    return fail(AHM_ERR_INVALID, "gated product: workspace not prepared");
  Workspace ws = own_workspace(h);
  for (int q = 0; q < 4; ++q) ws.gate[q] = Gate{frozen, seg ? seg[q] : nullptr};
  if (h->sym && seg) return fail(AHM_ERR_UNSUPPORTED, "gated product: symmetric matrices are single systems");
  if (h->storage == AHM_F64 && vec == AHM_F64)
    AHM_TRY((launch_matvec<double, double>(h, ws, alpha, (const double*)x, h->n_cols, (double*)y, h->n_rows,
                                           1)));
  else if (h->storage != AHM_F64 && vec == AHM_F32)
    AHM_TRY((launch_matvec<float, float>(h, ws, (float)alpha, (const float*)x, h->n_cols, (float*)y,
                                         h->n_rows, 1)));
  else
    return fail(AHM_ERR_UNSUPPORTED, "gated product: storage %d with vectors %d", h->storage, vec);
  if (h->sym && !h->tr.sym_built) return fail(AHM_ERR_INVALID, "gated product: symmetric part not prepared");
  return AHM_OK;
}

void matvec_signature(ahm_hmat h, std::vector<uint64_t>& sig) {
  sig = {(uint64_t)(uintptr_t)h->tasks_a.p, (uint64_t)(uintptr_t)h->tasks_b.p,
         (uint64_t)(uintptr_t)h->tasks_ra.p, (uint64_t)(uintptr_t)h->tasks_r.p,
         (uint64_t)(uintptr_t)h->zsrc.p,    (uint64_t)(uintptr_t)h->arena.p,
         (uint64_t)(uintptr_t)h->tbuf.p,    (uint64_t)(uintptr_t)h->pbuf.p,
         h->n_a, h->n_b, h->n_ra, h->n_r, h->t_len, h->part_len, h->n_rows, h->n_cols,
         (uint64_t)h->storage, h->chunk_values,
         // the transposed part of symmetric products
         (uint64_t)h->sym, (uint64_t)(uintptr_t)h->tr.contrib.p, (uint64_t)(uintptr_t)h->tr.tprime.p,
         (uint64_t)(uintptr_t)h->tr.tasks_a.p, (uint64_t)(uintptr_t)h->tr.tasks_b.p,
         (uint64_t)(uintptr_t)h->tr.sym_tasks.p, (uint64_t)h->tr.n_sym};
}

}  // namespace ahm

namespace ahm {

// re-cut the panels of both phases into chunks of about cv values (the arena
// is unchanged: chunks are column ranges of the same panels)
// BF16 copy of the phase-A tasks for AHM_BF16 storage (after every change
// of the tasks): each task's m x W values in ceil(W/8) k-groups of m rows of 8
// columns, at an offset that is a multiple of 8 values; zeroed slack after the
// last copy, since the last row tile of a task reads up to 15 rows past m.
static int build_bf16(ahm_hmat h) {
  const std::vector<GemvTask>& A = h->h_mm_tasks_a;
  std::vector<uint64_t> off(A.size());
  uint64_t tot = 0;
  for (size_t i = 0; i < A.size(); ++i) {
    const GemvTask& tk = A[i];
    if (tk.m > (uint32_t)kPanelThreads)
      return fail(AHM_ERR_UNSUPPORTED, "BF16 V^T X: phase-A task of %u rows", tk.m);
    for (uint32_t c = 0; c < tk.W; ++c)
      if (h->h_zsrc[tk.zoff + c] >= h->n_cols)
        return fail(AHM_ERR_INVALID, "BF16 V^T X: phase-A column source out of range");
    off[i] = tot;
    tot += (uint64_t)tk.m * ((tk.W + 7) & ~7u);
  }
  constexpr uint64_t kSlack = 256;
  hipStream_t s = h->ctx->stream;
  AHM_HIP(hipStreamSynchronize(s));
  AHM_TRY(h->bfarena.alloc((tot + kSlack) * sizeof(uint16_t)));
  AHM_HIP(hipMemsetAsync((uint16_t*)h->bfarena.p + tot, 0, kSlack * sizeof(uint16_t), s));
  AHM_TRY(upload_vec(h->bfoff, off, s));
  if (!A.empty())
    hipLaunchKernelGGL(bf16_panels_kernel, dim3((uint32_t)A.size()), dim3(256), 0, s,
                       (const GemvTask*)h->mm_tasks_a.p, (const uint64_t*)h->bfoff.p, (const float*)h->arena.p,
                       (uint16_t*)h->bfarena.p);
  AHM_HIP(hipGetLastError());
  AHM_HIP(hipStreamSynchronize(s));
  h->bf_vals = tot;
  return AHM_OK;
}

// Plan of the single-launch product: the unchunked row panels (largest
// work first), the unchunked V^T panels, and per row panel its groups of U
// columns that take consecutive rows of one V^T panel (one group per
// low-rank leaf).
static int build_fused(ahm_hmat h) {
  const std::vector<GemvTask>& B = h->h_panels;
  const std::vector<GemvTask>& A = h->h_panels_a;
  const uint32_t none = 0xFFFFFFFFu;
  std::vector<uint32_t> ta(h->t_len, none), tr(h->t_len, 0);
  for (size_t a = 0; a < A.size(); ++a) {
    const GemvTask& tk = A[a];
    if (tk.partial || (uint64_t)tk.out + tk.m > h->t_len)
      return fail(AHM_ERR_INVALID, "single launch: V^T panel %zu out of range", a);
    // the kernel takes the panel's x entries as a contiguous range
    const uint32_t x0 = tk.W ? h->h_zsrc[tk.zoff] : 0;
    if ((uint64_t)x0 + tk.W > h->n_cols) return fail(AHM_ERR_INVALID, "single launch: V^T panel %zu columns", a);
    for (uint32_t c = 0; c < tk.W; ++c)
      if (h->h_zsrc[tk.zoff + c] != x0 + c)
        return fail(AHM_ERR_UNSUPPORTED, "single launch: V^T panel %zu columns are not one range", a);
    for (uint32_t r = 0; r < tk.m; ++r) {
      ta[tk.out + r] = (uint32_t)a;
      tr[tk.out + r] = r;
    }
  }
  std::vector<uint64_t> cost(B.size(), 0);
  std::vector<std::vector<FusedGroup>> per(B.size());
  for (size_t b = 0; b < B.size(); ++b) {
    const GemvTask& tk = B[b];
    if (tk.partial || tk.m == 0 || tk.m > (uint32_t)kPanelThreads || (uint64_t)tk.out + tk.m > h->n_rows)
      return fail(AHM_ERR_INVALID, "single launch: row panel %zu", b);
    cost[b] = (uint64_t)tk.m * tk.W;
    for (uint32_t c = 0; c < tk.W; ++c) {
      const uint32_t src = h->h_zsrc[tk.zoff + c];
      if (!(src & kFromT)) continue;
      const uint32_t t = src & ~kFromT;
      if (t >= h->t_len || ta[t] == none) return fail(AHM_ERR_INVALID, "single launch: t source %u", t);
      std::vector<FusedGroup>& gv = per[b];
      if (!gv.empty()) {
        FusedGroup& l = gv.back();
        if (l.apanel == ta[t] && l.row0 + l.k == tr[t] && l.col + l.k == c) {
          ++l.k;
          continue;
        }
      }
      gv.push_back(FusedGroup{c, 1, ta[t], tr[t]});
    }
    for (const FusedGroup& fg : per[b]) cost[b] += (uint64_t)fg.k * A[fg.apanel].W;
  }
  std::vector<uint32_t> order(B.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  std::vector<GemvTask> panels(B.size());
  std::vector<FusedGroup> groups;
  std::vector<uint32_t> goff(B.size() + 1, 0);
  for (size_t q = 0; q < order.size(); ++q) {
    panels[q] = B[order[q]];
    goff[q] = (uint32_t)groups.size();
    groups.insert(groups.end(), per[order[q]].begin(), per[order[q]].end());
  }
  goff[B.size()] = (uint32_t)groups.size();
  hipStream_t s = h->ctx->stream;
  AHM_HIP(hipStreamSynchronize(s));
  AHM_TRY(upload_vec(h->fz_panels, panels, s));
  AHM_TRY(upload_vec(h->fz_apanels, A, s));
  AHM_TRY(upload_vec(h->fz_groups, groups, s));
  AHM_TRY(upload_vec(h->fz_goff, goff, s));
  AHM_HIP(hipStreamSynchronize(s));
  h->fz_n = (uint32_t)panels.size();
  h->fz_built = 1;
  return AHM_OK;
}

int set_chunking(ahm_hmat h, uint64_t cv) {
  std::vector<GemvTask> a, b;
  std::vector<ReduceTask> ra, r;
  uint32_t part_a = 0, part_b = 0;
  chunk_panels(h->h_panels_a, cv, a, ra, part_a);
  chunk_panels(h->h_panels, cv, b, r, part_b);
  hipStream_t s = h->ctx->stream;
  AHM_HIP(hipStreamSynchronize(s));
  AHM_TRY(upload_vec(h->tasks_a, a, s));
  AHM_TRY(upload_vec(h->tasks_ra, ra, s));
  AHM_TRY(upload_vec(h->tasks_b, b, s));
  AHM_TRY(upload_vec(h->tasks_r, r, s));
  AHM_HIP(hipStreamSynchronize(s));
  h->n_a = (uint32_t)a.size();
  h->n_ra = (uint32_t)ra.size();
  h->n_b = (uint32_t)b.size();
  h->n_r = (uint32_t)r.size();
  h->part_len = std::max(part_a, part_b);
  h->pbuf_cap = 0;  // partial-sum slots change with the chunking
  h->pbuf.release();
  h->h_tasks_a = std::move(a);
  h->h_tasks_ra = std::move(ra);
  h->h_tasks_b = std::move(b);
  h->h_tasks_r = std::move(r);
  AHM_TRY(build_mm_tasks(h, s));
  h->chunk_values = cv;
  h->tr.sym_built = false;   // the symmetric pass follows the phase-B tasks
  if (h->storage == AHM_BF16) AHM_TRY(build_bf16(h));
  return AHM_OK;
}

struct PlanJob {
  Plan P;
  int st = AHM_OK;
  std::string err;
  uint32_t n_rows = 0, n_cols = 0, nl = 0;
  const ahm_leaf_desc* leaves = nullptr;
  std::thread th;
  ~PlanJob() {
    if (th.joinable()) th.join();
  }
};

std::shared_ptr<PlanJob> plan_start(uint32_t n_rows, uint32_t n_cols, const std::vector<ahm_leaf_desc>& leaves) {
  auto job = std::make_shared<PlanJob>();
  PlanJob* j = job.get();
  j->n_rows = n_rows;
  j->n_cols = n_cols;
  j->nl = (uint32_t)leaves.size();
  j->leaves = leaves.data();
  j->th = std::thread([j]() {
    j->st = build_plan(j->n_rows, j->n_cols, j->nl, j->leaves, j->P);
    if (j->st) j->err = ahm_last_error();
  });
  return job;
}

int hmat_build(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols, std::vector<ahm_leaf_desc>&& leaves,
               DeviceBuf& raw, ahm_dtype data_dtype, ahm_dtype storage, int keep_leaves,
               ahm_hmat* out, std::shared_ptr<PlanJob> pre) {
  PhaseTimer timer("layout");
  Plan P;
  if (pre) {
    if (pre->th.joinable()) pre->th.join();
    if (pre->n_rows != n_rows || pre->n_cols != n_cols || pre->nl != leaves.size() || pre->leaves != leaves.data())
      return fail(AHM_ERR_INVALID, "layout plan of another leaf list");
    if (pre->st) return fail(pre->st, "%s", pre->err.c_str());
    P = std::move(pre->P);
  } else {
    AHM_TRY(build_plan(n_rows, n_cols, (uint32_t)leaves.size(), leaves.data(), P));
  }
  timer.mark("plan (panels, chunks, copy jobs)");
  for (const ahm_leaf_desc& d : leaves) {
    const uint64_t nv = leaf_values(d);
    if ((d.offset + nv) * dtype_size(data_dtype) > raw.bytes && nv)
      return fail(AHM_ERR_INVALID, "leaf values beyond the data buffer");
  }
  ahm_hmat h = new ahm_hmat_s;
  h->ctx = ctx;
  h->n_rows = n_rows;
  h->n_cols = n_cols;
  h->storage = storage;
  h->leaf_dtype = data_dtype;
  h->nvals = P.nvals;
  h->t_len = P.t_len;
  h->max_rank = P.max_rank;
  h->n_dense = P.n_dense;
  h->n_lowrank = P.n_lowrank;
  h->nv_dense = P.nv_dense;
  h->nv_u = P.nv_u;
  h->nv_v = P.nv_v;
  for (const ahm_leaf_desc& d : leaves)
    if (d.flags & AHM_LEAF_HEM) h->sym = 1;
  h->leaves = std::move(leaves);
  hipStream_t s = ctx->stream;
  const size_t ss = dtype_size(storage);
  int st = AHM_OK;
  DeviceBuf jobs;
  do {
    if ((st = h->arena.alloc((P.nvals + kArenaSlack) * ss))) break;
    if (hipMemsetAsync((char*)h->arena.p + P.nvals * ss, 0, kArenaSlack * ss, s) != hipSuccess) {
      st = fail(AHM_ERR_HIP, "arena slack");
      break;
    }
    timer.mark("arena allocation");
    if ((st = upload_vec(jobs, P.jobs, s))) break;
    if ((st = upload_vec(h->tasks_a, P.a, s))) break;
    if ((st = upload_vec(h->tasks_b, P.b, s))) break;
    if ((st = upload_vec(h->zsrc, P.zsrc, s))) break;
    if ((st = upload_vec(h->tasks_r, P.r, s))) break;
    if ((st = upload_vec(h->tasks_ra, P.ra, s))) break;
    h->n_a = (uint32_t)P.a.size();
    h->n_b = (uint32_t)P.b.size();
    h->n_r = (uint32_t)P.r.size();
    h->n_ra = (uint32_t)P.ra.size();
    h->part_len = P.part_len;
    h->n_zsrc = P.zsrc.size();
    if (timer.on && hipStreamSynchronize(s) == hipSuccess) timer.mark("task, source-list and job uploads");
    for (size_t j0 = 0; j0 < P.jobs.size(); j0 += (1u << 30)) {
      const uint32_t nj = (uint32_t)std::min<size_t>(1u << 30, P.jobs.size() - j0);
      const CopyJob* J = (const CopyJob*)jobs.p + j0;
      if (data_dtype == AHM_F64 && storage == AHM_F64)
        hipLaunchKernelGGL((relayout_kernel<double, double>), dim3(nj), dim3(256), 0, s, J,
                           (const double*)raw.p, (double*)h->arena.p);
      else if (data_dtype == AHM_F64)
        hipLaunchKernelGGL((relayout_kernel<float, double>), dim3(nj), dim3(256), 0, s, J,
                           (const double*)raw.p, (float*)h->arena.p);
      else
        hipLaunchKernelGGL((relayout_kernel<float, float>), dim3(nj), dim3(256), 0, s, J,
                           (const float*)raw.p, (float*)h->arena.p);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      st = fail(AHM_ERR_HIP, "H-matrix re-layout kernel failed");
      break;
    }
    timer.mark("re-layout kernel");
    h->h_tasks_a = std::move(P.a);
    h->h_tasks_b = std::move(P.b);
    h->h_tasks_r = std::move(P.r);
    h->h_tasks_ra = std::move(P.ra);
    h->h_panels = std::move(P.panels);
    h->h_panels_a = std::move(P.a_panels);
    h->chunk_values = kChunkValues;
    h->h_zsrc = std::move(P.zsrc);
    if ((st = build_mm_tasks(h, s))) break;
    if (keep_leaves) {
      h->leafdata.p = raw.p;
      h->leafdata.bytes = raw.bytes;
      h->leafdata.cap = raw.cap;
      h->leafdata.dev = raw.dev;
      raw.p = nullptr;
      raw.bytes = raw.cap = 0;
    }
  } while (0);
  if (st) {
    delete h;
    return st;
  }
  *out = h;
  return AHM_OK;
}

int hmat_from_device(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols,
                     std::vector<ahm_leaf_desc>&& leaves, DeviceBuf& raw, ahm_dtype storage,
                     int keep_leaves, ahm_hmat* out, std::shared_ptr<PlanJob> plan) {
  return hmat_build(ctx, n_rows, n_cols, std::move(leaves), raw, AHM_F64, storage, keep_leaves, out,
                    std::move(plan));
}

}  // namespace ahm

using namespace ahm;

extern "C" {

int ahm_hmat_upload(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols, uint32_t n_leaves,
                    const ahm_leaf_desc* leaves, const void* data, ahm_dtype data_dtype,
                    ahm_dtype storage, int keep_leaves, ahm_hmat* out) {
  if (!ctx || !out || (!leaves && n_leaves))
    return fail(AHM_ERR_INVALID, "ahm_hmat_upload: NULL argument");
  if ((data_dtype != AHM_F64 && data_dtype != AHM_F32) || (storage != AHM_F64 && storage != AHM_F32))
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_upload: dtypes %d -> %d", data_dtype, storage);
  if (data_dtype == AHM_F32 && storage == AHM_F64)
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_upload: float leaves cannot be stored as double");
  AHM_HIP(hipSetDevice(ctx->device));
  uint64_t total = 0;
  for (uint32_t l = 0; l < n_leaves; ++l) {
    const ahm_leaf_desc& d = leaves[l];
    const uint64_t nv = leaf_values(d);
    total = std::max(total, d.offset + nv);
  }
  if (total && !data) return fail(AHM_ERR_INVALID, "ahm_hmat_upload: data is NULL");
  DeviceBuf raw;
  AHM_TRY(raw.alloc(total * dtype_size(data_dtype)));
  if (total)
    AHM_HIP(hipMemcpyAsync(raw.p, data, total * dtype_size(data_dtype), hipMemcpyHostToDevice,
                           ctx->stream));
  std::vector<ahm_leaf_desc> L(leaves, leaves + n_leaves);
  return hmat_build(ctx, n_rows, n_cols, std::move(L), raw, data_dtype, storage, keep_leaves, out);
}

int ahm_hmat_convert(ahm_hmat src, ahm_dtype storage, ahm_hmat* out) {
  if (!src || !out) return fail(AHM_ERR_INVALID, "ahm_hmat_convert: NULL argument");
  if (src->storage != AHM_F64 || (storage != AHM_F32 && storage != AHM_BF16))
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_convert: only F64 -> F32 or BF16");
  AHM_HIP(hipSetDevice(src->ctx->device));
  hipStream_t s = src->ctx->stream;
  ahm_hmat h = new ahm_hmat_s;
  h->ctx = src->ctx;
  h->n_rows = src->n_rows;
  h->n_cols = src->n_cols;
  h->storage = storage;
  h->leaf_dtype = src->leaf_dtype;
  h->nvals = src->nvals;
  h->t_len = src->t_len;
  h->max_rank = src->max_rank;
  h->n_dense = src->n_dense;
  h->n_lowrank = src->n_lowrank;
  h->nv_dense = src->nv_dense;
  h->nv_u = src->nv_u;
  h->nv_v = src->nv_v;
  h->sym = src->sym;
  h->leaves = src->leaves;
  h->h_tasks_a = src->h_tasks_a;
  h->h_tasks_b = src->h_tasks_b;
  h->h_tasks_r = src->h_tasks_r;
  h->h_tasks_ra = src->h_tasks_ra;
  h->h_panels = src->h_panels;
  h->h_panels_a = src->h_panels_a;
  h->chunk_values = src->chunk_values;
  h->seg_off = src->seg_off;
  h->h_zsrc = src->h_zsrc;
  h->part_len = src->part_len;
  int st = AHM_OK;
  do {
    if ((st = h->arena.alloc((h->nvals + kArenaSlack) * 4))) break;
    if (hipMemsetAsync((char*)h->arena.p + h->nvals * 4, 0, kArenaSlack * 4, s) != hipSuccess) {
      st = fail(AHM_ERR_HIP, "arena slack");
      break;
    }
    if ((st = upload_vec(h->tasks_a, h->h_tasks_a, s))) break;
    if ((st = upload_vec(h->tasks_b, h->h_tasks_b, s))) break;
    if ((st = upload_vec(h->zsrc, h->h_zsrc, s))) break;
    if ((st = upload_vec(h->tasks_r, h->h_tasks_r, s))) break;
    if ((st = upload_vec(h->tasks_ra, h->h_tasks_ra, s))) break;
    h->n_a = src->n_a;
    h->n_b = src->n_b;
    h->n_r = src->n_r;
    h->n_ra = src->n_ra;
    h->n_zsrc = src->n_zsrc;
    if ((st = build_mm_tasks(h, s))) break;
    if (h->nvals)
      hipLaunchKernelGGL((cast_kernel<float, double>), dim3(4096), dim3(256), 0, s,
                         (const double*)src->arena.p, (float*)h->arena.p, h->nvals);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      st = fail(AHM_ERR_HIP, "ahm_hmat_convert: cast kernel failed");
      break;
    }
    if (storage == AHM_BF16 && (st = build_bf16(h))) break;
  } while (0);
  if (st) {
    delete h;
    return st;
  }
  *out = h;
  return AHM_OK;
}

int ahm_hmat_destroy(ahm_hmat h) {
  if (!h) return AHM_OK;
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  gmres_cache_free(h);
  delete h;
  return AHM_OK;
}

int ahm_hmat_stats_get(ahm_hmat h, ahm_hmat_stats* st) {
  if (!h || !st) return fail(AHM_ERR_INVALID, "ahm_hmat_stats_get: NULL argument");
  std::memset(st, 0, sizeof *st);
  st->n_leaves = h->leaves.size();
  st->n_dense = h->n_dense;
  st->n_lowrank = h->n_lowrank;
  st->nvals = h->nvals;
  st->nvals_dense = h->nv_dense;
  st->nvals_u = h->nv_u;
  st->nvals_v = h->nv_v;
  st->t_len = h->t_len;
  st->bytes_per_value = h->storage == AHM_BF16 ? 4 : dtype_size(h->storage);
  st->n_tasks_a = h->n_a;
  st->n_tasks_b = h->n_b;
  st->max_rank = h->max_rank;
  st->n_rows = h->n_rows;
  st->n_cols = h->n_cols;
  st->symmetric = (uint32_t)h->sym;
  return AHM_OK;
}

int ahm_hmat_segments(ahm_hmat h, uint32_t* n_seg, uint32_t* offsets) {
  if (!h || !n_seg) return fail(AHM_ERR_INVALID, "ahm_hmat_segments: NULL argument");
  if (h->seg_off.empty()) {
    *n_seg = 1;
    if (offsets) {
      offsets[0] = 0;
      offsets[1] = h->n_rows;
    }
    return AHM_OK;
  }
  *n_seg = (uint32_t)h->seg_off.size() - 1;
  if (offsets) std::memcpy(offsets, h->seg_off.data(), h->seg_off.size() * sizeof(uint32_t));
  return AHM_OK;
}

int ahm_hmat_download(ahm_hmat h, ahm_leaf_desc* leaves, void* data) {
  if (!h) return fail(AHM_ERR_INVALID, "ahm_hmat_download: NULL handle");
  if (leaves) std::memcpy(leaves, h->leaves.data(), h->leaves.size() * sizeof(ahm_leaf_desc));
  if (data) {
    if (!h->leafdata.p && h->leafdata.bytes == 0 && !h->leaves.empty())
      return fail(AHM_ERR_INVALID, "ahm_hmat_download: matrix was uploaded without keep_leaves");
    AHM_HIP(hipSetDevice(h->ctx->device));
    AHM_HIP(hipMemcpyAsync(data, h->leafdata.p, h->leafdata.bytes, hipMemcpyDeviceToHost,
                           h->ctx->stream));
    AHM_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  return AHM_OK;
}

int ahm_hmat_matvec(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y, uint32_t ldy,
                    uint32_t p, ahm_dtype vec) {
  if (!h || !X || !Y) return fail(AHM_ERR_INVALID, "ahm_hmat_matvec: NULL argument");
  if (p == 0) return AHM_OK;
  if (p > 65535) return fail(AHM_ERR_INVALID, "ahm_hmat_matvec: p=%u > 65535", p);
  if ((p > 1 && (ldx < h->n_cols || ldy < h->n_rows)))
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec: ldx=%u ldy=%u for %u x %u", ldx, ldy,
                h->n_rows, h->n_cols);
  AHM_HIP(hipSetDevice(h->ctx->device));
  if (h->storage == AHM_F64 && vec == AHM_F64) {
    AHM_TRY(ensure_tbuf(h, p, 8));
    AHM_TRY(ensure_pbuf(h, p, 8));
    AHM_TRY(sym_prepare(h, vec, p));
    AHM_TRY((launch_matvec<double, double>(h, own_workspace(h), alpha, (const double*)X, ldx, (double*)Y,
                                           ldy, p)));
    return p > 1 ? sym_transposed(h, alpha, X, ldx, Y, ldy, p, vec, Gate{}) : AHM_OK;
  }
  if ((h->storage == AHM_F32 || h->storage == AHM_BF16) && vec == AHM_F32) {
    AHM_TRY(ensure_tbuf(h, p, 4));
    AHM_TRY(ensure_pbuf(h, p, 4));
    AHM_TRY(sym_prepare(h, vec, p));
    AHM_TRY((launch_matvec<float, float>(h, own_workspace(h), (float)alpha, (const float*)X, ldx, (float*)Y,
                                         ldy, p)));
    return p > 1 ? sym_transposed(h, alpha, X, ldx, Y, ldy, p, vec, Gate{}) : AHM_OK;
  }
  return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_matvec: storage %d with vectors %d", h->storage, vec);
}

int ahm_hmat_matvec_profile(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                            uint32_t ldy, uint32_t p, ahm_dtype vec, uint32_t reps,
                            double* ms_a, double* ms_b) {
  if (!h || !X || !Y || !ms_a || !ms_b || reps == 0)
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_profile: bad arguments");
  if (!((h->storage == AHM_F64 && vec == AHM_F64) || ((h->storage == AHM_F32 || h->storage == AHM_BF16) && vec == AHM_F32)))
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_matvec_profile: storage %d with vectors %d",
                h->storage, vec);
  if (p == 0 || p > 65535 || (p > 1 && (ldx < h->n_cols || ldy < h->n_rows)))
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_profile: bad p / leading dimensions");
  AHM_HIP(hipSetDevice(h->ctx->device));
  AHM_TRY(ensure_tbuf(h, p, dtype_size(vec)));
  AHM_TRY(ensure_pbuf(h, p, dtype_size(vec)));
  hipEvent_t e0, e1, e2;
  AHM_HIP(hipEventCreate(&e0));
  AHM_HIP(hipEventCreate(&e1));
  AHM_HIP(hipEventCreate(&e2));
  double ta = 0.0, tb = 0.0;
  int st = AHM_OK;
  for (uint32_t r = 0; r < reps && st == AHM_OK; ++r) {
    if (hipEventRecord(e0, h->ctx->stream) != hipSuccess) { st = fail(AHM_ERR_HIP, "event"); break; }
    if (vec == AHM_F64)
      st = launch_matvec<double, double>(h, own_workspace(h), alpha, (const double*)X, ldx, (double*)Y,
                                         ldy, p, e1);
    else
      st = launch_matvec<float, float>(h, own_workspace(h), (float)alpha, (const float*)X, ldx, (float*)Y,
                                       ldy, p, e1);
    if (st) break;
    float a = 0.f, b = 0.f;
    if (hipEventRecord(e2, h->ctx->stream) != hipSuccess || hipEventSynchronize(e2) != hipSuccess ||
        hipEventElapsedTime(&a, e0, e1) != hipSuccess || hipEventElapsedTime(&b, e1, e2) != hipSuccess) {
      st = fail(AHM_ERR_HIP, "ahm_hmat_matvec_profile: event timing failed");
      break;
    }
    ta += a;
    tb += b;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  if (st) return st;
  *ms_a = ta / reps;
  *ms_b = tb / reps;
  return AHM_OK;
}

int ahm_hmat_set_fused(ahm_hmat h, int on) {
  if (!h) return fail(AHM_ERR_INVALID, "ahm_hmat_set_fused: NULL handle");
  if (on < 0 || on > 2) return fail(AHM_ERR_INVALID, "ahm_hmat_set_fused: form %d", on);
  if (on && h->sym) return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_set_fused: symmetric matrices");
  if (on == 1 && !h->fz_built) {
    AHM_HIP(hipSetDevice(h->ctx->device));
    AHM_TRY(build_fused(h));
  }
  h->fused = on;
  return AHM_OK;
}

int ahm_hmat_fused_status(ahm_hmat h, int* aborted) {
  if (!h || !aborted) return fail(AHM_ERR_INVALID, "ahm_hmat_fused_status: NULL argument");
  *aborted = 0;   // the single launch has no waits to give up
  return AHM_OK;
}

int ahm_hmat_autotune(ahm_hmat h, uint32_t p, ahm_dtype vec, uint32_t reps, uint64_t* chunk_values,
                      double* best_ms, int* fused) {
  if (!h || p == 0 || reps == 0) return fail(AHM_ERR_INVALID, "ahm_hmat_autotune: bad arguments");
  if (!((h->storage == AHM_F64 && vec == AHM_F64) ||
        ((h->storage == AHM_F32 || h->storage == AHM_BF16) && vec == AHM_F32)))
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_autotune: storage %d with vectors %d", h->storage, vec);
  AHM_HIP(hipSetDevice(h->ctx->device));
  hipStream_t s = h->ctx->stream;
  const size_t es = dtype_size(vec);
  DeviceBuf dx, dy;
  AHM_TRY(dx.alloc((size_t)h->n_cols * p * es));
  AHM_TRY(dy.alloc((size_t)h->n_rows * p * es));
  AHM_HIP(hipMemsetAsync(dx.p, 0, (size_t)h->n_cols * p * es, s));
  AHM_HIP(hipMemsetAsync(dy.p, 0, (size_t)h->n_rows * p * es, s));
  // tile sizes (values per row-panel / V^T task) times the launch form: three
  // kernels (phases A, B, chunk sums) or one persistent launch (p <= 8)
  const uint64_t cand[] = {4u << 10, 8u << 10, 16u << 10, 32u << 10, 48u << 10, 96u << 10, 192u << 10, 1ull << 40};
  double best = 1e300;
  uint64_t arg = h->chunk_values;
  int arg_fused = h->fused;
  hipEvent_t e0, e1;
  AHM_HIP(hipEventCreate(&e0));
  AHM_HIP(hipEventCreate(&e1));
  int st = AHM_OK;
  const int saved_fused = h->fused;
  for (int fu = 0; fu <= (p <= 8 && !h->sym ? 2 : 0) && !st; ++fu) {
    if (fu == 1 && !h->fz_built && (st = build_fused(h))) break;
    h->fused = fu;
    for (uint64_t cv : cand) {
      if (fu == 1 && cv != cand[0]) break;   // the row-panel single launch runs unchunked panels
      if ((st = set_chunking(h, cv))) break;
      if ((st = ensure_tbuf(h, p, es)) || (st = ensure_pbuf(h, p, es))) break;
      auto run = [&]() {
        return vec == AHM_F64
                   ? launch_matvec<double, double>(h, own_workspace(h), 1.0, (const double*)dx.p, h->n_cols,
                                                   (double*)dy.p, h->n_rows, p)
                   : launch_matvec<float, float>(h, own_workspace(h), 1.0f, (const float*)dx.p, h->n_cols,
                                                 (float*)dy.p, h->n_rows, p);
      };
      if ((st = run())) break;  // warm-up
      if (hipEventRecord(e0, s) != hipSuccess) { st = fail(AHM_ERR_HIP, "autotune: event"); break; }
      for (uint32_t r = 0; r < reps && !st; ++r) st = run();
      float ms = 0.f;
      if (st || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        if (!st) st = fail(AHM_ERR_HIP, "autotune: event timing");
        break;
      }
      if (ms / reps < best) {
        best = ms / reps;
        arg = cv;
        arg_fused = fu;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->fused = st ? saved_fused : arg_fused;
  if (st) return st;
  AHM_TRY(set_chunking(h, arg));
  if (chunk_values) *chunk_values = arg;
  if (best_ms) *best_ms = best;
  if (fused) *fused = arg_fused;
  return AHM_OK;
}

int ahm_hmat_matvec_host(ahm_hmat h, double alpha, const double* X, uint32_t ldx, double* Y,
                         uint32_t ldy, uint32_t p) {
  if (!h || !X || !Y) return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_host: NULL argument");
  if (h->storage != AHM_F64)
    return fail(AHM_ERR_UNSUPPORTED, "ahm_hmat_matvec_host: needs an F64 matrix");
  if (p == 0) return AHM_OK;
  if (p > 1 && (ldx < h->n_cols || ldy < h->n_rows))
    return fail(AHM_ERR_INVALID, "ahm_hmat_matvec_host: bad leading dimensions");
  AHM_HIP(hipSetDevice(h->ctx->device));
  hipStream_t s = h->ctx->stream;
  DeviceBuf dx, dy;
  AHM_TRY(dx.alloc((size_t)h->n_cols * p * 8));
  AHM_TRY(dy.alloc((size_t)h->n_rows * p * 8));
  AHM_HIP(hipMemcpy2DAsync(dx.p, (size_t)h->n_cols * 8, X, (size_t)ldx * 8, (size_t)h->n_cols * 8, p,
                           hipMemcpyHostToDevice, s));
  AHM_HIP(hipMemcpy2DAsync(dy.p, (size_t)h->n_rows * 8, Y, (size_t)ldy * 8, (size_t)h->n_rows * 8, p,
                           hipMemcpyHostToDevice, s));
  AHM_TRY(ahm_hmat_matvec(h, alpha, dx.p, h->n_cols, dy.p, h->n_rows, p, AHM_F64));
  AHM_HIP(hipMemcpy2DAsync(Y, (size_t)ldy * 8, dy.p, (size_t)h->n_rows * 8, (size_t)h->n_rows * 8, p,
                           hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  return AHM_OK;
}

}  // extern "C"
