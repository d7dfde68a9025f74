// gmres.hip -- restarted GMRES(m) with the Arnoldi step, the Givens
// rotations and the convergence test on the device, for one system or for
// the independent diagonal blocks of a batched H-matrix.
//
// Per system the iteration is solvers/GMRES.cpp:80-181: right-preconditioned
// (HLU_solve, single system only), modified Gram-Schmidt, Givens rotations
// (genPlRot / applPlRot, :33-55), convergence on the implicit residual
// |s_{i+1}| / |b| (:151), x updated at convergence or at the end of a cycle
// (update, :58-77), true residual recomputed per restart (:165-176).
//
// Device state per system: |b|, beta, the residual estimate, the (m+1) x m
// Hessenberg matrix, the rotations, the rotated right-hand side s, and
// counters (steps taken, steps of the current cycle, restarts) and flags
// (converged, frozen).  A system whose test passes is frozen on the device:
// every later kernel (zeroing, the H-matvec through its per-task gate, the
// MGS launches, the normalisation, the rotations) returns at once for it.
// The host only queues work.  It reads the systems' frozen flags kLookahead
// steps behind (an event per step and a copy of the flags to pinned memory),
// so step i+1 is queued while step i runs, and a converged system costs at
// most kLookahead queued steps of empty launches.  x is updated once per
// cycle: a system that converged inside the cycle gets the update the
// reference makes at once (k = its converged step), the others the
// end-of-cycle update; then the systems still running restart.
//
// Arnoldi step i launches the same kernels on the same buffers in every
// cycle and every call, so it is captured into a hipGraph on its second use
// (in a later cycle or call; capture and instantiation cost about as much as
// a few hundred launches) and replayed from then on.  The graphs, the Krylov
// basis and the state stay with the handle (GmresCache) until its product
// layout changes (matvec_signature), a call needs other sizes, or the handle
// is destroyed.
//
// Reductions are deterministic: each system's rows are cut into fixed
// chunks, one workgroup per chunk writes one partial, and each system's
// partials are summed in a fixed order.  The MGS step k fuses "h_k = <w, v_k>
// (from partials), w -= h_k v_k, partials of <w, v_{k+1}>" in one launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

namespace {

constexpr int kGmThreads = 256;
constexpr int kLookahead = 2;      // steps queued ahead of the host's flag read
constexpr uint32_t kLdsK = 64;     // back substitution staged in LDS up to this k
// rows per reduction chunk; a constant, so that a system cut from a batch
// gets the same chunks (and bit for bit the same iterates) as alone
constexpr uint32_t kChunk = 1024;

struct SegChunk {
  uint32_t seg, begin, end, pad;
};

// Device state, struct of arrays (S systems; H is (m+1) x m per system,
// column-major; cs, sn, sv m+1 per system; Y m per system).
struct State {
  double *normb, *beta, *resid, *H, *cs, *sn, *sv, *Y, *hist, *eps;
  uint32_t *frozen, *conv, *taken, *kcyc, *kupd, *rsteps, *restarts, *zero_b, *nh, *nhist, *flag;
};

// Wavefront sum by the xor butterfly v += v[lane ^ o], o = 32, 16, .., 1,
// valid in every lane, with VALU exchanges instead of LDS permutes: xor 32
// and 16 by the lane-half swaps, 8 by a row rotation, and 4 and 2 by row
// rotations too, which read a lane equal to lane ^ o once the earlier steps
// have made lanes l, l ^ 8 (and l ^ 4) equal; 1 by a quad permutation.
// Every lane adds the same two values as the butterfly's lane, so the sum
// is bit for bit the butterfly's.
template <int CTRL>
__device__ __forceinline__ double gm_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double gm_xor32(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const bool up = threadIdx.x & 32;
  return __hiloint2double((int)(up ? b[0] : b[1]), (int)(up ? a[0] : a[1]));
}
__device__ __forceinline__ double gm_xor16(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const bool up = threadIdx.x & 16;
  return __hiloint2double((int)(up ? b[0] : b[1]), (int)(up ? a[0] : a[1]));
}
__device__ __forceinline__ double gm_wsum(double v) {
  v += gm_xor32(v);
  v += gm_xor16(v);
  v += gm_dpp<0x128>(v);   // row_ror:8
  v += gm_dpp<0x124>(v);   // row_ror:4
  v += gm_dpp<0x122>(v);   // row_ror:2
  v += gm_dpp<0xB1>(v);    // quad_perm [1,0,3,2]
  return v;
}

// sum over a kGmThreads-wide workgroup, fixed order, valid in every thread
__device__ double gm_group_sum(double v) {
  __shared__ double sh[kGmThreads / 64];
  v = gm_wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < kGmThreads / 64; ++q) s += sh[q];
  return s;
}

// system s's partials (chunks cb[s] .. cb[s+1]) summed in a fixed order
__device__ double gm_seg_sum(const double* __restrict__ part, const uint32_t* __restrict__ cb, uint32_t s) {
  double v = 0.0;
  for (uint32_t c = cb[s] + threadIdx.x; c < cb[s + 1]; c += kGmThreads) v += part[c];
  return gm_group_sum(v);
}

__global__ __launch_bounds__(kGmThreads) void gm_reset_kernel(State st, uint32_t S, double eps,
                                                              uint32_t nhist) {
  for (uint32_t s = threadIdx.x; s < S; s += kGmThreads) {
    st.frozen[s] = st.conv[s] = st.taken[s] = st.kcyc[s] = st.kupd[s] = 0;
    st.rsteps[s] = st.restarts[s] = st.zero_b[s] = 0;
    st.resid[s] = 0.0;
  }
  if (threadIdx.x == 0) {
    *st.eps = eps;
    *st.nhist = nhist;
    *st.nh = 0;
    *st.flag = 0;
  }
}

// partials of <a, b> per chunk (frozen systems skipped)
__global__ __launch_bounds__(kGmThreads) void gm_dot_kernel(const double* __restrict__ a,
                                                            const double* __restrict__ b,
                                                            const SegChunk* __restrict__ ch,
                                                            const uint32_t* __restrict__ frozen,
                                                            double* __restrict__ part) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  double s = 0.0;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) s = fma(a[i], b[i], s);
  s = gm_group_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// partials of sum |a_i| (zero exactly when every a_i is zero)
__global__ __launch_bounds__(kGmThreads) void gm_asum_kernel(const double* __restrict__ a,
                                                             const SegChunk* __restrict__ ch,
                                                             double* __restrict__ part) {
  const SegChunk C = ch[blockIdx.x];
  double s = 0.0;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) s += fabs(a[i]);
  s = gm_group_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// h = sum of the system's partials of <w, v>; H entry = h (first chunk);
// w -= h v; partials of <w, next> (next = nullptr: <w, w>)
__global__ __launch_bounds__(kGmThreads) void gm_mgs_kernel(
    double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ next,
    const SegChunk* __restrict__ ch, const uint32_t* __restrict__ cb, const uint32_t* __restrict__ frozen,
    const double* __restrict__ part_in, double* __restrict__ hent, uint32_t hstride,
    double* __restrict__ part_out) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  const double h = gm_seg_sum(part_in, cb, C.seg);
  if (blockIdx.x == cb[C.seg] && threadIdx.x == 0) hent[(size_t)C.seg * hstride] = h;
  double s = 0.0;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) {
    const double wi = fma(-h, v[i], w[i]);
    w[i] = wi;
    s = fma(wi, next ? next[i] : wi, s);
  }
  s = gm_group_sum(s);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

// h = sqrt(sum of partials of <w, w>); H entry = h; w *= 1/h (h = 0: the
// Krylov space is exhausted and w is zero already)
__global__ __launch_bounds__(kGmThreads) void gm_normalize_kernel(
    double* __restrict__ w, const SegChunk* __restrict__ ch, const uint32_t* __restrict__ cb,
    const uint32_t* __restrict__ frozen, const double* __restrict__ part, double* __restrict__ hent,
    uint32_t hstride) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  const double h = sqrt(gm_seg_sum(part, cb, C.seg));
  if (blockIdx.x == cb[C.seg] && threadIdx.x == 0) hent[(size_t)C.seg * hstride] = h;
  if (h == 0.0) return;
  const double inv = 1.0 / h;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) w[i] *= inv;
}

__global__ __launch_bounds__(kGmThreads) void gm_zero_kernel(double* __restrict__ d,
                                                             const SegChunk* __restrict__ ch,
                                                             const uint32_t* __restrict__ frozen) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) d[i] = 0.0;
}

// d = 0 on the systems with mask set (b = 0: x = 0, solvers/GMRES.cpp:95-102)
__global__ __launch_bounds__(kGmThreads) void gm_zero_masked_kernel(double* __restrict__ d,
                                                                    const SegChunk* __restrict__ ch,
                                                                    const uint32_t* __restrict__ mask) {
  const SegChunk C = ch[blockIdx.x];
  if (!mask[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) d[i] = 0.0;
}

__global__ __launch_bounds__(kGmThreads) void gm_copy_kernel(const double* __restrict__ src,
                                                             double* __restrict__ dst,
                                                             const SegChunk* __restrict__ ch,
                                                             const uint32_t* __restrict__ frozen) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) dst[i] = src[i];
}

// FP32 operator: x32 = (float) x, y32 = 0
__global__ __launch_bounds__(kGmThreads) void gm_to_f32_kernel(const double* __restrict__ x,
                                                               float* __restrict__ x32,
                                                               float* __restrict__ y32,
                                                               const SegChunk* __restrict__ ch,
                                                               const uint32_t* __restrict__ frozen) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) {
    x32[i] = (float)x[i];
    y32[i] = 0.0f;
  }
}

// FP32 operator: y += (double) y32
__global__ __launch_bounds__(kGmThreads) void gm_add_f32_kernel(const float* __restrict__ y32,
                                                                double* __restrict__ y,
                                                                const SegChunk* __restrict__ ch,
                                                                const uint32_t* __restrict__ frozen) {
  const SegChunk C = ch[blockIdx.x];
  if (frozen[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) y[i] += (double)y32[i];
}

__global__ __launch_bounds__(kGmThreads) void gm_axpy_kernel(const double* __restrict__ d,
                                                             double* __restrict__ x,
                                                             const SegChunk* __restrict__ ch) {
  const SegChunk C = ch[blockIdx.x];
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) x[i] += d[i];
}

// Start of a cycle (solvers/GMRES.cpp:117-121): v_0 = r * (1/beta),
// s = (beta, 0, ..., 0), for the systems still running.
__global__ __launch_bounds__(kGmThreads) void gm_cycle_kernel(const double* __restrict__ r,
                                                              double* __restrict__ v0,
                                                              const SegChunk* __restrict__ ch,
                                                              const uint32_t* __restrict__ cb, State st,
                                                              uint32_t m) {
  const SegChunk C = ch[blockIdx.x];
  if (st.frozen[C.seg]) return;
  const double beta = st.beta[C.seg];
  const double a = 1.0 / beta;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) v0[i] = a * r[i];
  if (blockIdx.x == cb[C.seg]) {
    double* sv = st.sv + (size_t)C.seg * (m + 1);
    for (uint32_t q = threadIdx.x; q <= m; q += kGmThreads) sv[q] = q == 0 ? beta : 0.0;
  }
}

// genPlRot / applPlRot (solvers/GMRES.cpp:33-55), without contraction into
// fused multiply-adds so that the rotations round as the plain formulas do
__device__ void gm_gen_rot(double dx, double dy, double& cs, double& sn) {
#pragma clang fp contract(off)
  if (dy == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (fabs(dy) > fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / sqrt(1.0 + t * t);
    sn = t * cs;
  }
}
__device__ __forceinline__ void gm_app_rot(double& dx, double& dy, double cs, double sn) {
#pragma clang fp contract(off)
  const double t = cs * dx + sn * dy;
  dy = cs * dy - sn * dx;
  dx = t;
}

// Rotations of step i (solvers/GMRES.cpp:140-157) for system s, run by one
// workgroup: the new Hessenberg column and the rotations are staged in LDS
// (gl: column m+2, cs m+1, sn m+1), one lane applies the i old rotations,
// generates the new one, rotates s and tests |s_{i+1}| / |b| < eps.  A
// converged system is frozen with rsteps = its step count (the reference's
// j).  The history is kept when there is one system (one workgroup).
__device__ void gm_givens_sys(const State& st, uint32_t m, uint32_t i, uint32_t s, double* gl) {
  double* Hc = st.H + (size_t)s * (m + 1) * m + (size_t)i * (m + 1);
  double* cs = st.cs + (size_t)s * (m + 1);
  double* sn = st.sn + (size_t)s * (m + 1);
  double* col = gl;
  double* lc = gl + (m + 2);
  double* ls = lc + (m + 1);
  for (uint32_t k = threadIdx.x; k <= i + 1; k += blockDim.x) col[k] = Hc[k];
  for (uint32_t k = threadIdx.x; k < i; k += blockDim.x) {
    lc[k] = cs[k];
    ls[k] = sn[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double dx = col[0];
    for (uint32_t k = 0; k < i; ++k) {
      double dy = col[k + 1];
      gm_app_rot(dx, dy, lc[k], ls[k]);
      col[k] = dx;
      dx = dy;
    }
    double dy = col[i + 1];
    double c, sg;
    gm_gen_rot(dx, dy, c, sg);
    gm_app_rot(dx, dy, c, sg);
    col[i] = dx;
    col[i + 1] = dy;
    cs[i] = c;
    sn[i] = sg;
    double* sv = st.sv + (size_t)s * (m + 1);
    double s0 = sv[i], s1 = sv[i + 1];
    gm_app_rot(s0, s1, c, sg);
    sv[i] = s0;
    sv[i + 1] = s1;
    const double resid = fabs(s1 / st.normb[s]);
    const uint32_t taken = st.taken[s] + 1;
    st.taken[s] = taken;
    st.kcyc[s] = i + 1;
    st.resid[s] = resid;
    if (gridDim.x == 1 && *st.nh < *st.nhist) st.hist[(*st.nh)++] = resid;
    if (resid < *st.eps) {
      st.frozen[s] = 1;
      st.conv[s] = 1;
      st.rsteps[s] = taken;
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k <= i + 1; k += blockDim.x) Hc[k] = col[k];
}

__global__ __launch_bounds__(64) void gm_givens_kernel(State st, uint32_t m, uint32_t i) {
  extern __shared__ double gl[];
  if (st.frozen[blockIdx.x]) return;
  gm_givens_sys(st, m, i, blockIdx.x, gl);
}

// Arnoldi step i after the product, for systems of at most kArChunks
// chunks: one workgroup per system runs what the chunked launches do -- the
// partials of <w, v_0>, the MGS passes k = 0..i, the normalisation -- and
// then the rotations, in one launch instead of i + 4.
//
// Thread T works as thread T % kGmThreads of the chunked kernels'
// workgroup, on chunks T / kGmThreads + kArGroups j: the same rows in the
// same order and the same wavefront sums.  Every thread then adds the
// chunk partials itself in the order gm_group_sum and gm_seg_sum add them
// (the partial of chunk c in lane c of wavefront 0 and zeros elsewhere,
// the xor butterfly, the four wavefront sums), so the iterates are bit for
// bit those of the chunked form.  w and the next basis vector stay in
// registers; v_{k+2} is loaded while pass k's partials are summed, and one
// barrier per pass orders the partials (two LDS buffers in turn).
constexpr int kArGroups = 4;
constexpr int kArThreads = kArGroups * kGmThreads;
constexpr uint32_t kArChunks = 16;
constexpr int kWaves = kGmThreads / 64;
// auto choice: one workgroup per system for systems of up to this many chunks
constexpr uint32_t kArAutoChunks = 16;

// h from the pass's wavefront sums ws[c][q] of chunks c < nc (nc <= 64):
// every wavefront forms what wavefront 0 of gm_seg_sum holds (chunk c's
// partial in lane c, zeros elsewhere) and runs the same butterfly; the
// other three wavefront sums of gm_group_sum are zeros
__device__ __forceinline__ double ar_sum(const double (*ws)[kWaves], uint32_t nc) {
  const uint32_t l = threadIdx.x & 63;
  double v = 0.0;
  if (l < nc) {
    double p = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) p += ws[l][q];
    v += p;
  }
  v = gm_wsum(v);
  double h = 0.0;
  h += v;
#pragma unroll
  for (int q = 1; q < kWaves; ++q) h += 0.0;
  return h;
}

template <int CPT>
__global__ __launch_bounds__(kArThreads) void gm_arnoldi_kernel(double* __restrict__ V, size_t n,
                                                                const SegChunk* __restrict__ ch,
                                                                const uint32_t* __restrict__ cb, State st,
                                                                uint32_t m, uint32_t i) {
  constexpr int R = kChunk / kGmThreads;   // rows per thread and chunk
  constexpr int E = CPT * R;               // rows per thread
  constexpr bool kPrefetch = CPT <= 2;     // three vectors of E rows fit the registers
  extern __shared__ double gl[];
  __shared__ double ws[2][kArChunks][kWaves];
  const uint32_t s = blockIdx.x;
  if (st.frozen[s]) return;
  const uint32_t c0 = cb[s], nc = cb[s + 1] - c0;
  const uint32_t g = threadIdx.x / kGmThreads, t = threadIdx.x % kGmThreads;
  const int wave = (int)(t >> 6);
  const bool lead = (t & 63) == 0;
  // chunk j of this thread: rows beg[j] + kGmThreads q < end[j] (none past nc)
  uint32_t beg[CPT], end[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const uint32_t c = g + kArGroups * j;
    beg[j] = end[j] = 0;
    if (c < nc) {
      const SegChunk C = ch[c0 + c];
      beg[j] = C.begin + t;
      end[j] = C.end;
    }
  }
#define AR_ROW(j, q) (beg[j] + (uint32_t)(kGmThreads * (q)))
#define AR_OK(j, q) (AR_ROW(j, q) < end[j])
  auto load = [&](const double* __restrict__ src, double* dst) {
#pragma unroll
    for (int j = 0; j < CPT; ++j)
#pragma unroll
      for (int q = 0; q < R; ++q) dst[j * R + q] = AR_OK(j, q) ? src[AR_ROW(j, q)] : 0.0;
  };
  // each chunk's fma chain over its rows, wavefront sums into ws[buf]
  auto partials = [&](const double* a, const double* b, int buf) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const uint32_t c = g + kArGroups * j;
      if (c >= nc) continue;   // uniform in the wavefront
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (AR_OK(j, q)) acc = fma(a[j * R + q], b[j * R + q], acc);
      acc = gm_wsum(acc);
      if (lead) ws[buf][c][wave] = acc;
    }
  };
  double* __restrict__ w = V + (size_t)(i + 1) * n;
  double* Hc = st.H + (size_t)s * (m + 1) * m + (size_t)i * (m + 1);
  double wv[E], va[E];
  load(w, wv);
  load(V, va);
  partials(wv, va, 0);
  int buf = 0;
  if constexpr (kPrefetch) {
    // three vectors in registers, in turn: pass k updates by v_k, forms the
    // partials with v_{k+1}, and loads v_{k+3} into v_k's registers, two
    // passes before its first use
    double rb[E], rc[E];
    if (i >= 1) load(V + n, rb);
    if (i >= 2) load(V + 2 * n, rc);
    auto pass = [&](double* cur, double* nxt, uint32_t k) {
      __syncthreads();
      const double h = ar_sum(ws[buf], nc);
      if (threadIdx.x == 0) Hc[k] = h;
#pragma unroll
      for (int x = 0; x < E; ++x) wv[x] = fma(-h, cur[x], wv[x]);   // rows past end: unused
      buf ^= 1;
      if (k < i) {
        partials(wv, nxt, buf);
        if (k + 3 <= i) load(V + (size_t)(k + 3) * n, cur);
        return true;
      }
      partials(wv, wv, buf);
      return false;
    };
    for (uint32_t k = 0;; k += 3) {
      if (!pass(va, rb, k)) break;
      if (!pass(rb, rc, k + 1)) break;
      if (!pass(rc, va, k + 2)) break;
    }
  } else {
    // two vectors in registers: v_{k+1} is loaded after the update by v_k
    for (uint32_t k = 0; k <= i; ++k) {
      __syncthreads();
      const double h = ar_sum(ws[buf], nc);
      if (threadIdx.x == 0) Hc[k] = h;
#pragma unroll
      for (int x = 0; x < E; ++x) wv[x] = fma(-h, va[x], wv[x]);
      buf ^= 1;
      if (k < i) {
        load(V + (size_t)(k + 1) * n, va);
        partials(wv, va, buf);
      } else {
        partials(wv, wv, buf);
      }
    }
  }
  __syncthreads();
  const double h = sqrt(ar_sum(ws[buf], nc));
  if (threadIdx.x == 0) Hc[i + 1] = h;
  if (h != 0.0) {
    const double inv = 1.0 / h;
#pragma unroll
    for (int x = 0; x < E; ++x) wv[x] *= inv;
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j)
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (AR_OK(j, q)) w[AR_ROW(j, q)] = wv[j * R + q];
#undef AR_ROW
#undef AR_OK
  __syncthreads();   // thread 0's column entries visible to the workgroup
  gm_givens_sys(st, m, i, s, gl);
}

// update (solvers/GMRES.cpp:58-77), first half: y = H^{-1} s on the k x k
// upper triangle (k = the system's steps in this cycle; 0: nothing to do),
// by columns from the last (dtrtrs' back substitution order), one
// workgroup per system.  kupd = k for the apply kernel; kcyc = 0.
__global__ __launch_bounds__(64) void gm_solve_kernel(State st, uint32_t m) {
#pragma clang fp contract(off)
  __shared__ double Hl[kLdsK * kLdsK];
  __shared__ double y[kLdsK];
  const uint32_t s = blockIdx.x;
  const uint32_t k = st.kcyc[s];
  __syncthreads();
  if (threadIdx.x == 0) {
    st.kupd[s] = k;
    st.kcyc[s] = 0;
  }
  if (k == 0) return;
  const double* H = st.H + (size_t)s * (m + 1) * m;
  const double* sv = st.sv + (size_t)s * (m + 1);
  double* Y = st.Y + (size_t)s * m;
  if (k <= kLdsK) {
    for (uint32_t e = threadIdx.x; e < k * k; e += 64) {
      const uint32_t r = e % k, c = e / k;
      Hl[r + c * kLdsK] = r <= c ? H[r + (size_t)c * (m + 1)] : 0.0;
    }
    for (uint32_t q = threadIdx.x; q < k; q += 64) y[q] = sv[q];
    __syncthreads();
    for (int j = (int)k - 1; j >= 0; --j) {
      if (threadIdx.x == 0) y[j] = y[j] / Hl[j + j * kLdsK];
      __syncthreads();
      const double yj = y[j];
      for (uint32_t q = threadIdx.x; q < (uint32_t)j; q += 64) y[q] = y[q] - yj * Hl[q + j * kLdsK];
      __syncthreads();
    }
    for (uint32_t q = threadIdx.x; q < k; q += 64) Y[q] = y[q];
  } else {
    for (uint32_t q = threadIdx.x; q < k; q += 64) Y[q] = sv[q];
    __syncthreads();
    for (int j = (int)k - 1; j >= 0; --j) {
      if (threadIdx.x == 0) Y[j] = Y[j] / H[j + (size_t)j * (m + 1)];
      __syncthreads();
      const double yj = Y[j];
      for (uint32_t q = threadIdx.x; q < (uint32_t)j; q += 64)
        Y[q] = Y[q] - yj * H[q + (size_t)j * (m + 1)];
      __syncthreads();
    }
  }
}

// update, second half: x (+)= V_s y_s over the system's kupd columns
// (ASSIGN: x = V y, the preconditioned form's x_h before M is applied)
template <bool ASSIGN>
__global__ __launch_bounds__(kGmThreads) void gm_apply_kernel(const double* __restrict__ V, uint32_t n,
                                                              uint32_t m, const SegChunk* __restrict__ ch,
                                                              State st, double* __restrict__ x) {
  const SegChunk C = ch[blockIdx.x];
  const uint32_t k = st.kupd[C.seg];
  if (k == 0) return;
  const double* y = st.Y + (size_t)C.seg * m;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kGmThreads) {
    double s = 0.0;
    for (uint32_t l = 0; l < k; ++l) s = fma(V[(uint64_t)l * n + i], y[l], s);
    if (ASSIGN)
      x[i] = s;
    else
      x[i] += s;
  }
}

// Set-up and restart tests, one workgroup, the systems in turn.
//  mode 0: |b| from the partials of <b, b>; b = 0 freezes the system as
//          converged with x = 0 (solvers/GMRES.cpp:95-102); flag = 1 when
//          the partials of sum |x| (part2) are not all zero.
//  mode 1: beta = |r| of the initial residual; resid = beta/|b| <= eps:
//          converged with 0 steps (:107-113).
//  mode 2: restart (:170-176): beta = |r|; resid = beta/|b| < eps:
//          converged with rsteps = steps + 1 (the reference's j after the
//          cycle's last increment); the restart's product is counted.
__global__ __launch_bounds__(kGmThreads) void gm_test_kernel(const double* __restrict__ part,
                                                             const double* __restrict__ part2,
                                                             const uint32_t* __restrict__ cb, State st,
                                                             uint32_t S, int mode) {
  double any = 0.0;
  for (uint32_t s = 0; s < S; ++s) {
    if (st.frozen[s]) continue;   // uniform over the workgroup
    const double v = gm_seg_sum(part, cb, s);
    if (mode == 0) any += gm_seg_sum(part2, cb, s);
    if (threadIdx.x != 0) continue;
    const double nrm = sqrt(v);
    if (mode == 0) {
      st.normb[s] = nrm;
      if (nrm == 0.0) {
        st.zero_b[s] = st.frozen[s] = st.conv[s] = 1;
        st.resid[s] = 0.0;
      }
      continue;
    }
    st.beta[s] = nrm;
    const double resid = nrm / st.normb[s];
    st.resid[s] = resid;
    if (mode == 2) st.restarts[s] += 1;
    if (mode == 1 ? resid <= *st.eps : resid < *st.eps) {
      st.frozen[s] = st.conv[s] = 1;
      st.rsteps[s] = mode == 1 ? 0 : st.taken[s] + 1;
    }
  }
  if (mode == 0 && threadIdx.x == 0) *st.flag = any != 0.0 ? 1u : 0u;
}

}  // namespace

struct GmresCache {
  int dev = 0;
  uint32_t n = 0, m = 0, S = 0, nch = 0, nhist_cap = 0;
  uint32_t max_nc = 0;          // chunks of the largest system
  int ar_one = -1;              // form of the captured steps: 1 one workgroup per system, 0 chunked
  ahm_dtype vec = AHM_F64;
  bool batch = false, prec = false;
  std::vector<uint64_t> sig;
  hipStream_t gs = nullptr;
  DeviceBuf dch, dcb, dV, dr, dp0, dp1, dD, dU, dx32, dy32, dxh, dseg;
  const uint32_t* seg[4] = {nullptr, nullptr, nullptr, nullptr};
  State st{};
  uint32_t* hflags = nullptr;   // pinned: (m + 1) x S frozen flags, then 1 flag
  std::vector<hipGraphExec_t> graphs;
  std::vector<uint32_t> uses;   // runs of each step index since the last capture reset
  std::vector<hipEvent_t> ev;
  void drop_graphs() {
    std::fill(uses.begin(), uses.end(), 0u);
    for (auto& g : graphs)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }
  ~GmresCache() {
    if (gs) (void)hipStreamSynchronize(gs);
    drop_graphs();
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (hflags) (void)hipHostFree(hflags);
    if (gs) (void)hipStreamDestroy(gs);
  }
};

void gmres_cache_free(ahm_hmat h) {
  delete h->gm;
  h->gm = nullptr;
}

namespace {

// System of every product task of a batched H-matrix (all of a task's rows
// and columns lie in one diagonal block).  Phase A: the first column source;
// phase-A reductions: the chunk whose partial slot starts the reduction;
// phase B: the first source (a column, or a V^T x entry whose system is
// that of the phase-A task or reduction writing it); phase-B reductions:
// their first output row.
int task_segments(ahm_hmat h, const std::vector<uint32_t>& off, std::vector<uint32_t> out[4]) {
  auto seg_of = [&](uint32_t i) {
    return (uint32_t)(std::upper_bound(off.begin(), off.end(), i) - off.begin()) - 1;
  };
  const uint32_t S = (uint32_t)off.size() - 1;
  std::unordered_map<uint32_t, uint32_t> slot_seg;
  struct Range {
    uint32_t begin, end, seg;
  };
  std::vector<Range> tr;
  out[0].resize(h->h_tasks_a.size());
  for (size_t q = 0; q < h->h_tasks_a.size(); ++q) {
    const GemvTask& tk = h->h_tasks_a[q];
    if (tk.W == 0) return fail(AHM_ERR_INVALID, "batched GMRES: empty phase-A task");
    const uint32_t src = h->h_zsrc[tk.zoff];
    if ((src & kFromT) || src >= h->n_cols) return fail(AHM_ERR_INVALID, "batched GMRES: phase-A source");
    const uint32_t s = seg_of(src);
    out[0][q] = s;
    if (tk.partial)
      slot_seg[tk.out] = s;
    else
      tr.push_back(Range{tk.out, tk.out + tk.m, s});
  }
  out[2].resize(h->h_tasks_ra.size());
  for (size_t q = 0; q < h->h_tasks_ra.size(); ++q) {
    const ReduceTask& r = h->h_tasks_ra[q];
    auto it = slot_seg.find(r.part);
    if (it == slot_seg.end()) return fail(AHM_ERR_INVALID, "batched GMRES: phase-A reduction without chunks");
    out[2][q] = it->second;
    tr.push_back(Range{r.row, r.row + r.m, it->second});
  }
  std::sort(tr.begin(), tr.end(), [](const Range& a, const Range& b) { return a.begin < b.begin; });
  auto t_seg = [&](uint32_t t) -> int64_t {
    auto it = std::upper_bound(tr.begin(), tr.end(), t, [](uint32_t v, const Range& r) { return v < r.begin; });
    if (it == tr.begin()) return -1;
    --it;
    return t < it->end ? (int64_t)it->seg : -1;
  };
  out[1].resize(h->h_tasks_b.size());
  for (size_t q = 0; q < h->h_tasks_b.size(); ++q) {
    const GemvTask& tk = h->h_tasks_b[q];
    if (tk.W == 0) return fail(AHM_ERR_INVALID, "batched GMRES: empty phase-B task");
    const uint32_t src = h->h_zsrc[tk.zoff];
    int64_t s;
    if (src & kFromT)
      s = t_seg(src & ~kFromT);
    else
      s = src < h->n_cols ? (int64_t)seg_of(src) : -1;
    if (s < 0 || s >= (int64_t)S) return fail(AHM_ERR_INVALID, "batched GMRES: phase-B source");
    out[1][q] = (uint32_t)s;
  }
  out[3].resize(h->h_tasks_r.size());
  for (size_t q = 0; q < h->h_tasks_r.size(); ++q) {
    const uint32_t row = h->h_tasks_r[q].row;
    if (row >= h->n_rows) return fail(AHM_ERR_INVALID, "batched GMRES: phase-B reduction row");
    out[3][q] = seg_of(row);
  }
  return AHM_OK;
}

// The handle's cache for this call's sizes; rebuilt when they change.
int get_cache(ahm_hmat h, const GmresArgs& a, GmresCache*& out) {
  const uint32_t n = h->n_rows, m = a.m;
  std::vector<uint32_t> off = a.batch ? h->seg_off : std::vector<uint32_t>{};
  if (off.size() < 2) off = {0, n};
  const uint32_t S = (uint32_t)off.size() - 1;
  const ahm_dtype vec = h->storage == AHM_F64 ? AHM_F64 : AHM_F32;
  const bool prec = a.M && a.M->on();
  GmresCache* c = h->gm;
  if (!c || c->n != n || c->m != m || c->S != S || c->vec != vec || c->batch != a.batch || c->prec != prec ||
      c->nhist_cap < a.nhist) {
    gmres_cache_free(h);
    c = new GmresCache;
    struct Guard {   // a cache that failed to build is not kept
      GmresCache* c;
      ~Guard() { delete c; }
    } guard{c};
    c->dev = h->ctx->device;
    c->n = n;
    c->m = m;
    c->S = S;
    c->vec = vec;
    c->batch = a.batch;
    c->prec = prec;
    c->nhist_cap = std::max<uint32_t>(a.nhist, 1);
    std::vector<SegChunk> chunks;
    std::vector<uint32_t> cb(S + 1, 0);
    for (uint32_t s = 0; s < S; ++s) {
      cb[s] = (uint32_t)chunks.size();
      for (uint32_t b0 = off[s]; b0 < off[s + 1]; b0 += kChunk)
        chunks.push_back(SegChunk{s, b0, std::min(b0 + kChunk, off[s + 1]), 0});
    }
    cb[S] = (uint32_t)chunks.size();
    c->nch = cb[S];
    for (uint32_t s = 0; s < S; ++s) c->max_nc = std::max(c->max_nc, cb[s + 1] - cb[s]);
    AHM_HIP(hipStreamCreateWithFlags(&c->gs, hipStreamNonBlocking));
    AHM_TRY(c->dch.alloc(chunks.size() * sizeof(SegChunk)));
    AHM_TRY(c->dcb.alloc(cb.size() * 4));
    AHM_HIP(hipMemcpyAsync(c->dch.p, chunks.data(), chunks.size() * sizeof(SegChunk), hipMemcpyHostToDevice,
                           c->gs));
    AHM_HIP(hipMemcpyAsync(c->dcb.p, cb.data(), cb.size() * 4, hipMemcpyHostToDevice, c->gs));
    AHM_TRY(c->dV.alloc((size_t)n * (m + 1) * 8));
    AHM_TRY(c->dr.alloc((size_t)n * 8));
    AHM_TRY(c->dp0.alloc((size_t)c->nch * 8));
    AHM_TRY(c->dp1.alloc((size_t)c->nch * 8));
    if (vec == AHM_F32) {
      AHM_TRY(c->dx32.alloc((size_t)n * 4));
      AHM_TRY(c->dy32.alloc((size_t)n * 4));
    }
    if (prec) AHM_TRY(c->dxh.alloc((size_t)n * 8));
    const size_t nd = 3 * (size_t)S + (size_t)S * (m + 1) * m + 3 * (size_t)S * (m + 1) + (size_t)S * m +
                      c->nhist_cap + 1;
    const size_t nu = 8 * (size_t)S + 3;
    AHM_TRY(c->dD.alloc(nd * 8));
    AHM_TRY(c->dU.alloc(nu * 4));
    double* d = (double*)c->dD.p;
    State& st = c->st;
    st.normb = d;
    st.beta = d + S;
    st.resid = d + 2 * S;
    st.H = d + 3 * S;
    st.cs = st.H + (size_t)S * (m + 1) * m;
    st.sn = st.cs + (size_t)S * (m + 1);
    st.sv = st.sn + (size_t)S * (m + 1);
    st.Y = st.sv + (size_t)S * (m + 1);
    st.hist = st.Y + (size_t)S * m;
    st.eps = st.hist + c->nhist_cap;
    uint32_t* u = (uint32_t*)c->dU.p;
    st.frozen = u;
    st.conv = u + S;
    st.taken = u + 2 * S;
    st.kcyc = u + 3 * S;
    st.kupd = u + 4 * S;
    st.rsteps = u + 5 * S;
    st.restarts = u + 6 * S;
    st.zero_b = u + 7 * S;
    st.nh = u + 8 * S;
    st.nhist = st.nh + 1;
    st.flag = st.nh + 2;
    AHM_HIP(hipHostMalloc((void**)&c->hflags, ((size_t)(m + 1) * S + 1) * 4, hipHostMallocDefault));
    c->graphs.assign(m, nullptr);
    c->uses.assign(m, 0u);
    c->ev.assign(m, nullptr);
    for (auto& e : c->ev) AHM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (S > 1) {
      std::vector<uint32_t> ts[4];
      AHM_TRY(task_segments(h, off, ts));
      size_t tot = 0;
      for (auto& v : ts) tot += v.size();
      AHM_TRY(c->dseg.alloc(std::max<size_t>(tot, 1) * 4));
      size_t o = 0;
      for (int q = 0; q < 4; ++q) {
        c->seg[q] = (const uint32_t*)c->dseg.p + o;
        if (!ts[q].empty())
          AHM_HIP(hipMemcpyAsync((uint32_t*)c->dseg.p + o, ts[q].data(), ts[q].size() * 4, hipMemcpyHostToDevice,
                                 c->gs));
        o += ts[q].size();
      }
    }
    AHM_HIP(hipStreamSynchronize(c->gs));
    guard.c = nullptr;
    h->gm = c;
  }
  out = c;
  return AHM_OK;
}

// The Arnoldi form of a solve: mode 1 the chunked launches, 2 one workgroup
// per system where every system has at most kArChunks chunks, 0 (auto) the
// latter where it is also the faster
bool arnoldi_one(int mode, uint32_t S, uint32_t max_nc) {
  if (max_nc > kArChunks || mode == 1) return false;
  if (mode == 2) return true;
  (void)S;
  return max_nc <= kArAutoChunks;
}

struct Run {
  ahm_hmat h;
  GmresCache& c;
  const GmresArgs& a;
  hipStream_t gs;
  uint32_t n, m, S, nch;
  bool one;   // the orthogonalisation and rotations in gm_arnoldi_kernel
  const SegChunk* ch() const { return (const SegChunk*)c.dch.p; }
  const uint32_t* cb() const { return (const uint32_t*)c.dcb.p; }
  double* V() const { return (double*)c.dV.p; }
  double* p0() const { return (double*)c.dp0.p; }
  double* p1() const { return (double*)c.dp1.p; }
  const uint32_t* frozen() const { return c.st.frozen; }

  // y += alpha A x on the running systems, FP64 vectors
  int op(double alpha, const double* x, double* y) {
    const uint32_t* const* sg = S > 1 ? c.seg : nullptr;
    if (c.vec == AHM_F64) return matvec_gated(h, alpha, x, y, AHM_F64, frozen(), sg);
    hipLaunchKernelGGL(gm_to_f32_kernel, dim3(nch), dim3(kGmThreads), 0, gs, x, (float*)c.dx32.p,
                       (float*)c.dy32.p, ch(), frozen());
    AHM_TRY(matvec_gated(h, alpha, c.dx32.p, c.dy32.p, AHM_F32, frozen(), sg));
    hipLaunchKernelGGL(gm_add_f32_kernel, dim3(nch), dim3(kGmThreads), 0, gs, (const float*)c.dy32.p, y, ch(),
                       frozen());
    return AHM_OK;
  }

  // Arnoldi step i (solvers/GMRES.cpp:123-157) for every running system;
  // ends with a copy of the frozen flags to pinned row i
  int step(uint32_t i) {
    double* vi = V() + (size_t)i * n;
    double* w = V() + (size_t)(i + 1) * n;
    const uint32_t hs = (m + 1) * m;
    hipLaunchKernelGGL(gm_zero_kernel, dim3(nch), dim3(kGmThreads), 0, gs, w, ch(), frozen());
    const double* src = vi;
    if (a.M && a.M->on()) {
      double* xh = (double*)c.dxh.p;
      hipLaunchKernelGGL(gm_copy_kernel, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)vi, xh, ch(),
                         frozen());
      AHM_TRY(a.M->apply(xh));
      src = xh;
    }
    AHM_TRY(op(1.0, src, w));
    if (one) {
      const size_t lds = (size_t)(3 * m + 4) * 8;
      if (c.max_nc <= kArGroups)
        hipLaunchKernelGGL(gm_arnoldi_kernel<1>, dim3(S), dim3(kArThreads), lds, gs, V(), (size_t)n, ch(), cb(),
                           c.st, m, i);
      else if (c.max_nc <= 2 * kArGroups)
        hipLaunchKernelGGL(gm_arnoldi_kernel<2>, dim3(S), dim3(kArThreads), lds, gs, V(), (size_t)n, ch(), cb(),
                           c.st, m, i);
      else
        hipLaunchKernelGGL(gm_arnoldi_kernel<4>, dim3(S), dim3(kArThreads), lds, gs, V(), (size_t)n, ch(), cb(),
                           c.st, m, i);
      AHM_HIP(hipGetLastError());
      AHM_HIP(hipMemcpyAsync(c.hflags + (size_t)i * S, c.st.frozen, (size_t)S * 4, hipMemcpyDeviceToHost, gs));
      return AHM_OK;
    }
    hipLaunchKernelGGL(gm_dot_kernel, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)w, (const double*)V(),
                       ch(), frozen(), p0());
    double* pin = p0();
    double* pout = p1();
    double* Hc = c.st.H + (size_t)i * (m + 1);
    for (uint32_t k = 0; k <= i; ++k) {
      const double* next = k < i ? V() + (size_t)(k + 1) * n : nullptr;
      hipLaunchKernelGGL(gm_mgs_kernel, dim3(nch), dim3(kGmThreads), 0, gs, w, (const double*)(V() + (size_t)k * n),
                         next, ch(), cb(), frozen(), (const double*)pin, Hc + k, hs, pout);
      std::swap(pin, pout);
    }
    hipLaunchKernelGGL(gm_normalize_kernel, dim3(nch), dim3(kGmThreads), 0, gs, w, ch(), cb(), frozen(),
                       (const double*)pin, Hc + i + 1, hs);
    hipLaunchKernelGGL(gm_givens_kernel, dim3(S), dim3(64), (size_t)(3 * m + 4) * 8, gs, c.st, m, i);
    AHM_HIP(hipGetLastError());
    AHM_HIP(hipMemcpyAsync(c.hflags + (size_t)i * S, c.st.frozen, (size_t)S * 4, hipMemcpyDeviceToHost, gs));
    return AHM_OK;
  }

  int running_after(uint32_t row) const {
    volatile const uint32_t* f = c.hflags + (size_t)row * S;
    uint32_t r = 0;
    for (uint32_t s = 0; s < S; ++s) r += f[s] == 0;
    return (int)r;
  }

  // frozen flags into pinned row `row`, flag word after the rows; waits
  int read_flags(uint32_t row) {
    AHM_HIP(hipMemcpyAsync(c.hflags + (size_t)row * S, c.st.frozen, (size_t)S * 4, hipMemcpyDeviceToHost, gs));
    AHM_HIP(hipMemcpyAsync(c.hflags + (size_t)(m + 1) * S, c.st.flag, 4, hipMemcpyDeviceToHost, gs));
    AHM_HIP(hipStreamSynchronize(gs));
    return AHM_OK;
  }

  // update x with every system's steps of this cycle (solvers/GMRES.cpp:58-77)
  int update() {
    hipLaunchKernelGGL(gm_solve_kernel, dim3(S), dim3(64), 0, gs, c.st, m);
    if (a.M && a.M->on()) {
      double* xh = (double*)c.dxh.p;
      hipLaunchKernelGGL(gm_apply_kernel<true>, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)V(), n, m, ch(),
                         c.st, xh);
      AHM_TRY(a.M->apply(xh));
      hipLaunchKernelGGL(gm_axpy_kernel, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)xh, a.x, ch());
    } else {
      hipLaunchKernelGGL(gm_apply_kernel<false>, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)V(), n, m,
                         ch(), c.st, a.x);
    }
    AHM_HIP(hipGetLastError());
    return AHM_OK;
  }

  // r = b - A x for the running systems, |r|, and the test of `mode`
  int residual(int mode) {
    double* r = (double*)c.dr.p;
    hipLaunchKernelGGL(gm_copy_kernel, dim3(nch), dim3(kGmThreads), 0, gs, a.b, r, ch(), frozen());
    AHM_TRY(op(-1.0, a.x, r));
    hipLaunchKernelGGL(gm_dot_kernel, dim3(nch), dim3(kGmThreads), 0, gs, (const double*)r, (const double*)r, ch(),
                       frozen(), p0());
    hipLaunchKernelGGL(gm_test_kernel, dim3(1), dim3(kGmThreads), 0, gs, (const double*)p0(),
                       (const double*)p1(), cb(), c.st, S, mode);
    AHM_HIP(hipGetLastError());
    return AHM_OK;
  }
};

}  // namespace

int gmres_run(ahm_hmat h, const GmresArgs& a) {
  if (a.m == 0 || !a.b || !a.x) return fail(AHM_ERR_INVALID, "gmres: bad arguments");
  if ((size_t)(3 * a.m + 4) * 8 > 60 * 1024) return fail(AHM_ERR_UNSUPPORTED, "gmres: restart length %u", a.m);
  const bool prec = a.M && a.M->on();
  if (prec && a.batch) return fail(AHM_ERR_UNSUPPORTED, "gmres: preconditioned batched solve");
  AHM_HIP(hipSetDevice(h->ctx->device));
  const ahm_dtype vec = h->storage == AHM_F64 ? AHM_F64 : AHM_F32;
  // the product's workspace exists before any capture; graphs made for
  // another layout are dropped
  AHM_TRY(matvec_prepare(h, vec));
  GmresCache* cp = nullptr;
  AHM_TRY(get_cache(h, a, cp));
  GmresCache& c = *cp;
  std::vector<uint64_t> sig;
  matvec_signature(h, sig);
  if (sig != c.sig) {
    c.drop_graphs();
    c.sig = sig;
  }
  const int one = arnoldi_one(h->arnoldi, c.S, c.max_nc) ? 1 : 0;
  if (one != c.ar_one) {
    c.drop_graphs();
    c.ar_one = one;
  }
  Run R{h, c, a, c.gs, c.n, c.m, c.S, c.nch, one != 0};
  const uint32_t S = c.S, m = c.m, n = c.n;
  hipStream_t user = h->ctx->stream;
  hipEvent_t ev0;
  AHM_HIP(hipEventCreateWithFlags(&ev0, hipEventDisableTiming));
  struct Restore {
    ahm_hmat h;
    hipStream_t s;
    hipEvent_t e;
    ~Restore() {
      h->ctx->stream = s;
      (void)hipEventDestroy(e);
    }
  } restore{h, user, ev0};
  // order after the caller's work; the products below run on the cache's stream
  AHM_HIP(hipEventRecord(ev0, user));
  AHM_HIP(hipStreamWaitEvent(c.gs, ev0, 0));
  h->ctx->stream = c.gs;

  hipLaunchKernelGGL(gm_reset_kernel, dim3(1), dim3(kGmThreads), 0, c.gs, c.st, S, a.eps, a.nhist);
  hipLaunchKernelGGL(gm_dot_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, a.b, a.b, R.ch(), c.st.frozen, R.p0());
  hipLaunchKernelGGL(gm_asum_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, (const double*)a.x, R.ch(), R.p1());
  hipLaunchKernelGGL(gm_test_kernel, dim3(1), dim3(kGmThreads), 0, c.gs, (const double*)R.p0(),
                     (const double*)R.p1(), R.cb(), c.st, S, 0);
  hipLaunchKernelGGL(gm_zero_masked_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, a.x, R.ch(),
                     (const uint32_t*)c.st.zero_b);
  AHM_HIP(hipGetLastError());
  AHM_TRY(R.read_flags(m));
  uint32_t mv_init = 0;
  int running = R.running_after(m);
  if (running) {
    // r = b - A x; the product is skipped when x is zero (A*0 adds exact
    // zeros, r = b bit for bit)
    const bool x_nonzero = c.hflags[(size_t)(m + 1) * S] != 0;
    double* r = (double*)c.dr.p;
    hipLaunchKernelGGL(gm_copy_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, a.b, r, R.ch(), c.st.frozen);
    if (x_nonzero) {
      AHM_TRY(R.op(-1.0, a.x, r));
      mv_init = 1;
    }
    hipLaunchKernelGGL(gm_dot_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, (const double*)r, (const double*)r,
                       R.ch(), c.st.frozen, R.p0());
    hipLaunchKernelGGL(gm_test_kernel, dim3(1), dim3(kGmThreads), 0, c.gs, (const double*)R.p0(),
                       (const double*)R.p1(), R.cb(), c.st, S, 1);
    AHM_HIP(hipGetLastError());
    AHM_TRY(R.read_flags(m));
    running = R.running_after(m);
  }
  const int lag = prec ? 0 : kLookahead;
  const bool graph = a.use_graph && !prec;
  uint32_t j = 1;
  while (running && j <= a.maxsteps) {
    hipLaunchKernelGGL(gm_cycle_kernel, dim3(c.nch), dim3(kGmThreads), 0, c.gs, (const double*)c.dr.p, R.V(),
                       R.ch(), R.cb(), c.st, m);
    bool stop = false;
    for (uint32_t i = 0; i < m && j <= a.maxsteps && !stop; ++i, ++j) {
      // a step runs one by one on its first use and is captured on its
      // second (a solve that never reaches a step index twice pays no capture)
      if (graph && c.uses[i]++ > 0) {
        if (!c.graphs[i]) {
          hipGraph_t g;
          AHM_HIP(hipStreamBeginCapture(c.gs, hipStreamCaptureModeRelaxed));
          const int st = R.step(i);
          const hipError_t ce = hipStreamEndCapture(c.gs, &g);
          if (st) return st;
          if (ce != hipSuccess) return fail(AHM_ERR_HIP, "gmres: graph capture: %s", hipGetErrorString(ce));
          const hipError_t ie = hipGraphInstantiate(&c.graphs[i], g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (ie != hipSuccess) {
            c.graphs[i] = nullptr;
            return fail(AHM_ERR_HIP, "gmres: graph instantiate: %s", hipGetErrorString(ie));
          }
        }
        AHM_HIP(hipGraphLaunch(c.graphs[i], c.gs));
      } else {
        AHM_TRY(R.step(i));
      }
      AHM_HIP(hipEventRecord(c.ev[i], c.gs));
      if ((int)i >= lag) {
        AHM_HIP(hipEventSynchronize(c.ev[i - lag]));
        if (R.running_after(i - lag) == 0) stop = true;
      }
    }
    AHM_TRY(R.update());
    if (stop) break;
    AHM_TRY(R.residual(2));
    AHM_TRY(R.read_flags(m));
    running = R.running_after(m);
  }

  // results
  std::vector<double> resid(S);
  std::vector<uint32_t> u(8 * (size_t)S + 1);
  std::vector<double> hist(c.nhist_cap);
  AHM_HIP(hipMemcpyAsync(resid.data(), c.st.resid, S * 8, hipMemcpyDeviceToHost, c.gs));
  AHM_HIP(hipMemcpyAsync(u.data(), c.st.frozen, u.size() * 4, hipMemcpyDeviceToHost, c.gs));
  if (a.hist && a.nhist) AHM_HIP(hipMemcpyAsync(hist.data(), c.st.hist, (size_t)a.nhist * 8, hipMemcpyDeviceToHost, c.gs));
  AHM_HIP(hipEventRecord(ev0, c.gs));
  AHM_HIP(hipStreamWaitEvent(user, ev0, 0));
  AHM_HIP(hipStreamSynchronize(c.gs));
  const uint32_t* conv = u.data() + S;
  const uint32_t* taken = u.data() + 2 * S;
  const uint32_t* rsteps = u.data() + 5 * S;
  const uint32_t* restarts = u.data() + 6 * S;
  const uint32_t nh = u[8 * (size_t)S];
  int all = 1;
  for (uint32_t s = 0; s < S; ++s) {
    if (a.eps_out) a.eps_out[s] = resid[s];
    if (a.steps_out) a.steps_out[s] = conv[s] ? rsteps[s] : a.maxsteps;
    if (a.conv_out) a.conv_out[s] = conv[s] ? 1 : 0;
    all &= conv[s] ? 1 : 0;
  }
  if (a.nmatvec) *a.nmatvec = mv_init + taken[0] + restarts[0];
  if (a.hist)
    for (uint32_t q = 0; q < a.nhist; ++q) a.hist[q] = q < nh ? hist[q] : 0.0;
  (void)n;
  return all ? AHM_OK : AHM_ERR_NOT_CONVERGED;
}

}  // namespace ahm

extern "C" int ahm_gmres_batch(ahm_hmat h, const double* b, double* x, double eps, uint32_t m,
                               uint32_t nsteps, int use_graph, double* eps_out,
                               uint32_t* steps_out, int* converged) {
  using namespace ahm;
  if (!h || !b || !x || m == 0 || h->storage != AHM_F64 || h->n_rows != h->n_cols)
    return fail(AHM_ERR_INVALID, "ahm_gmres_batch: bad arguments");
  GmresArgs a;
  a.b = b;
  a.x = x;
  a.eps = eps;
  a.m = m;
  a.maxsteps = nsteps;
  a.use_graph = use_graph;
  a.batch = true;
  a.eps_out = eps_out;
  a.steps_out = steps_out;
  a.conv_out = converged;
  const int st = gmres_run(h, a);
  if (st == AHM_ERR_NOT_CONVERGED) set_error("ahm_gmres_batch: some systems did not converge");
  return st;
}
