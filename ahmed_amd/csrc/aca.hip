// aca.hip -- H-matrix assembly on the GPU: batched ACA, batched recompression
// and dense leaf generation, all leaves in flight at once.
//
// Follows the reference's per-leaf rule apprx_unsym (Include/apprx.h:136-172):
// admissible leaves try ACA (Include/ACA.h:217-350) with
// maxk = min(n1*n2/(n1+n2)+4, rankmax); on success the factors are
// recompressed as cpyLrM_cmpr -> addtrll does (H/mblock_D.cpp:357-465): QR of
// U and of V, R = R1 R2^T with entries below delta*|R|_F/sqrt(mmin*nmin)
// zeroed, SVD of R, truncation at delta*s_0, singular values folded into U.
// Otherwise (or on ACA failure) the leaf is generated dense.
//
// One workgroup (256 threads, 4 wavefronts) runs one leaf; the sequential
// pivot logic runs uniformly in every thread, the vector work (kernel rows /
// columns, the k-term updates, norms, arg-max searches, Householder and
// Jacobi dot products) is spread over the workgroup.  Reductions use a fixed
// order, so results are bit-reproducible run to run.
//
// The kernel entries (the reference's MATGEN_T plugin, contract
// Include/matgen_sqntl.h:21-29) are evaluated on the device from the
// permuted points: entry(i,j) = 1 on i == j, else c/r (Laplace) or
// exp(-kappa r) c / r (decaying kernels).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "ahm_internal.h"
#include "ahm_aca.h"

namespace ahm {

constexpr int kAcaThreads = 256;
constexpr int kWaves = kAcaThreads / 64;
// The ACA kernel runs leaves of up to kSmallLeaf rows + columns one
// wavefront each (no workgroup barriers to wait on, four times as many leaves
// per CU), larger leaves with kAcaThreads threads; its helpers take the
// workgroup size from blockDim.
constexpr uint32_t kSmallLeaf = 1024;
constexpr int kSmallThreads = 64;
// leaves beyond kLargeLeaf rows + columns run with kLargeThreads threads
constexpr uint32_t kLargeLeaf = 8192;
constexpr int kLargeThreads = 512;
__device__ __forceinline__ uint32_t nthr() { return blockDim.x; }
__device__ __forceinline__ int nwaves() { return (int)(blockDim.x >> 6); }

__device__ __forceinline__ double kentry(const KernelSpec& K, const double* __restrict__ P,
                                         uint32_t i, uint32_t j) {
  if (i == j) return 1.0;
  const double dx = P[3 * (size_t)i] - P[3 * (size_t)j];
  const double dy = P[3 * (size_t)i + 1] - P[3 * (size_t)j + 1];
  const double dz = P[3 * (size_t)i + 2] - P[3 * (size_t)j + 2];
  const double r = sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
  if (K.kind == AHM_KERNEL_LAPLACE) return K.c / r;
  return (exp(-K.kappa * r) * K.c) / r;
}

// ------------------------------------------------------------------------
// workgroup reductions (fixed order: lanes by xor butterfly, then waves 0..3)

// a double moved between lanes by a DPP pattern (two 32-bit moves)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes (all active), the same value in every lane: within
// each row of 16 lanes by DPP exchanges (lanes paired 1-0/3-2, then 2-3/0-1,
// then mirrored within 8 and within 16; each step adds a lane's partner's
// partial, so partners hold equal sums), then the four row sums added in a
// fixed order.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);   // row_half_mirror
  v += dpp_f64<0x140>(v);   // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = sh[0];
  for (int q = 1; q < nwaves(); ++q) s += sh[q];
  return s;
}

// arg-max of v (ties -> smallest index); candidates carry v >= 0, others -1
struct ArgMax {
  double v;
  int i;
};
__device__ __forceinline__ ArgMax amax(ArgMax a, ArgMax b) {
  return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}
__device__ ArgMax block_argmax(ArgMax a, double* shv, int* shi) {
  for (int o = 32; o > 0; o >>= 1) {
    ArgMax b;
    b.v = __shfl_xor(a.v, o, 64);
    b.i = __shfl_xor(a.i, o, 64);
    a = amax(a, b);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    shv[w] = a.v;
    shi[w] = a.i;
  }
  __syncthreads();
  ArgMax r{shv[0], shi[0]};
  for (int q = 1; q < nwaves(); ++q) r = amax(r, ArgMax{shv[q], shi[q]});
  return r;
}

// arg-min of integer keys (ties -> smallest index); non-candidates INT_MAX
__device__ void block_argmin_int(int& key, int& idx, int* shk, int* shi) {
  for (int o = 32; o > 0; o >>= 1) {
    const int k2 = __shfl_xor(key, o, 64), i2 = __shfl_xor(idx, o, 64);
    if (k2 < key || (k2 == key && i2 < idx)) {
      key = k2;
      idx = i2;
    }
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    shk[w] = key;
    shi[w] = idx;
  }
  __syncthreads();
  key = shk[0];
  idx = shi[0];
  for (int q = 1; q < nwaves(); ++q)
    if (shk[q] < key || (shk[q] == key && shi[q] < idx)) {
      key = shk[q];
      idx = shi[q];
    }
}

// Long loops over the rows of a large leaf are bound by memory latency when
// each step waits for its own loads; these helpers issue eight steps' loads
// before using them and keep the order of the sums.

// v - sum_{l<k} c[l] * M[l*ld + i], terms subtracted in l order
__device__ __forceinline__ double sub_terms(double v, const double* c, const double* M, size_t ld, uint32_t i,
                                            uint32_t k) {
  uint32_t l = 0;
  for (; l + 8 <= k; l += 8) {
    double m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = M[(size_t)(l + u) * ld + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) v = fma(-c[l + u], m[u], v);
  }
  for (; l < k; ++l) v = fma(-c[l], M[(size_t)l * ld + i], v);
  return v;
}

// acc + sum a[i] b[i] over i = i0, i0 + step, ... < n, in that order
__device__ __forceinline__ double strided_dot(const double* a, const double* b, uint32_t i0, uint32_t n,
                                              uint32_t step, double acc) {
  uint32_t i = i0;
  for (; i + 7 * step < n; i += 8 * step) {
    double x[8], y[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      x[u] = a[i + u * step];
      y[u] = b[i + u * step];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = fma(x[u], y[u], acc);
  }
  for (; i < n; i += step) acc = fma(a[i], b[i], acc);
  return acc;
}

// x[i] -= s * (i == j ? 1 : a[i]) for i = i0, i0 + step, ... < n (a Householder
// update of one column; a is not written)
__device__ __forceinline__ void house_axpy(double* x, const double* a, double s, uint32_t j, uint32_t i0,
                                           uint32_t n, uint32_t step) {
  uint32_t i = i0;
  for (; i + 7 * step < n; i += 8 * step) {
    double xv[8], av[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      xv[u] = x[i + u * step];
      av[u] = a[i + u * step];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t ii = i + u * step;
      x[ii] = xv[u] - s * (ii == j ? 1.0 : av[u]);
    }
  }
  for (; i < n; i += step) x[i] -= s * (i == j ? 1.0 : a[i]);
}

// ACA steps with at most kFuseK earlier rank columns form the dot products
// of the new vectors with the earlier ones (the cross term of the Frobenius
// norm update) inside the remainder passes, from the values those passes
// load anyway, so the factors are streamed once per step instead of twice.
constexpr uint32_t kFuseK = 16;

// per-wavefront partials of the workgroup reductions (up to kLargeThreads)
constexpr int kMaxWaves = kLargeThreads / 64;

struct Shared {
  double dv[kMaxWaves];
  double av[kMaxWaves];
  int ai[kMaxWaves];
  int ki[kMaxWaves];
  int kk[kMaxWaves];
};

// LDS of the fused dots (ACA kernel only)
struct DotShared {
  double dw[kMaxWaves][kFuseK];   // per-wave partial dots
  double dA[kFuseK], dB[kFuseK], dC[kFuseK];
};

// ------------------------------------------------------------------------
// ACA (Include/ACA.h:29-350), one leaf per workgroup

// next index >= j (cyclic) with Z != -1 (nexti, Include/ACA.h:29-36)
__device__ uint32_t nexti_dev(uint32_t n, const int* Z, uint32_t j) {
  for (;;) {
    if (Z[j] != -1) return j;
    j = (j + 1 < n) ? j + 1 : 0;
  }
}

struct AcaState {
  uint32_t no;
  double nrmlsk2;
};

// pv = row i0 of the remainder (Include/ACA.h:57-64); returns arg-max over S>=0
__device__ ArgMax remainder_row(const KernelSpec& K, const double* P, const AcaBlock& B,
                                uint32_t i0, uint32_t k, const double* U, const double* V,
                                double* pv, const int* S, double* ucoef, Shared& sh) {
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < k; l += nthr()) ucoef[l] = U[(size_t)l * B.n1 + i0];
  __syncthreads();
  ArgMax a{-1.0, 0x7fffffff};
  for (uint32_t j = threadIdx.x; j < B.n2; j += nthr()) {
    const double v = sub_terms(kentry(K, P, B.b1 + i0, B.b2 + j), ucoef, V, B.n2, j, k);
    pv[j] = v;
    if (S[j] >= 0) a = amax(a, ArgMax{fabs(v), (int)j});
  }
  return block_argmax(a, sh.av, sh.ai);
}

// pu = column j0 of the remainder (Include/ACA.h:143-158)
__device__ ArgMax remainder_col(const KernelSpec& K, const double* P, const AcaBlock& B,
                                uint32_t j0, uint32_t k, const double* U, const double* V,
                                double* pu, const int* Z, double* vcoef, Shared& sh) {
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < k; l += nthr()) vcoef[l] = V[(size_t)l * B.n2 + j0];
  __syncthreads();
  ArgMax a{-1.0, 0x7fffffff};
  for (uint32_t i = threadIdx.x; i < B.n1; i += nthr()) {
    const double v = sub_terms(kentry(K, P, B.b1 + i, B.b2 + j0), vcoef, U, B.n1, i, k);
    pu[i] = v;
    if (Z[i] >= 0) a = amax(a, ArgMax{fabs(v), (int)i});
  }
  return block_argmax(a, sh.av, sh.ai);
}

// Per-thread dot accumulators of the fused remainder passes, in dynamic LDS
// (as registers they would halve the kernel's occupancy): element
// acc[l * blockDim + thread], kFuseK x blockDim values.
typedef double* DotAcc;

// acc[l][thread] += M_l[i] * v for l < k <= kFuseK, eight loads in flight
__device__ __forceinline__ void dots_acc(DotAcc acc, double v, const double* M, size_t ld, uint32_t i,
                                         uint32_t k) {
  const uint32_t t = threadIdx.x, nt = nthr();
  uint32_t l = 0;
  for (; l + 8 <= k; l += 8) {
    double m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = M[(size_t)(l + u) * ld + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[(l + u) * nt + t] = fma(m[u], v, acc[(l + u) * nt + t]);
  }
  for (; l < k; ++l) acc[l * nt + t] = fma(M[(size_t)l * ld + i], v, acc[l * nt + t]);
}

__device__ __forceinline__ void dots_zero(DotAcc acc, uint32_t k) {
  for (uint32_t l = 0; l < k; ++l) acc[l * nthr() + threadIdx.x] = 0.0;
}

// out[l] = sum of acc[l][.] over the workgroup (l < k): lanes by wave_sum,
// then waves in order; ends with a barrier
__device__ void block_dots(DotAcc acc, uint32_t k, double* out, DotShared& sh) {
  const int w = threadIdx.x >> 6;
  for (uint32_t l = 0; l < k; ++l) {
    const double s = wave_sum(acc[l * nthr() + threadIdx.x]);
    if ((threadIdx.x & 63) == 0) sh.dw[w][l] = s;
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double s = sh.dw[0][threadIdx.x];
    for (int q = 1; q < nwaves(); ++q) s += sh.dw[q][threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// remainder_row for k <= kFuseK, with the dots of the unscaled row pv with
// the earlier V columns: dA[l] over the live columns (S >= 0, the ones the
// pivot scaling reaches), dB[l] over the retired ones.  pv and the arg-max
// are those of remainder_row bit for bit (same terms in the same order).
// The dots re-read the column's V entries (cache hits) after pv is stored.
__device__ ArgMax remainder_row_dots(const KernelSpec& K, const double* P, const AcaBlock& B,
                                     uint32_t i0, uint32_t k, const double* U, const double* V,
                                     double* pv, const int* S, double* ucoef, Shared& sh, DotAcc acc,
                                     DotShared& ds) {
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < k; l += nthr()) ucoef[l] = U[(size_t)l * B.n1 + i0];
  __syncthreads();
  ArgMax a{-1.0, 0x7fffffff};
  dots_zero(acc, k);
  bool retired = false;
  // two columns per iteration: both columns' loads are issued before either
  // is stored, so twice as many are in flight
  const uint32_t nt = nthr();
  for (uint32_t j = threadIdx.x; j < B.n2; j += 2 * nt) {
    const uint32_t j2 = j + nt;
    const bool two = j2 < B.n2;
    const int s1 = S[j], s2 = two ? S[j2] : -1;
    const double v = sub_terms(kentry(K, P, B.b1 + i0, B.b2 + j), ucoef, V, B.n2, j, k);
    const double v2 = two ? sub_terms(kentry(K, P, B.b1 + i0, B.b2 + j2), ucoef, V, B.n2, j2, k) : 0.0;
    pv[j] = v;
    if (two) pv[j2] = v2;
    if (s1 >= 0) {
      a = amax(a, ArgMax{fabs(v), (int)j});
      dots_acc(acc, v, V, B.n2, j, k);
    } else {
      retired = true;
    }
    if (two) {
      if (s2 >= 0) {
        a = amax(a, ArgMax{fabs(v2), (int)j2});
        dots_acc(acc, v2, V, B.n2, j2, k);
      } else {
        retired = true;
      }
    }
  }
  block_dots(acc, k, ds.dA, ds);
  // the retired columns (at most a few per leaf) in a second, sparse pass
  dots_zero(acc, k);
  if (retired)
    for (uint32_t j = threadIdx.x; j < B.n2; j += nthr())
      if (S[j] < 0) dots_acc(acc, pv[j], V, B.n2, j, k);
  block_dots(acc, k, ds.dB, ds);
  return block_argmax(a, sh.av, sh.ai);
}

// remainder_col for k <= kFuseK, with dC[l] = U_l . pu (unscaled pu)
__device__ ArgMax remainder_col_dots(const KernelSpec& K, const double* P, const AcaBlock& B,
                                     uint32_t j0, uint32_t k, const double* U, const double* V,
                                     double* pu, const int* Z, double* vcoef, Shared& sh, DotAcc acc,
                                     DotShared& ds) {
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < k; l += nthr()) vcoef[l] = V[(size_t)l * B.n2 + j0];
  __syncthreads();
  ArgMax a{-1.0, 0x7fffffff};
  dots_zero(acc, k);
  const uint32_t nt = nthr();
  for (uint32_t i = threadIdx.x; i < B.n1; i += 2 * nt) {   // two rows per iteration
    const uint32_t i2 = i + nt;
    const bool two = i2 < B.n1;
    const int z1 = Z[i], z2 = two ? Z[i2] : -1;
    const double v = sub_terms(kentry(K, P, B.b1 + i, B.b2 + j0), vcoef, U, B.n1, i, k);
    const double v2 = two ? sub_terms(kentry(K, P, B.b1 + i2, B.b2 + j0), vcoef, U, B.n1, i2, k) : 0.0;
    pu[i] = v;
    if (two) pu[i2] = v2;
    if (z1 >= 0) a = amax(a, ArgMax{fabs(v), (int)i});
    dots_acc(acc, v, U, B.n1, i, k);
    if (two) {
      if (z2 >= 0) a = amax(a, ArgMax{fabs(v2), (int)i2});
      dots_acc(acc, v2, U, B.n1, i2, k);
    }
  }
  block_dots(acc, k, ds.dC, ds);
  return block_argmax(a, sh.av, sh.ai);
}

// sum_{l<k} (U_l.U_k)(V_l.V_k) (Include/ACA.h:257-260) from the fused dots:
// U_k = dsca * pu everywhere, V_k = sca * pv on the live columns and pv on
// the retired ones
__device__ __forceinline__ double fused_cross(const DotShared& sh, uint32_t k, double sca, double dsca) {
  double sum = 0.0;
  for (uint32_t l = 0; l < k; ++l) sum += (dsca * sh.dC[l]) * fma(sca, sh.dA[l], sh.dB[l]);
  return sum;
}

// ACA_row_step (Include/ACA.h:46-130).  Returns false if no pivot was found.
// With k <= kFuseK the cross term of the norm update comes back in cross.
__device__ bool aca_row_step(const KernelSpec& K, const double* P, const AcaBlock& B,
                             uint32_t& klast, uint32_t& i0, uint32_t k, AcaState& st, int* Z,
                             int* S, double* U, double* V, double scale, double* coef,
                             Shared& sh, DotAcc dacc, DotShared& ds, double& cross) {
  double* pu = U + (size_t)k * B.n1;
  double* pv = V + (size_t)k * B.n2;
  const bool fuse = k <= kFuseK;
  ArgMax m;
  for (;;) {
    m = fuse ? remainder_row_dots(K, P, B, i0, k, U, V, pv, S, coef, sh, dacc, ds)
             : remainder_row(K, P, B, i0, k, U, V, pv, S, coef, sh);
    if (!(fmax(m.v, 0.0) < 1e-14 * scale)) break;
    __syncthreads();
    if (threadIdx.x == 0) Z[i0] = -1;
    __syncthreads();
    if (++st.no < B.n1) i0 = nexti_dev(B.n1, Z, i0);
    if (!(st.no < B.n1)) break;
  }
  // m.v < 0: no candidate column at all (the reference reads an unset j0 there)
  if (m.v < 0.0 || fmax(m.v, 0.0) < 1e-14 * scale) return false;
  const uint32_t j0 = (uint32_t)m.i;
  const double sqrtpiv = sqrt(m.v);
  const double thr = 1e-8 * scale * sqrtpiv;
  __syncthreads();
  const double sca = sqrtpiv / pv[j0];
  __syncthreads();
  double nv = 0.0;
  for (uint32_t j = threadIdx.x; j < B.n2; j += nthr())
    if (S[j] >= 0) {
      const double e = (pv[j] *= sca);
      nv += e * e;
      if (fabs(e) > thr) ++S[j];
    }
  const double nrmv2 = block_sum(nv, sh.dv);
  // column j0 of the remainder; pivot row/column retired afterwards
  if (fuse)
    remainder_col_dots(K, P, B, j0, k, U, V, pu, Z, coef, sh, dacc, ds);
  else
    remainder_col(K, P, B, j0, k, U, V, pu, Z, coef, sh);
  __syncthreads();
  klast = (uint32_t)Z[i0];
  __syncthreads();
  if (threadIdx.x == 0) {
    Z[i0] = -1;
    S[j0] = -1;
  }
  __syncthreads();
  ++st.no;
  const double dsca = 1.0 / sqrtpiv;
  double nu = 0.0;
  ArgMax a{-1.0, 0x7fffffff};
  for (uint32_t i = threadIdx.x; i < B.n1; i += nthr()) {
    const double eabs = fabs(pu[i] *= dsca);
    if (Z[i] >= 0 && eabs > thr) ++Z[i];
    nu += eabs * eabs;
    if (Z[i] >= 0) a = amax(a, ArgMax{eabs, (int)i});
  }
  const double nrmu2 = block_sum(nu, sh.dv);
  a = block_argmax(a, sh.av, sh.ai);
  if (a.v > 0.0) i0 = (uint32_t)a.i;
  st.nrmlsk2 = nrmu2 * nrmv2;
  if (fuse) cross = fused_cross(ds, k, sca, dsca);
  return true;
}

// ACA_col_step (Include/ACA.h:134-213)
__device__ bool aca_col_step(const KernelSpec& K, const double* P, const AcaBlock& B,
                             uint32_t& j0, uint32_t k, AcaState& st, int* Z, int* S, double* U,
                             double* V, double scale, double* coef, Shared& sh) {
  double* pu = U + (size_t)k * B.n1;
  double* pv = V + (size_t)k * B.n2;
  ArgMax m;
  for (;;) {
    m = remainder_col(K, P, B, j0, k, U, V, pu, Z, coef, sh);
    if (!(fmax(m.v, 0.0) < 1e-14 * scale)) break;
    __syncthreads();
    if (threadIdx.x == 0) S[j0] = -1;
    __syncthreads();
    if (++st.no < B.n2) j0 = nexti_dev(B.n2, S, j0);
    if (!(st.no < B.n2)) break;
  }
  if (m.v < 0.0 || fmax(m.v, 0.0) < 1e-14 * scale) return false;
  const uint32_t i0 = (uint32_t)m.i;
  const double sqrtpiv = sqrt(m.v);
  const double thr = 1e-8 * scale * sqrtpiv;
  __syncthreads();
  const double sca = sqrtpiv / pu[i0];
  __syncthreads();
  double nu = 0.0;
  for (uint32_t i = threadIdx.x; i < B.n1; i += nthr())
    if (Z[i] >= 0) {
      const double e = (pu[i] *= sca);
      nu += e * e;
      if (fabs(e) > thr) ++Z[i];
    }
  const double nrmu2 = block_sum(nu, sh.dv);
  remainder_row(K, P, B, i0, k, U, V, pv, S, coef, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    Z[i0] = -1;
    S[j0] = -1;
  }
  __syncthreads();
  ++st.no;
  const double dsca = 1.0 / sqrtpiv;
  double nv = 0.0;
  ArgMax a{-1.0, 0x7fffffff};
  for (uint32_t j = threadIdx.x; j < B.n2; j += nthr()) {
    const double eabs = fabs(pv[j] *= dsca);
    if (S[j] >= 0 && eabs > thr) ++S[j];
    nv += eabs * eabs;
    a = amax(a, ArgMax{eabs, (int)j});
  }
  const double nrmv2 = block_sum(nv, sh.dv);
  a = block_argmax(a, sh.av, sh.ai);
  if (a.v > 0.0) j0 = (uint32_t)a.i;
  st.nrmlsk2 = nrmu2 * nrmv2;
  return true;
}

// ACAr_step (Include/ACA.h:578-655): row pivoting only; every column is a
// pivot candidate (S stays zero) and no success counters are kept
__device__ bool acar_step(const KernelSpec& K, const double* P, const AcaBlock& B, uint32_t& i0,
                          uint32_t k, AcaState& st, int* Z, const int* S, double* U, double* V,
                          double scale, double* coef, Shared& sh, DotAcc dacc, DotShared& ds, double& cross) {
  double* pu = U + (size_t)k * B.n1;
  double* pv = V + (size_t)k * B.n2;
  const bool fuse = k <= kFuseK;
  ArgMax m;
  for (;;) {
    m = fuse ? remainder_row_dots(K, P, B, i0, k, U, V, pv, S, coef, sh, dacc, ds)
             : remainder_row(K, P, B, i0, k, U, V, pv, S, coef, sh);
    if (!(fmax(m.v, 0.0) < 1e-14 * scale)) break;
    __syncthreads();
    if (threadIdx.x == 0) Z[i0] = -1;
    __syncthreads();
    if (++st.no < B.n1) i0 = nexti_dev(B.n1, Z, i0);
    if (!(st.no < B.n1)) break;
  }
  if (m.v < 0.0 || fmax(m.v, 0.0) < 1e-14 * scale) return false;
  const uint32_t j0 = (uint32_t)m.i;
  const double sqrtpiv = sqrt(m.v);
  __syncthreads();
  const double sca = sqrtpiv / pv[j0];
  __syncthreads();
  double nv = 0.0;
  for (uint32_t j = threadIdx.x; j < B.n2; j += nthr()) {
    const double e = (pv[j] *= sca);
    nv += e * e;
  }
  const double nrmv2 = block_sum(nv, sh.dv);
  if (fuse)
    remainder_col_dots(K, P, B, j0, k, U, V, pu, Z, coef, sh, dacc, ds);
  else
    remainder_col(K, P, B, j0, k, U, V, pu, Z, coef, sh);
  __syncthreads();
  if (threadIdx.x == 0) Z[i0] = -1;
  __syncthreads();
  ++st.no;
  const double dsca = 1.0 / sqrtpiv;
  double nu = 0.0;
  ArgMax a{-1.0, 0x7fffffff};
  for (uint32_t i = threadIdx.x; i < B.n1; i += nthr()) {
    const double eabs = fabs(pu[i] *= dsca);
    nu += eabs * eabs;
    if (Z[i] >= 0) a = amax(a, ArgMax{eabs, (int)i});
  }
  const double nrmu2 = block_sum(nu, sh.dv);
  a = block_argmax(a, sh.av, sh.ai);
  if (a.v > 0.0) i0 = (uint32_t)a.i;
  st.nrmlsk2 = nrmu2 * nrmv2;
  if (fuse) cross = fused_cross(ds, k, sca, dsca);
  return true;
}

// sum_{l<k} (U_l.U_k)(V_l.V_k): one wavefront per term, lanes over rows
// (steps with more than kFuseK earlier columns)
__device__ double cross_sum(const AcaBlock& B, uint32_t k, const double* U, const double* V,
                            Shared& sh) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double* uk = U + (size_t)k * B.n1;
  const double* vk = V + (size_t)k * B.n2;
  double acc = 0.0;  // per-wave sum of its terms, in l order
  for (uint32_t l = w; l < k; l += nwaves()) {
    double du = strided_dot(U + (size_t)l * B.n1, uk, lane, B.n1, 64, 0.0);
    double dv = strided_dot(V + (size_t)l * B.n2, vk, lane, B.n2, 64, 0.0);
    du = wave_sum(du);
    dv = wave_sum(dv);
    acc += du * dv;
  }
  if (lane != 0) acc = 0.0;
  return block_sum(acc, sh.dv);
}

template <int MAXT>
__global__ __launch_bounds__(MAXT, 4) void aca_kernel(const AcaBlock* __restrict__ blocks,
                                                          AcaResult* __restrict__ res,
                                                          const double* __restrict__ P,
                                                          const KernelSpec* __restrict__ Ks,
                                                          double eps,
                                                          double* __restrict__ slab,
                                                          int* __restrict__ iscr) {
  __shared__ Shared sh;
  __shared__ double coef[kAcaCapMax + 1];
  __shared__ DotShared ds;
  extern __shared__ double dacc[];   // kFuseK x blockDim dot accumulators
  const AcaBlock B = blocks[blockIdx.x];
  const KernelSpec K = Ks[B.kidx];
  double* U = slab + B.uoff;
  double* V = slab + B.voff;
  int* Z = iscr + B.zoff;
  int* S = Z + B.n1;
  for (uint32_t i = threadIdx.x; i < B.n1; i += nthr()) Z[i] = 0;
  for (uint32_t j = threadIdx.x; j < B.n2; j += nthr()) S[j] = 0;
  __syncthreads();
  AcaState st{0, 0.0};
  int kmin = -2, keps = -1;
  uint32_t k = 0, klast = 0, inext, j0 = 0, i0 = B.i0;
  double nrms2 = 0.0;
  double scale = fabs(kentry(K, P, B.b1 + B.n1 / 2, B.b2 + B.n2 / 2));
  int status = 0;  // 0 fail, 1 success, 2 slab capacity reached
  const uint32_t kmax = B.kmax;
  if (K.rows_only) {
    // ACAr (Include/ACA.h:657-720); its loop compares the rows tried with n2
    do {
      if (k > B.kcap) {
        status = 2;
        goto DONE;
      }
      double cross = 0.0;
      if (!acar_step(K, P, B, i0, k, st, Z, S, U, V, scale, coef, sh, dacc, ds, cross)) {
        st.nrmlsk2 = 0.0;
        goto SUCC;
      }
      {
        const double sum = k <= kFuseK ? cross : cross_sum(B, k, U, V, sh);
        nrms2 += 2.0 * sum + st.nrmlsk2;
      }
      ++k;
      scale = sqrt(st.nrmlsk2 / ((double)B.n1 * B.n2));
    } while (st.no < B.n2 && k < kmax && st.nrmlsk2 >= eps * eps * nrms2);
    goto SUCC;
  }
  do {
    if (k > B.kcap) {
      status = 2;
      goto DONE;
    }
    inext = i0;
    double cross = 0.0;
    if (!aca_row_step(K, P, B, klast, inext, k, st, Z, S, U, V, scale, coef, sh, dacc, ds, cross)) {
      st.nrmlsk2 = 0.0;
      goto SUCC;
    }
    {
      const double sum = k <= kFuseK ? cross : cross_sum(B, k, U, V, sh);
      nrms2 += 2.0 * sum + st.nrmlsk2;
    }
    if (st.nrmlsk2 < eps * eps * nrms2 && keps == -1) keps = (int)(klast + 1u);
    ++k;
    scale = sqrt(st.nrmlsk2 / ((double)B.n1 * B.n2));
    if (keps > 0) {
      int key = 0x7fffffff, idx = 0x7fffffff;
      for (uint32_t i = threadIdx.x; i < B.n1; i += nthr())
        if (Z[i] >= 0 && (Z[i] < key || (Z[i] == key && (int)i < idx))) {
          key = Z[i];
          idx = (int)i;
        }
      block_argmin_int(key, idx, sh.kk, sh.ki);
      if (key == 0x7fffffff) {
        st.nrmlsk2 = 0.0;
        goto SUCC;
      }
      kmin = key;
      i0 = (uint32_t)idx;
      if ((uint32_t)kmin == k) i0 = inext;
    } else
      i0 = inext;
  } while (st.no < B.n1 && k < kmax && kmin < keps);

  if ((st.no == B.n1 || k == kmax) && kmin < keps) goto NSUCC;

  do {
    {
      int key = 0x7fffffff, idx = 0x7fffffff;
      for (uint32_t j = threadIdx.x; j < B.n2; j += nthr())
        if (S[j] >= 0 && (S[j] < key || (S[j] == key && (int)j < idx))) {
          key = S[j];
          idx = (int)j;
        }
      block_argmin_int(key, idx, sh.kk, sh.ki);
      if (key == 0x7fffffff) {
        st.nrmlsk2 = 0.0;
        goto SUCC;
      }
      kmin = key;
      j0 = (uint32_t)idx;
    }
    if (kmin >= keps) goto SUCC;
    if (k > B.kcap) {
      status = 2;
      goto DONE;
    }
    if (!aca_col_step(K, P, B, j0, k, st, Z, S, U, V, scale, coef, sh)) {
      st.nrmlsk2 = 0.0;
      goto SUCC;
    }
    ++k;
  } while (st.no < B.n2 && k < kmax);

NSUCC:
  status = 0;
  goto DONE;
SUCC : {
  const double epsa = sqrt(st.nrmlsk2 / nrms2);
  status = (epsa >= eps) ? 0 : 1;  // NaN (0/0) compares false: success
}
DONE:
  if (threadIdx.x == 0) {
    res[blockIdx.x].status = status;
    res[blockIdx.x].k = k;
    res[blockIdx.x].kt = 0;
    res[blockIdx.x].pad = 0;   // recompression method (set by recompress_kernel)
  }
}

// ------------------------------------------------------------------------
// recompression (addtrll, H/mblock_D.cpp:357-465)

// Householder QR in place of A (n x k, column-major, ld n): reflector j has
// v = [1; A[j+1:, j]], tau[j]; R in the upper triangle (LAPACK dgeqr2 form).
__device__ void house_qr(double* A, uint32_t n, uint32_t k, double* tau, Shared& sh,
                         double* wbuf) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t r = min(n, k);
  for (uint32_t j = 0; j < r; ++j) {
    double* a = A + (size_t)j * n;
    const double s = strided_dot(a, a, j + 1 + threadIdx.x, n, nthr(), 0.0);
    const double xn2 = block_sum(s, sh.dv);
    __syncthreads();
    const double alpha = a[j];
    double t = 0.0, beta = alpha, scal = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    if (xn2 > 0.0)
      for (uint32_t i = j + 1 + threadIdx.x; i < n; i += nthr()) a[i] *= scal;
    if (threadIdx.x == 0) {
      a[j] = beta;
      tau[j] = t;
    }
    __syncthreads();
    if (t == 0.0) continue;
    // w_c = v^T A[:, c] for c > j, one wavefront per column
    for (uint32_t c = j + 1 + w; c < k; c += nwaves()) {
      const double* ac = A + (size_t)c * n;
      double d = strided_dot(a, ac, j + 1 + lane, n, 64, 0.0);
      d = wave_sum(d);
      if (lane == 0) wbuf[c] = d + ac[j];
    }
    __syncthreads();
    for (uint32_t c = j + 1; c < k; ++c) {
      const double tw = t * wbuf[c];
      house_axpy(A + (size_t)c * n, a, tw, j, j + threadIdx.x, n, nthr());
    }
    __syncthreads();
  }
}

// X (n x m, ld n) <- Q X with Q = H_0 H_1 ... H_{r-1} from house_qr
__device__ void apply_q(const double* A, uint32_t n, uint32_t r, const double* tau, double* X,
                        uint32_t m, double* wbuf) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = (int)r - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;
    const double* a = A + (size_t)j * n;
    __syncthreads();
    for (uint32_t c = w; c < m; c += kWaves) {
      const double* xc = X + (size_t)c * n;
      double d = strided_dot(a, xc, (uint32_t)j + 1 + lane, n, 64, 0.0);
      d = wave_sum(d);
      if (lane == 0) wbuf[c] = d + xc[j];
    }
    __syncthreads();
    for (uint32_t c = 0; c < m; ++c) {
      const double tw = t * wbuf[c];
      house_axpy(X + (size_t)c * n, a, tw, (uint32_t)j, (uint32_t)j + threadIdx.x, n, kAcaThreads);
    }
  }
  __syncthreads();
}

// Q [C; 0] into O (n x kt, ld n) for Q = H_0 ... H_{r-1}, the reflectors of
// A (n x r, ld n; unit diagonal implicit, vector below it) and tau, and the
// core C (r x kt, ld r), in compact WY form: Q = I - Y T Y^T with
// T^{-1} = diag(1/tau) + the strictly upper part of Y^T Y (the UT transform
// of Joffrain, Low, Quintana-Orti, van de Geijn and Van Zee, 2006).  Then
// Q [C; 0] = [C; 0] - Y M with T^{-1} M = Y_top^T C (Y_top: rows 0..r-1),
// solved by back substitution; tau_j = 0 (H_j = I) gives a zero row j of M.
// Dots of Y^T Y all at once, a triangular solve per column and a row-parallel
// product take four barriers instead of two per reflector in apply_q.
// lds holds n*r + r*r + 2*r*kt + r values.
__device__ void wy_apply(const double* A, uint32_t n, uint32_t r, const double* tau,
                         const double* C, uint32_t kt, double* O, double* lds) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* Y = lds;                      // n x r
  double* G = Y + (size_t)n * r;        // r x r, strictly upper part
  double* Cl = G + (size_t)r * r;       // r x kt
  double* S = Cl + (size_t)r * kt;      // r x kt: Y_top^T C, then M
  double* tl = S + (size_t)r * kt;      // r: tau
  for (uint32_t e = threadIdx.x; e < n * r; e += kAcaThreads) Y[e] = A[e];
  for (uint32_t e = threadIdx.x; e < r; e += kAcaThreads) tl[e] = tau[e];
  for (uint32_t e = threadIdx.x; e < r * kt; e += kAcaThreads) Cl[e] = C[e];
  __syncthreads();
  // G[i + j r] = y_i . y_j (i < j) = A[j, i] + sum_{m > j} A[m, i] A[m, j]
  const uint32_t npair = r * (r - 1) / 2;
  auto pair_of = [](uint32_t p, uint32_t& i, uint32_t& j) {  // p = j (j - 1) / 2 + i
    j = (uint32_t)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
    while (j * (j - 1) / 2 > p) --j;
    while ((j + 1) * j / 2 <= p) ++j;
    i = p - j * (j - 1) / 2;
  };
  if (n <= 256) {  // one thread per pair
    for (uint32_t p = threadIdx.x; p < npair; p += kAcaThreads) {
      uint32_t i, j;
      pair_of(p, i, j);
      const double* yi = Y + (size_t)i * n;
      const double* yj = Y + (size_t)j * n;
      double d = yi[j];
      for (uint32_t m = j + 1; m < n; ++m) d = fma(yi[m], yj[m], d);
      G[i + (size_t)j * r] = d;
    }
  } else {  // one wavefront per pair
    for (uint32_t p = w; p < npair; p += kWaves) {
      uint32_t i, j;
      pair_of(p, i, j);
      const double* yi = Y + (size_t)i * n;
      const double* yj = Y + (size_t)j * n;
      double d = 0.0;
      for (uint32_t m = j + 1 + lane; m < n; m += 64) d = fma(yi[m], yj[m], d);
      d = wave_sum(d);
      if (lane == 0) G[i + (size_t)j * r] = yi[j] + d;
    }
  }
  // S[i + c r] = C[i, c] + sum_{i < m < r} A[m, i] C[m, c]  (r <= 65 here)
  {
    const uint32_t per = kAcaThreads / r, sub = threadIdx.x / r, i = threadIdx.x - sub * r;
    if (sub < per)
      for (uint32_t c = sub; c < kt; c += per) {
        const double* cc = Cl + (size_t)c * r;
        const double* yi = Y + (size_t)i * n;
        double v = cc[i];
        for (uint32_t m = i + 1; m < r; ++m) v = fma(yi[m], cc[m], v);
        S[i + (size_t)c * r] = v;
      }
  }
  __syncthreads();
  // M[i, c] = tau_i (S[i, c] - sum_{l > i} G[i, l] M[l, c]), in place over S
  for (uint32_t c = threadIdx.x; c < kt; c += kAcaThreads) {
    double* sc = S + (size_t)c * r;
    for (int i = (int)r - 1; i >= 0; --i) {
      double v = sc[i];
      for (uint32_t l = (uint32_t)i + 1; l < r; ++l) v = fma(-G[i + (size_t)l * r], sc[l], v);
      sc[i] = tl[i] * v;
    }
  }
  __syncthreads();
  // O[m, c] = [m < r] C[m, c] - sum_{i < min(m, r)} A[m, i] M[i, c] - [m < r] M[m, c]
  const uint32_t step = n >= kAcaThreads ? kAcaThreads : n, per = kAcaThreads / step;
  const uint32_t sub = threadIdx.x / step, m0 = threadIdx.x - sub * step;
  if (sub < per)
    for (uint32_t c = sub; c < kt; c += per) {
      const double* mc = S + (size_t)c * r;
      for (uint32_t m = m0; m < n; m += step) {
        const uint32_t ie = m < r ? m : r;
        double v = 0.0;
        for (uint32_t i = 0; i < ie; ++i) v = fma(Y[m + (size_t)i * n], mc[i], v);
        O[m + (size_t)c * n] = m < r ? (Cl[m + (size_t)c * r] - mc[m]) - v : -v;
      }
    }
  __syncthreads();
}

typedef double aca_f64x4 __attribute__((ext_vector_type(4)));

// wy_apply for reflectors too long for LDS: A stays in memory; the core, M,
// tau and the Gram matrix live in LDS (r*r + 2*r*kt + r values, r <= 16*NB).
// Y^T Y runs on v_mfma_f64_16x16x4f64, one wavefront per group of 16 rows
// (groups w, w+4, ...): lane l loads rows 4(l/16) .. 4(l/16)+3 of column
// 16b + l%16 of each column block b and feeds row 4(l/16)+s of the group to
// k-step s (A operand: lane l holds (row l%16, k l/16); B: (k l/16, column
// l%16); result value q of lane l: (l/16 + 4q, l%16)).  The four
// wavefronts' tiles are summed in LDS in wavefront order.  The product
// O = [C; 0] - Y M runs one row per thread, eight columns of M at a time.
template <int NB>
__device__ void wy_apply_big(const double* A, uint32_t n, uint32_t r, const double* tau,
                             const double* C, uint32_t kt, double* O, double* lds) {
  constexpr int NT = NB * (NB + 1) / 2;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* G = lds;                      // r x r: first the top r x r of A, then y_i . y_j
  double* S = G + (size_t)r * r;        // r x kt: Y_top^T C, then M
  double* Cl = S + (size_t)r * kt;      // r x kt
  double* tl = Cl + (size_t)r * kt;     // r
  for (uint32_t c = 0; c < r; ++c)
    for (uint32_t m = threadIdx.x; m < r; m += kAcaThreads) G[m + (size_t)c * r] = A[m + (size_t)c * n];
  for (uint32_t e = threadIdx.x; e < r * kt; e += kAcaThreads) Cl[e] = C[e];
  for (uint32_t e = threadIdx.x; e < r; e += kAcaThreads) tl[e] = tau[e];
  __syncthreads();
  {  // S[i + c r] = C[i, c] + sum_{i < m < r} A[m, i] C[m, c]
    const uint32_t per = kAcaThreads / r, sub = threadIdx.x / r, i = threadIdx.x - sub * r;
    if (sub < per)
      for (uint32_t c = sub; c < kt; c += per) {
        const double* cc = Cl + (size_t)c * r;
        const double* ai = G + (size_t)i * r;
        double v = cc[i];
        for (uint32_t m = i + 1; m < r; ++m) v = fma(ai[m], cc[m], v);
        S[i + (size_t)c * r] = v;
      }
  }
  aca_f64x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = aca_f64x4{0.0, 0.0, 0.0, 0.0};
  const uint32_t li = lane & 15, kk = lane >> 4, ng = (n + 15) / 16;
  for (uint32_t g = w; g < ng; g += kWaves) {
    double y[NB][4];
    const uint32_t mb = g * 16 + 4 * kk;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint32_t c = b * 16 + li;
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const uint32_t m = mb + st;
        double v = 0.0;
        if (c < r && m < n) v = m > c ? A[m + (size_t)c * n] : (m == c ? 1.0 : 0.0);
        y[b][st] = v;
      }
    }
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      int t = 0;
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = bi; bj < NB; ++bj, ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[bi][st], y[bj][st], acc[t], 0, 0, 0);
    }
  }
  __syncthreads();  // S done and the top of A no longer read: G is free
  for (int q = 0; q < kWaves; ++q) {
    if (w == q) {
      int t = 0;
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = bi; bj < NB; ++bj, ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const uint32_t i = bi * 16 + kk + 4 * v, j = bj * 16 + li;
            if (i < j && j < r) G[i + (size_t)j * r] = (q == 0 ? 0.0 : G[i + (size_t)j * r]) + acc[t][v];
          }
    }
    __syncthreads();
  }
  // M[i, c] = tau_i (S[i, c] - sum_{l > i} G[i, l] M[l, c]), in place over S
  for (uint32_t c = threadIdx.x; c < kt; c += kAcaThreads) {
    double* sc = S + (size_t)c * r;
    for (int i = (int)r - 1; i >= 0; --i) {
      double v = sc[i];
      for (uint32_t l = (uint32_t)i + 1; l < r; ++l) v = fma(-G[i + (size_t)l * r], sc[l], v);
      sc[i] = tl[i] * v;
    }
  }
  __syncthreads();
  // O[m, c] = [m < r] (C[m, c] - M[m, c]) - sum_{i < min(m, r)} A[m, i] M[i, c]
  for (uint32_t m = threadIdx.x; m < n; m += kAcaThreads) {
    const uint32_t ie = m < r ? m : r;
    for (uint32_t c0 = 0; c0 < kt; c0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = 0.0;
      for (uint32_t i = 0; i < ie; ++i) {
        const double a = A[m + (size_t)i * n];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (c0 + u < kt) v[u] = fma(a, S[i + (size_t)(c0 + u) * r], v[u]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u < kt) {
          const uint32_t c = c0 + u;
          O[m + (size_t)c * n] = m < r ? (Cl[m + (size_t)c * r] - S[m + (size_t)c * r]) - v[u] : -v[u];
        }
    }
  }
  __syncthreads();
}

// One-sided Jacobi SVD of R (mm x nn, ld mm): R <- R Z with orthogonal Z
// (nn x nn, ld nn) until the columns of R are mutually orthogonal.  Column
// pairs are visited in round-robin tournament order, one wavefront per pair.
// A pair counts as orthogonal when |p.q| <= mm * eps * |p| |q| (the rounding
// error of the computed p.q itself) or when either column is below
// eps * |R|_F; without the second test the columns a rank-deficient core
// drives to zero stay rounding noise that is rotated in every sweep (a third
// of the leaves of a 1M-point matrix swept until the sweep limit).
// The same iteration for mm, nn <= 16, on the first wavefront alone: four
// pairs at a time, sixteen lanes a pair (lane i owns row i of the pair's
// columns and of their rotations), sums over the sixteen lanes by a fixed
// xor butterfly; no workgroup barriers inside the sweeps.
__device__ void jacobi_svd16(double* R, uint32_t mm, uint32_t nn, double* Zm, Shared& sh) {
  const double tol = (double)max(mm, 2u) * 2.220446049250313e-16;
  for (uint32_t e = threadIdx.x; e < nn * nn; e += nthr()) Zm[e] = (e % nn == e / nn) ? 1.0 : 0.0;
  __syncthreads();
  if (nn < 2) return;
  double f = 0.0;
  for (uint32_t e = threadIdx.x; e < mm * nn; e += nthr()) f = fma(R[e], R[e], f);
  const double negl = 4.930380657631324e-32 * block_sum(f, sh.dv);   // eps^2 |R|_F^2
  if (threadIdx.x < 64) {
    const uint32_t g = threadIdx.x >> 4, i = threadIdx.x & 15;
    const uint32_t np = (nn + 1) & ~1u;
    auto sum16 = [](double v) {
      v += __shfl_xor(v, 8, 16);
      v += __shfl_xor(v, 4, 16);
      v += __shfl_xor(v, 2, 16);
      v += __shfl_xor(v, 1, 16);
      return v;
    };
    for (int sweep = 0; sweep < 40; ++sweep) {
      bool rotated = false;
      for (uint32_t round = 0; round < np - 1; ++round) {
        for (uint32_t base = 0; base < np / 2; base += 4) {
          const uint32_t pi = base + g;
          auto player = [&](uint32_t s) -> uint32_t {
            return s == 0 ? 0 : 1 + (s - 1 + round) % (np - 1);
          };
          uint32_t p = 0, q = 0;
          bool live = pi < np / 2;
          if (live) {
            p = player(pi);
            q = player(np - 1 - pi);
            live = p < nn && q < nn;
            if (p > q) {
              const uint32_t tt = p;
              p = q;
              q = tt;
            }
          }
          double* ap = R + (size_t)p * mm;
          double* aq = R + (size_t)q * mm;
          const double x = live && i < mm ? ap[i] : 0.0;
          const double y = live && i < mm ? aq[i] : 0.0;
          const double al = sum16(x * x), be = sum16(y * y), ga = sum16(x * y);
          if (!live || fabs(ga) <= tol * sqrt(al * be) || ga == 0.0 || fmin(al, be) <= negl) continue;
          const double zeta = (be - al) / (2.0 * ga);
          const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
          const double c = 1.0 / sqrt(fma(t, t, 1.0)), s = c * t;
          if (i < mm) {
            ap[i] = c * x - s * y;
            aq[i] = s * x + c * y;
          }
          if (i < nn) {
            double* zp = Zm + (size_t)p * nn;
            double* zq = Zm + (size_t)q * nn;
            const double zx = zp[i], zy = zq[i];
            zp[i] = c * zx - s * zy;
            zq[i] = s * zx + c * zy;
          }
          rotated = true;
        }
        // the next round reads rows that other lanes of this wavefront wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (!__any(rotated)) break;
    }
  }
  __syncthreads();
}

__device__ void jacobi_svd(double* R, uint32_t mm, uint32_t nn, double* Zm, Shared& sh) {
  if (mm <= 16 && nn <= 16) {
    jacobi_svd16(R, mm, nn, Zm, sh);
    return;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double tol = (double)max(mm, 2u) * 2.220446049250313e-16;
  for (uint32_t e = threadIdx.x; e < nn * nn; e += nthr()) Zm[e] = (e % nn == e / nn) ? 1.0 : 0.0;
  __syncthreads();
  if (nn < 2) return;
  // columns below eps * |R|_F are rounding noise (the rotations keep |R|_F):
  // they are left alone, or they would be rotated on noise in every sweep
  double f = 0.0;
  for (uint32_t e = threadIdx.x; e < mm * nn; e += nthr()) f = fma(R[e], R[e], f);
  const double negl = 4.930380657631324e-32 * block_sum(f, sh.dv);   // eps^2 |R|_F^2
  const uint32_t np = (nn + 1) & ~1u;  // even number of players (one bye if odd)
  __shared__ int rot;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (threadIdx.x == 0) rot = 0;
    __syncthreads();
    for (uint32_t round = 0; round < np - 1; ++round) {
      for (uint32_t pi = w; pi < np / 2; pi += nwaves()) {
        // tournament pairing: player 0 fixed, others rotate
        auto player = [&](uint32_t s) -> uint32_t {
          return s == 0 ? 0 : 1 + (s - 1 + round) % (np - 1);
        };
        uint32_t p = player(pi), q = player(np - 1 - pi);
        if (p >= nn || q >= nn) continue;
        if (p > q) {
          const uint32_t tt = p;
          p = q;
          q = tt;
        }
        double* ap = R + (size_t)p * mm;
        double* aq = R + (size_t)q * mm;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (uint32_t i = lane; i < mm; i += 64) {
          al = fma(ap[i], ap[i], al);
          be = fma(aq[i], aq[i], be);
          ga = fma(ap[i], aq[i], ga);
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (fabs(ga) <= tol * sqrt(al * be) || ga == 0.0 || fmin(al, be) <= negl) continue;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
        const double c = 1.0 / sqrt(fma(t, t, 1.0)), s = c * t;
        for (uint32_t i = lane; i < mm; i += 64) {
          const double x = ap[i], y = aq[i];
          ap[i] = c * x - s * y;
          aq[i] = s * x + c * y;
        }
        double* zp = Zm + (size_t)p * nn;
        double* zq = Zm + (size_t)q * nn;
        for (uint32_t i = lane; i < nn; i += 64) {
          const double x = zp[i], y = zq[i];
          zp[i] = c * x - s * y;
          zq[i] = s * x + c * y;
        }
        if (lane == 0) rot = 1;
      }
      __syncthreads();
    }
    const int any = rot;
    __syncthreads();
    if (!any) break;
  }
}

// QR of U and V, core R = R1 R2^T, Jacobi SVD, truncation.  Leaves the
// sorted, truncated core factors in the scratch: W = R Z (columns scaled by
// the singular values, mmin x kt, ld mmin) and Z (nmin x kt, ld nmin).  U, V,
// tau1, tau2, R and Zm are the leaf's working arrays, in LDS or in the scratch.
// The core of the recompression from the triangular factors R1 (rows
// 0..mmin-1 of the upper triangle of U, ld ld1) and R2 (of V, ld ld2).
__device__ __forceinline__ void rc_core(const AcaBlock& B, AcaResult* __restrict__ res, uint32_t k, double eps,
                                        const double* U, size_t ld1, uint32_t mmin, const double* V, size_t ld2,
                                        uint32_t nmin, double* R, double* Zm, double* Ws, double* Zs, Shared& sh,
                                        double* sig, int* order) {
  const uint32_t amin = min(mmin, nmin);
  // R = R1 R2^T (utrmmh, Include/blas.h:2767-2781)
  double fs = 0.0;
  for (uint32_t e = threadIdx.x; e < mmin * nmin; e += nthr()) {
    const uint32_t i = e % mmin, j = e / mmin;
    double d = 0.0;
    uint32_t l = max(i, j);
    for (; l + 8 <= k; l += 8) {
      double x[8], y[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        x[u] = U[(size_t)(l + u) * ld1 + i];
        y[u] = V[(size_t)(l + u) * ld2 + j];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) d += x[u] * y[u];
    }
    for (; l < k; ++l) d += U[(size_t)l * ld1 + i] * V[(size_t)l * ld2 + j];
    R[e] = d;
    fs += d * d;
  }
  const double fro = sqrt(block_sum(fs, sh.dv));
  const double thresh = eps * fro / (double)sqrtf((float)mmin * nmin);
  for (uint32_t e = threadIdx.x; e < mmin * nmin; e += nthr())
    if (fabs(R[e]) < thresh) R[e] = 0.0;
  __syncthreads();
  jacobi_svd(R, mmin, nmin, Zm, sh);
  // singular values = column norms; order descending (ties by index)
  for (uint32_t c = threadIdx.x >> 6; c < nmin; c += nwaves()) {
    double s = 0.0;
    for (uint32_t i = threadIdx.x & 63; i < mmin; i += 64) s = fma(R[(size_t)c * mmin + i], R[(size_t)c * mmin + i], s);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sig[c] = sqrt(s);
  }
  __syncthreads();
  for (uint32_t c = threadIdx.x; c < nmin; c += nthr()) {
    uint32_t rank = 0;
    for (uint32_t d = 0; d < nmin; ++d)
      if (sig[d] > sig[c] || (sig[d] == sig[c] && d < c)) ++rank;
    order[rank] = (int)c;
  }
  __syncthreads();
  // kt: the reference keeps S_i > delta*S_0 and >= 1e-64 among the first amin
  uint32_t kt = min(amin, k);
  const double s0 = sig[order[0]];
  while (kt > 0 && (sig[order[kt - 1]] <= eps * s0 || sig[order[kt - 1]] < 1e-64)) --kt;
  // sorted, truncated copies of the core factors
  for (uint32_t e = threadIdx.x; e < mmin * kt; e += nthr()) {
    const uint32_t i = e % mmin, c = e / mmin;
    Ws[e] = R[(size_t)order[c] * mmin + i];
  }
  for (uint32_t e = threadIdx.x; e < nmin * kt; e += nthr()) {
    const uint32_t i = e % nmin, c = e / nmin;
    Zs[e] = Zm[(size_t)order[c] * nmin + i];
  }
  if (threadIdx.x == 0) res[blockIdx.x].kt = kt;
}

// QR of U and V (Householder, in place), then the core: leaves the sorted,
// truncated core factors in the scratch: W = R Z (columns scaled by the
// singular values, mmin x kt, ld mmin) and Z (nmin x kt, ld nmin).
__device__ __forceinline__ void rc_body(const AcaBlock& B, AcaResult* __restrict__ res, uint32_t k, double eps,
                                        double* U, double* V, double* R, double* Zm, double* Ws, double* Zs,
                                        double* tau1, double* tau2, Shared& sh, double* wbuf, double* sig,
                                        int* order) {
  house_qr(U, B.n1, k, tau1, sh, wbuf);
  house_qr(V, B.n2, k, tau2, sh, wbuf);
  rc_core(B, res, k, eps, U, B.n1, min(B.n1, k), V, B.n2, min(B.n2, k), R, Zm, Ws, Zs, sh, sig, order);
}

// ------------------------------------------------------------------------
// Factors of at most kCq columns: Cholesky QR twice (CholeskyQR2; Fukaya,
// Nakatsukasa, Yanagisawa and Yamamoto, 2014).  F (n x k) -> F1 = F R1^{-1}
// with R1^T R1 = F^T F, then R2^T R2 = F1^T F1; Q = F1 R2^{-1} and
// R = R2 R1.  The factor is read three times and written once, against once
// per reflector for the Householder QR; each Gram matrix is one MFMA pass.
// A factor too close to rank deficiency for the Cholesky factor to be
// trusted (a pivot below 1e-12 of its column's squared norm) keeps the
// Householder QR.  The expansion then forms U' = F1 (R2^{-1} W) directly.
constexpr uint32_t kCq = 16;

// G (16 x 16, ld 16) = F^T F for F (n x k, ld n), k <= 16, on
// v_mfma_f64_16x16x4f64: wavefront w takes 16-row groups w, w+4, ...;
// lane l loads rows 4(l/16) .. 4(l/16)+3 of column l%16 of the group and
// feeds row 4(l/16)+s to k-step s (the operand layout of wy_apply_big);
// the wavefronts' tiles are summed in LDS in wavefront order.  part holds
// kWaves x 256 values.
__device__ void cq_gram(const double* F, uint32_t n, uint32_t k, double* G, double* part) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t li = lane & 15, kk = lane >> 4, ng = (n + 15) / 16;
  aca_f64x4 acc = aca_f64x4{0.0, 0.0, 0.0, 0.0};
  for (uint32_t g = w; g < ng; g += nwaves()) {
    const uint32_t mb = g * 16 + 4 * kk;
    double y[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const uint32_t m = mb + st;
      y[st] = (li < k && m < n) ? F[(size_t)li * n + m] : 0.0;
    }
#pragma unroll
    for (int st = 0; st < 4; ++st) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(y[st], y[st], acc, 0, 0, 0);
  }
  // value v of lane l: G[kk + 4v][li] of this wavefront's rows
#pragma unroll
  for (int v = 0; v < 4; ++v) part[w * 256 + (kk + 4 * v) + li * 16] = acc[v];
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < 256; e += nthr()) {
    double t = part[e];
    for (int q = 1; q < nwaves(); ++q) t += part[q * 256 + e];
    G[e] = t;
  }
  __syncthreads();
}

// R (k x k upper, ld 16, zero below) with R^T R = G; false when a pivot is
// not above 1e-12 of its column's diagonal entry of G.  One thread.
__device__ bool cq_chol(const double* G, uint32_t k, double* R) {
  for (uint32_t j = 0; j < k; ++j) {
    for (uint32_t i = j + 1; i < k; ++i) R[i + 16 * j] = 0.0;
    double d = G[j + 16 * j];
    for (uint32_t p = 0; p < j; ++p) d -= R[p + 16 * j] * R[p + 16 * j];
    if (!(d > 1e-12 * G[j + 16 * j])) return false;
    const double r = sqrt(d);
    R[j + 16 * j] = r;
    for (uint32_t i = j + 1; i < k; ++i) {
      double t = G[j + 16 * i];
      for (uint32_t p = 0; p < j; ++p) t -= R[p + 16 * j] * R[p + 16 * i];
      R[j + 16 * i] = t / r;
    }
  }
  return true;
}

// Ri = R^{-1} (k x k upper, ld 16).  One thread.
__device__ void cq_inv(const double* R, uint32_t k, double* Ri) {
  for (uint32_t j = 0; j < k; ++j) {
    for (uint32_t i = j + 1; i < k; ++i) Ri[i + 16 * j] = 0.0;
    Ri[j + 16 * j] = 1.0 / R[j + 16 * j];
    for (int i = (int)j - 1; i >= 0; --i) {
      double t = 0.0;
      for (uint32_t p = (uint32_t)i + 1; p <= j; ++p) t = fma(R[i + 16 * p], Ri[p + 16 * j], t);
      Ri[i + 16 * j] = -t / R[i + 16 * i];
    }
  }
}

// C = A B for upper triangular A, B (k x k, ld 16).  One thread.
__device__ void cq_mul_upper(const double* A, const double* Bm, uint32_t k, double* Cm) {
  for (uint32_t j = 0; j < k; ++j)
    for (uint32_t i = 0; i < k; ++i) {
      double t = 0.0;
      for (uint32_t p = i; p <= j; ++p) t = fma(A[i + 16 * p], Bm[p + 16 * j], t);
      Cm[i + 16 * j] = i <= j ? t : 0.0;
    }
}

// F <- F T in place for upper triangular T (k x k, ld 16 in LDS), one row
// per thread (row c of the result needs entries 0..c of the row: columns
// are rewritten from the last)
__device__ void cq_apply(double* F, uint32_t n, uint32_t k, const double* T) {
  for (uint32_t i = threadIdx.x; i < n; i += nthr()) {
    double f[kCq];
#pragma unroll
    for (uint32_t l = 0; l < kCq; ++l) f[l] = l < k ? F[(size_t)l * n + i] : 0.0;
#pragma unroll
    for (int c = (int)kCq - 1; c >= 0; --c) {
      if ((uint32_t)c >= k) continue;
      double t = 0.0;
#pragma unroll
      for (int l = 0; l <= c; ++l) t = fma(f[l], T[l + 16 * c], t);
      f[c] = t;
    }
#pragma unroll
    for (uint32_t l = 0; l < kCq; ++l)
      if (l < k) F[(size_t)l * n + i] = f[l];
  }
  __syncthreads();
}

// LDS of the CholeskyQR2 recompression (doubles): Gram partials of the
// workgroup's wavefronts, then eight 16 x 16 matrices
constexpr uint32_t kCqLds = kWaves * 256 + 8 * 256;
constexpr uint32_t kCqLdsSmall = 256 + 8 * 256;   // one wavefront

// CholeskyQR2 of U and V and the core; false (U, V as on entry) when either
// factor keeps the Householder QR.  On success U, V hold F1 and Ws, Zs the
// expansion's multipliers R2^{-1} W and R2^{-1} Z (k x kt, ld k).
__device__ bool cq_recompress(const AcaBlock& B, AcaResult* __restrict__ res, uint32_t k, double eps, double* U,
                              double* V, double* Ws, double* Zs, Shared& sh, double* sig, int* order,
                              double* lds) {
  // eight 16 x 16 matrices after the Gram partials, reused as they free up
  double* part = lds;
  double* M = part + (size_t)nwaves() * 256;
  double* G = M;             // Gram matrix, then R = R2 R1 of U
  double* R1u = M + 256;     // then the core
  double* R1v = M + 512;     // then the core's rotations
  double* Iu = M + 768;      // R1^{-1}, then R2^{-1}
  double* Iv = M + 1024;
  double* R2u = M + 1280;    // then scratch of the multipliers
  double* R2v = M + 1536;
  double* Rv = M + 1792;
  double* Ru = G;
  double* Rc = R1u;
  double* Zc = R1v;
  __shared__ int ok;
  // first pass: both Cholesky factors before either factor is changed
  cq_gram(U, B.n1, k, G, part);
  if (threadIdx.x == 0) {
    ok = cq_chol(G, k, R1u);
    if (ok) cq_inv(R1u, k, Iu);
  }
  __syncthreads();
  if (!ok) return false;
  cq_gram(V, B.n2, k, G, part);
  if (threadIdx.x == 0) {
    ok = cq_chol(G, k, R1v);
    if (ok) cq_inv(R1v, k, Iv);
  }
  __syncthreads();
  if (!ok) return false;
  cq_apply(U, B.n1, k, Iu);
  cq_apply(V, B.n2, k, Iv);
  // second pass
  cq_gram(U, B.n1, k, G, part);
  if (threadIdx.x == 0) ok = cq_chol(G, k, R2u);
  __syncthreads();
  if (ok) {
    cq_gram(V, B.n2, k, G, part);
    if (threadIdx.x == 0) ok = cq_chol(G, k, R2v);
    __syncthreads();
  }
  if (!ok) {   // back to the factors as they came (to rounding)
    cq_apply(U, B.n1, k, R1u);
    cq_apply(V, B.n2, k, R1v);
    return false;
  }
  if (threadIdx.x == 0) {
    cq_mul_upper(R2u, R1u, k, Ru);   // G is free: the last Gram is factored
    cq_mul_upper(R2v, R1v, k, Rv);
    cq_inv(R2u, k, Iu);
    cq_inv(R2v, k, Iv);
  }
  __syncthreads();
  rc_core(B, res, k, eps, Ru, 16, k, Rv, 16, k, Rc, Zc, Ws, Zs, sh, sig, order);
  __syncthreads();
  const uint32_t kt = res[blockIdx.x].kt;
  // multipliers M = R2^{-1} W (k x kt, ld k), through LDS
  for (uint32_t e = threadIdx.x; e < k * kt; e += nthr()) {
    const uint32_t i = e % k, c = e / k;
    double tu = 0.0, tv = 0.0;
    for (uint32_t l = i; l < k; ++l) {
      tu = fma(Iu[i + 16 * l], Ws[l + (size_t)k * c], tu);
      tv = fma(Iv[i + 16 * l], Zs[l + (size_t)k * c], tv);
    }
    R2u[e] = tu;
    R2v[e] = tv;
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < k * kt; e += nthr()) {
    Ws[e] = R2u[e];
    Zs[e] = R2v[e];
  }
  if (threadIdx.x == 0) res[blockIdx.x].pad = 1;
  return true;
}


// Doubles of dynamic LDS per workgroup of the recompression and expansion:
// a leaf whose working arrays fit runs on LDS copies, without the round trips
// to memory that every workgroup barrier would otherwise wait for.
constexpr uint32_t kRcLds = 4352;
static_assert(kCqLds <= kRcLds, "CholeskyQR2 working set exceeds the recompression LDS");

// small != 0: one-wavefront workgroups with kCqLdsSmall doubles of LDS
// (CholeskyQR2, else the Householder QR in memory); otherwise kRcLds
__global__ __launch_bounds__(kAcaThreads) void recompress_kernel(const AcaBlock* __restrict__ blocks,
                                                                 AcaResult* __restrict__ res,
                                                                 double eps, double* __restrict__ slab,
                                                                 int small) {
  __shared__ Shared sh;
  __shared__ double wbuf[kAcaCapMax + 1];
  __shared__ double sig[kAcaCapMax + 1];
  __shared__ int order[kAcaCapMax + 1];
  extern __shared__ double rc_lds[];
  const AcaBlock B = blocks[blockIdx.x];
  const AcaResult r0 = res[blockIdx.x];
  if (r0.status != 1) return;
  const uint32_t k = r0.k;
  if (k == 0) return;
  double* U = slab + B.uoff;
  double* V = slab + B.voff;
  const size_t sq = (size_t)(B.kcap + 1) * (B.kcap + 1);
  double* R = slab + B.roff;
  double* Zm = R + sq;
  double* Ws = Zm + sq;
  double* Zs = Ws + sq;
  double* tau1 = Zs + sq;
  double* tau2 = tau1 + (B.kcap + 1);
  const uint32_t mmin = min(B.n1, k), nmin = min(B.n2, k);
  const size_t nu = (size_t)B.n1 * k, nv = (size_t)B.n2 * k;
  const bool fits = !small && nu + nv + 2 * (size_t)k + (size_t)mmin * nmin + (size_t)nmin * nmin <= kRcLds;
  // factors that fit in LDS keep the Householder QR on LDS copies (no
  // memory round trips); longer ones with few columns take CholeskyQR2
  if (!fits && k <= kCq && B.n1 >= 2 * k && B.n2 >= 2 * k &&
      cq_recompress(B, res, k, eps, U, V, Ws, Zs, sh, sig, order, rc_lds))
    return;
  __syncthreads();
  if (fits) {
    double* Ul = rc_lds;
    double* Vl = Ul + nu;
    double* t1 = Vl + nv;
    double* t2 = t1 + k;
    double* Rl = t2 + k;
    double* Zl = Rl + (size_t)mmin * nmin;
    for (uint32_t e = threadIdx.x; e < nu; e += nthr()) Ul[e] = U[e];
    for (uint32_t e = threadIdx.x; e < nv; e += nthr()) Vl[e] = V[e];
    __syncthreads();
    rc_body(B, res, k, eps, Ul, Vl, Rl, Zl, Ws, Zs, t1, t2, sh, wbuf, sig, order);
    __syncthreads();
    // the reflectors and their factors are what the expansion reads
    for (uint32_t e = threadIdx.x; e < nu; e += nthr()) U[e] = Ul[e];
    for (uint32_t e = threadIdx.x; e < nv; e += nthr()) V[e] = Vl[e];
    for (uint32_t e = threadIdx.x; e < mmin; e += nthr()) tau1[e] = t1[e];
    for (uint32_t e = threadIdx.x; e < nmin; e += nthr()) tau2[e] = t2[e];
  } else {
    rc_body(B, res, k, eps, U, V, R, Zm, Ws, Zs, tau1, tau2, sh, wbuf, sig, order);
  }
}

// final U = Q1 [W; 0] (n1 x kt), V = Q2 [Z; 0] (n2 x kt) at the leaf's offset
template <bool BIG>
__global__ __launch_bounds__(kAcaThreads, 4) void expand_kernel(const AcaBlock* __restrict__ blocks,
                                                             const AcaResult* __restrict__ res,
                                                             const uint64_t* __restrict__ out_off,
                                                             const double* __restrict__ slab_c,
                                                             double* __restrict__ out) {
  __shared__ double wbuf[kAcaCapMax + 1];
  const AcaBlock B = blocks[blockIdx.x];
  const AcaResult r0 = res[blockIdx.x];
  if (r0.status != 1 || r0.kt == 0) return;
  const uint32_t k = r0.k, kt = r0.kt;
  if (r0.pad == 1) {
    // CholeskyQR2 leaves: U' = F1_u M_u, V' = F1_v M_v (multipliers k x kt)
    if constexpr (!BIG) {
      const size_t sq = (size_t)(B.kcap + 1) * (B.kcap + 1);
      const double* Mu = slab_c + B.roff + 2 * sq;
      const double* Mv = Mu + sq;
      double* Uo = out + out_off[blockIdx.x];
      double* Vo = Uo + (size_t)B.n1 * kt;
      extern __shared__ double rc_lds[];
      double* Ml = rc_lds;
      for (uint32_t e = threadIdx.x; e < k * kt; e += kAcaThreads) {
        Ml[e] = Mu[e];
        Ml[256 + e] = Mv[e];
      }
      __syncthreads();
      auto gemm = [&](const double* F, uint32_t n, const double* M, double* O) {
        for (uint32_t i = threadIdx.x; i < n; i += kAcaThreads) {
          double f[kCq];
#pragma unroll
          for (uint32_t l = 0; l < kCq; ++l) f[l] = l < k ? F[(size_t)l * n + i] : 0.0;
          for (uint32_t c = 0; c < kt; ++c) {
            double t = 0.0;
#pragma unroll
            for (uint32_t l = 0; l < kCq; ++l)
              if (l < k) t = fma(f[l], M[l + (size_t)k * c], t);
            O[(size_t)c * n + i] = t;
          }
        }
      };
      gemm(slab_c + B.uoff, B.n1, Ml, Uo);
      gemm(slab_c + B.voff, B.n2, Ml + 256, Vo);
    }
    return;
  }
  const uint32_t mmin = min(B.n1, k), nmin = min(B.n2, k);
  const double* U = slab_c + B.uoff;
  const double* V = slab_c + B.voff;
  const size_t sq = (size_t)(B.kcap + 1) * (B.kcap + 1);
  const double* Ws = slab_c + B.roff + 2 * sq;
  const double* Zs = Ws + sq;
  const double* tau1 = Zs + sq;
  const double* tau2 = tau1 + (B.kcap + 1);
  double* Uo = out + out_off[blockIdx.x];
  double* Vo = Uo + (size_t)B.n1 * kt;
  extern __shared__ double rc_lds[];
  // one factor: Q [core; 0] with Q's reflectors in A (n x r) and the core in C
  // (r x kt, ld r), into O (n x kt); on LDS copies when they fit
  auto factor = [&](const double* A, uint32_t n, uint32_t r, const double* tau, const double* C, double* O) {
    const size_t na = (size_t)n * r, no = (size_t)n * kt;
    const size_t nsm = (size_t)r * r + 2 * (size_t)r * kt + r;
    const bool big = na + nsm > kRcLds && nsm <= kRcLds && r <= 48;
    if (big != BIG) return;  // the other instance expands this factor
    if constexpr (BIG) {
      if (r <= 16) wy_apply_big<1>(A, n, r, tau, C, kt, O, rc_lds);
      else if (r <= 32) wy_apply_big<2>(A, n, r, tau, C, kt, O, rc_lds);
      else wy_apply_big<3>(A, n, r, tau, C, kt, O, rc_lds);
      return;
    } else if (na + nsm <= kRcLds) {
      wy_apply(A, n, r, tau, C, kt, O, rc_lds);
      return;
    }
    double* X = O;
    const double* Ar = A;
    if (na + no <= kRcLds) {
      double* Al = rc_lds;
      for (uint32_t e = threadIdx.x; e < na; e += kAcaThreads) Al[e] = A[e];
      Ar = Al;
      X = Al + na;
    }
    for (uint32_t c = 0; c < kt; ++c)
      for (uint32_t i = threadIdx.x; i < n; i += kAcaThreads) X[(size_t)c * n + i] = i < r ? C[(size_t)c * r + i] : 0.0;
    __syncthreads();
    apply_q(Ar, n, r, tau, X, kt, wbuf);   // ends with a barrier
    if (X != O) {
      for (uint32_t e = threadIdx.x; e < no; e += kAcaThreads) O[e] = X[e];
      __syncthreads();
    }
  };
  factor(U, B.n1, mmin, tau1, Ws, Uo);
  factor(V, B.n2, nmin, tau2, Zs, Vo);
}

// dense leaves (and failed ACA): cmpbl of the whole block, column-major;
// packed diagonal blocks of symmetric matrices: cmpblsym, the upper triangle
// column by column, (i, j) at i + j(j+1)/2 (Include/matgen_omp.h:30-32)
__global__ void dense_fill_kernel(const DenseJob* __restrict__ jobs, const double* __restrict__ P,
                                  const KernelSpec* __restrict__ Ks, double* __restrict__ out) {
  const DenseJob J = jobs[blockIdx.x];
  const KernelSpec K = Ks[J.kidx];
  double* D = out + J.off;
  if (J.packed) {
    for (uint32_t j = 0; j < J.n2; ++j) {
      double* c = D + (uint64_t)j * (j + 1) / 2;
      for (uint32_t i = threadIdx.x; i <= j; i += blockDim.x) c[i] = kentry(K, P, J.b1 + i, J.b2 + j);
    }
    return;
  }
  // column by column (no per-element division); a workgroup covers
  // blockDim.x / n1 columns at a time when a column is shorter than it
  const uint32_t per = J.n1 >= blockDim.x ? 1u : blockDim.x / J.n1;
  const uint32_t sub = threadIdx.x / (J.n1 >= blockDim.x ? blockDim.x : J.n1);
  const uint32_t i0 = threadIdx.x - sub * (J.n1 >= blockDim.x ? blockDim.x : J.n1);
  const uint32_t step = J.n1 >= blockDim.x ? blockDim.x : J.n1;
  if (sub >= per) return;
  for (uint32_t j = sub; j < J.n2; j += per)
    for (uint32_t i = i0; i < J.n1; i += step) D[(uint64_t)j * J.n1 + i] = kentry(K, P, J.b1 + i, J.b2 + j);
}

}  // namespace ahm

// ------------------------------------------------------------------------
// host driver

namespace ahm {

namespace {

template <typename T>
int to_device(DeviceBuf& b, const std::vector<T>& v, hipStream_t s) {
  AHM_TRY(b.alloc(v.size() * sizeof(T)));
  if (!v.empty()) AHM_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return AHM_OK;
}

struct Part {
  DeviceBuf buf;
  uint64_t len = 0;
};

// device scratch reused across batches: reallocated only when a batch needs
// more (allocating and freeing tens of GB per batch costs more than the
// batch's kernels)
struct Grow {
  DeviceBuf buf;
  size_t cap = 0;
  int need(size_t bytes) {
    if (bytes <= cap && buf.p) return AHM_OK;
    AHM_TRY(buf.alloc(std::max<size_t>(bytes, 1)));
    cap = std::max<size_t>(bytes, 1);
    return AHM_OK;
  }
  template <typename T>
  int put(const std::vector<T>& v, hipStream_t s) {
    AHM_TRY(need(v.size() * sizeof(T)));
    if (!v.empty()) AHM_HIP(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return AHM_OK;
  }
};

}  // namespace

KernelSpec make_kernel(const ahm_kernel& k) {
  KernelSpec K;
  K.kind = k.kind;
  // apprx_sym (Include/apprx.h:87-131) approximates with ACAr
  K.sym = (k.flags & AHM_ASSEMBLE_SYM) ? 1 : 0;
  K.rows_only = (k.flags & (AHM_ASSEMBLE_ACAR | AHM_ASSEMBLE_SYM)) ? 1 : 0;
  K.pad = 0;
  K.kappa = k.kappa;
  K.c = k.kind == AHM_KERNEL_EXPR ? k.weight : k.weight * 0.079577471545947667884;
  return K;
}

int assemble_leaves(ahm_ctx ctx, const double* pts, const uint32_t* op_perm, uint32_t N,
                    uint32_t nl, const uint32_t* b1, const uint32_t* n1, const uint32_t* b2,
                    const uint32_t* n2, const uint8_t* adm, const uint32_t* i0,
                    const KernelSpec* Ks, uint32_t nk, const uint32_t* leaf_kernel,
                    double eps, uint32_t rankmax, const uint32_t* kcap_hint,
                    std::vector<ahm_leaf_desc>& leaves, DeviceBuf& data,
                    ahm_assembly_stats* stats, uint32_t* aca_rank_out, std::shared_ptr<PlanJob>* plan) {
  hipStream_t s = ctx->stream;
  PhaseTimer timer("assembly");
  if (nk == 0) return fail(AHM_ERR_INVALID, "no kernel");
  for (uint32_t l = 0; leaf_kernel && l < nl; ++l)
    if (leaf_kernel[l] >= nk) return fail(AHM_ERR_INVALID, "leaf %u: kernel %u of %u", l, leaf_kernel[l], nk);
  DeviceBuf dK;
  {
    std::vector<KernelSpec> kv(Ks, Ks + nk);
    AHM_TRY(to_device(dK, kv, s));
  }
  std::vector<double> P(3 * (size_t)N);
  for (uint32_t i = 0; i < N; ++i) {
    if (op_perm[i] >= N) return fail(AHM_ERR_INVALID, "op_perm[%u] = %u out of range", i, op_perm[i]);
    for (int j = 0; j < 3; ++j) P[3 * (size_t)i + j] = pts[3 * (size_t)op_perm[i] + j];
  }
  DeviceBuf dP;
  AHM_TRY(to_device(dP, P, s));
  timer.mark("set-up: kernels and permuted points");
  // ACA candidates and rank limits (Include/apprx.h:143-151).  The product
  // n1*n2 is taken in 64 bits (the reference's 32-bit product wraps for
  // leaves with n1*n2 >= 2^32).
  std::vector<uint32_t> cand, maxk(nl, 0);
  for (uint32_t l = 0; l < nl; ++l) {
    if (n1[l] == 0 || n2[l] == 0 || (uint64_t)b1[l] + n1[l] > N || (uint64_t)b2[l] + n2[l] > N ||
        i0[l] >= n1[l])
      return fail(AHM_ERR_INVALID, "leaf %u (%u,%u,%u,%u,i0=%u) invalid for N=%u", l, b1[l], n1[l],
                  b2[l], n2[l], i0[l], N);
    uint64_t mk = (uint64_t)n1[l] * n2[l] / ((uint64_t)n1[l] + n2[l]) + 4;
    if (mk > rankmax) mk = rankmax;
    maxk[l] = (uint32_t)mk;
    if (adm[l] && mk) cand.push_back(l);
  }
  std::vector<int32_t> rank(nl, -1);
  std::vector<uint32_t> kaca(nl, 0);
  std::vector<uint32_t> cap(cand.size());
  for (size_t c = 0; c < cand.size(); ++c) {
    const uint32_t l = cand[c];
    const uint32_t lim = std::min(maxk[l], kAcaCapMax);
    const uint32_t h = kcap_hint ? std::max<uint32_t>(kcap_hint[l], 1) : kAcaCapDefault;
    cap[c] = std::min(h, lim);
  }
  timer.mark("set-up: candidates and rank caps");
  uint64_t retries = 0, slab_total = 0, slab_peak = 0;
  std::vector<size_t> todo(cand.size());
  std::iota(todo.begin(), todo.end(), 0);
  // bytes of ACA scratch per batch: a quarter of the free device memory,
  // between 4 and 96 GiB.  Every batch lasts about as long as its largest
  // leaves take, so fewer, larger batches finish sooner.
  uint64_t budget = (uint64_t)24 << 30;
  {
    size_t free_b = 0, total_b = 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b) {
      free_b += cache_bytes(dev);   // blocks the allocation cache holds are free to reuse
      // in steps of 16 GiB, so that the next assembly asks for the same size
      // and the allocation cache can hand it this one's block
      const uint64_t q = (uint64_t)16 << 30;
      budget = std::min<uint64_t>(std::max<uint64_t>(free_b / 4 / q * q, (uint64_t)4 << 30), (uint64_t)96 << 30);
    }
  }
  std::vector<Part*> parts;
  std::vector<uint32_t> part_of(nl, 0);
  std::vector<uint64_t> off_in_part(nl, 0);
  auto cleanup = [&]() {
    for (auto* p : parts) delete p;
    parts.clear();
  };
  Grow gblk, gres, gslab, gscr, goff;
  // second stream of the middle and small leaves and its ordering events
  struct Aux {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    ~Aux() {
      if (s) (void)hipStreamSynchronize(s);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (e2) (void)hipEventDestroy(e2);
      if (s) (void)hipStreamDestroy(s);
    }
  } aux;
  AHM_HIP(hipEventCreateWithFlags(&aux.e0, hipEventDisableTiming));
  AHM_HIP(hipEventCreateWithFlags(&aux.e1, hipEventDisableTiming));
  AHM_HIP(hipEventCreateWithFlags(&aux.e2, hipEventDisableTiming));
  // the slab scratch is sized once, for the largest batch there can be (the
  // budget, or all first attempts if they take less): growing it batch by
  // batch re-allocates tens of GiB, which is slow right after large frees
  {
    uint64_t all = 0;
    for (size_t c = 0; c < cand.size(); ++c) {
      const uint32_t l = cand[c];
      const uint64_t kc = cap[c], sq = (kc + 1) * (kc + 1);
      all += (kc + 1) * ((uint64_t)n1[l] + n2[l]) + 4 * sq + 2 * (kc + 1);
    }
    if (all && gslab.need(std::min<uint64_t>(all * 8, budget))) return fail(AHM_ERR_HIP, "ACA scratch");
    // largest slabs first: the leaves most likely to need a rerun go early,
    // so their reruns share the following batches with many small first
    // attempts (a stable counting sort: caps are at most kAcaCapMax)
    std::vector<size_t> first(kAcaCapMax + 2, 0);
    for (size_t c = 0; c < cand.size(); ++c) first[kAcaCapMax - std::min(cap[c], kAcaCapMax) + 1]++;
    for (uint32_t q = 0; q <= kAcaCapMax; ++q) first[q + 1] += first[q];
    for (size_t c = 0; c < cand.size(); ++c) todo[first[kAcaCapMax - std::min(cap[c], kAcaCapMax)]++] = c;
  }
  timer.mark("set-up (points, candidates)");
  const auto t_aca0 = std::chrono::steady_clock::now();
  // one pass over a growing list: a leaf whose slab was too small goes back
  // into the list, just ahead of the next batch, with a larger slab, so reruns
  // share batches with first attempts instead of forming a thin pass of their
  // own
  {
    size_t pos = 0;
    while (pos < todo.size()) {
      std::vector<AcaBlock> blk;
      uint64_t slab = 0, iscr = 0;
      size_t end = pos;
      auto slab_need = [&](size_t c) {
        const uint32_t l = cand[c];
        const uint64_t kc = cap[c], sq = (kc + 1) * (kc + 1);
        return (kc + 1) * ((uint64_t)n1[l] + n2[l]) + 4 * sq + 2 * (kc + 1);
      };
      for (uint64_t acc = 0; end < todo.size(); ++end) {
        const uint64_t need = slab_need(todo[end]);
        if (end > pos && (acc + need) * 8 > budget) break;
        acc += need;
      }
      // the batch's leaves by size class (large, middle, small), one launch
      // of the ACA kernel per class
      auto size_of = [&](size_t c) { return (uint64_t)n1[cand[c]] + n2[cand[c]]; };
      const size_t nlarge = (size_t)(std::stable_partition(todo.begin() + pos, todo.begin() + end,
                                                           [&](size_t c) { return size_of(c) > kLargeLeaf; }) -
                                     (todo.begin() + pos));
      const size_t nbig = nlarge + (size_t)(std::stable_partition(todo.begin() + pos + nlarge, todo.begin() + end,
                                                                  [&](size_t c) { return size_of(c) > kSmallLeaf; }) -
                                            (todo.begin() + pos + nlarge));
      for (size_t e = pos; e < end; ++e) {
        const size_t c = todo[e];
        const uint32_t l = cand[c];
        const uint64_t kc = cap[c];
        const uint64_t need = slab_need(c);
        AcaBlock B{};
        B.b1 = b1[l];
        B.n1 = n1[l];
        B.b2 = b2[l];
        B.n2 = n2[l];
        B.i0 = i0[l];
        B.kmax = maxk[l];
        B.kcap = (uint32_t)kc;
        B.kidx = leaf_kernel ? leaf_kernel[l] : 0;
        B.uoff = slab;
        B.voff = slab + (kc + 1) * n1[l];
        B.roff = B.voff + (kc + 1) * n2[l];
        B.zoff = iscr;
        slab += need;
        iscr += (uint64_t)n1[l] + n2[l];
        blk.push_back(B);
      }
      slab_total += slab * 8;
      slab_peak = std::max<uint64_t>(slab_peak, slab * 8);
      // the previous batch's kernels finished before its results were read,
      // so the scratch can be reused
      int st = gblk.put(blk, s);
      if (!st) st = gres.need(blk.size() * sizeof(AcaResult));
      if (!st) st = gslab.need(slab * 8);
      if (!st) st = gscr.need(iscr * 4);
      DeviceBuf& dblk = gblk.buf;
      DeviceBuf& dres = gres.buf;
      DeviceBuf& dslab = gslab.buf;
      DeviceBuf& discr = gscr.buf;
      if (st) {
        cleanup();
        return st;
      }
      const uint32_t nb = (uint32_t)blk.size();
      // the large leaves run on the context's stream, the middle and small
      // ones after each other on a second stream beside them, so that they
      // fill the CUs the large leaves' tail leaves idle
      const bool two = nlarge && nb > nlarge;
      if (two) {
        if (!aux.s && hipStreamCreateWithFlags(&aux.s, hipStreamNonBlocking) != hipSuccess) aux.s = nullptr;
        if (aux.s && (hipEventRecord(aux.e0, s) != hipSuccess || hipStreamWaitEvent(aux.s, aux.e0, 0) != hipSuccess)) {
          cleanup();
          return fail(AHM_ERR_HIP, "ACA stream ordering");
        }
      }
      hipStream_t s_small = two && aux.s ? aux.s : s;
      auto launch = [&](size_t b0, size_t b1, int nt, hipStream_t st) {
        if (b1 <= b0) return;
        const size_t lds = kFuseK * (size_t)nt * sizeof(double);
        if (nt > kAcaThreads)
          hipLaunchKernelGGL(aca_kernel<kLargeThreads>, dim3((uint32_t)(b1 - b0)), dim3(nt), lds, st,
                             (const AcaBlock*)dblk.p + b0, (AcaResult*)dres.p + b0, (const double*)dP.p,
                             (const KernelSpec*)dK.p, eps, (double*)dslab.p, (int*)discr.p);
        else
          hipLaunchKernelGGL(aca_kernel<kAcaThreads>, dim3((uint32_t)(b1 - b0)), dim3(nt), lds, st,
                             (const AcaBlock*)dblk.p + b0, (AcaResult*)dres.p + b0, (const double*)dP.p,
                             (const KernelSpec*)dK.p, eps, (double*)dslab.p, (int*)discr.p);
      };
      launch(0, nlarge, kLargeThreads, s);
      launch(nlarge, nbig, kAcaThreads, s_small);
      // the recompression of the large and middle leaves waits for both
      if (s_small != s &&
          (hipEventRecord(aux.e1, s_small) != hipSuccess || hipStreamWaitEvent(s, aux.e1, 0) != hipSuccess)) {
        cleanup();
        return fail(AHM_ERR_HIP, "ACA stream ordering");
      }
      launch(nbig, nb, kSmallThreads, s_small);
      if (nbig)
        hipLaunchKernelGGL(recompress_kernel, dim3((uint32_t)nbig), dim3(kAcaThreads), kRcLds * sizeof(double), s,
                           (const AcaBlock*)dblk.p, (AcaResult*)dres.p, eps, (double*)dslab.p, 0);
      if (nb > nbig)
        hipLaunchKernelGGL(recompress_kernel, dim3(nb - (uint32_t)nbig), dim3(kSmallThreads),
                           kCqLdsSmall * sizeof(double), s_small, (const AcaBlock*)dblk.p + nbig,
                           (AcaResult*)dres.p + nbig, eps, (double*)dslab.p, 1);
      if (s_small != s &&
          (hipEventRecord(aux.e2, s_small) != hipSuccess || hipStreamWaitEvent(s, aux.e2, 0) != hipSuccess)) {
        cleanup();
        return fail(AHM_ERR_HIP, "ACA stream ordering");
      }
      std::vector<AcaResult> rr(nb);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(rr.data(), dres.p, nb * sizeof(AcaResult), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        return fail(AHM_ERR_HIP, "ACA / recompression kernels failed");
      }
      std::vector<uint64_t> off(nb, 0);
      uint64_t tot = 0;
      std::vector<size_t> rerun;
      for (uint32_t q = 0; q < nb; ++q) {
        const size_t c = todo[pos + q];
        const uint32_t l = cand[c];
        if (rr[q].status == 2) {
          // slab too small: ACA continues with a larger slab
          const uint32_t lim = std::min(maxk[l], kAcaCapMax);
          if (cap[c] >= lim) {
            rank[l] = -1;  // rank beyond the device ceiling: dense leaf
            continue;
          }
          cap[c] = std::min(std::max(cap[c] * 4, cap[c] + 1), lim);
          rerun.push_back(c);
          ++retries;
          continue;
        }
        kaca[l] = rr[q].k;
        if (rr[q].status == 1) {
          rank[l] = (int32_t)rr[q].kt;
          off[q] = tot;
          off_in_part[l] = tot;
          part_of[l] = (uint32_t)parts.size();
          tot += (uint64_t)rr[q].kt * ((uint64_t)n1[l] + n2[l]);
        }
      }
      Part* part = new Part;
      parts.push_back(part);
      part->len = tot;
      if (tot) {
        DeviceBuf& doff = goff.buf;
        // sizes of 1 GiB and more in steps of 4 GiB: blocks of one assembly's
        // batches fit the next assembly's (the allocation cache reuses a block
        // for requests of at least half its size)
        const uint64_t g4 = (uint64_t)4 << 30;
        st = part->buf.alloc(tot * 8 >= ((uint64_t)1 << 30) ? (tot * 8 + g4 - 1) / g4 * g4 : tot * 8);
        if (!st) st = goff.put(off, s);
        if (st) {
          cleanup();
          return st;
        }
        // factors whose reflectors outgrow LDS but whose small matrices fit
        // run in a second instance (MFMA Gram, more registers)
        hipLaunchKernelGGL(expand_kernel<false>, dim3(nb), dim3(kAcaThreads), kRcLds * sizeof(double), s,
                           (const AcaBlock*)dblk.p,
                           (const AcaResult*)dres.p, (const uint64_t*)doff.p,
                           (const double*)dslab.p, (double*)part->buf.p);
        hipLaunchKernelGGL(expand_kernel<true>, dim3(nb), dim3(kAcaThreads), kRcLds * sizeof(double), s,
                           (const AcaBlock*)dblk.p,
                           (const AcaResult*)dres.p, (const uint64_t*)doff.p,
                           (const double*)dslab.p, (double*)part->buf.p);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
          cleanup();
          return fail(AHM_ERR_HIP, "expand kernel failed");
        }
      }
      pos = end;
      // reruns join the next batch
      todo.insert(todo.begin() + (ptrdiff_t)pos, rerun.begin(), rerun.end());
    }
  }
  // the batches' small scratch goes back now; the slab scratch is kept for
  // the leaf buffer below (a fresh allocation of tens of GB costs seconds on
  // MI355X right after large frees)
  for (Grow* g : {&gblk, &gres, &gscr, &goff}) {
    g->buf.release();
    g->cap = 0;
  }
  timer.mark("ACA batches (kernels, result copies, slab decisions)");
  const double aca_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_aca0).count();
  // final buffer: the batch parts back to back, then the dense leaves
  std::vector<uint64_t> part_base(parts.size(), 0);
  uint64_t total = 0;
  for (size_t q = 0; q < parts.size(); ++q) {
    part_base[q] = total;
    total += parts[q]->len;
  }
  leaves.resize(nl);
  std::vector<DenseJob> dj;
  for (uint32_t l = 0; l < nl; ++l) {
    ahm_leaf_desc& d = leaves[l];
    d.b1 = b1[l];
    d.n1 = n1[l];
    d.b2 = b2[l];
    d.n2 = n2[l];
    d.rank = rank[l];
    d.flags = 0;
    const uint32_t kx = leaf_kernel ? leaf_kernel[l] : 0u;
    if (rank[l] > 0) d.offset = part_base[part_of[l]] + off_in_part[l];
    else if (rank[l] == 0) d.offset = 0;
    else if (Ks[kx].sym && b1[l] == b2[l] && n1[l] == n2[l]) {
      // diagonal leaf of a symmetric matrix: mblock::setHeM + cmpblsym
      d.offset = total;
      d.flags = AHM_LEAF_HEM;
      dj.push_back(DenseJob{b1[l], n1[l], b2[l], n2[l], total, kx, 1u});
      total += (uint64_t)n1[l] * (n1[l] + 1) / 2;
    } else {
      d.offset = total;
      dj.push_back(DenseJob{b1[l], n1[l], b2[l], n2[l], total, kx, 0u});
      total += (uint64_t)n1[l] * n2[l];
    }
  }
  timer.mark("leaf list");
  // the layout plan needs the leaf list only: it is built on a host thread
  // while the leaf buffer below is copied and the dense leaves are filled
  if (plan) *plan = plan_start(N, N, leaves);
  int st = AHM_OK;
  if (gslab.buf.p && gslab.buf.cap >= total * 8 && total) {
    // the slab scratch (no longer needed) becomes the leaf buffer
    data.release();
    data.p = gslab.buf.p;
    data.bytes = total * 8;
    data.cap = gslab.buf.cap;
    data.dev = gslab.buf.dev;
    gslab.buf.p = nullptr;
    gslab.buf.bytes = gslab.buf.cap = 0;
    gslab.cap = 0;
  } else {
    gslab.buf.release();
    gslab.cap = 0;
    st = data.alloc(total * 8);
  }
  timer.mark("leaf buffer allocation");
  if (st) {
    cleanup();
    return st;
  }
  for (size_t q = 0; q < parts.size(); ++q)
    if (parts[q]->len &&
        hipMemcpyAsync((double*)data.p + part_base[q], parts[q]->buf.p, parts[q]->len * 8,
                       hipMemcpyDeviceToDevice, s) != hipSuccess) {
      cleanup();
      return fail(AHM_ERR_HIP, "copy of assembled leaves failed");
    }
  DeviceBuf ddj;
  st = to_device(ddj, dj, s);
  if (st) {
    cleanup();
    return st;
  }
  for (size_t j0 = 0; j0 < dj.size(); j0 += (1u << 30)) {
    const uint32_t nj = (uint32_t)std::min<size_t>(1u << 30, dj.size() - j0);
    hipLaunchKernelGGL(dense_fill_kernel, dim3(nj), dim3(256), 0, s, (const DenseJob*)ddj.p + j0,
                       (const double*)dP.p, (const KernelSpec*)dK.p, (double*)data.p);
  }
  const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  cleanup();
  if (!ok) return fail(AHM_ERR_HIP, "dense leaf kernel failed");
  timer.mark("leaf buffer (low-rank parts, dense leaves)");
  if (aca_rank_out)
    for (uint32_t l = 0; l < nl; ++l) aca_rank_out[l] = kaca[l];
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->n_candidates = cand.size();
    stats->n_retries = retries;
    stats->slab_bytes = slab_total;
    stats->peak_slab_bytes = slab_peak;
    stats->aca_ms = aca_ms;
    stats->n_dense = dj.size();
    for (uint32_t l = 0; l < nl; ++l)
      if (rank[l] >= 0) {
        stats->n_lowrank++;
        stats->sum_rank += (uint64_t)rank[l];
        stats->sum_aca_rank += kaca[l];
      }
  }
  return AHM_OK;
}

}  // namespace ahm

extern "C" int ahm_assemble(ahm_ctx ctx, const double* pts, const uint32_t* op_perm, uint32_t N,
                            uint32_t n_leaves, const uint32_t* b1, const uint32_t* n1,
                            const uint32_t* b2, const uint32_t* n2, const uint8_t* adm,
                            const uint32_t* i0, const ahm_kernel* kernel, double eps,
                            uint32_t rankmax, const uint32_t* rank_hint, ahm_dtype storage,
                            int keep_leaves, ahm_assembly_stats* stats, uint32_t* aca_rank,
                            ahm_hmat* out) {
  using namespace ahm;
  if (!ctx || !pts || !op_perm || !kernel || !out || !b1 || !n1 || !b2 || !n2 || !adm || !i0)
    return fail(AHM_ERR_INVALID, "ahm_assemble: NULL argument");
  if (kernel->kind < AHM_KERNEL_LAPLACE || kernel->kind > AHM_KERNEL_EXPR)
    return fail(AHM_ERR_INVALID, "ahm_assemble: unknown kernel %d", kernel->kind);
  if (!(eps > 0.0)) return fail(AHM_ERR_INVALID, "ahm_assemble: eps must be > 0");
  AHM_HIP(hipSetDevice(ctx->device));
  std::vector<ahm_leaf_desc> leaves;
  DeviceBuf data;
  std::shared_ptr<PlanJob> plan;   // after leaves: destroyed (its thread joined) first
  const KernelSpec K = make_kernel(*kernel);
  AHM_TRY(assemble_leaves(ctx, pts, op_perm, N, n_leaves, b1, n1, b2, n2, adm, i0, &K, 1, nullptr,
                          eps, rankmax, rank_hint, leaves, data, stats, aca_rank, &plan));
  return hmat_from_device(ctx, N, N, std::move(leaves), data, storage, keep_leaves, out, plan);
}
