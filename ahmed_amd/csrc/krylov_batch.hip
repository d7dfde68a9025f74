// krylov_batch.hip -- GMRES(m) on many independent systems at once.
//
// The systems are the diagonal blocks of a batched H-matrix
// (ahm_assemble_batch).  Each keeps its own Krylov space, Hessenberg matrix,
// rotations and convergence test (solvers/GMRES.cpp:80-181 per system); the
// Arnoldi steps run in lockstep so that one H-matvec launch per phase serves
// every system, and the orthogonalisation uses segmented reductions (each
// reduction chunk belongs to one system; partials are summed per system in a
// fixed order).  A converged system is frozen: its x is final and later
// steps do not touch it.
//
// Arnoldi step i (zero w, H-matvec, i+2 MGS / normalisation launches, copy
// of the new Hessenberg column to pinned host memory) is captured once per i
// into a hipGraph on an internal stream and replayed in every later restart
// cycle.  Host work per step is O(systems * m) (rotations, convergence).
#include <cmath>
#include <cstring>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

constexpr uint32_t kSegChunk = 4096;
constexpr int kSegThreads = 256;

struct SegChunk {
  uint32_t seg, begin, end, pad;
};

__device__ __forceinline__ double seg_wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ double seg_group_sum(double v) {
  __shared__ double sh[kSegThreads / 64];
  v = seg_wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < kSegThreads / 64; ++q) s += sh[q];
  return s;
}

// sum of the partials of system s (chunks cb[s] .. cb[s+1]), fixed order
__device__ double seg_sum(const double* __restrict__ part, const uint32_t* __restrict__ cb,
                          uint32_t s) {
  double v = 0.0;
  for (uint32_t c = cb[s] + threadIdx.x; c < cb[s + 1]; c += kSegThreads) v += part[c];
  return seg_group_sum(v);
}

__global__ __launch_bounds__(kSegThreads) void sdot_kernel(const double* __restrict__ a,
                                                           const double* __restrict__ b,
                                                           const SegChunk* __restrict__ ch,
                                                           double* __restrict__ part) {
  const SegChunk C = ch[blockIdx.x];
  double s = 0.0;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kSegThreads) s = fma(a[i], b[i], s);
  s = seg_group_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kSegThreads) void smgs_kernel(
    double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ next,
    const SegChunk* __restrict__ ch, const uint32_t* __restrict__ cb,
    const double* __restrict__ part_in, double* __restrict__ hcol, double* __restrict__ part_out) {
  const SegChunk C = ch[blockIdx.x];
  const double h = seg_sum(part_in, cb, C.seg);
  if (blockIdx.x == cb[C.seg] && threadIdx.x == 0) hcol[C.seg] = h;
  double s = 0.0;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kSegThreads) {
    const double wi = fma(-h, v[i], w[i]);
    w[i] = wi;
    s = fma(wi, next ? next[i] : wi, s);
  }
  s = seg_group_sum(s);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kSegThreads) void snormalize_kernel(
    double* __restrict__ w, const SegChunk* __restrict__ ch, const uint32_t* __restrict__ cb,
    const double* __restrict__ part, double* __restrict__ hcol) {
  const SegChunk C = ch[blockIdx.x];
  const double h = sqrt(seg_sum(part, cb, C.seg));
  if (blockIdx.x == cb[C.seg] && threadIdx.x == 0) hcol[C.seg] = h;
  if (h == 0.0) return;  // breakdown: w is zero already
  const double inv = 1.0 / h;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += kSegThreads) w[i] *= inv;
}

// one workgroup per system
__global__ __launch_bounds__(kSegThreads) void sfinish_kernel(const double* __restrict__ part,
                                                              const uint32_t* __restrict__ cb,
                                                              double* __restrict__ out) {
  const double h = seg_sum(part, cb, blockIdx.x);
  if (threadIdx.x == 0) out[blockIdx.x] = h;
}

__global__ void sscale_kernel(const double* __restrict__ r, double* __restrict__ v,
                              const SegChunk* __restrict__ ch, const double* __restrict__ scale) {
  const SegChunk C = ch[blockIdx.x];
  const double a = scale[C.seg];
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += blockDim.x) v[i] = a * r[i];
}

__global__ void sfill_kernel(double* __restrict__ d, const SegChunk* __restrict__ ch,
                             const uint8_t* __restrict__ mask) {
  const SegChunk C = ch[blockIdx.x];
  if (!mask[C.seg]) return;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += blockDim.x) d[i] = 0.0;
}

__global__ void zero_kernel(double* __restrict__ d, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = 0.0;
}

// x_s += V_s y_s with k_s columns (k_s = 0: no change)
__global__ void supdate_kernel(const double* __restrict__ V, uint32_t n, uint32_t m,
                               const SegChunk* __restrict__ ch, const uint32_t* __restrict__ kseg,
                               const double* __restrict__ Y, double* __restrict__ x) {
  const SegChunk C = ch[blockIdx.x];
  const uint32_t k = kseg[C.seg];
  if (k == 0) return;
  const double* y = Y + (size_t)C.seg * m;
  for (uint32_t i = C.begin + threadIdx.x; i < C.end; i += blockDim.x) {
    double s = 0.0;
    for (uint32_t l = 0; l < k; ++l) s = fma(V[(uint64_t)l * n + i], y[l], s);
    x[i] += s;
  }
}

namespace {

void grot(double dx, double dy, double& cs, double& sn) {
  if (dy == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (std::fabs(dy) > std::fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / std::sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / std::sqrt(1.0 + t * t);
    sn = t * cs;
  }
}
void arot(double& dx, double& dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = cs * dy - sn * dx;
  dx = t;
}

struct Batch {
  ahm_hmat h = nullptr;
  uint32_t S = 0, n = 0, m = 0, nch = 0;
  hipStream_t gs = nullptr;
  DeviceBuf dch, dcb, dV, dr, dp0, dp1, dH, dY, dk, dscale, dmask, dnorm;
  double* hH = nullptr;        // pinned: (m+2) x S Hessenberg column
  std::vector<hipGraphExec_t> graphs;
  ~Batch() {
    for (auto g : graphs)
      if (g) (void)hipGraphExecDestroy(g);
    if (hH) (void)hipHostFree(hH);
    if (gs) (void)hipStreamDestroy(gs);
  }
  double* V() { return (double*)dV.p; }
  const SegChunk* ch() { return (const SegChunk*)dch.p; }
  const uint32_t* cb() { return (const uint32_t*)dcb.p; }

  int matvec(double alpha, const double* x, double* y) {
    return ahm_hmat_matvec(h, alpha, x, n, y, n, 1, AHM_F64);
  }
  // per-system |v| into host out[S]
  int norms(const double* v, std::vector<double>& out) {
    hipLaunchKernelGGL(sdot_kernel, dim3(nch), dim3(kSegThreads), 0, gs, v, v, ch(),
                       (double*)dp0.p);
    hipLaunchKernelGGL(sfinish_kernel, dim3(S), dim3(kSegThreads), 0, gs, (const double*)dp0.p,
                       cb(), (double*)dnorm.p);
    out.resize(S);
    AHM_HIP(hipMemcpyAsync(out.data(), dnorm.p, S * 8, hipMemcpyDeviceToHost, gs));
    AHM_HIP(hipStreamSynchronize(gs));
    for (auto& v2 : out) v2 = std::sqrt(v2);
    return AHM_OK;
  }
  // Arnoldi step i for every system (the captured sequence)
  int step(uint32_t i) {
    double* vi = V() + (size_t)i * n;
    double* w = V() + (size_t)(i + 1) * n;
    double* Hd = (double*)dH.p;
    hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, gs, w, n);
    AHM_TRY(matvec(1.0, vi, w));
    hipLaunchKernelGGL(sdot_kernel, dim3(nch), dim3(kSegThreads), 0, gs, w, V(), ch(),
                       (double*)dp0.p);
    double* pin = (double*)dp0.p;
    double* pout = (double*)dp1.p;
    for (uint32_t k = 0; k <= i; ++k) {
      const double* next = (k < i) ? V() + (size_t)(k + 1) * n : nullptr;
      hipLaunchKernelGGL(smgs_kernel, dim3(nch), dim3(kSegThreads), 0, gs, w, V() + (size_t)k * n,
                         next, ch(), cb(), (const double*)pin, Hd + (size_t)k * S, pout);
      std::swap(pin, pout);
    }
    hipLaunchKernelGGL(snormalize_kernel, dim3(nch), dim3(kSegThreads), 0, gs, w, ch(), cb(),
                       (const double*)pin, Hd + (size_t)(i + 1) * S);
    AHM_HIP(hipMemcpyAsync(hH, Hd, (size_t)(i + 2) * S * 8, hipMemcpyDeviceToHost, gs));
    return AHM_OK;
  }
};

}  // namespace

int gmres_batch_device(ahm_hmat h, const double* b, double* x, double eps, uint32_t m,
                       uint32_t nsteps, int use_graph, double* eps_out, uint32_t* steps_out,
                       int* conv_out) {
  Batch B;
  B.h = h;
  B.n = h->n_rows;
  B.m = m;
  std::vector<uint32_t> off = h->seg_off;
  if (off.empty()) off = {0, h->n_rows};
  B.S = (uint32_t)off.size() - 1;
  const uint32_t S = B.S, n = B.n;
  std::vector<SegChunk> chunks;
  std::vector<uint32_t> cb(S + 1, 0);
  for (uint32_t s = 0; s < S; ++s) {
    cb[s] = (uint32_t)chunks.size();
    for (uint32_t a = off[s]; a < off[s + 1]; a += kSegChunk)
      chunks.push_back(SegChunk{s, a, std::min(a + kSegChunk, off[s + 1]), 0});
  }
  cb[S] = (uint32_t)chunks.size();
  B.nch = cb[S];
  hipStream_t user = h->ctx->stream;
  AHM_HIP(hipStreamCreateWithFlags(&B.gs, hipStreamNonBlocking));
  AHM_TRY(B.dch.alloc(chunks.size() * sizeof(SegChunk)));
  AHM_TRY(B.dcb.alloc(cb.size() * 4));
  AHM_TRY(B.dV.alloc((size_t)n * (m + 1) * 8));
  AHM_TRY(B.dr.alloc((size_t)n * 8));
  AHM_TRY(B.dp0.alloc((size_t)B.nch * 8));
  AHM_TRY(B.dp1.alloc((size_t)B.nch * 8));
  AHM_TRY(B.dH.alloc((size_t)(m + 2) * S * 8));
  AHM_TRY(B.dY.alloc((size_t)S * m * 8));
  AHM_TRY(B.dk.alloc((size_t)S * 4));
  AHM_TRY(B.dscale.alloc((size_t)S * 8));
  AHM_TRY(B.dmask.alloc(S));
  AHM_TRY(B.dnorm.alloc((size_t)S * 8));
  AHM_HIP(hipHostMalloc((void**)&B.hH, (size_t)(m + 2) * S * 8, hipHostMallocDefault));
  B.graphs.assign(m, nullptr);
  // order after the caller's work, run on the internal stream
  hipEvent_t ev;
  AHM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  AHM_HIP(hipEventRecord(ev, user));
  AHM_HIP(hipStreamWaitEvent(B.gs, ev, 0));
  AHM_HIP(hipMemcpyAsync(B.dch.p, chunks.data(), chunks.size() * sizeof(SegChunk),
                         hipMemcpyHostToDevice, B.gs));
  AHM_HIP(hipMemcpyAsync(B.dcb.p, cb.data(), cb.size() * 4, hipMemcpyHostToDevice, B.gs));
  h->ctx->stream = B.gs;
  struct Restore {
    ahm_hmat h;
    hipStream_t s;
    ~Restore() { h->ctx->stream = s; }
  } restore{h, user};
  // the matvec workspace must exist before any capture
  AHM_TRY(ahm_hmat_matvec(h, 0.0, b, n, B.dr.p, n, 1, AHM_F64));

  std::vector<double> normb, beta;
  AHM_TRY(B.norms(b, normb));
  std::vector<uint8_t> active(S, 0), zero(S, 0);
  std::vector<uint32_t> steps(S, 0);
  std::vector<double> resid(S, 0.0);
  // r = b - A x
  AHM_HIP(hipMemcpyAsync(B.dr.p, b, (size_t)n * 8, hipMemcpyDeviceToDevice, B.gs));
  AHM_TRY(B.matvec(-1.0, x, (double*)B.dr.p));
  AHM_TRY(B.norms((const double*)B.dr.p, beta));
  bool any_zero = false;
  for (uint32_t s = 0; s < S; ++s) {
    if (normb[s] == 0.0) {
      zero[s] = 1;
      any_zero = true;
      resid[s] = 0.0;
      continue;
    }
    resid[s] = beta[s] / normb[s];
    active[s] = resid[s] > eps;
  }
  if (any_zero) {  // b_s = 0: x_s = 0 (solvers/GMRES.cpp:95-102)
    AHM_HIP(hipMemcpyAsync(B.dmask.p, zero.data(), S, hipMemcpyHostToDevice, B.gs));
    hipLaunchKernelGGL(sfill_kernel, dim3(B.nch), dim3(256), 0, B.gs, x, B.ch(),
                       (const uint8_t*)B.dmask.p);
  }
  auto n_active = [&]() {
    uint32_t c = 0;
    for (uint8_t a : active) c += a;
    return c;
  };
  std::vector<double> H((size_t)S * (m + 1) * m, 0.0), cs((size_t)S * (m + 1)),
      sn((size_t)S * (m + 1)), sv((size_t)S * (m + 1)), Y((size_t)S * m, 0.0);
  std::vector<uint32_t> kseg(S, 0);
  auto Hs = [&](uint32_t s, uint32_t r, uint32_t c) -> double& {
    return H[(size_t)s * (m + 1) * m + r + (size_t)c * (m + 1)];
  };
  // y_s = H_s^{-1} s_s (k x k upper triangle) into Y
  auto solve = [&](uint32_t s, uint32_t k) {
    double* y = Y.data() + (size_t)s * m;
    for (uint32_t q = 0; q < k; ++q) y[q] = sv[(size_t)s * (m + 1) + q];
    for (int i = (int)k - 1; i >= 0; --i) {
      double t = y[i];
      for (uint32_t l = i + 1; l < k; ++l) t -= Hs(s, i, l) * y[l];
      y[i] = t / Hs(s, i, i);
    }
    kseg[s] = k;
  };
  auto apply_updates = [&]() -> int {
    AHM_HIP(hipMemcpyAsync(B.dY.p, Y.data(), Y.size() * 8, hipMemcpyHostToDevice, B.gs));
    AHM_HIP(hipMemcpyAsync(B.dk.p, kseg.data(), S * 4, hipMemcpyHostToDevice, B.gs));
    hipLaunchKernelGGL(supdate_kernel, dim3(B.nch), dim3(256), 0, B.gs, (const double*)B.V(), n, m,
                       B.ch(), (const uint32_t*)B.dk.p, (const double*)B.dY.p, x);
    AHM_HIP(hipStreamSynchronize(B.gs));  // Y, kseg are host buffers
    std::fill(kseg.begin(), kseg.end(), 0u);
    return AHM_OK;
  };
  uint32_t j = 1;
  while (n_active() && j <= nsteps) {
    std::vector<double> scale(S, 0.0);
    for (uint32_t s = 0; s < S; ++s)
      if (active[s]) {
        scale[s] = 1.0 / beta[s];
        sv[(size_t)s * (m + 1)] = beta[s];
        for (uint32_t q = 1; q <= m; ++q) sv[(size_t)s * (m + 1) + q] = 0.0;
      }
    AHM_HIP(hipMemcpyAsync(B.dscale.p, scale.data(), S * 8, hipMemcpyHostToDevice, B.gs));
    hipLaunchKernelGGL(sscale_kernel, dim3(B.nch), dim3(256), 0, B.gs, (const double*)B.dr.p, B.V(),
                       B.ch(), (const double*)B.dscale.p);
    AHM_HIP(hipStreamSynchronize(B.gs));  // scale is a host buffer
    uint32_t i;
    for (i = 0; i < m && j <= nsteps && n_active(); i++, j++) {
      if (use_graph) {
        if (!B.graphs[i]) {
          hipGraph_t g;
          AHM_HIP(hipStreamBeginCapture(B.gs, hipStreamCaptureModeRelaxed));
          const int st = B.step(i);
          const hipError_t ce = hipStreamEndCapture(B.gs, &g);
          if (st) return st;
          if (ce != hipSuccess) return fail(AHM_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
          const hipError_t ie = hipGraphInstantiate(&B.graphs[i], g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (ie != hipSuccess) return fail(AHM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
        }
        AHM_HIP(hipGraphLaunch(B.graphs[i], B.gs));
      } else {
        AHM_TRY(B.step(i));
      }
      AHM_HIP(hipStreamSynchronize(B.gs));
      bool upd = false;
      for (uint32_t s = 0; s < S; ++s) {
        if (!active[s]) continue;
        for (uint32_t k = 0; k <= i + 1; ++k) Hs(s, k, i) = B.hH[(size_t)k * S + s];
        double* c = cs.data() + (size_t)s * (m + 1);
        double* sno = sn.data() + (size_t)s * (m + 1);
        double* svs = sv.data() + (size_t)s * (m + 1);
        for (uint32_t k = 0; k < i; ++k) arot(Hs(s, k, i), Hs(s, k + 1, i), c[k], sno[k]);
        grot(Hs(s, i, i), Hs(s, i + 1, i), c[i], sno[i]);
        arot(Hs(s, i, i), Hs(s, i + 1, i), c[i], sno[i]);
        arot(svs[i], svs[i + 1], c[i], sno[i]);
        resid[s] = std::fabs(svs[i + 1] / normb[s]);
        if (resid[s] < eps) {
          solve(s, i + 1);
          steps[s] = j;
          active[s] = 0;
          upd = true;
        }
      }
      if (upd) AHM_TRY(apply_updates());
    }
    // end of the cycle: update the systems still running, true residuals
    bool upd = false;
    for (uint32_t s = 0; s < S; ++s)
      if (active[s] && i > 0) {
        solve(s, i);
        upd = true;
      }
    if (upd) AHM_TRY(apply_updates());
    if (!n_active()) break;
    AHM_HIP(hipMemcpyAsync(B.dr.p, b, (size_t)n * 8, hipMemcpyDeviceToDevice, B.gs));
    AHM_TRY(B.matvec(-1.0, x, (double*)B.dr.p));
    AHM_TRY(B.norms((const double*)B.dr.p, beta));
    for (uint32_t s = 0; s < S; ++s)
      if (active[s]) {
        resid[s] = beta[s] / normb[s];
        if (resid[s] < eps) {
          active[s] = 0;
          steps[s] = j;
        }
      }
  }
  AHM_HIP(hipEventRecord(ev, B.gs));
  AHM_HIP(hipStreamWaitEvent(user, ev, 0));
  AHM_HIP(hipStreamSynchronize(B.gs));
  (void)hipEventDestroy(ev);
  int all = 1;
  for (uint32_t s = 0; s < S; ++s) {
    const bool conv = zero[s] || (!active[s] && resid[s] <= eps);
    if (eps_out) eps_out[s] = resid[s];
    if (steps_out) steps_out[s] = conv ? steps[s] : nsteps;
    if (conv_out) conv_out[s] = conv ? 1 : 0;
    all &= conv ? 1 : 0;
  }
  return all ? AHM_OK : AHM_ERR_NOT_CONVERGED;
}

}  // namespace ahm

extern "C" int ahm_gmres_batch(ahm_hmat h, const double* b, double* x, double eps, uint32_t m,
                               uint32_t nsteps, int use_graph, double* eps_out,
                               uint32_t* steps_out, int* converged) {
  using namespace ahm;
  if (!h || !b || !x || m == 0 || h->storage != AHM_F64 || h->n_rows != h->n_cols)
    return fail(AHM_ERR_INVALID, "ahm_gmres_batch: bad arguments");
  AHM_HIP(hipSetDevice(h->ctx->device));
  const int st = gmres_batch_device(h, b, x, eps, m, nsteps, use_graph, eps_out, steps_out, converged);
  if (st == AHM_ERR_NOT_CONVERGED) set_error("ahm_gmres_batch: some systems did not converge");
  return st;
}
