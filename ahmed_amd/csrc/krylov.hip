// krylov.hip -- the GMRES entry points, CG and FP64 iterative refinement on
// the GPU.
//
// GMRes (solvers/GMRES.cpp:80-181) runs in gmres.hip, with the Hessenberg
// matrix, the Givens rotations and the convergence test on the device; this
// file maps the reference's in/out arguments onto it.  CG follows
// solvers/CG.cpp:31-113.
//
// Vectors, the orthogonalisation and the rotations stay FP64 on the device
// for every matrix precision; with an FP32 matrix only the H-matvec runs in
// FP32 (the Krylov vector is rounded to FP32 on the way in, the product
// widened to FP64 on the way out).
//
// Reductions are deterministic: a fixed grid of kRedBlocks workgroups, each
// with a fixed element set, writes one partial per workgroup; consumers sum
// the partials in a fixed order.
#include <cmath>
#include <cstring>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum of a kRedThreads-wide workgroup, fixed order; valid in every thread
__device__ double group_sum(double v) {
  __shared__ double sh[kRedThreads / 64];
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < kRedThreads / 64; ++q) s += sh[q];
  return s;
}

// sum of the kRedBlocks partials, fixed order, in every thread
__device__ double sum_partials(const double* __restrict__ part) {
  __shared__ double sp;
  double v = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v += part[i];
  v = group_sum(v);
  if (threadIdx.x == 0) sp = v;
  __syncthreads();
  return sp;
}

__global__ __launch_bounds__(kRedThreads) void dot_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b, uint32_t n,
                                                          double* __restrict__ part) {
  double s = 0.0;
  for (uint32_t i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += kRedBlocks * kRedThreads)
    s = fma(a[i], b[i], s);
  s = group_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// partials of sum |a_i| (zero exactly when every a_i is zero)
__global__ __launch_bounds__(kRedThreads) void asum_kernel(const double* __restrict__ a, uint32_t n,
                                                           double* __restrict__ part) {
  double s = 0.0;
  for (uint32_t i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += kRedBlocks * kRedThreads)
    s += fabs(a[i]);
  s = group_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// result = sum(part) into out (single workgroup)
__global__ __launch_bounds__(kRedThreads) void finish_kernel(const double* __restrict__ part,
                                                             double* __restrict__ out) {
  const double h = sum_partials(part);
  if (threadIdx.x == 0) *out = h;
}

__global__ void fill_kernel(double* __restrict__ d, uint32_t n, double v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = v;
}

__global__ void to_f32_kernel(const double* __restrict__ s, float* __restrict__ d, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}

__global__ void add_f32_kernel(const float* __restrict__ s, double* __restrict__ d, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] += (double)s[i];
}

// y = a*x + y
__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y,
                            uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = fma(a, x[i], y[i]);
}

// p = rhat + beta*p (CG direction update, solvers/CG.cpp:80-83: scal then axpy)
__global__ void cg_dir_kernel(double beta, const double* __restrict__ r, double* __restrict__ p,
                              uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = r[i] + beta * p[i];
}

namespace {

constexpr int kEwBlocks = 1024;

// The operator y += alpha A x with FP64 vectors, for either matrix precision.
struct Op {
  ahm_hmat h;
  uint32_t n;
  DeviceBuf x32, y32;
  int init() {
    if (h->storage != AHM_F64) {   // FP32 values (AHM_F32, AHM_BF16)
      AHM_TRY(x32.alloc((size_t)n * 4));
      AHM_TRY(y32.alloc((size_t)n * 4));
    }
    return AHM_OK;
  }
  int amux(double alpha, const double* x, double* y) {
    hipStream_t s = h->ctx->stream;
    if (h->storage == AHM_F64) return ahm_hmat_matvec(h, alpha, x, n, y, n, 1, AHM_F64);
    hipLaunchKernelGGL(to_f32_kernel, dim3(kEwBlocks), dim3(256), 0, s, x, (float*)x32.p, n);
    AHM_HIP(hipMemsetAsync(y32.p, 0, (size_t)n * 4, s));
    AHM_TRY(ahm_hmat_matvec(h, alpha, x32.p, n, y32.p, n, 1, AHM_F32));
    hipLaunchKernelGGL(add_f32_kernel, dim3(kEwBlocks), dim3(256), 0, s, (const float*)y32.p, y, n);
    return AHM_OK;
  }
};

struct Red {
  DeviceBuf part, scal;   // kRedBlocks partials (x2), scalars
  double* host = nullptr;
  int init(uint32_t nscal) {
    AHM_TRY(part.alloc(2 * kRedBlocks * sizeof(double)));
    AHM_TRY(scal.alloc((size_t)nscal * sizeof(double)));
    return AHM_OK;
  }
  double* p0() { return (double*)part.p; }
  double* p1() { return (double*)part.p + kRedBlocks; }
  double* s() { return (double*)scal.p; }
};

// |v|_2 on the host (one sync)
int norm2(hipStream_t s, Red& R, const double* v, uint32_t n, double* out) {
  hipLaunchKernelGGL(dot_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, v, v, n, R.p0());
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, s, R.p0(), R.s());
  AHM_HIP(hipMemcpyAsync(out, R.s(), sizeof(double), hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  *out = std::sqrt(*out);
  return AHM_OK;
}

// r = b - A x.  When x is zero on entry (every element exactly 0) the product
// is skipped: A*0 contributes exact zeros, so r = b bit for bit.
int initial_residual(hipStream_t s, Red& R, Op& A, const double* b, const double* x, double* r,
                     uint32_t n, uint32_t& mv) {
  hipLaunchKernelGGL(asum_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, n, R.p0());
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, s, R.p0(), R.s());
  double ax = 0.0;
  AHM_HIP(hipMemcpyAsync(&ax, R.s(), sizeof(double), hipMemcpyDeviceToHost, s));
  AHM_HIP(hipMemcpyAsync(r, b, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  AHM_HIP(hipStreamSynchronize(s));
  if (ax == 0.0) return AHM_OK;
  AHM_TRY(A.amux(-1.0, x, r));
  ++mv;
  return AHM_OK;
}

int dot(hipStream_t s, Red& R, const double* a, const double* b, uint32_t n, double* out) {
  hipLaunchKernelGGL(dot_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, a, b, n, R.p0());
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kRedThreads), 0, s, R.p0(), R.s());
  AHM_HIP(hipMemcpyAsync(out, R.s(), sizeof(double), hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  return AHM_OK;
}

}  // namespace

// CG, solvers/CG.cpp:31-113; with M, r^ = M r (precond_apply) each step.
int cg_device(ahm_hmat h, const double* b, double* x, double* eps, uint32_t* nsteps,
              const Prec& M = Prec()) {
  const uint32_t n = h->n_rows;
  hipStream_t s = h->ctx->stream;
  Op A{h, n};
  AHM_TRY(A.init());
  Red R;
  AHM_TRY(R.init(4));
  DeviceBuf buf;
  AHM_TRY(buf.alloc((size_t)n * 4 * 8));
  double* p = (double*)buf.p;
  double* q = p + n;
  double* r = q + n;
  double* rhat = M.on() ? r + n : r;
  double nrmb;
  AHM_TRY(norm2(s, R, b, n, &nrmb));
  if (nrmb < 1e-16) {
    hipLaunchKernelGGL(fill_kernel, dim3(kEwBlocks), dim3(256), 0, s, x, n, 0.0);
    AHM_HIP(hipStreamSynchronize(s));
    *eps = 0.0;
    *nsteps = 0;
    return AHM_OK;
  }
  uint32_t mv0 = 0;
  AHM_TRY(initial_residual(s, R, A, b, x, r, n, mv0));
  double resid;
  AHM_TRY(norm2(s, R, r, n, &resid));
  if (resid <= *eps * nrmb) {
    *eps = resid / nrmb;
    *nsteps = 0;
    return AHM_OK;
  }
  double rho, rho1 = 0.0;
  const uint32_t maxsteps = *nsteps;
  for (uint32_t l = 1; l <= maxsteps; ++l) {
    if (M.on()) {
      AHM_HIP(hipMemcpyAsync(rhat, r, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
      AHM_TRY(M.apply(rhat));
    }
    AHM_TRY(dot(s, R, r, rhat, n, &rho));
    if (l > 1)
      hipLaunchKernelGGL(cg_dir_kernel, dim3(kEwBlocks), dim3(256), 0, s, rho / rho1, rhat, p, n);
    else
      AHM_HIP(hipMemcpyAsync(p, rhat, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(fill_kernel, dim3(kEwBlocks), dim3(256), 0, s, q, n, 0.0);
    AHM_TRY(A.amux(1.0, p, q));
    double pq;
    AHM_TRY(dot(s, R, p, q, n, &pq));
    const double alpha = rho / pq;
    hipLaunchKernelGGL(axpy_kernel, dim3(kEwBlocks), dim3(256), 0, s, alpha, p, x, n);
    hipLaunchKernelGGL(axpy_kernel, dim3(kEwBlocks), dim3(256), 0, s, -alpha, q, r, n);
    AHM_TRY(norm2(s, R, r, n, &resid));
    if (resid <= *eps * nrmb) {
      *eps = resid / nrmb;
      *nsteps = l;
      return AHM_OK;
    }
    rho1 = rho;
  }
  *eps = resid / nrmb;
  return AHM_ERR_NOT_CONVERGED;
}

// One system through gmres_run (gmres.hip) with the reference's in/out
// arguments: eps, nsteps in; achieved residual and steps out.
int gmres_single(ahm_hmat h, const double* b, double* x, double* eps, uint32_t m, uint32_t* nsteps,
                 double* hist, uint32_t nhist, uint32_t* nmatvec, const Prec* M = nullptr) {
  GmresArgs a;
  a.b = b;
  a.x = x;
  a.eps = *eps;
  a.m = m;
  a.maxsteps = *nsteps;
  a.use_graph = h->solver_graphs;
  a.M = M;
  a.hist = hist;
  a.nhist = hist ? nhist : 0;
  double e = 0.0;
  uint32_t steps = 0;
  int conv = 0;
  a.eps_out = &e;
  a.steps_out = &steps;
  a.conv_out = &conv;
  a.nmatvec = nmatvec;
  const int st = gmres_run(h, a);
  if (st != AHM_OK && st != AHM_ERR_NOT_CONVERGED) return st;
  *eps = e;
  *nsteps = steps;
  return st;
}

}  // namespace ahm

using namespace ahm;

extern "C" {

int ahm_hmat_set_graphs(ahm_hmat h, int on) {
  if (!h) return fail(AHM_ERR_INVALID, "ahm_hmat_set_graphs: NULL handle");
  h->solver_graphs = on ? 1 : 0;
  return AHM_OK;
}

int ahm_hmat_set_arnoldi(ahm_hmat h, int mode) {
  if (!h) return fail(AHM_ERR_INVALID, "ahm_hmat_set_arnoldi: NULL handle");
  if (mode < 0 || mode > 2) return fail(AHM_ERR_INVALID, "ahm_hmat_set_arnoldi: mode %d", mode);
  h->arnoldi = mode;
  return AHM_OK;
}

int ahm_gmres(ahm_hmat h, const double* b, double* x, double* eps, uint32_t m, uint32_t* nsteps,
              double* hist, uint32_t nhist, uint32_t* nmatvec) {
  if (!h || !b || !x || !eps || !nsteps || m == 0)
    return fail(AHM_ERR_INVALID, "ahm_gmres: bad arguments");
  if (h->n_rows != h->n_cols) return fail(AHM_ERR_INVALID, "ahm_gmres: matrix is not square");
  AHM_HIP(hipSetDevice(h->ctx->device));
  const int st = gmres_single(h, b, x, eps, m, nsteps, hist, nhist, nmatvec);
  if (st == AHM_ERR_NOT_CONVERGED)
    set_error("ahm_gmres: no convergence in %u steps (residual %.3e)", *nsteps, *eps);
  return st;
}

int ahm_cg(ahm_hmat h, const double* b, double* x, double* eps, uint32_t* nsteps) {
  if (!h || !b || !x || !eps || !nsteps) return fail(AHM_ERR_INVALID, "ahm_cg: bad arguments");
  if (h->n_rows != h->n_cols) return fail(AHM_ERR_INVALID, "ahm_cg: matrix is not square");
  AHM_HIP(hipSetDevice(h->ctx->device));
  const int st = cg_device(h, b, x, eps, nsteps);
  if (st == AHM_ERR_NOT_CONVERGED)
    set_error("ahm_cg: no convergence in %u steps (residual %.3e)", *nsteps, *eps);
  return st;
}

static int prec_of(ahm_hmat h, ahm_trih L, ahm_trih U, Prec& M, const char* who) {
  for (ahm_trih t : {L, U}) {
    if (!t) continue;
    uint32_t nd = 0;
    uint64_t ni = 0, nv = 0;
    AHM_TRY(ahm_trih_info(t, &nd, &ni, &nv));
    if (t->ctx != h->ctx) return fail(AHM_ERR_INVALID, "%s: the factors must share the matrix's context", who);
    if (t->n != h->n_rows) return fail(AHM_ERR_INVALID, "%s: factor of size %u for a matrix of %u rows", who, t->n,
                                       h->n_rows);
  }
  M.L = L;
  M.U = U;
  M.n = h->n_rows;
  return AHM_OK;
}

int ahm_gmres_prec(ahm_hmat h, ahm_trih L, ahm_trih U, const double* b, double* x, double* eps,
                   uint32_t m, uint32_t* nsteps, double* hist, uint32_t nhist, uint32_t* nmatvec) {
  if (!h || !b || !x || !eps || !nsteps || m == 0)
    return fail(AHM_ERR_INVALID, "ahm_gmres_prec: bad arguments");
  if (h->n_rows != h->n_cols) return fail(AHM_ERR_INVALID, "ahm_gmres_prec: matrix is not square");
  Prec M;
  AHM_TRY(prec_of(h, L, U, M, "ahm_gmres_prec"));
  AHM_HIP(hipSetDevice(h->ctx->device));
  const int st = gmres_single(h, b, x, eps, m, nsteps, hist, nhist, nmatvec, &M);
  if (st == AHM_ERR_NOT_CONVERGED)
    set_error("ahm_gmres_prec: no convergence in %u steps (residual %.3e)", *nsteps, *eps);
  return st;
}

int ahm_cg_prec(ahm_hmat h, ahm_trih L, ahm_trih U, const double* b, double* x, double* eps,
                uint32_t* nsteps) {
  if (!h || !b || !x || !eps || !nsteps) return fail(AHM_ERR_INVALID, "ahm_cg_prec: bad arguments");
  if (h->n_rows != h->n_cols) return fail(AHM_ERR_INVALID, "ahm_cg_prec: matrix is not square");
  Prec M;
  AHM_TRY(prec_of(h, L, U, M, "ahm_cg_prec"));
  AHM_HIP(hipSetDevice(h->ctx->device));
  const int st = cg_device(h, b, x, eps, nsteps, M);
  if (st == AHM_ERR_NOT_CONVERGED)
    set_error("ahm_cg_prec: no convergence in %u steps (residual %.3e)", *nsteps, *eps);
  return st;
}

// FP64 iterative refinement around GMRES on a lower-precision matrix:
//   r = b - A64 x;  solve A_lo d = r by GMRES(m) to inner_eps;  x += d
// until |r| / |b| < eps.  The residual uses the FP64 H-matvec.
int ahm_gmres_ir(ahm_hmat hi, ahm_hmat lo, const double* b, double* x, double* eps, uint32_t m,
                 double inner_eps, uint32_t inner_steps, uint32_t max_outer, uint32_t* outer,
                 uint32_t* total_inner, double* hist, uint32_t nhist) {
  if (!hi || !lo || !b || !x || !eps || m == 0 || hi->storage != AHM_F64 ||
      hi->n_rows != lo->n_rows || hi->n_rows != hi->n_cols)
    return fail(AHM_ERR_INVALID, "ahm_gmres_ir: bad arguments");
  AHM_HIP(hipSetDevice(hi->ctx->device));
  const uint32_t n = hi->n_rows;
  hipStream_t s = hi->ctx->stream;
  Red R;
  AHM_TRY(R.init(2));
  DeviceBuf dr, dd;
  AHM_TRY(dr.alloc((size_t)n * 8));
  AHM_TRY(dd.alloc((size_t)n * 8));
  double* r = (double*)dr.p;
  double* d = (double*)dd.p;
  double normb, resid = 1.0;
  AHM_TRY(norm2(s, R, b, n, &normb));
  uint32_t it = 0, inner_total = 0, nh = 0;
  if (normb == 0.0) {
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s, x, n, 0.0);
    AHM_HIP(hipStreamSynchronize(s));
    *eps = 0.0;
    if (outer) *outer = 0;
    if (total_inner) *total_inner = 0;
    return AHM_OK;
  }
  int st = AHM_ERR_NOT_CONVERGED;
  Op Ahi{hi, n};
  for (;;) {
    // FP64 residual r = b - A x (skipped product when x starts at zero)
    uint32_t mvh = 0;
    if (it == 0) {
      AHM_TRY(initial_residual(s, R, Ahi, b, x, r, n, mvh));
    } else {
      AHM_HIP(hipMemcpyAsync(r, b, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
      AHM_TRY(ahm_hmat_matvec(hi, -1.0, x, n, r, n, 1, AHM_F64));
    }
    double nr;
    AHM_TRY(norm2(s, R, r, n, &nr));
    resid = nr / normb;
    if (hist && nh < nhist) hist[nh++] = resid;
    if (resid < *eps) {
      st = AHM_OK;
      break;
    }
    if (it >= max_outer) break;
    // inner solve A_lo d = r, d0 = 0, relative tolerance inner_eps
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s, d, n, 0.0);
    double ie = std::max(inner_eps, 0.5 * (*eps) / resid);
    uint32_t ns = inner_steps, mvs = 0;
    const int gs = gmres_single(lo, r, d, &ie, m, &ns, nullptr, 0, &mvs);
    if (gs != AHM_OK && gs != AHM_ERR_NOT_CONVERGED) return gs;
    inner_total += ns;
    hipLaunchKernelGGL(axpy_kernel, dim3(1024), dim3(256), 0, s, 1.0, (const double*)d, x, n);
    ++it;
  }
  AHM_HIP(hipStreamSynchronize(s));
  *eps = resid;
  if (outer) *outer = it;
  if (total_inner) *total_inner = inner_total;
  if (st) set_error("ahm_gmres_ir: no convergence after %u refinements (residual %.3e)", it, resid);
  return st;
}

}  // extern "C"
