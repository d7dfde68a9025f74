// ahm_aca.h -- device assembly declarations (aca.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

// ACA slab capacity: default rank columns per leaf and the hard ceiling
constexpr uint32_t kAcaCapDefault = 48;
constexpr uint32_t kAcaCapMax = 256;

struct KernelSpec {
  int kind;      // AHM_KERNEL_*
  int rows_only; // ACAr instead of ACA on the admissible leaves
  double kappa;
  double c;      // weight/(4 pi) (Laplace, Yukawa) or weight (EXPR)
  int sym;       // symmetric storage: diagonal dense leaves packed (AHM_LEAF_HEM)
  int pad;
};

struct AcaBlock {
  uint32_t b1, n1, b2, n2;
  uint32_t i0;        // ACA start row (relative to b1)
  uint32_t kmax;      // reference rank limit n1*n2/(n1+n2)+4 (or rankmax)
  uint32_t kcap;      // slab capacity in rank columns (index 0..kcap)
  uint32_t kidx;      // kernel of this leaf (index into the kernel table)
  uint64_t uoff, voff, roff;  // slab offsets (doubles)
  uint64_t zoff;              // Z/S offsets (ints)
};

struct AcaResult {
  int32_t status;     // 0 ACA failed, 1 success, 2 slab capacity reached
  uint32_t k;         // ACA rank
  uint32_t kt;        // rank after recompression
  uint32_t pad;
};

struct DenseJob {
  uint32_t b1, n1, b2, n2;
  uint64_t off;
  uint32_t kidx;
  uint32_t packed;   // 1: diagonal block, upper triangle packed column by column
};

KernelSpec make_kernel(const ahm_kernel& k);

// Ks[0..nk) kernel table; leaf_kernel (may be NULL: all 0) selects per leaf
// The layout plan of a leaf list (panels, chunks, copy jobs; hmat.hip)
// built on a host thread, so that the assembly's last GPU work runs
// meanwhile.  The leaves must stay alive and unchanged until the plan is
// used or the job is destroyed (which waits for the thread).
struct PlanJob;
std::shared_ptr<PlanJob> plan_start(uint32_t n_rows, uint32_t n_cols, const std::vector<ahm_leaf_desc>& leaves);

int assemble_leaves(ahm_ctx ctx, const double* pts, const uint32_t* op_perm, uint32_t N,
                    uint32_t nl, const uint32_t* b1, const uint32_t* n1, const uint32_t* b2,
                    const uint32_t* n2, const uint8_t* adm, const uint32_t* i0,
                    const KernelSpec* Ks, uint32_t nk, const uint32_t* leaf_kernel,
                    double eps, uint32_t rankmax, const uint32_t* kcap_hint,
                    std::vector<ahm_leaf_desc>& leaves, DeviceBuf& data,
                    ahm_assembly_stats* stats, uint32_t* aca_rank_out = nullptr,
                    std::shared_ptr<PlanJob>* plan = nullptr);

// builds an H-matrix from leaves whose values are already on the device
// (mblock layout, F64); takes ownership of `raw`.  plan: the layout plan of
// these leaves started by plan_start (built here when null)
int hmat_from_device(ahm_ctx ctx, uint32_t n_rows, uint32_t n_cols,
                     std::vector<ahm_leaf_desc>&& leaves, DeviceBuf& raw, ahm_dtype storage,
                     int keep_leaves, ahm_hmat* out, std::shared_ptr<PlanJob> plan = nullptr);

}  // namespace ahm
