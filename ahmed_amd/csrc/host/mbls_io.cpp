// mbls_io.cpp -- the reference's on-disk H-matrix format (savembls /
// loadmbls, Include/H.h:47-103; mblock::save / load, Include/mblock.h:737-804)
// read and written from leaf descriptors and mblock-layout values.  Host only.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <vector>

#include "ahm.h"
#include "ahm_internal.h"

namespace {

constexpr uint32_t kStatusLowRank = 0, kStatusDense = 4;

size_t esize(ahm_dtype t) { return t == AHM_F64 ? 8 : t == AHM_F32 ? 4 : 0; }

struct Reader {
  std::ifstream is;
  uint64_t left = 0;  // bytes not yet read
  explicit Reader(const char* path) : is(path, std::ios::binary | std::ios::ate) {
    if (is) {
      left = (uint64_t)is.tellg();
      is.seekg(0);
    }
  }
  bool get(void* p, uint64_t n) {
    if (n > left) return false;
    is.read((char*)p, (std::streamsize)n);
    left -= n;
    return (bool)is;
  }
};

// walk a file: header, then per leaf (n1, n2, rank, values); values are
// copied into data (when given) at the running offset
int walk(const char* path, ahm_dtype dtype, double* eta, uint32_t* bmin, uint32_t* n_leaves,
         uint64_t* n_values, uint32_t want_leaves, ahm_leaf_desc* leaves, void* data) {
  const size_t es = esize(dtype);
  if (!path || !es) return ahm::fail(AHM_ERR_INVALID, "mbls: bad path or dtype");
  Reader r(path);
  if (!r.is) return ahm::fail(AHM_ERR_INVALID, "mbls: cannot open %s", path);
  double e;
  uint32_t b, n;
  if (!r.get(&e, 8) || !r.get(&b, 4) || !r.get(&n, 4))
    return ahm::fail(AHM_ERR_INVALID, "mbls: %s: truncated header", path);
  if (eta) *eta = e;
  if (bmin) *bmin = b;
  if (n_leaves) *n_leaves = n;
  if (leaves && want_leaves != n)
    return ahm::fail(AHM_ERR_INVALID, "mbls: %s holds %u leaves, %u expected", path, n, want_leaves);
  uint64_t off = 0;
  for (uint32_t l = 0; l < n; ++l) {
    uint8_t flag = 0;
    uint32_t h[3];
    if (!r.get(&flag, 1)) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: truncated at leaf %u", path, l);
    if (flag != 1) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: leaf %u absent", path, l);
    if (!r.get(h, 12)) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: truncated at leaf %u", path, l);
    const uint32_t n1 = h[0], n2 = h[1], status = h[2];
    int32_t rank;
    uint64_t nv;
    if (status == kStatusLowRank) {
      uint32_t k;
      if (!r.get(&k, 4)) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: truncated at leaf %u", path, l);
      if (k > (uint32_t)INT32_MAX) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: leaf %u rank %u", path, l, k);
      rank = (int32_t)k;
      nv = (uint64_t)k * ((uint64_t)n1 + n2);
    } else if (status == kStatusDense) {
      rank = -1;
      nv = (uint64_t)n1 * n2;
    } else {
      return ahm::fail(AHM_ERR_UNSUPPORTED,
                       "mbls: %s: leaf %u has status %u (Hermitian/triangular/symmetric storage)", path,
                       l, status);
    }
    if (nv > r.left / es) return ahm::fail(AHM_ERR_INVALID, "mbls: %s: truncated in leaf %u", path, l);
    if (leaves) {
      ahm_leaf_desc& d = leaves[l];
      std::memset(&d, 0, sizeof d);
      d.n1 = n1;
      d.n2 = n2;
      d.rank = rank;
      d.offset = off;
    }
    if (data) {
      if (!r.get((char*)data + off * es, nv * es))
        return ahm::fail(AHM_ERR_INVALID, "mbls: %s: read error in leaf %u", path, l);
    } else {
      r.is.seekg((std::streamoff)(nv * es), std::ios::cur);
      r.left -= nv * es;
    }
    off += nv;
  }
  if (n_values) *n_values = off;
  return AHM_OK;
}

}  // namespace

extern "C" {

int ahm_mbls_save(const char* path, double eta, uint32_t bmin, uint32_t n_leaves,
                  const ahm_leaf_desc* leaves, const void* data, ahm_dtype dtype) {
  const size_t es = esize(dtype);
  if (!path || !es || (n_leaves && (!leaves || !data)))
    return ahm::fail(AHM_ERR_INVALID, "ahm_mbls_save: bad arguments");
  std::ofstream os(path, std::ios::binary | std::ios::trunc);
  if (!os) return ahm::fail(AHM_ERR_INVALID, "ahm_mbls_save: cannot create %s", path);
  os.write((const char*)&eta, 8);
  os.write((const char*)&bmin, 4);
  os.write((const char*)&n_leaves, 4);
  const uint8_t present = 1;
  for (uint32_t l = 0; l < n_leaves; ++l) {
    const ahm_leaf_desc& d = leaves[l];
    const uint32_t h[3] = {d.n1, d.n2, d.rank < 0 ? kStatusDense : kStatusLowRank};
    os.write((const char*)&present, 1);
    os.write((const char*)h, 12);
    uint64_t nv;
    if (d.rank < 0) {
      nv = (uint64_t)d.n1 * d.n2;
    } else {
      const uint32_t k = (uint32_t)d.rank;
      os.write((const char*)&k, 4);
      nv = (uint64_t)k * ((uint64_t)d.n1 + d.n2);
    }
    if (nv) os.write((const char*)data + d.offset * es, (std::streamsize)(nv * es));
  }
  os.flush();
  if (!os) return ahm::fail(AHM_ERR_INVALID, "ahm_mbls_save: write to %s failed", path);
  return AHM_OK;
}

int ahm_mbls_info(const char* path, ahm_dtype dtype, double* eta, uint32_t* bmin,
                  uint32_t* n_leaves, uint64_t* n_values) {
  return walk(path, dtype, eta, bmin, n_leaves, n_values, 0, nullptr, nullptr);
}

int ahm_mbls_read(const char* path, ahm_dtype dtype, uint32_t n_leaves, ahm_leaf_desc* leaves,
                  void* data) {
  if (!leaves || !data) return ahm::fail(AHM_ERR_INVALID, "ahm_mbls_read: NULL argument");
  return walk(path, dtype, nullptr, nullptr, nullptr, nullptr, n_leaves, leaves, data);
}

}  // extern "C"
