// Host-side planning: validation, index fusion, tile choice.  No HIP calls,
// so plans can be made (and inspected) without a device.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/ttc_hip.h"
#include "ttc_params.h"

namespace ttc {

// Vector widths (elements per lane) instantiated for an element of `bytes`
// bytes: 1, 2 and the 16-byte width.  The kernel generator uses the same rule.
inline std::vector<int> vec_widths(int bytes) {
  std::vector<int> w{1};
  if (bytes <= 8) w.push_back(2);
  const int wide = 16 / bytes;
  if (wide > 2) w.push_back(wide);
  return w;
}

// Vector widths of the row-copy kernel: one width for both sides, at most
// 16 bytes on each (1, 2 and the widest).  The kernel generator uses the
// same rule.
inline std::vector<int> row_widths(int bytes_a, int bytes_b) {
  std::vector<int> w{1};
  const int wide = 16 / (bytes_a > bytes_b ? bytes_a : bytes_b);
  if (wide >= 2) w.push_back(2);
  if (wide > 2) w.push_back(wide);
  return w;
}

int dtype_bytes(int t);
// Bytes of the LDS staging type (alpha * a in A's arithmetic) for a type pair.
int lds_elem_bytes(int ta, int tb);
bool pair_supported(int ta, int tb);

struct Geometry {
  int rank = 0;
  int perm[kMaxDim] = {0};
  int size[kMaxDim] = {0};
  int64_t sa[kMaxDim] = {0};  // A stride of A-dim k
  int64_t sb[kMaxDim] = {0};  // B stride of A-dim k
  int lda[kMaxDim] = {0};     // fused outer extents
  int ldb[kMaxDim] = {0};
  int64_t elements = 0;
};

// Restatement of ttc/ttc_util.py:214-276 on fixed arrays.
ttc_status_t fuse_indices(int dim, const int* perm, const int* size, const int* lda, const int* ldb, int* out_dim,
                          int* out_perm, int* out_size, int* out_lda, int* out_ldb);

// Validation (ttc/ttc.py:733-768) + fusion + strides.
ttc_status_t make_geometry(int dim, const int* perm, const int* size, const int* lda, const int* ldb, Geometry& g);

// Kernel families (ttc_config_t.family).
constexpr int kFamilyBox = 1;
constexpr int kFamilyRC = 2;
constexpr int kFamilyRow = 3;

struct PlanChoice {
  int tile[kMaxDim];
  int va, vb;
  int lds_pad;
  int run_a, run_b;
};

int kernel_grid_order(int cfg);

// Fills TileParams for a geometry and a tile; returns TTC_ERROR_UNSUPPORTED
// when the tile violates a kernel limit.
ttc_status_t build_params(const Geometry& g, const int* tile, int ta, int tb, int va, int vb, int grid_order,
                          TileParams& p, int& lds_bytes, int& lds_pad);

// Contiguous element runs of a full tile in A and in B.
void tile_runs(const Geometry& g, const int* tile, long long& ra, long long& rb);
double tile_score(const Geometry& g, const int* tile, int szA, int szB);
// Heuristic tile for a geometry.
void choose_tile(const Geometry& g, int ta, int tb, int va, int vb, int run_a, int run_b, int* tile);

// Widest legal vector on each side for the geometry (alignment of the
// pointers is checked at execution).
int widest_vec_a(const Geometry& g, int ta);
int widest_vec_b(const Geometry& g, int tb);
// Largest tile volume the kernel takes for these vector widths and types.
int max_tile_volume(int ta, int tb, int va, int vb);

// Row x column family: largest tile volume, the group split of a tile (false
// when the tile does not have the row x column shape), its launch parameters
// and the heuristic tile (false when no row x column tile fits).
int max_rc_volume(int ta, int tb);
bool rc_groups(const Geometry& g, const int* tile, int& na, int* ga, int& nb, int* gb);
ttc_status_t build_rc_params(const Geometry& g, const int* tile, int ta, int tb, int va, int vb, int grid_order,
                             int tile_order, RCParams& p, int& lds_bytes, int nbuf = 1);
bool choose_rc_tile(const Geometry& g, int ta, int tb, int va, int vb, int run_a, int run_b, int* tile,
                    double* score);
// The launch's LDS offset tables (p.ntab entries, the layout of RCParams),
// as the kernel's setup() builds them.
void rc_fill_tables(const RCParams& p, int32_t* out);

// Row-copy family (perm[0] == 0): launch parameters for vectors of v
// elements; TTC_ERROR_UNSUPPORTED when the shape does not fit the kernel.
ttc_status_t build_row_params(const Geometry& g, int ta, int tb, int v, RowParams& p);

}  // namespace ttc

struct ttc_plan_s {
  int type_a, type_b, beta_is_zero;
  ttc::Geometry g;
  int va, vb;
  int nt_store;
  int wg_per_cu;
  int grid_order;
  int tile[ttc::kMaxDim];
  ttc::TileParams p;
  int lds_bytes, lds_pad;
  int run_a, run_b;
  int family;            // 1 box tile kernel (p), 2 row x column kernel (rc), 3 row copy (row)
  int tile_order;        // 1 A order, 2 B order, 3 / 4 grouped B order (row x column tiles only)
  int lds_buffers;       // LDS tile buffers of the row x column kernel
  ttc::RCParams rc;
  ttc::RowParams row;
  ttc_plan_s* fallback;  // scalar-vector plan for misaligned pointers (may be null)
};
