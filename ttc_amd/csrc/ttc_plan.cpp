// Host-side planning for libttc_hip.so: validation, index fusion, tile choice
// and kernel launch parameters.  No HIP runtime calls here.
#include "ttc_plan.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

namespace ttc {

int dtype_bytes(int t) {
  switch (t) {
    case TTC_S: return 4;
    case TTC_D: return 8;
    case TTC_C: return 8;
    case TTC_Z: return 16;
    case TTC_BF16:
    case TTC_F16: return 2;
    case TTC_E4M3:
    case TTC_E5M2: return 1;
    default: return 0;
  }
}

static bool f32_class(int t) { return t == TTC_S || t == TTC_BF16 || t == TTC_F16; }

bool pair_supported(int ta, int tb) {
  if (f32_class(ta) && f32_class(tb)) return !(ta != TTC_S && tb != TTC_S && ta != tb);  // no bf16 <-> f16
  if ((ta == TTC_S || ta == TTC_D) && (tb == TTC_S || tb == TTC_D)) return true;
  if ((ta == TTC_C || ta == TTC_Z) && (tb == TTC_C || tb == TTC_Z)) return true;
  if (f32_class(ta) && (tb == TTC_E4M3 || tb == TTC_E5M2)) return true;
  return false;
}

int lds_elem_bytes(int ta, int /*tb*/) {
  switch (ta) {
    case TTC_D: return 8;
    case TTC_C: return 8;
    case TTC_Z: return 16;
    default: return 4;  // s, bf16, f16: alpha*a is held in fp32
  }
}

// ------------------------------------------------------------------ fusion
// Restatement of ttc/ttc_util.py:214-276.  For every A index i, find the B
// position j holding it and extend the match (A indices i, i+1, ... equal to
// perm[j], perm[j+1], ...) while the previous dimension is unpadded in both A
// (lda == size) and B (ldb == size); the matched run becomes one index whose
// extent, lda and ldb are the products.  Afterwards the surviving index
// labels are renumbered to 0..rank-1 keeping their order.
ttc_status_t fuse_indices(int dim, const int* perm_in, const int* size_in, const int* lda_in, const int* ldb_in,
                          int* out_dim, int* out_perm, int* out_size, int* out_lda, int* out_ldb) {
  if (dim < 1 || dim > TTC_MAX_INPUT_DIM || !perm_in || !size_in) return TTC_ERROR_INVALID_VALUE;
  std::vector<long long> size(dim), lda(dim), ldb(dim);
  std::vector<int> perm(perm_in, perm_in + dim), idx(dim);
  for (int k = 0; k < dim; ++k) {
    if (perm[k] < 0 || perm[k] >= dim) return TTC_ERROR_INVALID_VALUE;
    size[k] = size_in[k];
    idx[k] = k;
  }
  for (int k = 0; k < dim; ++k) {
    lda[k] = lda_in ? lda_in[k] : size[k];
    ldb[k] = ldb_in ? ldb_in[k] : size[perm[k]];
  }
  for (int i = 0; i < dim; ++i) {
    for (int j = 0; j < dim; ++j) {
      int l = 0;
      long long fused = 1;
      while (i + l < (int)idx.size() && j + l < (int)perm.size() && idx[i + l] == perm[j + l] &&
             (l == 0 || lda[i + l - 1] == size[i + l - 1]) && (l == 0 || ldb[j + l - 1] == size[i + l - 1])) {
        fused *= size[i + l];
        ++l;
      }
      if (l >= 1) size[i] = fused;
      for (; l > 1; --l) {
        idx.erase(idx.begin() + i + 1);
        perm.erase(perm.begin() + j + 1);
        size.erase(size.begin() + i + 1);
        lda[i] *= lda[i + 1];
        lda.erase(lda.begin() + i + 1);
        ldb[j] *= ldb[j + 1];
        ldb.erase(ldb.begin() + j + 1);
      }
    }
  }
  // Renumber perm labels to 0..rank-1 in increasing order (the labels left
  // are distinct, so each becomes its rank among them).
  const int r = (int)perm.size();
  std::vector<int> sorted(perm);
  std::sort(sorted.begin(), sorted.end());
  for (int q = 0; q < r; ++q)
    perm[q] = (int)(std::lower_bound(sorted.begin(), sorted.end(), perm[q]) - sorted.begin());
  for (int k = 0; k < r; ++k)
    if (size[k] > INT_MAX || lda[k] > INT_MAX || ldb[k] > INT_MAX) return TTC_ERROR_UNSUPPORTED;
  *out_dim = r;
  for (int k = 0; k < r; ++k) {
    out_perm[k] = perm[k];
    out_size[k] = (int)size[k];
    if (out_lda) out_lda[k] = (int)lda[k];
    if (out_ldb) out_ldb[k] = (int)ldb[k];
  }
  return TTC_SUCCESS;
}

ttc_status_t make_geometry(int dim, const int* perm, const int* size, const int* lda, const int* ldb, Geometry& g) {
  if (dim < 1 || dim > TTC_MAX_INPUT_DIM || !perm || !size) return TTC_ERROR_INVALID_VALUE;
  bool seen[TTC_MAX_INPUT_DIM] = {false};
  for (int k = 0; k < dim; ++k) {  // ttc/ttc.py:753-757
    if (perm[k] < 0 || perm[k] >= dim || seen[perm[k]]) return TTC_ERROR_INVALID_VALUE;
    seen[perm[k]] = true;
    if (size[k] < 0) return TTC_ERROR_INVALID_VALUE;
  }
  for (int k = 0; k < dim; ++k) {  // ttc/ttc.py:739-746
    if (lda && lda[k] < size[k]) return TTC_ERROR_INVALID_VALUE;
    if (ldb && ldb[k] < size[perm[k]]) return TTC_ERROR_INVALID_VALUE;
  }
  int r = 0;
  int fperm[TTC_MAX_INPUT_DIM], fsize[TTC_MAX_INPUT_DIM], flda[TTC_MAX_INPUT_DIM], fldb[TTC_MAX_INPUT_DIM];
  ttc_status_t st = fuse_indices(dim, perm, size, lda, ldb, &r, fperm, fsize, flda, fldb);
  if (st != TTC_SUCCESS) return st;
  // strides of the fused problem
  int64_t sa[TTC_MAX_INPUT_DIM], sb[TTC_MAX_INPUT_DIM], sbpos[TTC_MAX_INPUT_DIM];
  sa[0] = 1;
  sbpos[0] = 1;
  for (int k = 1; k < r; ++k) {
    sa[k] = sa[k - 1] * flda[k - 1];
    sbpos[k] = sbpos[k - 1] * fldb[k - 1];
  }
  for (int q = 0; q < r; ++q) sb[fperm[q]] = sbpos[q];
  // Extent-1 dimensions index nothing: drop them (their padding is already
  // in the strides of the others); a tensor of one element keeps dimension 0.
  // The new leading dimensions may then have strides other than 1, which the
  // vector-width checks see.
  int keep[TTC_MAX_INPUT_DIM], newid[TTC_MAX_INPUT_DIM], nk = 0;
  bool all_one = true;
  for (int k = 0; k < r; ++k) all_one &= fsize[k] == 1;
  for (int k = 0; k < r; ++k) {
    keep[k] = fsize[k] != 1 || (all_one && k == 0);
    newid[k] = keep[k] ? nk++ : -1;
  }
  if (nk > kMaxDim) return TTC_ERROR_UNSUPPORTED;
  g = Geometry();
  g.rank = nk;
  g.elements = 1;
  for (int k = 0; k < r; ++k) {
    g.elements *= fsize[k];
    if (!keep[k]) continue;
    const int n = newid[k];
    g.size[n] = fsize[k];
    g.lda[n] = flda[k];
    g.sa[n] = sa[k];
    g.sb[n] = sb[k];
  }
  int q2 = 0;
  for (int q = 0; q < r; ++q) {
    if (!keep[fperm[q]]) continue;
    g.perm[q2] = newid[fperm[q]];
    g.ldb[q2] = fldb[q];
    ++q2;
  }
  return TTC_SUCCESS;
}

// ------------------------------------------------------------ vector widths
static bool vec_ok_a(const Geometry& g, int w) {
  if (w > 1 && g.sa[0] != 1) return false;
  if (g.size[0] % w) return false;
  for (int k = 1; k < g.rank; ++k)
    if (g.sa[k] % w) return false;
  return true;
}
static bool vec_ok_b(const Geometry& g, int w) {
  const int p0 = g.perm[0];
  if (w > 1 && g.sb[p0] != 1) return false;
  if (g.size[p0] % w) return false;
  for (int k = 0; k < g.rank; ++k)
    if (k != p0 && g.sb[k] % w) return false;
  return true;
}
int widest_vec_a(const Geometry& g, int ta) {
  int best = 1;
  if (g.elements == 0) return 1;
  for (int w : vec_widths(dtype_bytes(ta)))
    if (vec_ok_a(g, w)) best = std::max(best, w);
  return best;
}
int widest_vec_b(const Geometry& g, int tb) {
  int best = 1;
  if (g.elements == 0) return 1;
  for (int w : vec_widths(dtype_bytes(tb)))
    if (vec_ok_b(g, w)) best = std::max(best, w);
  return best;
}

int max_tile_volume(int ta, int tb, int va, int vb) {
  const int kLdsBudget = 32768;  // bytes of LDS per work-group (5 work-groups per 160 KiB CU)
  const int v = kThreads * std::min(slots_for(va, dtype_bytes(ta)) * va, slots_for(vb, dtype_bytes(tb)) * vb);
  return std::min(v, kLdsBudget / lds_elem_bytes(ta, tb));
}

// -------------------------------------------------------------- tile choice
static int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Split extent n into pieces of about `want`, balanced so the last piece is
// not a sliver, then rounded up to `align` (which divides n).
static int split_extent(int n, int want, int align) {
  want = std::max(want, 1);
  if (want >= n) return n;
  const int pieces = (n + want - 1) / want;
  int t = (n + pieces - 1) / pieces;
  t = round_up(t, align);
  return std::min(t, n);
}

static int align_of(const Geometry& g, int k, int va, int vb) {
  int a = 1;
  if (k == 0) a = std::max(a, va);
  if (k == g.perm[0]) a = std::max(a, vb);
  return a;
}

// Largest tile-relative byte offset on each side must fit the kernels'
// unsigned 32-bit offsets (kept below 2^31).
static bool span_ok(const Geometry& g, const int* tile, int szA, int szB) {
  long long sa = 0, sb = 0;
  for (int k = 0; k < g.rank; ++k) {
    sa += (long long)(tile[k] - 1) * g.sa[k];
    sb += (long long)(tile[k] - 1) * g.sb[k];
  }
  return sa * szA < INT_MAX && sb * szB < INT_MAX;
}

// Tile for given run targets (elements) on A's and B's side.
static void tile_for_runs(const Geometry& g, int va, int vb, int vmax, int szA, int szB, int run_a, int run_b,
                          int* tile) {
  const int r = g.rank;
  for (int k = 0; k < kMaxDim; ++k) tile[k] = 1;
  // A side: cover A's leading dimensions while they stay one contiguous run.
  long long prod = 1;
  for (int k = 0; k < r; ++k) {
    if (k > 0 && g.sa[k] != g.sa[k - 1] * g.size[k - 1]) break;
    const int n = g.size[k];
    if (prod * n <= run_a) {
      tile[k] = n;
      prod *= n;
      continue;
    }
    tile[k] = std::max(tile[k], split_extent(n, (int)((run_a + prod - 1) / prod), align_of(g, k, va, vb)));
    break;
  }
  // B side: the same along B's order.
  prod = 1;
  for (int q = 0; q < r; ++q) {
    const int k = g.perm[q];
    if (q > 0 && g.sb[k] != g.sb[g.perm[q - 1]] * g.size[g.perm[q - 1]]) break;
    const int n = g.size[k];
    if (tile[k] == n || prod * n <= run_b) {
      tile[k] = n;
      prod *= n;
      continue;
    }
    tile[k] = std::max(tile[k], split_extent(n, (int)((run_b + prod - 1) / prod), align_of(g, k, va, vb)));
    break;
  }
  for (int k = 0; k < r; ++k)
    tile[k] = std::max(1, std::min(round_up(tile[k], align_of(g, k, va, vb)), std::max(g.size[k], 1)));
  // Shrink to the volume limit: halve the largest extent that can shrink.
  for (;;) {
    long long v = 1;
    for (int k = 0; k < r; ++k) v *= tile[k];
    if (v <= vmax) break;
    int best = -1;
    for (int k = 0; k < r; ++k) {
      const int a = align_of(g, k, va, vb);
      if (tile[k] > a && (best < 0 || tile[k] >= tile[best])) best = k;
    }
    if (best < 0) break;
    const int a = align_of(g, best, va, vb);
    tile[best] = std::max(a, round_up((tile[best] + 1) / 2, a));
  }
  // Shrink to the offset limit: halve the extent with the largest span.
  while (!span_ok(g, tile, szA, szB)) {
    int best = -1;
    long long bspan = 0;
    for (int k = 0; k < r; ++k) {
      const long long sp = (long long)(tile[k] - 1) * std::max(g.sa[k] * szA, g.sb[k] * szB);
      if (tile[k] > align_of(g, k, va, vb) && sp > bspan) {
        bspan = sp;
        best = k;
      }
    }
    if (best < 0) break;
    const int a = align_of(g, best, va, vb);
    tile[best] = std::max(a, round_up((tile[best] + 1) / 2, a));
  }
  // At most two dimensions may have a partial last tile; make any further
  // split extent a divisor of its dimension.
  int partial = 0;
  for (int k = 0; k < r; ++k) {
    if (g.size[k] % tile[k] == 0) continue;
    if (++partial <= 2) continue;
    const int a = align_of(g, k, va, vb);
    int t = tile[k];
    while (t > a && (g.size[k] % t != 0 || t % a != 0)) --t;
    tile[k] = std::max(t, a);
  }
}

void tile_runs(const Geometry& g, const int* tile, long long& ra, long long& rb) {
  ra = 1;
  for (int k = 0; k < g.rank; ++k) {
    ra *= tile[k];
    if (tile[k] != g.size[k] || (k + 1 < g.rank && g.sa[k + 1] != g.sa[k] * g.size[k])) break;
  }
  rb = 1;
  for (int q = 0; q < g.rank; ++q) {
    const int k = g.perm[q];
    rb *= tile[k];
    if (tile[k] != g.size[k] || (q + 1 < g.rank && g.sb[g.perm[q + 1]] != g.sb[k] * g.size[k])) break;
  }
}

// Model of a tile's efficiency: the fraction of launched tile slots that hold
// data (partial edge tiles waste the rest), times b/(b+64) per side for a
// contiguous run of b bytes (a misaligned run drags in about half a 128-byte
// line at each end), times V/(V+256) for the per-tile fixed cost.
double tile_score(const Geometry& g, const int* tile, int szA, int szB) {
  long long V = 1, ntiles = 1;
  for (int k = 0; k < g.rank; ++k) {
    V *= tile[k];
    ntiles *= (g.size[k] + tile[k] - 1) / std::max(tile[k], 1);
  }
  if (V <= 0 || ntiles <= 0) return 0.0;
  long long ra, rb;
  tile_runs(g, tile, ra, rb);
  const double eff = (double)g.elements / ((double)ntiles * (double)V);
  const double ba = (double)ra * szA, bb = (double)rb * szB;
  // each side's run efficiency b/(b+64), weighted by its share of the bytes
  // moved (a weighted geometric mean; equal element sizes give the product)
  const double wa = 2.0 * szA / (szA + szB), wb = 2.0 * szB / (szA + szB);
  return eff * std::pow(ba / (ba + 64.0), wa) * std::pow(bb / (bb + 64.0), wb) * ((double)V / ((double)V + 256.0));
}

void choose_tile(const Geometry& g, int ta, int tb, int va, int vb, int run_a, int run_b, int* tile) {
  const int szA = dtype_bytes(ta), szB = dtype_bytes(tb);
  const int vmax = max_tile_volume(ta, tb, va, vb);
  if (run_a > 0 && run_b > 0) {
    tile_for_runs(g, va, vb, vmax, szA, szB, run_a, run_b, tile);
    return;
  }
  // Search the run targets (powers of two) and keep the best-scoring tile.
  double best = -1.0;
  int cand[kMaxDim];
  for (int ra = 1; ra <= vmax; ra *= 2) {
    for (int rb = 1; rb <= vmax; rb *= 2) {
      if ((long long)ra * rb > 4LL * vmax) continue;
      tile_for_runs(g, va, vb, vmax, szA, szB, run_a > 0 ? run_a : ra, run_b > 0 ? run_b : rb, cand);
      const double sc = tile_score(g, cand, szA, szB);
      if (sc > best * (1.0 + 1e-9)) {
        best = sc;
        for (int k = 0; k < kMaxDim; ++k) tile[k] = cand[k];
      }
    }
  }
}

// ------------------------------------------------------------ launch params
// ttc_config_t.grid_order -> the kernels' tile_range order: strided walk (0),
// contiguous ranges (1), strided walk with XCD-grouped work-groups (2), one
// contiguous share of the tiles per XCD walked by its work-groups (3).
int kernel_grid_order(int cfg) {
  switch (cfg) {
    case 1: return 0;
    case 2: return 1;
    case 3: return 2;
    default: return 3;
  }
}

ttc_status_t build_params(const Geometry& g, const int* tile, int ta, int tb, int va, int vb, int grid_order,
                          TileParams& p, int& lds_bytes, int& lds_pad) {
  const int lds_elem = lds_elem_bytes(ta, tb);
  std::memset(&p, 0, sizeof(p));
  const int r = g.rank;
  p.rank = r;
  long long V = 1, ntiles = 1, spanA = 0, spanB = 0;
  int nsplit = 0;
  p.split0 = p.split1 = -1;
  for (int k = 0; k < r; ++k) {
    const int n = g.size[k], t = tile[k];
    if (n == 0) {
      ntiles = 0;
    }
    if (t < 1 || (n > 0 && t > n) || t > 65535) return TTC_ERROR_INVALID_VALUE;
    V *= t;
    spanA += (long long)(t - 1) * g.sa[k];
    spanB += (long long)(t - 1) * g.sb[k];
    p.t[k] = t;
    p.n[k] = n;
    p.perm[k] = g.perm[k];
    p.sa[k] = t > 1 ? (int32_t)g.sa[k] : 0;
    p.sb[k] = t > 1 ? (int32_t)g.sb[k] : 0;
    p.tsa[k] = (int64_t)t * g.sa[k];
    p.tsb[k] = (int64_t)t * g.sb[k];
    p.tdiv[k] = make_fastdiv((uint32_t)t);
    const long long gk = n > 0 ? (n + t - 1) / t : 1;
    ntiles *= gk;
    if (ntiles > INT_MAX) return TTC_ERROR_UNSUPPORTED;
    p.gdiv[k] = make_fastdiv((uint32_t)gk);
    if (n > 0 && n % t != 0) {
      if (nsplit == 0) { p.split0 = k; p.tsplit0 = t; }
      else if (nsplit == 1) { p.split1 = k; p.tsplit1 = t; }
      else return TTC_ERROR_UNSUPPORTED;
      ++nsplit;
    }
  }
  // tile-relative byte offsets are unsigned 32-bit in the kernels
  if (spanA * dtype_bytes(ta) >= INT_MAX || spanB * dtype_bytes(tb) >= INT_MAX) return TTC_ERROR_UNSUPPORTED;
  if (V > max_tile_volume(ta, tb, va, vb)) return TTC_ERROR_UNSUPPORTED;
  const int p0 = g.perm[0];
  if (tile[0] % va || tile[p0] % vb) return TTC_ERROR_INVALID_VALUE;
  if (g.size[0] % va || g.size[p0] % vb) return TTC_ERROR_INVALID_VALUE;
  p.V = (int32_t)V;
  p.ntiles = (uint32_t)ntiles;
  // LDS layout: A order; the stride of B's stride-1 dimension is made odd so
  // that lane-consecutive reads of the write phase hit distinct banks.
  lds_pad = 0;
  p.ls[0] = 1;
  for (int k = 1; k < r; ++k) {
    p.ls[k] = p.ls[k - 1] * tile[k - 1];
    if (k == p0 && (p.ls[k] % 2) == 0) {
      lds_pad = 1;
      p.ls[k] += 1;
    }
  }
  long long elems = 1;
  for (int k = 0; k < r; ++k) elems += (long long)(tile[k] - 1) * p.ls[k];
  p.lds_elems = (int32_t)elems;
  p.lds_stride_b = p.ls[p0];
  p.splitB0 = p.splitB1 = -1;
  for (int q = 0; q < r; ++q) {
    const int k = g.perm[q];
    p.tdivB[q] = p.tdiv[k];
    p.tB[q] = p.t[k];
    p.sbB[q] = p.sb[k];
    p.lsB[q] = p.ls[k];
    if (k == p.split0) p.splitB0 = q;
    if (k == p.split1) p.splitB1 = q;
  }
  // prefix products for the digits of the split dimensions
  auto prefix_a = [&](int k) {
    uint32_t P = 1;
    for (int m = 0; m < k; ++m) P *= (uint32_t)tile[m];
    return P;
  };
  auto prefix_b = [&](int k) {
    uint32_t P = 1;
    for (int q = 0; q < r && g.perm[q] != k; ++q) P *= (uint32_t)tile[g.perm[q]];
    return P;
  };
  p.pA0 = make_fastdiv(p.split0 >= 0 ? prefix_a(p.split0) : 1);
  p.pA1 = make_fastdiv(p.split1 >= 0 ? prefix_a(p.split1) : 1);
  p.pB0 = make_fastdiv(p.split0 >= 0 ? prefix_b(p.split0) : 1);
  p.pB1 = make_fastdiv(p.split1 >= 0 ? prefix_b(p.split1) : 1);
  p.ts0 = make_fastdiv(p.split0 >= 0 ? (uint32_t)tile[p.split0] : 1);
  p.ts1 = make_fastdiv(p.split1 >= 0 ? (uint32_t)tile[p.split1] : 1);
  lds_bytes = (int)((elems * lds_elem + 15) / 16 * 16);
  if (lds_bytes > 65536) return TTC_ERROR_UNSUPPORTED;
  p.grid_order = kernel_grid_order(grid_order);
  return TTC_SUCCESS;
}

// ------------------------------------------------------ row x column family
// LDS per work-group (tile data plus offset tables): four 256-thread
// work-groups per CU (160 KiB of LDS).
static constexpr int kRcLdsBudget = 40960 * (kRcThreads / 256);
static constexpr long long kRcFastTiles = 2048;  // fewest tiles of a launch on the lane-fixed walk

int max_rc_volume(int ta, int tb) {
  const int v = kRcThreads * std::min(rc_elems(dtype_bytes(ta)), rc_elems(dtype_bytes(tb)));
  return std::min(v, 32768 * (kRcThreads / 256) / lds_elem_bytes(ta, tb));
}

// LDS of a row x column tile: the data region (odd column pitch RP = R | 1,
// since the write phase reads LDS with stride RP; everything either walk can
// reach, with a spare vector slot at RP * C where padding slots stage) in
// bytes, and the offset-table entries.  The lane-fixed write walk reads rows
// up to capB / Cw - 1 of its lane's column (values of padding slots are never
// stored), which can reach a little past RP * C.
static void rc_layout(long long R, long long C, long long Rw, long long Cw, long long capA, long long capB,
                      int lds_elem, bool fast, long long& RP, long long& data, long long& tab) {
  RP = R | 1;
  long long reach = RP * C + 16;
  if (fast) reach = std::max(reach, capB / Cw + (C - 1) * RP);
  data = (reach * lds_elem + 15) / 16 * 16;
  tab = fast ? Rw + R + C + Cw + capA / Rw + capB / Cw : 2 * R + 2 * C;
}

// Rows: A's leading dimensions 0, 1, ... up to B's stride-1 dimension;
// columns: B's leading dimensions perm[0], perm[1], ... up to a row
// dimension, then any other tiled dimension (in A order).  A group's walk
// stops after a dimension whose extent the tile does not divide and before a
// dimension the tile does not cover (t == 1); only a group's outermost
// dimension may have a partial last tile.
bool rc_groups(const Geometry& g, const int* tile, int& na, int* ga, int& nb, int* gb) {
  const int r = g.rank;
  const int p0 = g.perm[0];
  na = nb = 0;
  if (p0 == 0 || r < 2) return false;
  bool in_a[kMaxDim] = {false}, in_b[kMaxDim] = {false};
  for (int k = 0; k < r && k != p0; ++k) {
    if (k > 0 && (tile[k] == 1 || g.size[k - 1] % tile[k - 1] != 0)) break;
    ga[na++] = k;
    in_a[k] = true;
  }
  for (int q = 0; q < r; ++q) {
    const int k = g.perm[q];
    if (in_a[k]) break;
    if (q > 0) {
      const int kp = g.perm[q - 1];
      if (tile[k] == 1 || g.size[kp] % tile[kp] != 0) break;
    }
    gb[nb++] = k;
    in_b[k] = true;
  }
  for (int k = 0; k < r; ++k) {
    if (tile[k] <= 1 || in_a[k] || in_b[k]) continue;
    gb[nb++] = k;  // an extra column dimension
    in_b[k] = true;
  }
  for (int m = 0; m + 1 < nb; ++m)
    if (g.size[gb[m]] % tile[gb[m]] != 0) return false;
  return true;
}

// Grows a row x column tile over the dimensions it leaves at t == 1, in A
// order, each by its largest divisor that keeps the volume within vmax (the
// column group's partial dimension, if any, must stay its outermost).
static void rc_grow(const Geometry& g, int* tile, long long vmax) {
  long long V = 1;
  for (int k = 0; k < g.rank; ++k) V *= tile[k];
  int na, nb, ga[kMaxDim], gb[kMaxDim];
  if (!rc_groups(g, tile, na, ga, nb, gb)) return;
  if (g.size[gb[nb - 1]] % tile[gb[nb - 1]] != 0) return;
  for (int k = 0; k < g.rank; ++k) {
    if (tile[k] != 1 || g.size[k] <= 1) continue;
    const long long room = vmax / V;
    if (room < 2) break;
    int best = 1;
    for (long long t = std::min<long long>(room, g.size[k]); t >= 2; --t)
      if (g.size[k] % t == 0) {
        best = (int)t;
        break;
      }
    tile[k] = best;
    V *= best;
  }
}

ttc_status_t build_rc_params(const Geometry& g, const int* tile, int ta, int tb, int va, int vb, int grid_order,
                             int tile_order, RCParams& p, int& lds_bytes, int nbuf) {
  int na, nb, ga[kMaxDim], gb[kMaxDim];
  if (!rc_groups(g, tile, na, ga, nb, gb)) return TTC_ERROR_INVALID_VALUE;
  std::memset(&p, 0, sizeof(p));
  const int r = g.rank;
  const int szA = dtype_bytes(ta), szB = dtype_bytes(tb);
  long long ntiles = 1, spanA = 0, spanB = 0;
  long long gk[kMaxDim];
  for (int k = 0; k < r; ++k) {
    const int n = g.size[k], t = tile[k];
    if (t < 1 || (n > 0 && t > n)) return TTC_ERROR_INVALID_VALUE;
    spanA += (long long)(t - 1) * g.sa[k];
    spanB += (long long)(t - 1) * g.sb[k];
    gk[k] = n > 0 ? (n + t - 1) / t : 1;
    if (n == 0) ntiles = 0;
    ntiles *= gk[k];
    if (ntiles > INT_MAX) return TTC_ERROR_UNSUPPORTED;
  }
  // byte offsets (plus one vector) stay below the buffer range limit 2^31
  if (spanA * szA + 64 >= INT_MAX || spanB * szB + 64 >= INT_MAX) return TTC_ERROR_UNSUPPORTED;
  const int p0 = g.perm[0];
  if (tile[0] % va || tile[p0] % vb || g.size[0] % va || g.size[p0] % vb) return TTC_ERROR_INVALID_VALUE;
  long long R = 1, C = 1;
  p.rowA_dense = 1;
  p.na = na;
  for (int m = 0; m < na; ++m) {
    const int k = ga[m];
    if (g.sa[k] != R) p.rowA_dense = 0;
    p.ta[m] = tile[k];
    p.dta[m] = make_fastdiv((uint32_t)tile[k]);
    p.saA[m] = (int32_t)(g.sa[k] * szA);
    p.sbA[m] = (int32_t)(g.sb[k] * szB);
    R *= tile[k];
  }
  p.colB_dense = 1;
  p.nb = nb;
  for (int m = 0; m < nb; ++m) {
    const int k = gb[m];
    if (g.sb[k] != C) p.colB_dense = 0;
    p.tb[m] = tile[k];
    p.dtb[m] = make_fastdiv((uint32_t)tile[k]);
    p.saB[m] = (int32_t)(g.sa[k] * szA);
    p.sbB[m] = (int32_t)(g.sb[k] * szB);
    C *= tile[k];
  }
  if (R * C > max_rc_volume(ta, tb)) return TTC_ERROR_UNSUPPORTED;
  if (nbuf < 1 || nbuf > 2) return TTC_ERROR_INVALID_VALUE;
  // Lane-fixed walk of full tiles: the walks' row / column extents rounded
  // up to powers of two must divide the walk steps and still cover the tile
  // within each side's slots, and a side whose walk has padding slots (its
  // tables then hold kRcTabOOB) must keep its table entries below
  // kRcFastSpan; without padding slots the buffer range bounds them (a
  // 64-column fp32 tile of a tensor whose column stride is 25 MB spans 1.6 GB).
  // Padding lanes cost memory instructions: at most 1/16 of either walk may
  // be padding (measured: rows 192 -> 256 or 100 -> 128 lost 8-10%).  Its
  // longer setup pays off only over a few tiles per work-group: launches of
  // fewer than kRcFastTiles tiles keep the general walk.  It pays where the
  // general walk's per-element index arithmetic weighs most, with 2-byte
  // elements on a side (same-box A/B: fp32 -> bf16 (3,1,0,2) +7%, bf16 ->
  // fp8 8192 x 14336 +7%); on 4- and 8-byte pairs it measured neutral over
  // the 64 ops of the 200 MiB fp32/fp64 set and -3% on fp32 -> fp8.
  const long long capA = (long long)kRcThreads * rc_elems(szA), capB = (long long)kRcThreads * rc_elems(szB);
  long long Rw = 1, Cw = 1;
  while (Rw < R) Rw <<= 1;
  while (Cw < C) Cw <<= 1;
  const bool padA = Rw > R || Rw * C < capA, padB = Cw > C || R * Cw < capB;  // sentinels in a side's tables
  bool fast = Rw <= (long long)kRcThreads * va && Cw <= (long long)kRcThreads * vb && Rw * C <= capA &&
              R * Cw <= capB && (!padA || spanA * szA < kRcFastSpan) && (!padB || spanB * szB < kRcFastSpan) &&
              16 * Rw <= 17 * R && 16 * Cw <= 17 * C && ntiles >= kRcFastTiles && (szA == 2 || szB == 2);
  const int le = lds_elem_bytes(ta, tb);
  long long RP = 0, data = 0, tab = 0;
  rc_layout(R, C, Rw, Cw, capA, capB, le, fast, RP, data, tab);
  if (fast && data + tab * 4 > kRcLdsBudget) {
    fast = false;
    rc_layout(R, C, Rw, Cw, capA, capB, le, fast, RP, data, tab);
  }
  if (data + tab * 4 > kRcLdsBudget) return TTC_ERROR_UNSUPPORTED;
  const long long lds = nbuf * data + tab * 4;
  const int ka = ga[na - 1], kb = gb[nb - 1];
  const bool partA = g.size[ka] > 0 && g.size[ka] % tile[ka];
  const bool partB = g.size[kb] > 0 && g.size[kb] % tile[kb];
  p.R = (int32_t)R;
  p.C = (int32_t)C;
  p.Rl = partA ? (int32_t)(R / tile[ka] * (g.size[ka] % tile[ka])) : p.R;
  p.Cl = partB ? (int32_t)(C / tile[kb] * (g.size[kb] % tile[kb])) : p.C;
  p.dR = make_fastdiv((uint32_t)p.R);
  p.dRl = make_fastdiv((uint32_t)p.Rl);
  p.dC = make_fastdiv((uint32_t)p.C);
  p.dCl = make_fastdiv((uint32_t)p.Cl);
  p.RP = (int32_t)RP;
  p.fast = fast ? 1 : 0;
  p.Rw = (int32_t)(fast ? Rw : R);
  p.Cw = (int32_t)(fast ? Cw : C);
  p.Rw_shift = fast ? __builtin_ctzll((unsigned long long)Rw) : 0;
  p.Cw_shift = fast ? __builtin_ctzll((unsigned long long)Cw) : 0;
  p.kA = fast ? (int32_t)((long long)kRcThreads * va / Rw) : 0;
  p.kB = fast ? (int32_t)((long long)kRcThreads * vb / Cw) : 0;
  p.nia = rc_elems(szA) / va;
  p.nib = rc_elems(szB) / vb;
  p.ntab = (int32_t)tab;
  // Tile grid: dimensions with more than one tile, in A order (1) or B
  // order (2).  Grouped B order (3: groups of 8, 4: groups of 16): the row
  // group's outermost tiled dimension is split into an inner position,
  // walked first, of G tiles and an outer position at its place in B
  // order, so that the tiles a round takes together extend each other's
  // runs on both sides (G along A's rows, the rest along B's columns).
  p.splitA = p.splitB = -1;
  p.grank = 0;
  p.gouter = -1;
  p.group = 1;
  p.gcount = 0;
  const int G = tile_order == 3 ? 8 : (tile_order == 4 ? 16 : 0);
  int kg = -1;
  if (G && ntiles > 0) {
    int tiled = 0;
    for (int k = 0; k < r; ++k) tiled += gk[k] > 1;
    for (int m = na - 1; m >= 0 && kg < 0; --m)
      if (gk[ga[m]] > 1) kg = ga[m];
    if (tiled + 1 > kMaxDim) kg = -1;  // no position left for the split
  }
  long long grid = ntiles > 0 ? 1 : 0;
  long long gi = 1, outer = 1;
  if (kg >= 0) {
    gi = std::min<long long>(G, gk[kg]);
    outer = (gk[kg] + gi - 1) / gi;
    const int pos = p.grank++;
    p.n[pos] = outer > 1 ? (int32_t)(gi * tile[kg]) : g.size[kg];  // never partial when split
    p.t[pos] = tile[kg];
    p.gdiv[pos] = make_fastdiv((uint32_t)gi);
    p.tsa[pos] = (int64_t)tile[kg] * g.sa[kg];
    p.tsb[pos] = (int64_t)tile[kg] * g.sb[kg];
    if (outer == 1 && partA && kg == ka) p.splitA = pos;
    grid *= gi;
  }
  for (int q = 0; q < r; ++q) {
    const int k = tile_order >= 2 ? g.perm[q] : q;
    if (gk[k] <= 1) continue;
    if (k == kg) {
      if (outer <= 1) continue;
      const int pos = p.grank++;
      p.n[pos] = g.size[k];
      p.t[pos] = tile[k];
      p.gdiv[pos] = make_fastdiv((uint32_t)outer);
      p.tsa[pos] = gi * tile[k] * g.sa[k];
      p.tsb[pos] = gi * tile[k] * g.sb[k];
      p.gouter = pos;
      p.group = (uint32_t)gi;
      p.gcount = (uint32_t)gk[k];
      if (partA && k == ka) p.splitA = pos;
      grid *= outer;
      continue;
    }
    const int pos = p.grank++;
    p.n[pos] = g.size[k];
    p.t[pos] = tile[k];
    p.gdiv[pos] = make_fastdiv((uint32_t)gk[k]);
    p.tsa[pos] = (int64_t)tile[k] * g.sa[k];
    p.tsb[pos] = (int64_t)tile[k] * g.sb[k];
    if (partA && k == ka) p.splitA = pos;
    if (partB && k == kb) p.splitB = pos;
    grid *= gk[k];
  }
  if (grid > INT_MAX) return TTC_ERROR_UNSUPPORTED;
  p.ntiles = (uint32_t)grid;
  p.grid_order = kernel_grid_order(grid_order);
  p.lds_data_elems = (int32_t)(data / le);
  p.lds_spare = (int32_t)(p.RP * C);
  p.nbuf = nbuf;
  lds_bytes = (int)lds;
  return TTC_SUCCESS;
}

// Offsets of tile row i (A and B) and tile column j: the mixed-radix
// decomposition the kernel's setup() performs.
static void rc_row(const RCParams& p, int i, int32_t& oa, int32_t& ob) {
  oa = ob = 0;
  for (int m = 0; m < p.na; ++m) {
    const int d = i % p.ta[m];
    i /= p.ta[m];
    oa += d * p.saA[m];
    ob += d * p.sbA[m];
  }
}
static void rc_col(const RCParams& p, int j, int32_t& oa, int32_t& ob) {
  oa = ob = 0;
  for (int m = 0; m < p.nb; ++m) {
    const int d = j % p.tb[m];
    j /= p.tb[m];
    oa += d * p.saB[m];
    ob += d * p.sbB[m];
  }
}

void rc_fill_tables(const RCParams& p, int32_t* out) {
  int32_t* rA = out;
  int32_t* rB = rA + p.Rw;
  int32_t* cA = rB + p.R;
  int32_t* cB = cA + p.C;
  for (int i = 0; i < p.Rw; ++i) {
    int32_t oa = (int32_t)kRcTabOOB, ob = 0;
    if (i < p.R) rc_row(p, i, oa, ob);
    rA[i] = oa;
    if (i < p.R) rB[i] = ob;
  }
  for (int j = 0; j < p.Cw; ++j) {
    int32_t oa = 0, ob = (int32_t)kRcTabOOB;
    if (j < p.C) rc_col(p, j, oa, ob);
    if (j < p.C) cA[j] = oa;
    cB[j] = ob;
  }
  if (!p.fast) return;
  int32_t* cS = cB + p.Cw;
  int32_t* rS = cS + p.kA * p.nia;
  for (int e = 0; e < p.kA * p.nia; ++e) {
    const int j = e / p.nia + (e % p.nia) * p.kA;
    cS[e] = j < p.C ? cA[j] : (int32_t)kRcTabOOB;
  }
  for (int e = 0; e < p.kB * p.nib; ++e) {
    const int i = e / p.nib + (e % p.nib) * p.kB;
    rS[e] = i < p.R ? rB[i] : (int32_t)kRcTabOOB;
  }
}

// Row copy: rows are the positions of dimensions 1..rank-1 in B order
// (perm[1], perm[2], ...); a work item is one vector of v elements of
// dimension 0.  Offsets are 64-bit in the kernel, so no span limit applies;
// the vector count must stay below 2^31 for the 32-bit index walk.
ttc_status_t build_row_params(const Geometry& g, int ta, int tb, int v, RowParams& p) {
  std::memset(&p, 0, sizeof(p));
  if (g.rank < 1 || g.perm[0] != 0 || v < 1 || g.size[0] % v) return TTC_ERROR_UNSUPPORTED;
  // rows are contiguous on both sides; a dropped size-1 leading dimension
  // with a padded leading dimension leaves dimension 0 strided (e.g. B of
  // size 1 x 4 with ldb[0] = 3): such ops take the tile kernels
  if (g.sa[0] != 1 || g.sb[0] != 1) return TTC_ERROR_UNSUPPORTED;
  if (g.elements == 0) return TTC_SUCCESS;  // total = 0: nothing is launched
  const int szA = dtype_bytes(ta), szB = dtype_bytes(tb);
  const long long nv = g.size[0] / v;
  long long rows = 1;
  p.rank = g.rank - 1;
  for (int q = 1; q < g.rank; ++q) {
    const int k = g.perm[q];
    p.dn[q - 1] = make_fastdiv((uint32_t)g.size[k]);
    p.sa[q - 1] = g.sa[k] * szA;
    p.sb[q - 1] = g.sb[k] * szB;
    rows *= g.size[k];
  }
  const long long total = rows * nv;
  if (total >= (1ll << 31)) return TTC_ERROR_UNSUPPORTED;
  p.total = (uint32_t)total;
  p.nv = (uint32_t)nv;
  p.dnv = make_fastdiv((uint32_t)nv);
  return TTC_SUCCESS;
}

static int rc_whole_extent_tiles(const Geometry& g, int va, int vb, int vmax, int (*out)[kMaxDim], int max_out) {
  const int r = g.rank, p0 = g.perm[0];
  int n = 0;
  if (r < 2 || p0 == 0) return 0;
  auto emit = [&](const int* t) {
    if (n < max_out)
      for (int k = 0; k < kMaxDim; ++k) out[n][k] = k < r ? t[k] : 1;
    ++n;
  };
  // side 0: whole leading B dimensions as columns, rows (A from dimension 0)
  // filling the rest of the volume; side 1: the mirror image
  for (int side = 0; side < 2; ++side) {
    long long whole = 1;
    for (int m = 1; m <= r; ++m) {
      int t[kMaxDim];
      for (int k = 0; k < kMaxDim; ++k) t[k] = 1;
      bool ok = true;
      whole = 1;
      for (int q = 0; q < m && ok; ++q) {
        const int k = side == 0 ? g.perm[q] : q;
        if (side == 0 ? k < p0 : k >= p0) ok = false;  // would take the other group's leading dimension
        t[k] = g.size[k];
        whole *= g.size[k];
      }
      if (!ok || whole * 8 > vmax) break;
      // the other group: its leading dimensions in order, whole while they
      // fit, then one split dimension with balanced pieces
      long long vol = whole;
      for (int q = 0; q < r; ++q) {
        const int k = side == 0 ? q : g.perm[q];
        if (t[k] != 1 || g.size[k] == 1) {
          if (t[k] != 1) break;
          continue;
        }
        const long long room = vmax / vol;
        if (room < 2) break;
        if (g.size[k] <= room) {
          t[k] = g.size[k];
          vol *= g.size[k];
          continue;
        }
        t[k] = split_extent(g.size[k], (int)room, align_of(g, k, va, vb));
        if ((long long)t[k] * vol > vmax) t[k] = std::max(1, (int)(room / align_of(g, k, va, vb)) * align_of(g, k, va, vb));
        break;
      }
      emit(t);
    }
  }
  return std::min(n, max_out);
}

bool choose_rc_tile(const Geometry& g, int ta, int tb, int va, int vb, int run_a, int run_b, int* tile,
                    double* score) {
  if (g.elements == 0 || g.rank < 2 || g.perm[0] == 0) return false;
  const int szA = dtype_bytes(ta), szB = dtype_bytes(tb);
  const int vmax = max_rc_volume(ta, tb);
  double best = -1.0;
  int cand[kMaxDim];
  RCParams p;
  int lds;
  auto consider = [&](const int* c) {
    if (build_rc_params(g, c, ta, tb, va, vb, 0, 1, p, lds) != TTC_SUCCESS) return;
    // Every tile slot is issued whether or not it holds an element: a
    // mostly empty tile costs nearly a full one (measured: tiles filling
    // 1% of the slots ran ~50x slower per byte, half-full ones ~15%).
    long long V = 1;
    for (int k = 0; k < g.rank; ++k) V *= c[k];
    const double u = (double)V / (double)vmax;
    const double sc = tile_score(g, c, szA, szB) * (u / (0.3 + 0.7 * u));
    if (sc > best * (1.0 + 1e-9)) {
      best = sc;
      for (int k = 0; k < kMaxDim; ++k) tile[k] = c[k];
    }
  };
  for (int ra = 1; ra <= vmax; ra *= 2) {
    for (int rb = 1; rb <= vmax; rb *= 2) {
      if ((long long)ra * rb > 4LL * vmax) continue;
      tile_for_runs(g, va, vb, vmax, szA, szB, run_a > 0 ? run_a : ra, run_b > 0 ? run_b : rb, cand);
      rc_grow(g, cand, vmax);
      consider(cand);
    }
  }
  // Tiles covering short leading extents whole (e.g. all 155 rows of B's
  // leading dimension): run targets never produce them, since an
  // overflowing tile halves its largest extent.  Measured +6-17% on such
  // shapes over the run-target search's best.
  if (run_a == 0 && run_b == 0) {
    int whole[2 * kMaxDim][kMaxDim];
    const int nw = rc_whole_extent_tiles(g, va, vb, vmax, whole, 2 * kMaxDim);
    for (int i = 0; i < nw; ++i) consider(whole[i]);
  }
  if (score) *score = best;
  return best >= 0.0;
}

}  // namespace ttc

// ================================================================== C ABI
using namespace ttc;

// ttc_version() is generated with a digest of the sources (gen/ttc_version.cpp).

extern "C" const char* ttc_status_string(ttc_status_t s) {
  switch (s) {
    case TTC_SUCCESS: return "success";
    case TTC_ERROR_INVALID_VALUE: return "invalid value";
    case TTC_ERROR_UNSUPPORTED: return "unsupported type pair or shape";
    case TTC_ERROR_HIP: return "HIP runtime error";
    case TTC_ERROR_ALLOC: return "allocation failed";
    case TTC_ERROR_NO_DEVICE: return "no usable gfx950 device";
    case TTC_ERROR_WORKSPACE: return "work table larger than the fixed device workspace (TTC_BATCH_WORKSPACE_BYTES)";
  }
  return "unknown status";
}

extern "C" int ttc_dtype_size(ttc_dtype_t t) { return dtype_bytes((int)t); }

extern "C" ttc_status_t ttc_fuse_indices(int dim, const int* perm, const int* size, const int* lda, const int* ldb,
                                         int* out_dim, int* out_perm, int* out_size, int* out_lda, int* out_ldb) {
  if (!out_dim || !out_perm || !out_size) return TTC_ERROR_INVALID_VALUE;
  return fuse_indices(dim, perm, size, lda, ldb, out_dim, out_perm, out_size, out_lda, out_ldb);
}

static ttc_status_t plan_fill(ttc_plan_s* pl, int va, int vb, const int* tile_in, int run_a, int run_b,
                              int family) {
  const Geometry& g = pl->g;
  int tile[kMaxDim];
  bool explicit_tile = false;
  if (tile_in)
    for (int k = 0; k < g.rank; ++k) explicit_tile |= tile_in[k] != 0;
  int fam = family;
  if (fam == kFamilyRow && g.perm[0] != 0) return TTC_ERROR_UNSUPPORTED;
  if (fam == kFamilyRow || (fam == 0 && !explicit_tile && g.perm[0] == 0)) {
    // A and B share the stride-1 dimension: rows move without LDS staging.
    // One vector width for both sides: the widest instantiated one within
    // both sides' widths (legal on both, since they are).
    int v = 1;
    for (int w : row_widths(dtype_bytes(pl->type_a), dtype_bytes(pl->type_b)))
      if (w <= va && w <= vb && va % w == 0 && vb % w == 0) v = std::max(v, w);
    RowParams rp;
    const ttc_status_t st = build_row_params(g, pl->type_a, pl->type_b, v, rp);
    if (st == TTC_SUCCESS) {
      pl->va = pl->vb = v;
      pl->family = kFamilyRow;
      for (int k = 0; k < kMaxDim; ++k) pl->tile[k] = k == 0 ? g.size[0] : 1;
      std::memset(&pl->p, 0, sizeof(pl->p));
      std::memset(&pl->rc, 0, sizeof(pl->rc));
      pl->row = rp;
      const long long chunk = (long long)kThreads * kRowUnroll;
      pl->p.ntiles = (uint32_t)((rp.total + chunk - 1) / chunk);
      pl->p.V = g.size[0];
      pl->lds_bytes = 0;
      pl->lds_pad = 0;
      pl->run_a = pl->run_b = g.size[0];
      return TTC_SUCCESS;
    }
    if (fam == kFamilyRow) return st;
    fam = 0;  // too many vectors for the row walk: tile kernels
  }
  if (explicit_tile) {
    for (int k = 0; k < kMaxDim; ++k) tile[k] = k < g.rank ? tile_in[k] : 1;
    if (fam == 0) fam = kFamilyBox;
  } else {
    int rc_tile[kMaxDim];
    double rc_score = -1.0;
    const bool rc_ok = fam != kFamilyBox &&
                       choose_rc_tile(g, pl->type_a, pl->type_b, va, vb, run_a, run_b, rc_tile, &rc_score);
    if (fam == kFamilyRC && !rc_ok) return TTC_ERROR_UNSUPPORTED;
    if (fam != kFamilyRC) choose_tile(g, pl->type_a, pl->type_b, va, vb, run_a, run_b, tile);
    if (fam == 0) {
      // The row x column kernel issues all of a tile's slots: it wins when
      // its tile fills at least half of them or is no smaller than the box.
      long long vr = 1, vbox = 1;
      for (int k = 0; k < g.rank; ++k) {
        vr *= rc_ok ? rc_tile[k] : 1;
        vbox *= tile[k];
      }
      fam = rc_ok && (2 * vr >= max_rc_volume(pl->type_a, pl->type_b) || vr >= vbox) ? kFamilyRC : kFamilyBox;
    }
    if (fam == kFamilyRC)
      for (int k = 0; k < kMaxDim; ++k) tile[k] = rc_tile[k];
  }
  pl->va = va;
  pl->vb = vb;
  pl->family = fam;
  for (int k = 0; k < kMaxDim; ++k) pl->tile[k] = k < g.rank ? tile[k] : 1;
  ttc_status_t st;
  if (fam == kFamilyRC) {
    std::memset(&pl->p, 0, sizeof(pl->p));
    st = build_rc_params(g, tile, pl->type_a, pl->type_b, va, vb, pl->grid_order, pl->tile_order, pl->rc,
                         pl->lds_bytes, pl->lds_buffers);
    pl->lds_pad = pl->rc.RP != pl->rc.R;
    pl->p.ntiles = pl->rc.ntiles;
    pl->p.V = pl->rc.R * pl->rc.C;
    pl->p.grid_order = pl->rc.grid_order;
  } else {
    std::memset(&pl->rc, 0, sizeof(pl->rc));
    st = build_params(g, tile, pl->type_a, pl->type_b, va, vb, pl->grid_order, pl->p, pl->lds_bytes, pl->lds_pad);
  }
  if (st != TTC_SUCCESS) return st;
  long long ra, rb;
  tile_runs(g, tile, ra, rb);
  pl->run_a = (int)std::min<long long>(ra, INT_MAX);
  pl->run_b = (int)std::min<long long>(rb, INT_MAX);
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_plan_create(ttc_plan_t* plan, ttc_dtype_t type_a, ttc_dtype_t type_b, int dim,
                                        const int* perm, const int* size, const int* lda, const int* ldb,
                                        int beta_is_zero, const ttc_config_t* cfg) {
  if (!plan) return TTC_ERROR_INVALID_VALUE;
  *plan = nullptr;
  if (!pair_supported(type_a, type_b)) return TTC_ERROR_UNSUPPORTED;
  const bool fp8 = type_b == TTC_E4M3 || type_b == TTC_E5M2;
  if (fp8 && !beta_is_zero) return TTC_ERROR_UNSUPPORTED;
  ttc_config_t c;
  std::memset(&c, 0, sizeof(c));
  if (cfg) c = *cfg;
  if (c.nt_store < 0 || c.nt_store > 2 || c.grid_order < 0 || c.grid_order > 4 || c.wg_per_cu < 0 || c.vec_a < 0 ||
      c.vec_b < 0 || c.family < 0 || c.family > 3 || c.tile_order < 0 || c.tile_order > 4 ||
      c.lds_buffers < 0 || c.lds_buffers > 2)
    return TTC_ERROR_INVALID_VALUE;
  ttc_plan_s* pl = new (std::nothrow) ttc_plan_s();
  if (!pl) return TTC_ERROR_ALLOC;
  pl->type_a = type_a;
  pl->type_b = type_b;
  pl->beta_is_zero = beta_is_zero ? 1 : 0;
  pl->fallback = nullptr;
  ttc_status_t st = make_geometry(dim, perm, size, lda, ldb, pl->g);
  if (st != TTC_SUCCESS) {
    delete pl;
    return st;
  }
  // Non-temporal stores measured slower than cached ones on MI355X for most
  // shapes (the autotuner still tries them); never with beta != 0, where B
  // is read back.
  pl->nt_store = c.nt_store == 1 && pl->beta_is_zero ? 1 : 0;
  pl->grid_order = c.grid_order;
  // B-order numbering: consecutive tiles of a round share B's cache lines
  // (measured the better default together with XCD-grouped work-groups)
  pl->tile_order = c.tile_order == 0 ? 2 : c.tile_order;
  // two LDS tile buffers: one barrier per tile instead of two, at the cost
  // of resident work-groups (measured +1.4% over the C2 permutations)
  pl->lds_buffers = c.lds_buffers == 0 ? 2 : c.lds_buffers;
  pl->wg_per_cu = c.wg_per_cu;
  auto in_list = [](int w, int bytes) {
    for (int x : vec_widths(bytes))
      if (x == w) return true;
    return false;
  };
  int va = c.vec_a ? c.vec_a : widest_vec_a(pl->g, type_a);
  int vb = c.vec_b ? c.vec_b : widest_vec_b(pl->g, type_b);
  if (c.family == kFamilyRow) {
    // the row copy has one width for both sides (its own instantiated list)
    auto in_row = [&](int w) {
      for (int x : row_widths(dtype_bytes(type_a), dtype_bytes(type_b)))
        if (x == w) return true;
      return false;
    };
    if (c.vec_a && c.vec_b && c.vec_a != c.vec_b) {
      delete pl;
      return TTC_ERROR_INVALID_VALUE;
    }
    if (c.vec_a || c.vec_b) {
      const int w = c.vec_a ? c.vec_a : c.vec_b;
      if (!in_row(w) || !vec_ok_a(pl->g, w) || !vec_ok_b(pl->g, w)) {
        delete pl;
        return TTC_ERROR_INVALID_VALUE;
      }
      va = vb = w;
    }
  } else if (!in_list(va, dtype_bytes(type_a)) || !in_list(vb, dtype_bytes(type_b)) || !vec_ok_a(pl->g, va) ||
             !vec_ok_b(pl->g, vb)) {
    delete pl;
    return TTC_ERROR_INVALID_VALUE;
  }
  st = plan_fill(pl, va, vb, c.tile, c.run_a, c.run_b, c.family);
  if (st != TTC_SUCCESS) {
    delete pl;
    return st;
  }
  if (va > 1 || vb > 1) {
    // Scalar-vector variant for pointers that are not aligned to the vector.
    ttc_plan_s* fb = new (std::nothrow) ttc_plan_s(*pl);
    if (!fb) {
      delete pl;
      return TTC_ERROR_ALLOC;
    }
    fb->fallback = nullptr;
    const int* t = pl->tile;
    long long v = 1;
    for (int k = 0; k < pl->g.rank; ++k) v *= t[k];
    const bool reuse = pl->family == kFamilyRC || v <= max_tile_volume(type_a, type_b, 1, 1);
    st = plan_fill(fb, 1, 1, reuse ? t : nullptr, c.run_a, c.run_b, pl->family);
    if (st != TTC_SUCCESS) {
      delete fb;
      delete pl;
      return st;
    }
    pl->fallback = fb;
  }
  *plan = pl;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_plan_get_info(ttc_plan_t pl, ttc_plan_info_t* info) {
  if (!pl || !info) return TTC_ERROR_INVALID_VALUE;
  std::memset(info, 0, sizeof(*info));
  const Geometry& g = pl->g;
  info->rank = g.rank;
  for (int k = 0; k < g.rank; ++k) {
    info->perm[k] = g.perm[k];
    info->size[k] = g.size[k];
    info->stride_a[k] = g.sa[k];
    info->stride_b[k] = g.sb[k];
    info->tile[k] = pl->tile[k];
  }
  info->vec_a = pl->va;
  info->vec_b = pl->vb;
  info->threads = pl->family == kFamilyRC ? kRcThreads : kThreads;
  info->slots_a = (pl->p.V / pl->va + kThreads - 1) / kThreads;
  info->slots_b = (pl->p.V / pl->vb + kThreads - 1) / kThreads;
  info->tiles = pl->p.ntiles;
  info->tile_volume = pl->p.V;
  info->lds_bytes = pl->lds_bytes;
  info->lds_pad = pl->lds_pad;
  info->nt_store = pl->nt_store;
  info->grid_order = pl->p.grid_order + 1;  // config numbering
  info->wg_per_cu = pl->wg_per_cu;
  info->elements = g.elements;
  const int sa = dtype_bytes(pl->type_a), sb = dtype_bytes(pl->type_b);
  info->bytes_algorithmic = g.elements * (int64_t)sa + g.elements * (int64_t)sb * (pl->beta_is_zero ? 1 : 2);
  info->type_a = pl->type_a;
  info->type_b = pl->type_b;
  info->beta_is_zero = pl->beta_is_zero;
  info->run_a = pl->run_a;
  info->run_b = pl->run_b;
  info->family = pl->family;
  info->tile_order = pl->family == kFamilyRC ? pl->tile_order : 1;
  if (pl->family == kFamilyRC) {
    info->slots_a = rc_elems(sa) / pl->va;
    info->slots_b = rc_elems(sb) / pl->vb;
    info->rows = pl->rc.R;
    info->cols = pl->rc.C;
  }
  info->lds_buffers = pl->family == kFamilyRC ? pl->rc.nbuf : 1;
  info->lane_fixed = pl->family == kFamilyRC ? pl->rc.fast : 0;
  if (pl->family == kFamilyRow) {
    info->slots_a = info->slots_b = kRowUnroll;
    info->tiles = g.size[0] > 0 ? g.elements / g.size[0] : 0;  // rows
  }
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_plan_destroy(ttc_plan_t pl) {
  if (!pl) return TTC_ERROR_INVALID_VALUE;
  delete pl->fallback;
  delete pl;
  return TTC_SUCCESS;
}
