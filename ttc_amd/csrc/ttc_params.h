// Launch parameters shared by the host planner (ttc_plan.cpp) and the device
// kernels (ttc_kernels.hpp).  Plain C++ so the planner builds without HIP.
//
// Geometry is that of the FUSED problem: rank r <= TTC_MAX_DIM, A-dimension k
// has extent n[k], element stride sa in A and sb in B.  A tile is the box
// t[0] x ... x t[r-1] (t[k] == 1 for dimensions the grid walks over, t[k] ==
// n[k] for dimensions a tile covers fully, 1 < t[k] < n[k] for "split"
// dimensions).  Tiles are numbered in A order (dimension 0 fastest) over the
// tile grid g[k] = ceil(n[k] / t[k]).
#pragma once
#include <stdint.h>

namespace ttc {

constexpr int kMaxDim = 8;     // fused rank limit
constexpr int kThreads = 256;  // work-group size (4 wave64)
constexpr int kRcThreads = 256;  // work-group size of the row x column kernels

// Vectors a thread owns per tile side for vectors of `vec` elements of
// `bytes` bytes: as many as fit 64 bytes of data (16 VGPRs), between 4 and
// 16.  A tile holds at most kThreads * slots_for(vec, bytes) * vec elements
// on each side; planner and kernels use this one rule.
constexpr int slots_for(int vec, int bytes) {
  return 64 / (vec * bytes) > 16 ? 16 : (64 / (vec * bytes) < 4 ? 4 : 64 / (vec * bytes));
}

// n / d for n < 2^31 via one multiply-high (round-up method): with
// s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1,
// n / d = (mulhi(n, m) + n) >> s.
struct FastDiv {
  uint32_t d, m, s, pad_;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{};
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  f.s = s;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
  return f;
}

struct TileParams {
  int32_t rank;
  int32_t V;            // elements in a full tile (prod t)
  int32_t t[kMaxDim];   // tile extents
  int32_t n[kMaxDim];   // tensor extents
  int32_t perm[kMaxDim];       // B position q holds A dimension perm[q]
  int32_t sa[kMaxDim];         // A element stride of dim k, used within a tile (int32 checked on host)
  int32_t sb[kMaxDim];         // B element stride of dim k, used within a tile
  int32_t ls[kMaxDim];         // LDS element stride of dim k
  int64_t tsa[kMaxDim];        // t[k] * A stride: tile-to-tile step
  int64_t tsb[kMaxDim];        // t[k] * B stride
  FastDiv tdiv[kMaxDim];       // division by t[k]
  FastDiv gdiv[kMaxDim];       // division by g[k] (tiles along dim k)
  // The same quantities in B order (position q describes A dimension
  // perm[q]), so that device code never indexes this struct with a value
  // known only at run time (that would force a copy into scratch memory).
  FastDiv tdivB[kMaxDim];
  int32_t tB[kMaxDim];
  int32_t sbB[kMaxDim];
  int32_t lsB[kMaxDim];
  int32_t splitB0, splitB1;    // B positions of split0 / split1 (-1 = none)
  // Digit of a split dimension in a tile-linear index l (A or B order):
  // (l / P) mod t, P = product of the tile extents before it in that order.
  FastDiv pA0, pA1, pB0, pB1;  // division by P
  FastDiv ts0, ts1;            // division by t[split]
  uint32_t ntiles;
  int32_t split0, split1;      // dims with partial tiles (-1 = none)
  int32_t tsplit0, tsplit1;    // t[split] (0 when none)
  int32_t lds_stride_b;        // ls[perm[0]]: LDS distance of consecutive B-order elements
  int32_t lds_elems;           // LDS elements used by one tile (with padding)
  int32_t grid_order;          // tile_range order: 0 strided, 1 contiguous, 2 XCD-grouped, 3 XCD shares
};

// Row x column tile (ttc_rc.hpp): the tile's dimensions split into an A
// group (A's leading dimensions, the rows) and a disjoint B group (B's
// leading dimensions, the columns); only the outermost dimension of each
// group may have a partial last tile.
struct RCParams {
  int32_t R, C;          // rows (A-group volume) and columns (B-group volume) of a full tile
  int32_t Rl, Cl;        // rows / columns of the last tile along the split dims (== R / C if none)
  FastDiv dR, dRl, dC, dCl;
  int32_t RP;            // LDS pitch of a column (R | 1)
  int32_t rowA_dense;    // rowA[i] == i * sizeof(TA): the rows are one contiguous run of A
  int32_t colB_dense;    // colB[j] == j * sizeof(TB): the columns are one contiguous run of B
  int32_t na, nb;        // dimensions in the A group / B group
  int32_t ta[kMaxDim];   // tile extents of the A group, A order
  FastDiv dta[kMaxDim];
  int32_t saA[kMaxDim], sbA[kMaxDim];  // A and B byte strides of the A-group dims
  int32_t tb[kMaxDim];   // tile extents of the B group, B order
  FastDiv dtb[kMaxDim];
  int32_t saB[kMaxDim], sbB[kMaxDim];  // A and B byte strides of the B-group dims
  // Tile grid: only the dimensions with more than one tile, in walk order
  // (A order or B order); position g describes one such dimension.
  int32_t grank;
  int32_t n[kMaxDim], t[kMaxDim];
  FastDiv gdiv[kMaxDim];   // division by the number of tiles along the dimension
  int64_t tsa[kMaxDim], tsb[kMaxDim];  // element offset of the next tile along it, in A and B
  int32_t splitA, splitB;  // grid positions of the row group's / column group's partial dimension (-1 none)
  // Grouped walk: one dimension is split into an inner position 0 of
  // `group` tiles and an outer position `gouter` (-1: no split); `gcount`
  // is its number of real tiles, ids past it are empty padding.
  int32_t gouter;
  uint32_t group, gcount;
  int32_t nbuf;            // LDS tile buffers: 2 lets a tile's staging overlap the previous tile's reads
  uint32_t ntiles;         // grid size (with the padding of a grouped walk)
  int32_t grid_order;
  int32_t lds_data_elems;  // tile data region in LDS elements (both walks' reach, a spare vector slot, 16-byte rounding)
  int32_t lds_spare;       // RP * C: where lanes past a partial tile stage their (unused) values
  // Lane-fixed walk of full tiles (fast != 0).  The read walk runs over Rw
  // rows (R rounded up to a power of two that divides the walk step 256 *
  // VA), so a lane keeps one row i for the whole tile and its slot `it`
  // takes column j0 + it * kA; the write walk mirrors it over Cw columns
  // with row step kB.  Positions past R or C read offset tables holding
  // kRcTabOOB, which sends the access out of the buffer range, and stage
  // into the spare slot (lds_spare), so no per-slot bounds test remains.
  // Partial and empty tiles take the general walk.
  int32_t fast;
  int32_t Rw, Cw;          // walk extents (powers of two >= R, C); R, C when fast == 0
  int32_t Rw_shift, Cw_shift;
  int32_t kA, kB;          // column step of the read walk, row step of the write walk
  int32_t nia, nib;        // slots per lane of the read and write walks (rc_elems / vector width)
  int32_t ntab;            // offset-table entries in LDS (layout below)
  int32_t pad_;
};

// LDS offset tables of a row x column launch, int32 byte offsets after the
// tile data: rowA[Rw] rowB[R] colA[C] colB[Cw], then (fast walk only)
// colS[kA * NIA] and rowS[kB * NIB], the read walk's column offsets and the
// write walk's row offsets in slot order (colS[jl * NIA + it] = colA[jl + it
// * kA]); entries past R or C hold kRcTabOOB.  On a side whose walk has such
// slots a table entry is below 2^30 - 64, so the sum of two entries is
// either a real offset (< 2^31) or, if either is kRcTabOOB, in [2^31, 2^32):
// past the buffer range 2^31.  A side without them holds real offsets only,
// whose sums stay below 2^31 (the tile's span).
constexpr uint32_t kRcTabOOB = 0xC0000000u;
constexpr long long kRcFastSpan = (1ll << 30) - 64;

// Row copy for perm[0] == 0, where A and B share the stride-1 dimension:
// a row is one position of the other dimensions, the rows are walked in B
// order, and each lane moves one vector of V elements of dimension 0.
constexpr int kRowUnroll = 4;  // vectors in flight per lane in the row-copy kernel

struct RowParams {
  uint32_t total;         // rows * nv (< 2^31)
  uint32_t nv;            // vectors per row: extent of dimension 0 / V
  FastDiv dnv;
  int32_t rank;           // row dimensions, in B order
  int32_t pad_;
  FastDiv dn[kMaxDim];    // their extents
  int64_t sa[kMaxDim];    // A byte strides
  int64_t sb[kMaxDim];    // B byte strides
};

// One row x column tile of a grouped launch: element offsets of its origin
// in A and B; flags bit 0: last (partial) tile along the row group, bit 1:
// along the column group.
struct RCTile {
  int64_t a, b;
  uint32_t flags, pad_;
};

// Elements per thread per tile side of the row x column kernels: as many as
// fit 128 bytes, at most 32.
constexpr int rc_elems(int bytes) { return 128 / bytes > 32 ? 32 : 128 / bytes; }

}  // namespace ttc
