// CPU check of the row x column kernels' index arithmetic, against the
// planner's own parameters (ttc_plan.cpp): for many geometries, type pairs,
// vector widths and tiles it replays, lane by lane and slot by slot, the
// general walk's LDS positions for full and partial tiles (check_general, on
// a sample) and the lane-fixed walk of full tiles that ttc_rc.hpp runs
// (setup tables, read walk, staging, write walk), checking for the latter that
//   - every LDS data index stays inside the data region and every table
//     index inside the table region the launch allocates;
//   - the read walk visits each element of the tile exactly once with its
//     true A offset, the write walk each element exactly once with its true
//     B offset, both at the same LDS position;
//   - every other slot's offset is out of the buffer range (>= 2^31), and no
//     such slot stages into an LDS position a valid slot reads.
// Prints one summary line; exit status 1 on the first violation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <vector>

#include "ttc_plan.h"

using namespace ttc;

static int g_fail = 0;
static long long big_a = 0;  // lane-fixed launches whose A offsets pass 2^30 (no padding slots on A's side)
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "FAIL %s: ", #cond);           \
      std::fprintf(stderr, __VA_ARGS__);                  \
      std::fprintf(stderr, "\n");                         \
      g_fail = 1;                                         \
      return false;                                       \
    }                                                     \
  } while (0)

// Offsets of tile row i and column j (the kernel's setup()).
static void tables(const RCParams& p, std::vector<long long>& rA, std::vector<long long>& rB,
                   std::vector<long long>& cA, std::vector<long long>& cB) {
  rA.assign(p.R, 0);
  rB.assign(p.R, 0);
  cA.assign(p.C, 0);
  cB.assign(p.C, 0);
  for (int i = 0; i < p.R; ++i) {
    long long l = i;
    for (int m = 0; m < p.na; ++m) {
      const long long d = l % p.ta[m];
      l /= p.ta[m];
      rA[i] += d * p.saA[m];
      rB[i] += d * p.sbA[m];
    }
  }
  for (int j = 0; j < p.C; ++j) {
    long long l = j;
    for (int m = 0; m < p.nb; ++m) {
      const long long d = l % p.tb[m];
      l /= p.tb[m];
      cA[j] += d * p.saB[m];
      cB[j] += d * p.sbB[m];
    }
  }
}

// The general walk (every tile of a launch without the lane-fixed walk,
// partial and empty tiles of one with it): replays finish(), stage() and
// walk_b() for a tile of rows x cols and checks that the staging positions
// of the read walk (element (i, j) at i + j * RP) and the positions the
// write walk reads agree element by element, stay inside the data region,
// and that padding slots stage only at or past the spare slot.
static bool check_general(const RCParams& p, int szA, int szB, int va, int vb, int rows, int cols) {
  const int NIA = rc_elems(szA) / va, NIB = rc_elems(szB) / vb;
  const uint32_t SA = (uint32_t)kRcThreads * va, SB = (uint32_t)kRcThreads * vb;
  const uint32_t RP = (uint32_t)p.RP, spare = (uint32_t)p.lds_spare;
  CHECK(rows <= p.R && cols <= p.C && rows % va == 0 && cols % vb == 0, "tile %d x %d", rows, cols);
  const long long n = (long long)rows * cols;
  std::vector<int> seen_r(n, 0), seen_w(n, 0);
  for (uint32_t tid = 0; tid < (uint32_t)kRcThreads; ++tid) {
    // finish()
    const int32_t limA = (int32_t)n - (int32_t)(tid * va), limB = (int32_t)n - (int32_t)(tid * vb);
    const uint32_t aj = tid * va / rows, ai = tid * va - aj * rows, adj = SA / rows, adi = SA - adj * rows;
    const uint32_t bi = tid * vb / cols, bj = tid * vb - bi * cols, bdi = SB / cols, bdj = SB - bdi * cols;
    // stage()
    uint32_t i = ai, j = aj, li = ai + aj * RP;
    const uint32_t dli = adi + adj * RP, wrap = RP - rows;
    for (int it = 0; it < NIA; ++it) {
      const bool ok = (int32_t)(it * SA) < limA;
      const uint32_t a0 = ok ? li : spare;
      for (int v = 0; v < va; ++v) {
        const uint32_t a = a0 + (uint32_t)v;
        CHECK(a < (uint32_t)p.lds_data_elems, "stage index %u of %d", a, p.lds_data_elems);
        if (ok) {
          CHECK(i + v < (uint32_t)rows && j < (uint32_t)cols, "slot (%u, %u) outside %d x %d", i + v, j, rows, cols);
          CHECK(a == (i + v) + j * RP, "stage position of (%u, %u)", i + v, j);
          CHECK(seen_r[(i + v) + (long long)j * rows]++ == 0, "staged twice (%u, %u)", i + v, j);
        } else {
          CHECK(a >= spare, "padding staged at %u", a);
        }
      }
      i += adi;
      j += adj;
      li += dli;
      if (i >= (uint32_t)rows) {
        i -= rows;
        ++j;
        li += wrap;
      }
    }
    // walk_b() and the reads of store()
    const uint32_t wdli = bdj * RP + bdi, wwrap = 1u - cols * RP;
    uint32_t wi = bi, wj = bj, wl = bi + bj * RP;
    for (int it = 0; it < NIB; ++it) {
      const bool ok = (int32_t)(it * SB) < limB;
      const uint32_t base = ok ? wl : 0u;
      for (int v = 0; v < vb; ++v) {
        const uint32_t a = base + (uint32_t)v * RP;
        CHECK(a < (uint32_t)p.lds_data_elems, "write-walk index %u of %d", a, p.lds_data_elems);
        if (ok) {
          CHECK(wi < (uint32_t)rows && wj + v < (uint32_t)cols, "slot (%u, %u) outside", wi, wj + v);
          CHECK(a == wi + (wj + v) * RP, "write-walk position of (%u, %u)", wi, wj + v);
          CHECK(seen_w[wi + (long long)(wj + v) * rows]++ == 0, "read twice (%u, %u)", wi, wj + v);
        }
      }
      wj += bdj;
      wi += bdi;
      wl += wdli;
      if (wj >= (uint32_t)cols) {
        wj -= cols;
        ++wi;
        wl += wwrap;
      }
    }
  }
  for (long long e = 0; e < n; ++e) CHECK(seen_r[e] == 1 && seen_w[e] == 1, "element %lld missed", e);
  return true;
}

static bool check_walk(const RCParams& p, int szA, int szB, int va, int vb, int lds_bytes, int lds_elem) {
  if (!p.fast) return true;
  const int NIA = rc_elems(szA) / va, NIB = rc_elems(szB) / vb;
  const long long SA = (long long)kRcThreads * va, SB = (long long)kRcThreads * vb;
  CHECK((1ll << p.Rw_shift) == p.Rw && (1ll << p.Cw_shift) == p.Cw, "Rw %d Cw %d", p.Rw, p.Cw);
  CHECK(p.Rw >= p.R && p.Cw >= p.C && p.RP == (p.R | 1), "R %d Rw %d C %d Cw %d RP %d", p.R, p.Rw, p.C, p.Cw, p.RP);
  CHECK(p.lds_spare == p.RP * p.C, "spare %d", p.lds_spare);
  CHECK(SA % p.Rw == 0 && SB % p.Cw == 0 && p.kA == SA / p.Rw && p.kB == SB / p.Cw, "kA %d kB %d", p.kA, p.kB);
  const long long ntab = (long long)p.Rw + p.R + p.C + p.Cw + (long long)p.kA * NIA + (long long)p.kB * NIB;
  CHECK(ntab == p.ntab, "ntab %d vs %lld", p.ntab, ntab);
  CHECK((long long)p.nbuf * p.lds_data_elems * lds_elem + 4ll * p.ntab == lds_bytes, "lds %d", lds_bytes);
  CHECK(p.lds_spare + 16 <= p.lds_data_elems, "spare %d data %d", p.lds_spare, p.lds_data_elems);
  std::vector<long long> rA, rB, cA, cB;
  tables(p, rA, rB, cA, cB);
  // a side whose walk has padding slots keeps its entries below kRcFastSpan
  // (sentinel sums stay out of range); entries of the other side are real
  // offsets, bounded by the span check below
  const long long capA = (long long)kRcThreads * rc_elems(szA), capB = (long long)kRcThreads * rc_elems(szB);
  const bool padA = p.Rw > p.R || (long long)p.Rw * p.C < capA, padB = p.Cw > p.C || (long long)p.R * p.Cw < capB;
  const long long limA = padA ? kRcFastSpan : 0x80000000ll - 64, limB = padB ? kRcFastSpan : 0x80000000ll - 64;
  for (long long x : rA) CHECK(x < limA, "rowA %lld", x);
  for (long long x : cA) CHECK(x < limA, "colA %lld", x);
  for (long long x : rB) CHECK(x < limB, "rowB %lld", x);
  for (long long x : cB) CHECK(x < limB, "colB %lld", x);
  if (!padA) big_a += rA.back() + cA.back() >= kRcFastSpan;
  // LDS table contents (32-bit, as the kernel holds them)
  const long long oRB = p.Rw, oCA = oRB + p.R, oCB = oCA + p.C, oCS = oCB + p.Cw, oRS = oCS + (long long)p.kA * NIA;
  std::vector<uint32_t> tab(p.ntab, 0xdeadbeefu);
  for (int i = 0; i < p.Rw; ++i) tab[i] = i < p.R ? (uint32_t)rA[i] : kRcTabOOB;
  for (int i = 0; i < p.R; ++i) tab[oRB + i] = (uint32_t)rB[i];
  for (int j = 0; j < p.C; ++j) tab[oCA + j] = (uint32_t)cA[j];
  for (int j = 0; j < p.Cw; ++j) tab[oCB + j] = j < p.C ? (uint32_t)cB[j] : kRcTabOOB;
  for (int e = 0; e < p.kA * NIA; ++e) {
    const int j = e / NIA + (e % NIA) * p.kA;
    tab[oCS + e] = j < p.C ? tab[oCA + j] : kRcTabOOB;
  }
  for (int e = 0; e < p.kB * NIB; ++e) {
    const int i = e / NIB + (e % NIB) * p.kB;
    tab[oRS + e] = i < p.R ? tab[oRB + i] : kRcTabOOB;
  }
  // the planner's copy of the same tables (what grouped launches upload)
  CHECK(p.nia == NIA && p.nib == NIB, "nia %d nib %d", p.nia, p.nib);
  std::vector<int32_t> host(p.ntab, 0x5a5a5a5a);
  rc_fill_tables(p, host.data());
  for (int e = 0; e < p.ntab; ++e) CHECK((uint32_t)host[e] == tab[e], "rc_fill_tables entry %d", e);
  const long long n = (long long)p.R * p.C;
  std::vector<int> seen_r(n, 0), seen_w(n, 0);
  std::vector<long long> lds_r(n, -1);
  std::set<long long> junk;  // LDS positions staged by padding slots
  for (uint32_t tid = 0; tid < (uint32_t)kRcThreads; ++tid) {
    const uint32_t la = tid * va;
    const uint32_t ai = la & (uint32_t)(p.Rw - 1), jl = la >> p.Rw_shift;
    CHECK(jl < (uint32_t)p.kA, "jl %u", jl);
    const uint32_t row = tab[ai];
    const long long colS = oCS + (long long)jl * NIA;
    const bool row_ok = ai < (uint32_t)p.R;
    const uint32_t lds0 = row_ok ? ai + jl * (uint32_t)p.RP : (uint32_t)p.lds_spare;
    const uint32_t step = row_ok ? (uint32_t)p.kA * (uint32_t)p.RP : 0u;
    for (int it = 0; it < NIA; ++it) {
      CHECK(colS + it < oRS, "colS index");
      const uint32_t off = row + tab[colS + it];
      const uint32_t li = std::min(lds0 + (uint32_t)it * step, (uint32_t)p.lds_spare);
      for (int v = 0; v < va; ++v) {
        const long long i = ai + v, j = jl + (long long)it * p.kA;
        CHECK(li + v < (uint32_t)p.lds_data_elems, "stage index %u of %d", li + v, p.lds_data_elems);
        if (i < p.R && j < p.C) {
          CHECK(li + v == i + j * p.RP, "stage position");
          CHECK(off < 0x80000000u, "valid read offset %u", off);
          CHECK((long long)off + (long long)v * szA == rA[i] + cA[j], "A offset of (%lld, %lld)", i, j);
          CHECK(seen_r[i + j * p.R]++ == 0, "read twice (%lld, %lld)", i, j);
          lds_r[i + j * p.R] = li + v;
        } else {
          CHECK(off >= 0x80000000u && (uint64_t)off + 64 <= 0x100000000ull, "padding read offset %u", off);
          CHECK(li >= (uint32_t)p.lds_spare, "padding staged at %u below the spare slot", li);
          junk.insert(li + v);
        }
      }
    }
  }
  for (long long e = 0; e < n; ++e) CHECK(seen_r[e] == 1, "read walk misses element %lld", e);
  for (uint32_t tid = 0; tid < (uint32_t)kRcThreads; ++tid) {
    const uint32_t lb = tid * vb;
    const uint32_t bj = lb & (uint32_t)(p.Cw - 1), il = lb >> p.Cw_shift;
    CHECK(il < (uint32_t)p.kB, "il %u", il);
    const uint32_t col = tab[oCB + bj];
    const long long rowS = oRS + (long long)il * NIB;
    const bool col_ok = bj < (uint32_t)p.C;
    const uint32_t lds0 = col_ok ? il + bj * (uint32_t)p.RP : 0u, step = col_ok ? (uint32_t)p.kB : 0u;
    for (int it = 0; it < NIB; ++it) {
      CHECK(rowS + it < p.ntab, "rowS index");
      const uint32_t off = tab[rowS + it] + col;
      const uint32_t li = lds0 + (uint32_t)it * step;
      for (int v = 0; v < vb; ++v) {
        const long long i = il + (long long)it * p.kB, j = bj + v;
        const uint32_t lv = li + (uint32_t)v * (uint32_t)p.RP;
        CHECK(lv < (uint32_t)p.lds_data_elems, "write-walk index %u of %d", lv, p.lds_data_elems);
        if (i < p.R && j < p.C) {
          CHECK(off < 0x80000000u, "valid write offset %u", off);
          CHECK((long long)off + (long long)v * szB == rB[i] + cB[j], "B offset of (%lld, %lld)", i, j);
          CHECK(seen_w[i + j * p.R]++ == 0, "written twice (%lld, %lld)", i, j);
          CHECK(lds_r[i + j * p.R] == (long long)lv, "LDS position of (%lld, %lld)", i, j);
          CHECK(junk.count(lv) == 0, "padding staged over (%lld, %lld)", i, j);
        } else {
          CHECK(off >= 0x80000000u && (uint64_t)off + 64 <= 0x100000000ull, "padding write offset %u", off);
        }
      }
    }
  }
  for (long long e = 0; e < n; ++e) CHECK(seen_w[e] == 1, "write walk misses element %lld", e);
  return true;
}

int main(int argc, char** argv) {
  const int ncases = argc > 1 ? std::atoi(argv[1]) : 400;
  std::mt19937_64 rng(12345);
  const int pairs[][2] = {{TTC_S, TTC_S},    {TTC_D, TTC_D},       {TTC_C, TTC_C},       {TTC_Z, TTC_Z},
                          {TTC_BF16, TTC_BF16}, {TTC_S, TTC_BF16}, {TTC_D, TTC_S},       {TTC_S, TTC_D},
                          {TTC_C, TTC_Z},    {TTC_BF16, TTC_E4M3}, {TTC_S, TTC_E5M2},    {TTC_F16, TTC_S}};
  long long checked = 0, fast = 0, general = 0;
  for (int c = 0; c < ncases && !g_fail; ++c) {
    const int r = 2 + (int)(rng() % 5);
    int perm[8], size[8], lda[8], ldb[8];
    for (int k = 0; k < r; ++k) perm[k] = k;
    std::shuffle(perm, perm + r, rng);
    long long vol = 1;
    for (int k = 0; k < r; ++k) {
      size[k] = 1 + (int)(rng() % (k < 2 ? 300 : 40));
      vol *= size[k];
    }
    // large enough for the lane-fixed walk's tile-count threshold
    while (vol < (1ll << 24)) {
      const int k = (int)(rng() % r);
      vol = vol / size[k] * size[k] * 2;
      size[k] *= 2;
    }
    while (vol > (1ll << 26)) {
      int big = 0;
      for (int k = 1; k < r; ++k)
        if (size[k] > size[big]) big = k;
      vol /= size[big];
      size[big] = (size[big] + 1) / 2;
      vol *= size[big];
    }
    const bool pad = rng() % 4 == 0;
    for (int k = 0; k < r; ++k) {
      lda[k] = size[k] + (pad ? (int)(rng() % 3) : 0);
      ldb[k] = size[perm[k]] + (pad ? (int)(rng() % 3) : 0);
    }
    // one case in eight: a padded outer dimension of A large enough that
    // tile spans pass 2^30 bytes (sub-tensor of a much larger array)
    if (rng() % 8 == 0 && r >= 3) lda[r - 2] = size[r - 2] * (40 + (int)(rng() % 40));
    Geometry g;
    if (make_geometry(r, perm, size, lda, ldb, g) != TTC_SUCCESS || g.elements == 0 || g.rank < 2 || g.perm[0] == 0)
      continue;
    for (const auto& pr : pairs) {
      const int ta = pr[0], tb = pr[1];
      if (!pair_supported(ta, tb)) continue;
      const int wa = widest_vec_a(g, ta), wb = widest_vec_b(g, tb);
      const int vas[] = {1, wa}, vbs[] = {1, wb};
      for (int va : vas)
        for (int vb : vbs) {
          std::vector<std::vector<int>> tiles;
          int t[kMaxDim];
          for (int ra : {0, 16, 64, 256})
            for (int rb : {0, 16, 64, 256})
              if (choose_rc_tile(g, ta, tb, va, vb, ra, rb, t, nullptr)) tiles.emplace_back(t, t + kMaxDim);
          for (const auto& tile : tiles)
            for (int nbuf = 1; nbuf <= 2; ++nbuf) {
              RCParams p;
              int lds = 0;
              if (build_rc_params(g, tile.data(), ta, tb, va, vb, 0, 2, p, lds, nbuf) != TTC_SUCCESS) continue;
              ++checked;
              fast += p.fast != 0;
              bool ok = check_walk(p, dtype_bytes(ta), dtype_bytes(tb), va, vb, lds, lds_elem_bytes(ta, tb));
              if (checked % 8 == 0 || va * vb > 2) {  // a sample of the narrow launches
                ++general;
                for (int rows : {p.R, p.Rl})
                  for (int cols : {p.C, p.Cl})
                    ok = ok && check_general(p, dtype_bytes(ta), dtype_bytes(tb), va, vb, rows, cols);
              }
              if (!ok) {
                std::fprintf(stderr, "case %d types %d->%d va %d vb %d tile", c, ta, tb, va, vb);
                for (int k = 0; k < g.rank; ++k) std::fprintf(stderr, " %d", tile[k]);
                std::fprintf(stderr, " R %d C %d nbuf %d\n", p.R, p.C, nbuf);
                return 1;
              }
            }
        }
    }
  }
  // fixed cases whose lane-fixed walk runs with A offsets past 2^30 (no
  // padding slots on A's side): fp32 -> bf16 and bf16 -> bf16 of a
  // (256, 192, 128, 96) tensor, perm (3, 1, 0, 2), whose dimension 3 has a
  // 25 MB (12.6 MB) stride in A
  for (const int tb : {TTC_BF16}) {
    for (const int ta : {TTC_S, TTC_BF16}) {
      const int perm[4] = {3, 1, 0, 2}, size[4] = {256, 192, 128, 96};
      const int lda[4] = {256, 192, 128, 96}, ldb[4] = {96, 192, 256, 128};
      Geometry g;
      if (make_geometry(4, perm, size, lda, ldb, g) != TTC_SUCCESS) return 1;
      const int tiles[][4] = {{128, 1, 1, 64}, {256, 1, 1, 32}, {64, 1, 1, 64}};
      for (const auto& t4 : tiles)
        for (int nbuf = 1; nbuf <= 2 && !g_fail; ++nbuf) {
          int tile[kMaxDim] = {1, 1, 1, 1, 1, 1, 1, 1};
          for (int k = 0; k < 4; ++k) tile[k] = t4[k];
          const int va = widest_vec_a(g, ta), vb = widest_vec_b(g, tb);
          RCParams p;
          int lds = 0;
          if (build_rc_params(g, tile, ta, tb, va, vb, 0, 2, p, lds, nbuf) != TTC_SUCCESS) continue;
          ++checked;
          fast += p.fast != 0;
          bool ok = check_walk(p, dtype_bytes(ta), dtype_bytes(tb), va, vb, lds, lds_elem_bytes(ta, tb));
          for (int rows : {p.R, p.Rl})
            for (int cols : {p.C, p.Cl})
              ok = ok && check_general(p, dtype_bytes(ta), dtype_bytes(tb), va, vb, rows, cols);
          if (!ok) {
            std::fprintf(stderr, "fixed case types %d->%d tile %d %d %d %d nbuf %d\n", ta, tb, t4[0], t4[1], t4[2],
                         t4[3], nbuf);
            return 1;
          }
        }
    }
  }
  std::printf("rc walk check: %lld launches, %lld with the lane-fixed walk, %lld general walks replayed, "
              "%lld lane-fixed with A offsets past 2^30: %s\n", checked, fast, general, big_a, g_fail ? "FAIL" : "ok");
  return g_fail;
}
