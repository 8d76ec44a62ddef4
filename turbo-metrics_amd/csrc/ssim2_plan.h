// Host-side memory and work planning for one image pair (shared by the
// single-pair context and the batch engine).
#pragma once

#include <algorithm>
#include <vector>

#include "ssim2_internal.h"

namespace ssim2 {

constexpr int kMaxDim = 32768;
constexpr long long kMaxPixels = 1LL << 28;
// Device result record per pair: score, 108 sums, 12 dims.
constexpr int kResultDoubles = 1 + kSumsPerPair + 2 * kMaxScales;
// Planes per scale in the arena: 6 XYB + 15 row-pass + 9 cached-reference.
constexpr int kPlanesPerScale = 6 + 3 * kNumQ + 9;

inline bool size_supported(long long w, long long h) {
  return w >= 8 && h >= 8 && w <= kMaxDim && h <= kMaxDim && w * h <= kMaxPixels;
}

// Scale count and sizes of crates/ssimulacra2-cuda/examples/cpu.rs:358-368:
// the size test runs before downscaling, so the first scale below 8 px is
// still scored.
inline int scale_dims(int w, int h, int* sw, int* sh) {
  int n = 0;
  for (int s = 0; s < kMaxScales; ++s) {
    if (w < 8 || h < 8) break;
    if (s > 0) {
      w = (w + 1) / 2;
      h = (h + 1) / 2;
    }
    sw[s] = w;
    sh[s] = h;
    ++n;
  }
  return n;
}

struct PairPlan {
  PairDesc desc{};               // arena/partials/result pointers left null
  long long arena_elems = 0;     // floats
  int nparts = 0;                // partial-sum records
  std::vector<WorkItem> items;   // pyramid items, row items per sides, column items
  int off_pyr = 0, n_pyr = 0, off_cols = 0, n_cols = 0;
  int off_rows[3] = {0, 0, 0}, n_rows[3] = {0, 0, 0};  // indexed by Sides
};

// Work lists. A pair's items are the pyramid items, then the row items of
// each Sides value, then the column items; item k of a section is a function
// of the pair's scale sizes alone, so the host planner and a device-side
// planner (the no-allocation entry point) produce the same list.

// Pyramid: 32x32 scale-0 tiles, row-major.
__host__ __device__ inline int pyramid_items(const PairDesc& d) {
  return ((d.w + kPyrTile - 1) / kPyrTile) * ((d.h + kPyrTile - 1) / kPyrTile);
}
__host__ __device__ inline WorkItem pyramid_item(const PairDesc& d, int k, int pair) {
  const int tx = (d.w + kPyrTile - 1) / kPyrTile;
  return {pair, k % tx, k / tx, 0};
}

// Row pass: one wave per (scale, channel*5+quantity, band of kRowsPerWave
// rows), largest scale first; each Sides value runs its quantities only.
// The waves of one band read the same two XYB rows, so they are placed on
// one XCD: workgroups go to the 8 XCDs round-robin by id, and in each group
// of 8 bands the quantity waves of band j get ids = j (mod 8), so they share
// that XCD's L2.
__host__ __device__ inline int row_quantities(int sides) { return sides == kSidesBoth ? 5 : (sides == kSidesRef ? 2 : 3); }
// Quantity index (ref^2, dis^2, ref*dis, ref, dis) of the qi-th quantity a Sides value runs.
__host__ __device__ inline int row_quantity_id(int sides, int qi) {
  if (sides == kSidesBoth) return qi;
  if (sides == kSidesRef) return qi ? 3 : 0;
  return qi == 0 ? 1 : (qi == 1 ? 2 : 4);
}
__host__ __device__ inline int row_bands(const PairDesc& d) {
  int n = 0;
  for (int s = 0; s < d.nscales; ++s) n += 3 * ((d.sc[s].h + kRowsPerWave - 1) / kRowsPerWave);
  return n;
}
__host__ __device__ inline int row_items(const PairDesc& d, int sides) { return row_quantities(sides) * row_bands(d); }
// Bands are numbered scale by scale, channel by channel, band by band.
__host__ __device__ inline WorkItem row_item(const PairDesc& d, int sides, int k, int pair) {
  const int nq = row_quantities(sides), total = row_bands(d);
  const int g = k / (8 * nq);             // group of 8 bands (the last may hold fewer)
  const int n = total - 8 * g < 8 ? total - 8 * g : 8;
  const int t = k - 8 * nq * g, qi = t / n;
  int band = 8 * g + t % n;
  for (int s = 0; s < d.nscales; ++s) {
    const int nb = (d.sc[s].h + kRowsPerWave - 1) / kRowsPerWave;
    if (band < 3 * nb) return {pair, s, (band / nb) * kNumQ + row_quantity_id(sides, qi), band % nb};
    band -= 3 * nb;
  }
  return {pair, 0, 0, 0};
}

// Column pass + maps: (scale, channel, kColBlock-column strip).
__host__ __device__ inline int column_items(const PairDesc& d) {
  int n = 0;
  for (int s = 0; s < d.nscales; ++s) n += 3 * d.sc[s].colblocks;
  return n;
}
__host__ __device__ inline WorkItem column_item(const PairDesc& d, int k, int pair) {
  for (int s = 0; s < d.nscales; ++s) {
    const int nb = d.sc[s].colblocks;
    if (k < 3 * nb) return {pair, s, k / nb, k % nb};
    k -= 3 * nb;
  }
  return {pair, 0, 0, 0};
}

// Scale sizes, pitches and plane offsets of one pair whose planes start at
// element `arena_base` and whose partial records start at `part_base`;
// `planes` planes per scale (kPlanesPerScale, or 6 + 3 * kNumQ without the
// cached-reference planes). Returns the arena elements used.
inline long long layout_pair(int w, int h, long long arena_base, int part_base, int planes, PairDesc* out,
                             int* nparts) {
  PairDesc& d = *out;
  d = PairDesc{};
  d.w = w;
  d.h = h;
  int sw[kMaxScales], sh[kMaxScales];
  d.nscales = scale_dims(w, h, sw, sh);
  long long off = arena_base;
  int part = part_base;
  for (int s = 0; s < d.nscales; ++s) {
    ScaleDesc& S = d.sc[s];
    S.w = sw[s];
    S.h = sh[s];
    // >= 8 zero columns right of every row (the row scan reads them as the
    // zero inputs past the image), rounded to 256 bytes
    S.pitch = (sw[s] + 8 + 63) / 64 * 64;
    S.colblocks = (sw[s] + kColBlock - 1) / kColBlock;
    // + 4 zero rows under every plane: the column pass reads rows h .. h+3 as
    // the zero inputs below the image
    S.plane = (long long)S.pitch * (S.h + kPadRows);
    S.xyb = off;
    S.hrow = off + 6 * S.plane;
    S.cache = off + (6 + 3 * kNumQ) * S.plane;
    off += planes * S.plane;
    S.part = part;
    part += 3 * S.colblocks;
  }
  *nparts = part - part_base;
  return off - arena_base;
}

// Lay out one pair's planes starting at element `arena_base` and its partial
// records starting at `part_base`; `pair` is written into every work item.
inline void plan_pair(int w, int h, long long arena_base, int part_base, PairPlan* out, int pair = 0) {
  PairPlan& p = *out;
  p = PairPlan{};
  p.arena_elems = layout_pair(w, h, arena_base, part_base, kPlanesPerScale, &p.desc, &p.nparts);
  const PairDesc& d = p.desc;
  p.off_pyr = 0;
  p.n_pyr = pyramid_items(d);
  for (int k = 0; k < p.n_pyr; ++k) p.items.push_back(pyramid_item(d, k, pair));
  for (int sd = 0; sd < 3; ++sd) {
    p.off_rows[sd] = (int)p.items.size();
    p.n_rows[sd] = row_items(d, sd);
    for (int k = 0; k < p.n_rows[sd]; ++k) p.items.push_back(row_item(d, sd, k, pair));
  }
  p.off_cols = (int)p.items.size();
  p.n_cols = column_items(d);
  for (int k = 0; k < p.n_cols; ++k) p.items.push_back(column_item(d, k, pair));
}

}  // namespace ssim2
