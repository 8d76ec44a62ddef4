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

// Lay out one pair's planes starting at element `arena_base` and its partial
// records starting at `part_base`; `pair` is written into every work item.
inline void plan_pair(int w, int h, long long arena_base, int part_base, PairPlan* out, int pair = 0) {
  PairPlan& p = *out;
  p = PairPlan{};
  PairDesc& d = p.desc;
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
    off += kPlanesPerScale * S.plane;
    S.part = part;
    part += 3 * S.colblocks;
  }
  p.arena_elems = off - arena_base;
  p.nparts = part - part_base;
  // Pyramid: 32x32 scale-0 tiles.
  p.off_pyr = 0;
  for (int ty = 0; ty < (h + kPyrTile - 1) / kPyrTile; ++ty)
    for (int tx = 0; tx < (w + kPyrTile - 1) / kPyrTile; ++tx) p.items.push_back({pair, tx, ty, 0});
  p.n_pyr = (int)p.items.size();
  // Row pass: one wave per (scale, channel*5+quantity, band of kRowsPerWave
  // rows), largest scale first; one list per Sides (its quantities only).
  // The waves of one band read the same two XYB rows, so they are placed on
  // one XCD: workgroups go to the 8 XCDs round-robin by id, and in each group
  // of 8 bands the quantity waves of band j get ids = j (mod 8), so they
  // share that XCD's L2.
  static const int kQ[3][5] = {{0, 1, 2, 3, 4}, {0, 3, -1, -1, -1}, {1, 2, 4, -1, -1}};
  for (int sd = 0; sd < 3; ++sd) {
    p.off_rows[sd] = (int)p.items.size();
    int nq = 0;
    while (nq < 5 && kQ[sd][nq] >= 0) ++nq;
    std::vector<WorkItem> bands;  // (scale, channel, band), largest scale first
    for (int s = 0; s < d.nscales; ++s)
      for (int c = 0; c < 3; ++c)
        for (int b = 0; b < (sh[s] + kRowsPerWave - 1) / kRowsPerWave; ++b) bands.push_back({pair, s, c, b});
    for (size_t g = 0; g < bands.size(); g += 8) {
      const size_t n = std::min<size_t>(8, bands.size() - g);
      for (int qi = 0; qi < nq; ++qi)
        for (size_t j = 0; j < n; ++j) {
          const WorkItem& B = bands[g + j];
          p.items.push_back({pair, B.a, B.b * kNumQ + kQ[sd][qi], B.c});
        }
    }
    p.n_rows[sd] = (int)p.items.size() - p.off_rows[sd];
  }
  // Column pass + maps: (scale, channel, 16-column strip).
  p.off_cols = (int)p.items.size();
  for (int s = 0; s < d.nscales; ++s)
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < d.sc[s].colblocks; ++b) p.items.push_back({pair, s, c, b});
  p.n_cols = (int)p.items.size() - p.off_cols;
}

}  // namespace ssim2
