// Grouped MFMA GEMM: precision dispatch and host-side tile planning (the kernels are in
// gemm_kernels.hpp, instantiated by gemm_fp64.hip, gemm_fp32.hip, gemm_fp32x3.hip and
// gemm_mix.hip).
#define CXA_GEMM_TU main
#include "gemm_kernels.hpp"

namespace cxa {

int launch_grouped_gemm(const GemmLaunch &g, hipStream_t s) {
  if (g.outs && g.dot_part)
    return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: no dot epilogue on mix launches");
  if (g.outs) return launch_mix(g, s);
  switch (g.prec) {
    case CXA_PREC_FP64: return launch_gemm_fp64(g, s);
    case CXA_PREC_FP32: return launch_gemm_fp32(g, s);
    case CXA_PREC_FP32X3: return launch_gemm_fp32x3(g, s);
    default: break;
  }
  return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: unsupported precision %d", g.prec);
}

int default_kind(int M, int N) {
  if (M <= 16 && N <= 16) return TK_1x1;
  if (M <= 16) return TK_1x2;
  if (N <= 16) return TK_2x1;
  return TK_2x2;
}

int make_tiles(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
               std::vector<GTile> &tiles, int *per_xcd, int kind_override, bool allow_rows) {
  struct Item {
    GTile t;
    int64_t cost;
  };
  // the row class is taken when every block is narrow (all of a launch's tiles share a class)
  bool all_narrow = allow_rows;
  for (const GBlk &b : blks)
    if (b.M > 0 && b.N > kRowsN) all_narrow = false;
  int cls = TC_MIXED;
  if (kind_override == TK_4x2 && allow_rows) cls = TC_BIG;  // allow_rows: real dtypes only
  else if (kind_override == TK_ROWS ? all_narrow : (kind_override < 0 && all_narrow))
    cls = TC_ROWS;
  // blocks in operand order: first segment's A offset, then its B offset (the scratch
  // vectors keep their capacity between calls: plans of tens of thousands of tiles are made
  // per DMRG bond, and fresh pages cost more than the planning itself)
  thread_local std::vector<int> order;
  order.clear();
  for (size_t bi = 0; bi < blks.size(); ++bi)
    if (blks[bi].M > 0 && blks[bi].N > 0) order.push_back((int)bi);
  auto key = [&](int bi) {
    const GBlk &b = blks[bi];
    return b.nseg > 0 ? std::make_pair(segs[b.seg0].a_off, segs[b.seg0].b_off)
                      : std::make_pair(INT64_MAX, INT64_MAX);
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
  thread_local std::vector<Item> items;
  items.clear();
  for (int bi : order) {
    const GBlk &b = blks[bi];
    int kind;
    if (cls == TC_ROWS)
      kind = TK_ROWS;
    else if (cls == TC_BIG)
      kind = TK_4x2;
    else if (kind_override >= 0 && kind_override <= TK_1x2)
      kind = kind_override;
    else
      kind = default_kind(b.M, b.N);
    const int tr = tile_rows(kind), tc = kind == TK_ROWS ? b.N : tile_cols(kind);
    int64_t ksum = 0;
    for (int s = 0; s < b.nseg; ++s) ksum += segs[b.seg0 + s].K;
    for (int m0 = 0; m0 < b.M; m0 += tr)
      for (int n0 = 0; n0 < b.N; n0 += tc)
        items.push_back({GTile{(int32_t)bi, m0, n0, kind}, (ksum + 16) * tr * tc});
  }
  // kXcds contiguous ranges of equal tile count (at these sizes a tile's time is mostly its
  // fixed chain of dependent loads, so equal counts balance better than equal flops, and the
  // ranges need no padding beyond the last few tiles)
  // Within a range the tiles go in decreasing cost, ties in list order.  A block's tiles share
  // one cost and sit together, so the range is sorted as runs of equal cost (one per block
  // piece): the same order as sorting the tiles, for hundreds of runs instead of thousands of
  // tiles.
  const size_t per = (items.size() + kXcds - 1) / kXcds;
  const size_t P = per ? (per + kWaves - 1) / kWaves * kWaves : 0;
  tiles.assign(kXcds * P, GTile{-1, 0, 0, 0});
  struct Run {
    size_t at, len;
    int64_t cost;
  };
  thread_local std::vector<Run> runs;
  for (int x = 0; x < kXcds; ++x) {
    const size_t lo = std::min(items.size(), x * per), hi = std::min(items.size(), lo + per);
    runs.clear();
    for (size_t i = lo; i < hi; ++i) {
      if (!runs.empty() && runs.back().cost == items[i].cost) ++runs.back().len;
      else runs.push_back({i, 1, items[i].cost});
    }
    std::stable_sort(runs.begin(), runs.end(),
                     [](const Run &a, const Run &b) { return a.cost > b.cost; });
    size_t o = x * P;
    for (const Run &r : runs)
      for (size_t i = 0; i < r.len; ++i) tiles[o++] = items[r.at + i].t;
  }
  *per_xcd = (int)P;
  return cls;
}

bool prefer_splitk(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
                   const std::vector<GTile> &tiles, int tile_class) {
  if (tiles.empty() || tile_class == TC_BIG) return false;
  // a tile's work is its block's: summed once per block, then once per tile
  thread_local std::vector<int64_t> bwork;
  bwork.assign(blks.size(), -1);
  int64_t work = 0, n = 0;
  for (const GTile &t : tiles) {
    if (t.blk < 0) continue;
    int64_t &w = bwork[t.blk];
    if (w < 0) {
      const GBlk &b = blks[t.blk];
      w = 0;
      for (int s = 0; s < b.nseg; ++s)
        w += tile_class == TC_ROWS ? segs[b.seg0 + s].K : (segs[b.seg0 + s].K + 15) / 16;
    }
    ++n;
    work += w;
  }
  if (n == 0) return false;
  // few tiles (under two per SIMD of a 256-CU device) that each walk a long K range
  const double avg = (double)work / (double)n;
  return n < 2048 && avg >= (tile_class == TC_ROWS ? 64.0 : 8.0);
}

}  // namespace cxa
