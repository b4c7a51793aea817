// Internal interface of the grouped MFMA GEMM (gemm.hip) shared by every caller in the
// library: dense matmul, batched GEMM, tensordot, block-sparse contraction and the
// effective-Hamiltonian pipeline.
//
// Work is described as output blocks (GBlk), each the sum of one or more segments (GSeg):
//   C_blk[m, n] = alpha * sum_seg sum_k A_seg[m, k] * B_seg[k, n]  (+ beta * C_blk[m, n])
// with every operand a strided matrix view (any row/column strides), so permutes that keep an
// operand a strided matrix are fused into the loads.  Output blocks are cut into wave tiles
// (GTile); one wave owns one tile and loops over all of its segments, so accumulation order
// is fixed and no atomics are used.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cxa {

struct GSeg {
  int64_t a_off, b_off;  // element offsets of A[0,0] / B[0,0] from the A / B base pointers
  int64_t sam, sak;      // element strides of A along m and k
  int64_t sbk, sbn;      // element strides of B along k and n
  int32_t K;
  int32_t pad;
};

struct GBlk {
  int64_t c_off;  // element offset of C[0,0] from the C base pointer
  int64_t ldc;    // element stride of C along m (C is unit-stride along n)
  double alpha, beta;
  int32_t M, N;
  int32_t seg0, nseg;
};

// tile shapes: MFMA wave tiles (rows x cols) and TK_ROWS, 64 rows of a block with N <= 8
// streamed one row per lane
enum TileKind : int32_t {
  TK_1x1 = 0, TK_2x2 = 1, TK_2x1 = 2, TK_1x2 = 3, TK_4x2 = 4, TK_ROWS = 5, TK_COUNT
};
inline int tile_rows(int kind) {
  static const int r[TK_COUNT] = {16, 32, 32, 16, 64, 64};
  return r[kind];
}
inline int tile_cols(int kind) {
  static const int c[TK_COUNT] = {16, 32, 16, 32, 32, 8};
  return c[kind];
}

struct GTile {
  int32_t blk, m0, n0, kind;
};

struct GemmLaunch {
  int prec = 0;                     // CXA_PREC_*
  int a_dt = 0, b_dt = 0, c_dt = 0; // CXA dtype ids of the A / B / C buffers
  const void *A = nullptr, *B = nullptr;
  void *C = nullptr;
  const GTile *tiles = nullptr;  // device arrays
  const GBlk *blks = nullptr;
  const GSeg *segs = nullptr;
  int ntiles = 0;
  int grid = 0;  // workgroups; 0 = pick from ntiles
  bool splitk = false;  // one workgroup per tile, K chunks interleaved over its waves
  int tile_class = 0;   // the class make_tiles returned
  // single-block mode: one block with one segment passed by value, no device task arrays
  bool single = false;
  GBlk blk0{};
  GSeg seg0{};
  int kind0 = -1;
};

// Enqueue one grouped launch.  Returns a CXA status.
int launch_grouped_gemm(const GemmLaunch &g, hipStream_t s);

// Host-side planning helper: cut every block into tiles of the chosen kind and order tiles by
// decreasing work (longest K first) so the persistent loop balances.  All tiles of a launch
// share one class (its kernel instantiation), returned: 0 = MFMA tiles up to 32x32, 1 = row
// streaming (taken when every block has N <= 8 and allow_rows), 2 = 64x32 MFMA tiles (only on
// request).  allow_rows is false for complex dtypes.
int make_tiles(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
               std::vector<GTile> &tiles, int kind_override = -1, bool allow_rows = true);

// Whether a tile list is better run with workgroup split-K (few tiles, long K).
bool prefer_splitk(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
                   const std::vector<GTile> &tiles, int tile_class);

// Tile kind used for a block of M x N when no tuned choice is given.
int default_kind(int M, int N);

}  // namespace cxa
