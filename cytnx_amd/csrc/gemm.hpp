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
  int32_t out0, nout;  // mix launches only: the item's outputs (MixOut)
  double alpha_im, beta_im;  // imaginary parts of alpha / beta (complex dtypes only)
};

// Mix launches (a grouped GEMM with a linear-combination epilogue).  An item is a GBlk whose
// segments are separate products ("slices") X_s = A_s B_s of one M x N shape; the item writes
// nout output matrices  C_j[m, n] = sum_s coef_j[s] * X_s[m, n]  (C_j at c_off + m*ldc + n).
// The slices are never stored: one workgroup computes all of an item's slices for its tile.
constexpr int kMixSlices = 8;
struct MixOut {
  int64_t c_off, ldc;
  double coef[kMixSlices];
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
  int ntiles = 0;   // = kXcds * per_xcd
  int per_xcd = 0;  // tiles per XCD range (make_tiles)
  int grid = 0;     // workgroups (a multiple of kXcds); 0 = pick from per_xcd
  // 0: one wave per tile; 4 or 8: one workgroup of that many waves per tile, K chunks
  // interleaved over its waves (mix launches: nonzero = four-wave workgroups)
  int splitk = 0;
  int tile_class = 0;   // the class make_tiles returned
  const MixOut *outs = nullptr;  // non-null: mix launch (blks are items, segs are slices)
  int conj = 0;  // complex dtypes: bit 0 conjugates A, bit 1 conjugates B
  // B is k4-packed (real dtypes, B = FP32 in FP32 mode or FP64 otherwise): element (k, n) of a
  // segment sits at b_off + (k >> 2) * sbk + n * 4 + (k & 3), b_off and sbk are multiples of
  // 4, sbn is 4, and every k-quad of a segment is whole (zero past K), so a lane reads the 4
  // consecutive k of one column with one vector load
  bool b_k4 = false;
  // A is k4-packed (needs b_k4): element (m, k) at a_off + (k >> 2) * sak + m * 4 + (k & 3)
  bool a_k4 = false;
  // mix launches (needs b_k4): outputs written k4-packed, each MixOut's c_off the region base
  // and ldc the k of its first row: output (row, col) at c_off + (k >> 2) * 4N + col * 4 +
  // (k & 3) with k = ldc + row and N the item's columns
  bool c_k4 = false;
  // Dot-product epilogue (real grouped launches with an FP64 C, MFMA tile classes): with dot_part
  // set (2 * ntiles doubles), the tile in slot i of the tile list (i = x * per_xcd + t) also
  // writes dot_part[i] = sum of C[e] * dot_x[e] and dot_part[ntiles + i] = sum of dot_x[e]^2
  // over its stored elements e (element offsets in C; a padding slot writes zeros), in a fixed
  // order: summed over all slots, C . dot_x and |dot_x|^2 where the tiles cover C.
  const double *dot_x = nullptr;
  double *dot_part = nullptr;
  // single-block mode: one block with one segment passed by value, no device task arrays
  bool single = false;
  GBlk blk0{};
  GSeg seg0{};
  int kind0 = -1;
};

// Enqueue one grouped launch.  Returns a CXA status.
int launch_grouped_gemm(const GemmLaunch &g, hipStream_t s);

// Workgroups are dispatched to the device's XCDs round-robin (workgroup w runs on XCD
// w % kXcds), and each XCD has its own L2.  Tile lists are therefore cut into kXcds ranges of
// equal length; workgroup w takes tiles from range w % kXcds only.
constexpr int kXcds = 8;

// Host-side planning helper: cut every block into tiles of the chosen kind.  Blocks are taken
// in operand order (offset of their first A, then B, operand), which keeps blocks that read the
// same operand blocks together, and the sequence is split into kXcds contiguous ranges of equal
// tile count, so each XCD's L2 holds about 1/kXcds of the operands.  Within a range tiles run in
// decreasing work (longest K first) so the persistent loop balances; ranges are padded to a
// common length *per_xcd (a multiple of 4) with tiles of blk = -1, which the kernels skip.
// All tiles of a launch share one class (its kernel instantiation), returned: 0 = MFMA tiles
// up to 32x32, 1 = row streaming (taken when every block has N <= 8 and allow_rows), 2 = 64x32
// MFMA tiles (only on request).  allow_rows is false for complex dtypes.
int make_tiles(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
               std::vector<GTile> &tiles, int *per_xcd, int kind_override = -1,
               bool allow_rows = true);

// Whether a tile list is better run with workgroup split-K (few tiles, long K); padding
// tiles are ignored.
bool prefer_splitk(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
                   const std::vector<GTile> &tiles, int tile_class);

// Tile kind used for a block of M x N when no tuned choice is given.
int default_kind(int M, int N);

}  // namespace cxa
