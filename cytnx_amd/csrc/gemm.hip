// Grouped, persistent MFMA GEMM for gfx950.
//
// One launch runs a list of tiles drawn from any number of independent output blocks.  Three
// execution shapes share the task format:
//
//  * MFMA wave tiles: one wave owns a 16*MI x 16*NI tile and walks the K dimension of every
//    segment of its block in chunks of 16 with MFMA 16x16x4 instructions.  Fragments are loaded
//    straight from global memory (the operands are L2/MALL-resident at these sizes) with the k
//    index assigned so that each lane reads 4 consecutive k values: MFMA number j of a chunk
//    takes k = k0 + 4*(lane/16) + j for both A (row lane%16) and B (column lane%16).  The MFMA
//    sums over its 4 k slots, so any assignment identical for A and B is exact; this one makes
//    row-major A loads 32-byte contiguous per lane and B loads 128-byte contiguous per 16 lanes.
//  * Workgroup split-K: the four waves of a workgroup take the same MFMA tile and interleave its
//    K chunks; the partial accumulators meet in LDS and are summed in fixed wave order, so the
//    result is deterministic.  Used when a launch has few tiles with long K.
//  * Row streaming: for blocks with N <= 8 (the MPO steps, K and N of a few MPO states) each
//    lane owns one output row and accumulates in FP64 registers; loads of A are coalesced across
//    lanes, B is a broadcast, C rows are contiguous.
//
// Precision modes: FP64 (v_mfma_f64_16x16x4_f64), FP32 (v_mfma_f32_16x16x4_f32, operands
// rounded to FP32), FP32X3 (FP64 operands split x = hi + lo in FP32; hi*hi + hi*lo + lo*hi on
// FP32 MFMA, the two cross terms in their own accumulator, combined in FP64 at the end).
#include <algorithm>

#include "cxa_common.hpp"
#include "gemm.hpp"

namespace cxa {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct alignas(16) cd {
  double re, im;
};
struct alignas(8) cf {
  float re, im;
};

constexpr int kWaves = 4;
constexpr int kRowsN = 8;  // widest block the row-streaming path takes

__device__ __forceinline__ f64x4 mma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Accumulator row held in register v by lane group g = lane/16 (column = lane%16), measured on
// gfx950: v_mfma_f64_16x16x4_f64 interleaves rows (g + 4v), v_mfma_f32_16x16x4_f32 blocks
// them (4g + v).
__device__ __forceinline__ int acc_row(bool f64, int g, int v) {
  return f64 ? g + 4 * v : 4 * g + v;
}

// base + element offset computed on integers: a batched call passes absolute addresses (in
// elements) as offsets from a null base, which pointer arithmetic may not express.  Offsets in
// elements let one task list serve FP64 and FP32 copies of the same block layout.
template <typename T, typename B>
__device__ __forceinline__ T *at(B *base, int64_t elem_off) {
  return reinterpret_cast<T *>(reinterpret_cast<uintptr_t>(base) +
                               (uintptr_t)elem_off * (uintptr_t)sizeof(T));
}

struct SegArray {
  const GSeg *p;
  __device__ __forceinline__ GSeg operator()(int i) const { return p[i]; }
};
struct SegOne {
  GSeg s;
  __device__ __forceinline__ GSeg operator()(int) const { return s; }
};

// ---------------------------------------------------------------------------------------
// real MFMA tiles
// ---------------------------------------------------------------------------------------
template <int PREC, int MI, int NI>
struct RealAcc {
  static constexpr bool F64 = (PREC == CXA_PREC_FP64);
  using acc_t = typename std::conditional<F64, f64x4, f32x4>::type;
  acc_t hi[MI][NI];
  acc_t lo[MI][NI];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        hi[i][j] = acc_t{0, 0, 0, 0};
        lo[i][j] = acc_t{0, 0, 0, 0};
      }
  }
  __device__ __forceinline__ double get(int i, int j, int v) const {
    double x = (double)hi[i][j][v];
    if constexpr (PREC == CXA_PREC_FP32X3) x += (double)lo[i][j][v];
    return x;
  }
};

// Operand fragments of one 16-deep K chunk (per lane: 4 k values per 16-row / 16-column
// group), split into hi/lo FP32 parts in the FP32X3 mode.
template <int PREC, int MI, int NI>
struct Frag {
  using op_t = typename std::conditional<PREC == CXA_PREC_FP64, double, float>::type;
  op_t a[MI][4], b[NI][4], alo[MI][4], blo[NI][4];
};

template <int PREC, typename TA, typename TB, int MI, int NI>
__device__ __forceinline__ void load_frag(Frag<PREC, MI, NI> &f, const TA *Ap, const TB *Bp,
                                          const GSeg &sg, int k0, int mt, int nt, int r, int g) {
  constexpr bool X3 = (PREC == CXA_PREC_FP32X3);
  using op_t = typename Frag<PREC, MI, NI>::op_t;
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    const int row = 16 * i + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * g + j;
      const double v = (row < mt && k < sg.K)
                           ? (double)Ap[(int64_t)row * sg.sam + (int64_t)k * sg.sak]
                           : 0.0;
      f.a[i][j] = (op_t)v;
      if constexpr (X3) f.alo[i][j] = (float)(v - (double)(float)v);
    }
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int col = 16 * i + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * g + j;
      const double v = (col < nt && k < sg.K)
                           ? (double)Bp[(int64_t)k * sg.sbk + (int64_t)col * sg.sbn]
                           : 0.0;
      f.b[i][j] = (op_t)v;
      if constexpr (X3) f.blo[i][j] = (float)(v - (double)(float)v);
    }
  }
}

template <int PREC, int MI, int NI>
__device__ __forceinline__ void mma_frag(RealAcc<PREC, MI, NI> &acc,
                                         const Frag<PREC, MI, NI> &f) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int jj = 0; jj < NI; ++jj) {
        if constexpr (PREC == CXA_PREC_FP64) {
          acc.hi[i][jj] = mma64(f.a[i][j], f.b[jj][j], acc.hi[i][jj]);
        } else {
          acc.hi[i][jj] = mma32(f.a[i][j], f.b[jj][j], acc.hi[i][jj]);
          if constexpr (PREC == CXA_PREC_FP32X3) {
            acc.lo[i][jj] = mma32(f.a[i][j], f.blo[jj][j], acc.lo[i][jj]);
            acc.lo[i][jj] = mma32(f.alo[i][j], f.b[jj][j], acc.lo[i][jj]);
          }
        }
      }
}

// Accumulate the chunks c with c % nslices == slice of the block's concatenated K range.
// Within a segment the next chunk's fragments are loaded before the current chunk's MFMAs
// issue, so global-memory latency overlaps the matrix work.
template <int PREC, typename TA, typename TB, int MI, int NI, typename SEGS>
__device__ __forceinline__ void mfma_accumulate(RealAcc<PREC, MI, NI> &acc,
                                                const char *__restrict__ Abase,
                                                const char *__restrict__ Bbase, const GBlk &blk,
                                                SEGS segs, int m0, int n0, int lane, int slice,
                                                int nslices) {
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0);
  const int nt = min(16 * NI, blk.N - n0);
  const int step = 16 * nslices;
  int chunk_base = 0;
  for (int s = 0; s < blk.nseg; ++s) {
    const GSeg sg = segs(blk.seg0 + s);
    const TA *Ap = at<const TA>(Abase, sg.a_off) + (int64_t)m0 * sg.sam;
    const TB *Bp = at<const TB>(Bbase, sg.b_off) + (int64_t)n0 * sg.sbn;
    const int nchunks = (sg.K + 15) >> 4;
    const int first = ((slice - chunk_base) % nslices + nslices) % nslices;
    chunk_base += nchunks;
    int k0 = 16 * first;
    if (k0 >= sg.K) continue;
    Frag<PREC, MI, NI> cur, nxt;
    load_frag<PREC, TA, TB, MI, NI>(cur, Ap, Bp, sg, k0, mt, nt, r, g);
    for (; k0 < sg.K; k0 += step) {
      const bool more = k0 + step < sg.K;
      if (more) load_frag<PREC, TA, TB, MI, NI>(nxt, Ap, Bp, sg, k0 + step, mt, nt, r, g);
      mma_frag<PREC, MI, NI>(acc, cur);
      if (more) cur = nxt;
    }
  }
}

template <typename TC>
__device__ __forceinline__ void store_c(TC *dst, double val, const GBlk &blk) {
  val *= blk.alpha;
  if (blk.beta != 0.0) val += blk.beta * (double)(*dst);
  *dst = (TC)val;
}

template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, typename SEGS>
__device__ __forceinline__ void wave_tile_real(const char *__restrict__ Abase,
                                               const char *__restrict__ Bbase,
                                               char *__restrict__ Cbase, const GBlk &blk,
                                               SEGS segs, int m0, int n0, int lane) {
  RealAcc<PREC, MI, NI> acc;
  acc.zero();
  mfma_accumulate<PREC, TA, TB, MI, NI>(acc, Abase, Bbase, blk, segs, m0, n0, lane, 0, 1);
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0), nt = min(16 * NI, blk.N - n0);
  TC *Cp = at<TC>(Cbase, blk.c_off);
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int jj = 0; jj < NI; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + acc_row(RealAcc<PREC, MI, NI>::F64, g, v);
        const int col = 16 * jj + r;
        if (row < mt && col < nt)
          store_c(Cp + (int64_t)(m0 + row) * blk.ldc + (n0 + col), acc.get(i, jj, v), blk);
      }
}

// Split-K tile: every wave of the workgroup calls this for the same tile.
template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, typename SEGS>
__device__ __forceinline__ void wg_tile_real(const char *__restrict__ Abase,
                                             const char *__restrict__ Bbase,
                                             char *__restrict__ Cbase, const GBlk &blk,
                                             SEGS segs, int m0, int n0, int lane, int wave,
                                             double (*red)[16][64]) {
  RealAcc<PREC, MI, NI> acc;
  acc.zero();
  mfma_accumulate<PREC, TA, TB, MI, NI>(acc, Abase, Bbase, blk, segs, m0, n0, lane, wave,
                                        kWaves);
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int jj = 0; jj < NI; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave][(i * NI + jj) * 4 + v][lane] = acc.get(i, jj, v);
  __syncthreads();
  if (wave == 0) {
    const int r = lane & 15, g = lane >> 4;
    const int mt = min(16 * MI, blk.M - m0), nt = min(16 * NI, blk.N - n0);
    TC *Cp = at<TC>(Cbase, blk.c_off);
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int jj = 0; jj < NI; ++jj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * i + acc_row(RealAcc<PREC, MI, NI>::F64, g, v);
          const int col = 16 * jj + r;
          const int idx = (i * NI + jj) * 4 + v;
          const double s = ((red[0][idx][lane] + red[1][idx][lane]) + red[2][idx][lane]) +
                           red[3][idx][lane];
          if (row < mt && col < nt)
            store_c(Cp + (int64_t)(m0 + row) * blk.ldc + (n0 + col), s, blk);
        }
  }
  __syncthreads();
}

// Row streaming for N <= kRowsN: lane = output row, FP64 accumulation.  The K loop runs the
// k values with k % nslices == slice, four at a time with the loads issued ahead of the FMAs.
template <typename TA, typename TB, typename SEGS>
__device__ __forceinline__ void rows_accumulate(double (&acc)[kRowsN], const char *Abase,
                                                const char *Bbase, const GBlk &blk, SEGS segs,
                                                int row, bool valid, int slice, int nslices) {
  const int N = blk.N;
#pragma unroll
  for (int n = 0; n < kRowsN; ++n) acc[n] = 0.0;
  int kbase = 0;
  for (int s = 0; s < blk.nseg; ++s) {
    const GSeg sg = segs(blk.seg0 + s);
    const TA *Ap = at<const TA>(Abase, sg.a_off) + (valid ? (int64_t)row * sg.sam : 0);
    const TB *Bp = at<const TB>(Bbase, sg.b_off);
    int k = ((slice - kbase) % nslices + nslices) % nslices;
    kbase += sg.K;
    for (; k < sg.K; k += 4 * nslices) {
      double a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u * nslices;
        a[u] = (valid && kk < sg.K) ? (double)Ap[(int64_t)kk * sg.sak] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u * nslices;
        if (kk < sg.K) {
          const TB *bk = Bp + (int64_t)kk * sg.sbk;
#pragma unroll
          for (int n = 0; n < kRowsN; ++n)
            if (n < N) acc[n] += a[u] * (double)bk[(int64_t)n * sg.sbn];
        }
      }
    }
  }
}

template <typename TA, typename TB, typename TC, typename SEGS>
__device__ __forceinline__ void wave_rows(const char *__restrict__ Abase,
                                          const char *__restrict__ Bbase,
                                          char *__restrict__ Cbase, const GBlk &blk, SEGS segs,
                                          int m0, int lane) {
  const int row = m0 + lane;
  const bool valid = row < blk.M;
  double acc[kRowsN];
  rows_accumulate<TA, TB>(acc, Abase, Bbase, blk, segs, row, valid, 0, 1);
  if (!valid) return;
  TC *Cp = at<TC>(Cbase, blk.c_off) + (int64_t)row * blk.ldc;
#pragma unroll
  for (int n = 0; n < kRowsN; ++n)
    if (n < blk.N) store_c(Cp + n, acc[n], blk);
}

// Split-K row tile: the four waves take interleaved k values of the same 64 rows.
template <typename TA, typename TB, typename TC, typename SEGS>
__device__ __forceinline__ void wg_rows(const char *__restrict__ Abase,
                                        const char *__restrict__ Bbase, char *__restrict__ Cbase,
                                        const GBlk &blk, SEGS segs, int m0, int lane, int wave,
                                        double (*red)[16][64]) {
  const int row = m0 + lane;
  const bool valid = row < blk.M;
  double acc[kRowsN];
  rows_accumulate<TA, TB>(acc, Abase, Bbase, blk, segs, row, valid, wave, kWaves);
#pragma unroll
  for (int n = 0; n < kRowsN; ++n) red[wave][n][lane] = acc[n];
  __syncthreads();
  if (wave == 0 && valid) {
    TC *Cp = at<TC>(Cbase, blk.c_off) + (int64_t)row * blk.ldc;
#pragma unroll
    for (int n = 0; n < kRowsN; ++n)
      if (n < blk.N)
        store_c(Cp + n, ((red[0][n][lane] + red[1][n][lane]) + red[2][n][lane]) + red[3][n][lane],
                blk);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------------
// complex MFMA tiles (C128 on FP64 MFMA, C64 on FP32 MFMA): four real products per step
// ---------------------------------------------------------------------------------------
template <typename CT, int MI, int NI, typename SEGS>
__device__ __forceinline__ void wave_tile_cplx(const char *__restrict__ Abase,
                                               const char *__restrict__ Bbase,
                                               char *__restrict__ Cbase, const GBlk &blk,
                                               SEGS segs, int m0, int n0, int lane) {
  constexpr bool F64 = std::is_same<CT, cd>::value;
  using acc_t = typename std::conditional<F64, f64x4, f32x4>::type;
  using op_t = typename std::conditional<F64, double, float>::type;
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0);
  const int nt = min(16 * NI, blk.N - n0);
  acc_t accr[MI][NI], acci[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      accr[i][j] = acc_t{0, 0, 0, 0};
      acci[i][j] = acc_t{0, 0, 0, 0};
    }
  for (int s = 0; s < blk.nseg; ++s) {
    const GSeg sg = segs(blk.seg0 + s);
    const CT *Ap = at<const CT>(Abase, sg.a_off) + (int64_t)m0 * sg.sam;
    const CT *Bp = at<const CT>(Bbase, sg.b_off) + (int64_t)n0 * sg.sbn;
    const int K = sg.K;
    for (int k0 = 0; k0 < K; k0 += 16) {
      op_t ar[MI][4], ai[MI][4], br[NI][4], bi[NI][4];
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int row = 16 * i + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + 4 * g + j;
          CT v{0, 0};
          if (row < mt && k < K) v = Ap[(int64_t)row * sg.sam + (int64_t)k * sg.sak];
          ar[i][j] = v.re;
          ai[i][j] = v.im;
        }
      }
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int col = 16 * i + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + 4 * g + j;
          CT v{0, 0};
          if (col < nt && k < K) v = Bp[(int64_t)k * sg.sbk + (int64_t)col * sg.sbn];
          br[i][j] = v.re;
          bi[i][j] = v.im;
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
          for (int jj = 0; jj < NI; ++jj) {
            if constexpr (F64) {
              accr[i][jj] = mma64(ar[i][j], br[jj][j], accr[i][jj]);
              accr[i][jj] = mma64(-ai[i][j], bi[jj][j], accr[i][jj]);
              acci[i][jj] = mma64(ar[i][j], bi[jj][j], acci[i][jj]);
              acci[i][jj] = mma64(ai[i][j], br[jj][j], acci[i][jj]);
            } else {
              accr[i][jj] = mma32(ar[i][j], br[jj][j], accr[i][jj]);
              accr[i][jj] = mma32(-ai[i][j], bi[jj][j], accr[i][jj]);
              acci[i][jj] = mma32(ar[i][j], bi[jj][j], acci[i][jj]);
              acci[i][jj] = mma32(ai[i][j], br[jj][j], acci[i][jj]);
            }
          }
    }
  }
  CT *Cp = at<CT>(Cbase, blk.c_off);
  const double alpha = blk.alpha, beta = blk.beta;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int jj = 0; jj < NI; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + acc_row(F64, g, v);
        const int col = 16 * jj + r;
        if (row < mt && col < nt) {
          double re = alpha * (double)accr[i][jj][v], im = alpha * (double)acci[i][jj][v];
          CT *dst = Cp + (int64_t)(m0 + row) * blk.ldc + (n0 + col);
          if (beta != 0.0) {
            const CT o = *dst;
            re += beta * (double)o.re;
            im += beta * (double)o.im;
          }
          CT w;
          w.re = (decltype(w.re))re;
          w.im = (decltype(w.im))im;
          *dst = w;
        }
      }
}

// Tile classes: each class is its own kernel instantiation so that it carries only its own
// register budget (occupancy matters most for the memory-bound row tiles).
enum TileClass { TC_MIXED = 0, TC_ROWS = 1, TC_BIG = 2 };

template <int PREC, typename TA, typename TB, typename TC, bool CPLX, int CLS, typename SEGS>
__device__ __forceinline__ void run_tile(int kind, const char *A, const char *B, char *C,
                                         const GBlk &blk, SEGS segs, int m0, int n0, int lane) {
  if constexpr (CPLX) {
    switch (kind) {
      case TK_1x1: wave_tile_cplx<TA, 1, 1>(A, B, C, blk, segs, m0, n0, lane); break;
      case TK_2x1: wave_tile_cplx<TA, 2, 1>(A, B, C, blk, segs, m0, n0, lane); break;
      case TK_1x2: wave_tile_cplx<TA, 1, 2>(A, B, C, blk, segs, m0, n0, lane); break;
      default: wave_tile_cplx<TA, 2, 2>(A, B, C, blk, segs, m0, n0, lane); break;
    }
  } else if constexpr (CLS == TC_ROWS) {
    wave_rows<TA, TB, TC>(A, B, C, blk, segs, m0, lane);
  } else if constexpr (CLS == TC_BIG) {
    wave_tile_real<PREC, TA, TB, TC, 4, 2>(A, B, C, blk, segs, m0, n0, lane);
  } else {
    switch (kind) {
      case TK_1x1: wave_tile_real<PREC, TA, TB, TC, 1, 1>(A, B, C, blk, segs, m0, n0, lane); break;
      case TK_2x1: wave_tile_real<PREC, TA, TB, TC, 2, 1>(A, B, C, blk, segs, m0, n0, lane); break;
      case TK_1x2: wave_tile_real<PREC, TA, TB, TC, 1, 2>(A, B, C, blk, segs, m0, n0, lane); break;
      default: wave_tile_real<PREC, TA, TB, TC, 2, 2>(A, B, C, blk, segs, m0, n0, lane); break;
    }
  }
}

// grouped launch: tiles, blocks and segments come from device arrays; one wave per tile
template <int PREC, typename TA, typename TB, typename TC, bool CPLX, int CLS>
__global__ __launch_bounds__(64 * kWaves) void k_grouped(const GTile *__restrict__ tiles,
                                                        int ntiles, const GBlk *__restrict__ blks,
                                                        const GSeg *__restrict__ segs,
                                                        const char *A, const char *B, char *C) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int t = blockIdx.x * kWaves + wave; t < ntiles; t += gridDim.x * kWaves) {
    const GTile tl = tiles[t];
    const GBlk blk = blks[tl.blk];
    run_tile<PREC, TA, TB, TC, CPLX, CLS>(tl.kind, A, B, C, blk, SegArray{segs}, tl.m0, tl.n0,
                                          lane);
  }
}

// grouped launch with workgroup split-K: one workgroup per tile (real dtypes; the mixed MFMA
// class or the row class)
template <int PREC, typename TA, typename TB, typename TC, int CLS>
__global__ __launch_bounds__(64 * kWaves) void k_grouped_split(
    const GTile *__restrict__ tiles, int ntiles, const GBlk *__restrict__ blks,
    const GSeg *__restrict__ segs, const char *A, const char *B, char *C) {
  __shared__ double red[kWaves][16][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GTile tl = tiles[t];
    const GBlk blk = blks[tl.blk];
    const SegArray sa{segs};
    if constexpr (CLS == TC_ROWS) {
      wg_rows<TA, TB, TC>(A, B, C, blk, sa, tl.m0, lane, wave, red);
    } else {
      switch (tl.kind) {
        case TK_1x1:
          wg_tile_real<PREC, TA, TB, TC, 1, 1>(A, B, C, blk, sa, tl.m0, tl.n0, lane, wave, red);
          break;
        case TK_2x1:
          wg_tile_real<PREC, TA, TB, TC, 2, 1>(A, B, C, blk, sa, tl.m0, tl.n0, lane, wave, red);
          break;
        case TK_1x2:
          wg_tile_real<PREC, TA, TB, TC, 1, 2>(A, B, C, blk, sa, tl.m0, tl.n0, lane, wave, red);
          break;
        default:
          wg_tile_real<PREC, TA, TB, TC, 2, 2>(A, B, C, blk, sa, tl.m0, tl.n0, lane, wave, red);
          break;
      }
    }
  }
}

// single-block launch: one block with one segment, tiles derived from the launch index, so a
// plain GEMM needs no device-side task list (mixed MFMA class)
template <int PREC, typename TA, typename TB, typename TC, bool CPLX>
__global__ __launch_bounds__(64 * kWaves) void k_single(GBlk blk, GSeg seg, int kind,
                                                       int tiles_n, int ntiles, const char *A,
                                                       const char *B, char *C) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tr = kind == TK_1x1 || kind == TK_1x2 ? 16 : 32;
  const int tc = kind == TK_1x1 || kind == TK_2x1 ? 16 : 32;
  for (int t = blockIdx.x * kWaves + wave; t < ntiles; t += gridDim.x * kWaves) {
    const int m0 = (t / tiles_n) * tr, n0 = (t % tiles_n) * tc;
    run_tile<PREC, TA, TB, TC, CPLX, TC_MIXED>(kind, A, B, C, blk, SegOne{seg}, m0, n0, lane);
  }
}

template <int PREC, typename TA, typename TB, typename TC, bool CPLX>
int launch_t(const GemmLaunch &g, hipStream_t s) {
  if (g.single) {
    const GBlk &b = g.blk0;
    if (b.M <= 0 || b.N <= 0) return CXA_OK;
    int kind = g.kind0 >= 0 ? g.kind0 : default_kind(b.M, b.N);
    if (kind == TK_ROWS || kind == TK_4x2) kind = TK_2x2;  // single launches: mixed class
    const int tiles_m = (b.M + tile_rows(kind) - 1) / tile_rows(kind);
    const int tiles_n = (b.N + tile_cols(kind) - 1) / tile_cols(kind);
    const int64_t nt64 = (int64_t)tiles_m * tiles_n;
    if (nt64 > INT32_MAX)
      return fail(CXA_ERR_UNSUPPORTED, "GEMM too large: %lld tiles", (long long)nt64);
    const int ntiles = (int)nt64;
    const int grid = std::min((ntiles + kWaves - 1) / kWaves, 8192);
    hipLaunchKernelGGL((k_single<PREC, TA, TB, TC, CPLX>), dim3(grid), dim3(64 * kWaves), 0, s,
                       b, g.seg0, kind, tiles_n, ntiles, (const char *)g.A, (const char *)g.B,
                       (char *)g.C);
    CXA_HIP(hipGetLastError());
    return CXA_OK;
  }
  if (g.ntiles <= 0) return CXA_OK;
  const int cls = g.tile_class;
  if (CPLX && cls != TC_MIXED)
    return fail(CXA_ERR_UNSUPPORTED, "complex GEMM tiles must be MFMA tiles up to 32x32");
  const int wg_grid = g.grid > 0 ? g.grid : std::min(g.ntiles, 8192);
  const int wave_grid = g.grid > 0 ? g.grid : std::min((g.ntiles + kWaves - 1) / kWaves, 4096);
  if constexpr (!CPLX) {
    if (g.splitk && cls == TC_ROWS)
      hipLaunchKernelGGL((k_grouped_split<PREC, TA, TB, TC, TC_ROWS>), dim3(wg_grid),
                         dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                         (const char *)g.A, (const char *)g.B, (char *)g.C);
    else if (g.splitk && cls == TC_MIXED)
      hipLaunchKernelGGL((k_grouped_split<PREC, TA, TB, TC, TC_MIXED>), dim3(wg_grid),
                         dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                         (const char *)g.A, (const char *)g.B, (char *)g.C);
    else if (cls == TC_ROWS)
      hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_ROWS>), dim3(wave_grid),
                         dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                         (const char *)g.A, (const char *)g.B, (char *)g.C);
    else if (cls == TC_BIG)
      hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_BIG>), dim3(wave_grid),
                         dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                         (const char *)g.A, (const char *)g.B, (char *)g.C);
    else
      hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_MIXED>), dim3(wave_grid),
                         dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                         (const char *)g.A, (const char *)g.B, (char *)g.C);
  } else {
    hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, true, TC_MIXED>), dim3(wave_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, g.ntiles, g.blks, g.segs,
                       (const char *)g.A, (const char *)g.B, (char *)g.C);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

}  // namespace

int launch_grouped_gemm(const GemmLaunch &g, hipStream_t s) {
  const int p = g.prec, a = g.a_dt, b = g.b_dt, c = g.c_dt;
  if (p == CXA_PREC_FP64) {
    if (a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
      return launch_t<CXA_PREC_FP64, double, double, double, false>(g, s);
    if (a == CXA_C128 && b == CXA_C128 && c == CXA_C128)
      return launch_t<CXA_PREC_FP64, cd, cd, cd, true>(g, s);
  } else if (p == CXA_PREC_FP32) {
#define CXA_F32_CASE(TA_, TB_, TC_, IA, IB, IC) \
  if (a == IA && b == IB && c == IC) return launch_t<CXA_PREC_FP32, TA_, TB_, TC_, false>(g, s);
    CXA_F32_CASE(float, float, float, CXA_F32, CXA_F32, CXA_F32)
    CXA_F32_CASE(float, float, double, CXA_F32, CXA_F32, CXA_F64)
    CXA_F32_CASE(float, double, float, CXA_F32, CXA_F64, CXA_F32)
    CXA_F32_CASE(float, double, double, CXA_F32, CXA_F64, CXA_F64)
    CXA_F32_CASE(double, float, float, CXA_F64, CXA_F32, CXA_F32)
    CXA_F32_CASE(double, float, double, CXA_F64, CXA_F32, CXA_F64)
    CXA_F32_CASE(double, double, float, CXA_F64, CXA_F64, CXA_F32)
    CXA_F32_CASE(double, double, double, CXA_F64, CXA_F64, CXA_F64)
#undef CXA_F32_CASE
    if (a == CXA_C64 && b == CXA_C64 && c == CXA_C64)
      return launch_t<CXA_PREC_FP32, cf, cf, cf, true>(g, s);
  } else if (p == CXA_PREC_FP32X3) {
    if (a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
      return launch_t<CXA_PREC_FP32X3, double, double, double, false>(g, s);
  }
  return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: unsupported precision %d with dtypes (%d,%d,%d)",
              p, a, b, c);
}

int default_kind(int M, int N) {
  if (M <= 16 && N <= 16) return TK_1x1;
  if (M <= 16) return TK_1x2;
  if (N <= 16) return TK_2x1;
  return TK_2x2;
}

int make_tiles(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
               std::vector<GTile> &tiles, int kind_override, bool allow_rows) {
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
  std::vector<Item> items;
  for (size_t bi = 0; bi < blks.size(); ++bi) {
    const GBlk &b = blks[bi];
    if (b.M <= 0 || b.N <= 0) continue;
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
  std::stable_sort(items.begin(), items.end(),
                   [](const Item &x, const Item &y) { return x.cost > y.cost; });
  tiles.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) tiles[i] = items[i].t;
  return cls;
}

bool prefer_splitk(const std::vector<GBlk> &blks, const std::vector<GSeg> &segs,
                   const std::vector<GTile> &tiles, int tile_class) {
  if (tiles.empty() || tile_class == TC_BIG) return false;
  int64_t work = 0;
  for (const GTile &t : tiles) {
    const GBlk &b = blks[t.blk];
    for (int s = 0; s < b.nseg; ++s)
      work += tile_class == TC_ROWS ? segs[b.seg0 + s].K : (segs[b.seg0 + s].K + 15) / 16;
  }
  // few tiles (under two per SIMD of a 256-CU device) that each walk a long K range
  const double avg = (double)work / (double)tiles.size();
  return tiles.size() < 2048 && avg >= (tile_class == TC_ROWS ? 64.0 : 8.0);
}

}  // namespace cxa
