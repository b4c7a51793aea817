#pragma once

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
//
// This header holds the kernel templates; the translation units gemm_fp64.hip, gemm_fp32.hip,
// gemm_fp32_a64.hip, gemm_fp32x3.hip and gemm_mix.hip each instantiate one precision family
// (so they compile in parallel), and gemm.hip holds the dispatch and the host-side tile
// planning.  Define
// CXA_GEMM_TU (a distinct token per translation unit) before including it: each unit carries
// its own zero cell.
#include <algorithm>
#include <type_traits>

#include "cxa_common.hpp"
#include "gemm.hpp"

#ifndef CXA_GEMM_TU
#error "define CXA_GEMM_TU before including gemm_kernels.hpp"
#endif
#define CXA_GEMM_CAT2(a, b) a##b
#define CXA_GEMM_CAT(a, b) CXA_GEMM_CAT2(a, b)
#define CXA_GEMM_ZERO_CELL CXA_GEMM_CAT(gemm_zero_cell_, CXA_GEMM_TU)

namespace cxa {


// Thirty-two zero bytes in device memory (zero-initialised, never written).  An operand element
// outside its block is read from here (the address is selected, not the value), so fragment
// loads need no branch and no wait at the load site; the MFMA that consumes them sees exact
// zeros.  External linkage keeps the compiler from folding these loads to constants, which
// would turn the selects back into branches around the loads.
__device__ __align__(32) double CXA_GEMM_ZERO_CELL[4];

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

// Operand layouts of a launch (template bit set PK): B, A k4-packed (GemmLaunch::b_k4 / a_k4),
// mix outputs written k4-packed (GemmLaunch::c_k4).
constexpr int kPackB = 1, kPackA = 2, kPackC = 4;

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

// Wait for every outstanding vector-memory load (s_waitcnt vmcnt(0); expcnt and lgkmcnt left
// at their maxima).  Issued where a load pipeline ends, when every load it issued has already
// been consumed, so it costs nothing at run time; it tells the compiler's wait analysis that no
// fragment load is still in flight, which it cannot prove by itself across the ring's
// conditional prefetches, and which it would otherwise answer with waits on the next tile's
// first loads.
__device__ __forceinline__ void drain_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Operand and result pointers live in the global address space: loads through them are
// global_load (counted by vmcnt alone), where generic pointers would give flat loads that also
// hold lgkmcnt, so every wait for a scalar descriptor load would drain the operand pipeline.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;

// base + element offset computed on integers: a batched call passes absolute addresses (in
// elements) as offsets from a null base, which pointer arithmetic may not express.  Offsets in
// elements let one task list serve FP64 and FP32 copies of the same block layout.
template <typename T, typename B>
__device__ __forceinline__ gptr<T> at(B *base, int64_t elem_off) {
  return (gptr<T>)(reinterpret_cast<uintptr_t>(base) +
                   (uintptr_t)elem_off * (uintptr_t)sizeof(T));
}

template <typename T>
__device__ __forceinline__ gptr<const T> zero_cell() {
  static_assert(sizeof(T) <= sizeof(CXA_GEMM_ZERO_CELL), "zero cell too small");
  return (gptr<const T>)(&CXA_GEMM_ZERO_CELL[0]);
}

// *p when `valid`, else 0, as one unconditional load.  The selected address passes through an
// empty asm operand so that the compiler cannot split the load into "valid ? *p : 0" (a branch
// around the load, which costs the load pipeline its in-order wait counts).
template <typename T>
__device__ __forceinline__ T load_or_zero(gptr<const T> p, bool valid) {
  uint64_t a = reinterpret_cast<uint64_t>(valid ? p : zero_cell<T>());
  __asm__("" : "+v"(a));
  return *reinterpret_cast<gptr<const T>>(a);
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
// group), held as loaded (operand dtype).  Rounding to FP32 and the FP32X3 hi/lo split happen
// when the chunk is consumed, so the loads of the chunks in flight are not waited for early.
template <typename TA, typename TB, int MI, int NI>
struct Frag {
  TA a[MI][4];
  TB b[NI][4];
};

// `live` is false once the cursor has passed the end of its K range: every element then reads
// the zero cell (the pipelines load only chunks inside the range, so this is a guard).  With
// kPackA / kPackB in PK the A / B operand is k4-packed (element (m, k) at (k >> 2) * sak +
// m * 4 + (k & 3), element (k, n) at (k >> 2) * sbk + n * 4 + (k & 3) from the segment's
// origin, whole quads): the 4 k of a lane's row or column are one vector load.
template <typename TA, typename TB, int MI, int NI, int PK = 0>
__device__ __forceinline__ void load_frag(Frag<TA, TB, MI, NI> &f, gptr<const TA> Ap,
                                          gptr<const TB> Bp, const GSeg &sg, int k0, int mt,
                                          int nt, int r, int g, bool live) {
  if constexpr ((PK & kPackA) != 0) {
    typedef TA ta4 __attribute__((ext_vector_type(4)));
    const bool quad = live && k0 + 4 * g < sg.K;
    const int64_t qoff = (int64_t)((k0 >> 2) + g) * sg.sak;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int row = 16 * i + r;
      const ta4 v = load_or_zero<ta4>((gptr<const ta4>)(Ap + (qoff + (int64_t)row * 4)),
                                      quad && row < mt);
      f.a[i][0] = v.x;
      f.a[i][1] = v.y;
      f.a[i][2] = v.z;
      f.a[i][3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int row = 16 * i + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + 4 * g + j;
        f.a[i][j] = load_or_zero<TA>(Ap + ((int64_t)row * sg.sam + (int64_t)k * sg.sak),
                                     live && row < mt && k < sg.K);
      }
    }
  }
  if constexpr ((PK & kPackB) != 0) {
    typedef TB tb4 __attribute__((ext_vector_type(4)));
    const bool quad = live && k0 + 4 * g < sg.K;  // quads past K are not stored
    const int64_t qoff = (int64_t)((k0 >> 2) + g) * sg.sbk;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int col = 16 * i + r;
      const tb4 v = load_or_zero<tb4>((gptr<const tb4>)(Bp + (qoff + (int64_t)col * 4)),
                                      quad && col < nt);
      f.b[i][0] = v.x;
      f.b[i][1] = v.y;
      f.b[i][2] = v.z;
      f.b[i][3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int col = 16 * i + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + 4 * g + j;
        f.b[i][j] = load_or_zero<TB>(Bp + ((int64_t)k * sg.sbk + (int64_t)col * sg.sbn),
                                     live && col < nt && k < sg.K);
      }
    }
  }
}

template <int PREC, typename TA, typename TB, int MI, int NI>
__device__ __forceinline__ void mma_frag(RealAcc<PREC, MI, NI> &acc,
                                         const Frag<TA, TB, MI, NI> &f) {
  constexpr bool X3 = (PREC == CXA_PREC_FP32X3);
  using op_t = typename std::conditional<PREC == CXA_PREC_FP64, double, float>::type;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    op_t a[MI], b[NI];
    float alo[MI], blo[NI];
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      a[i] = (op_t)f.a[i][j];
      if constexpr (X3) alo[i] = (float)((double)f.a[i][j] - (double)a[i]);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      b[i] = (op_t)f.b[i][j];
      if constexpr (X3) blo[i] = (float)((double)f.b[i][j] - (double)b[i]);
    }
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int jj = 0; jj < NI; ++jj) {
        if constexpr (PREC == CXA_PREC_FP64) {
          acc.hi[i][jj] = mma64(a[i], b[jj], acc.hi[i][jj]);
        } else {
          acc.hi[i][jj] = mma32(a[i], b[jj], acc.hi[i][jj]);
          if constexpr (X3) {
            acc.lo[i][jj] = mma32(a[i], blo[jj], acc.lo[i][jj]);
            acc.lo[i][jj] = mma32(alo[i], b[jj], acc.lo[i][jj]);
          }
        }
      }
  }
}

// Chunks in flight per wave: enough loads outstanding to cover an L2/MALL round trip with a
// tile's whole K range at the usual block sizes (K <= 64 is one round trip), within a register
// budget that keeps several waves per SIMD.
template <int PREC, int MI, int NI>
constexpr int pipe_depth() {
  return (MI * NI >= 8 || (PREC != CXA_PREC_FP32 && MI * NI >= 4)) ? 2 : 4;
}

// Walks the 16-deep K chunks c of a block's concatenated segments with c % nslices == slice,
// in order; all fields are wave-uniform.
template <typename TA, typename TB, typename SEGS>
struct ChunkCursor {
  SEGS segs;
  int seg0, nseg, nslices, m0, n0;
  int s, k0;  // current segment and chunk start within it
  GSeg sg;
  gptr<const TA> Ap;
  gptr<const TB> Bp;
  __device__ __forceinline__ void enter(const char *Abase, const char *Bbase) {
    sg = segs(seg0 + s);
    Ap = at<const TA>(Abase, sg.a_off) + (int64_t)m0 * sg.sam;
    Bp = at<const TB>(Bbase, sg.b_off) + (int64_t)n0 * sg.sbn;
  }
  __device__ __forceinline__ bool live() const { return s < nseg; }
  template <int MI, int NI, int PK>
  __device__ __forceinline__ void load(Frag<TA, TB, MI, NI> &f, int mt, int nt, int r,
                                       int g) const {
    load_frag<TA, TB, MI, NI, PK>(f, Ap, Bp, sg, k0, mt, nt, r, g, live());
  }
  // skip `nchunk` chunks forward (crossing segment ends); past the last segment the cursor
  // stays exhausted (live() false)
  __device__ __forceinline__ void skip(int nchunk, const char *Abase, const char *Bbase) {
    k0 += 16 * nchunk;
    bool moved = false;
    while (s < nseg) {
      const int span = 16 * ((sg.K + 15) >> 4);
      if (k0 < span) break;
      k0 -= span;
      ++s;
      if (s < nseg) {
        sg = segs(seg0 + s);
        moved = true;
      }
    }
    if (moved && s < nseg) enter(Abase, Bbase);
  }
};

// Accumulate the chunks c with c % nslices == slice of the block's concatenated K range.
// A ring of pipe_depth() fragment sets keeps that many chunks' loads in flight across the
// whole range (segment boundaries included): the MFMAs of chunk c issue while chunks
// c+1 .. c+depth-1 are still arriving, and chunk c+depth is loaded into the set just consumed.
// Only real chunks are loaded: the launches are bound by the vector-memory address path, so
// padding a wave's chunk count to whole rounds of the ring (which would let every wait count
// be exact) costs more than the waits it saves.
template <int PREC, typename TA, typename TB, int MI, int NI, int PK, typename SEGS>
__device__ __forceinline__ void mfma_accumulate(RealAcc<PREC, MI, NI> &acc,
                                                const char *__restrict__ Abase,
                                                const char *__restrict__ Bbase, const GBlk &blk,
                                                SEGS segs, int m0, int n0, int lane, int slice,
                                                int nslices) {
  constexpr int D = pipe_depth<PREC, MI, NI>();
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0);
  const int nt = min(16 * NI, blk.N - n0);
  // chunks of this slice over all segments
  int total = 0, cb = 0;
  for (int s = 0; s < blk.nseg; ++s) {
    const int nc = (segs(blk.seg0 + s).K + 15) >> 4;
    const int first = ((slice - cb) % nslices + nslices) % nslices;
    if (first < nc) total += (nc - first + nslices - 1) / nslices;
    cb += nc;
  }
  if (total == 0) return;
  ChunkCursor<TA, TB, SEGS> cur{segs, blk.seg0, blk.nseg, nslices, m0, n0, 0, 0};
  cur.enter(Abase, Bbase);
  cur.skip(slice, Abase, Bbase);  // also steps over empty leading segments
  Frag<TA, TB, MI, NI> f[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (d < total) {
      cur.template load<MI, NI, PK>(f[d], mt, nt, r, g);
      cur.skip(nslices, Abase, Bbase);
    }
  }
  for (int base = 0; base < total; base += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int c = base + d;
      if (c < total) {
        mma_frag<PREC, TA, TB, MI, NI>(acc, f[d]);
        if (c + D < total) {
          cur.template load<MI, NI, PK>(f[d], mt, nt, r, g);
          cur.skip(nslices, Abase, Bbase);
        }
      }
    }
  }
  drain_loads();
}

template <typename TC>
__device__ __forceinline__ void store_c(gptr<TC> dst, double val, const GBlk &blk) {
  val *= blk.alpha;
  if (blk.beta != 0.0) val += blk.beta * (double)(*dst);
  *dst = (TC)val;
}

// Sum over the 64 lanes of a wave (butterfly; every lane gets the same value).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The dot-product epilogue of one tile: lane partials of C[e] * x[e] and of x[e]^2 summed over
// the wave into slot[0] and slot[nslots] (lane 0).  The C values come from the accumulators.
__device__ __forceinline__ void dot_epilogue(double cx, double xx, double *slot, int64_t nslots,
                                             int lane) {
  cx = wave_sum(cx);
  xx = wave_sum(xx);
  if (lane == 0) {
    slot[0] = cx;
    slot[nslots] = xx;
  }
}

template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, int PK,
          typename SEGS>
__device__ __forceinline__ void wave_tile_real(const char *__restrict__ Abase,
                                               const char *__restrict__ Bbase,
                                               char *__restrict__ Cbase, const GBlk &blk,
                                               SEGS segs, int m0, int n0, int lane,
                                               const double *__restrict__ dotx = nullptr,
                                               double *dslot = nullptr, int64_t nslots = 0) {
  RealAcc<PREC, MI, NI> acc;
  acc.zero();
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0), nt = min(16 * NI, blk.N - n0);
  // FP32 launches load the dot vector's values of this tile before the K loop, so that their
  // latency overlaps it instead of following it (FP64 launches load them after it: holding them
  // across the loop would cost FP64 tiles a wave per SIMD of occupancy)
  constexpr bool kPre = PREC == CXA_PREC_FP32;
  double xpre[MI][NI][4];
  if (kPre && dslot) {
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int jj = 0; jj < NI; ++jj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * i + acc_row(RealAcc<PREC, MI, NI>::F64, g, v);
          const int col = 16 * jj + r;
          // branch-free (the zero cell outside the tile), so that the K loop's wait counts
          // stay exact
          xpre[i][jj][v] = load_or_zero<double>(
              (gptr<const double>)(dotx + blk.c_off + (int64_t)(m0 + row) * blk.ldc + (n0 + col)),
              row < mt && col < nt);
        }
  }
  mfma_accumulate<PREC, TA, TB, MI, NI, PK>(acc, Abase, Bbase, blk, segs, m0, n0, lane, 0, 1);
  gptr<TC> Cp = at<TC>(Cbase, blk.c_off);
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
  if (dslot) {  // uniform over the wave; only launches with an FP64 C and alpha 1, beta 0
    double cx = 0.0, xx = 0.0;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int jj = 0; jj < NI; ++jj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * i + acc_row(RealAcc<PREC, MI, NI>::F64, g, v);
          const int col = 16 * jj + r;
          if (row < mt && col < nt) {
            const double xe =
                kPre ? xpre[i][jj][v]
                     : dotx[blk.c_off + (int64_t)(m0 + row) * blk.ldc + (n0 + col)];
            cx += acc.get(i, jj, v) * xe;
            xx += xe * xe;
          }
        }
    dot_epilogue(cx, xx, dslot, nslots, lane);
  }
}

// Split-K tile: every wave of the workgroup (WAVES of them) calls this for the same tile; wave
// w takes the K chunks c with c % WAVES == w, and the partial tiles are summed in wave order.
template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, int WAVES, int PK,
          typename SEGS>
__device__ __forceinline__ void wg_tile_real(const char *__restrict__ Abase,
                                             const char *__restrict__ Bbase,
                                             char *__restrict__ Cbase, const GBlk &blk,
                                             SEGS segs, int m0, int n0, int lane, int wave,
                                             double (*red)[16][64],
                                             const double *__restrict__ dotx = nullptr,
                                             double *dslot = nullptr, int64_t nslots = 0) {
  RealAcc<PREC, MI, NI> acc;
  acc.zero();
  // The waves share the epilogue: wave w sums, stores and (with a dot vector) multiplies the
  // tile values idx = w, w + WAVES, ... of every lane; the dot-vector entries of those values
  // are loaded before the K loop, so that their latency overlaps it instead of following it.
  constexpr int E = MI * NI * 4;
  constexpr int PER = (E + WAVES - 1) / WAVES;
  constexpr bool F64L = RealAcc<PREC, MI, NI>::F64;
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, blk.M - m0), nt = min(16 * NI, blk.N - n0);
  auto where = [&](int idx, int *row, int *col) {  // position of value idx of this lane
    const int ij = idx >> 2;
    *row = 16 * (ij / NI) + acc_row(F64L, g, idx & 3);
    *col = 16 * (ij % NI) + r;
  };
  double xpre[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) xpre[t] = 0.0;
  if (dslot) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int idx = wave + t * WAVES;
      int row, col;
      where(idx, &row, &col);
      const bool in = idx < E && row < mt && col < nt;
      if constexpr (PREC == CXA_PREC_FP32) {
        // branch-free (the zero cell outside the tile), so that the K loop's wait counts stay
        // exact (measured: FP32 launch 8.8 -> 8.2 us; the same form made the FP64 launch
        // 11 -> 14 us, which keeps the branch)
        xpre[t] = load_or_zero<double>(
            (gptr<const double>)(dotx + blk.c_off + (int64_t)(m0 + row) * blk.ldc + (n0 + col)),
            in);
      } else if (in) {
        xpre[t] = dotx[blk.c_off + (int64_t)(m0 + row) * blk.ldc + (n0 + col)];
      }
    }
  }
  mfma_accumulate<PREC, TA, TB, MI, NI, PK>(acc, Abase, Bbase, blk, segs, m0, n0, lane, wave,
                                        WAVES);
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int jj = 0; jj < NI; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave][(i * NI + jj) * 4 + v][lane] = acc.get(i, jj, v);
  __syncthreads();
  gptr<TC> Cp = at<TC>(Cbase, blk.c_off);
  double cx = 0.0, xx = 0.0;
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int idx = wave + t * WAVES;
    int row, col;
    where(idx, &row, &col);
    if (idx < E && row < mt && col < nt) {
      double s = red[0][idx][lane];
#pragma unroll
      for (int q = 1; q < WAVES; ++q) s += red[q][idx][lane];
      store_c(Cp + (int64_t)(m0 + row) * blk.ldc + (n0 + col), s, blk);
      cx += s * xpre[t];
      xx += xpre[t] * xpre[t];
    }
  }
  if (dslot) {  // uniform over the workgroup: wave partials, then their sum in wave order
    __shared__ double dred[2][WAVES];
    cx = wave_sum(cx);
    xx = wave_sum(xx);
    if (lane == 0) {
      dred[0][wave] = cx;
      dred[1][wave] = xx;
    }
    __syncthreads();
    if (wave == 0 && lane == 0) {
      double a = dred[0][0], b = dred[1][0];
#pragma unroll
      for (int q = 1; q < WAVES; ++q) {
        a += dred[0][q];
        b += dred[1][q];
      }
      dslot[0] = a;
      dslot[nslots] = b;
    }
  }
  __syncthreads();
}

// Walks the k values kk of a block's concatenated segments with kk % nslices == slice, in
// order (wave-uniform).
template <typename SEGS>
struct RowCursor {
  SEGS segs;
  int seg0, nseg;
  int s, k;
  GSeg sg;
  __device__ __forceinline__ void skip(int n) {
    k += n;
    while (s < nseg && k >= sg.K) {
      k -= sg.K;
      ++s;
      if (s < nseg) sg = segs(seg0 + s);
    }
  }
};

// Row streaming for N <= kRowsN: lane = output row, FP64 accumulation.  The k values of all
// segments form one sequence; kRowsDepth loads of A run ahead of the FMAs that consume them
// (B rows are wave-uniform reads).
constexpr int kRowsDepth = 8;

// element offset of k along a row of A (k4-packed: quad stride sak, 4 consecutive k adjacent)
template <int PK>
__device__ __forceinline__ int64_t a_koff(int k, int64_t sak) {
  return (PK & kPackA) ? (int64_t)(k >> 2) * sak + (k & 3) : (int64_t)k * sak;
}

template <typename TA, typename TB, int PK, typename SEGS>
__device__ __forceinline__ void rows_accumulate(double (&acc)[kRowsN], const char *Abase,
                                                const char *Bbase, const GBlk &blk, SEGS segs,
                                                int row, bool valid, int slice, int nslices) {
  const int N = blk.N;
#pragma unroll
  for (int n = 0; n < kRowsN; ++n) acc[n] = 0.0;
  int total = 0, kb = 0;
  for (int s = 0; s < blk.nseg; ++s) {
    const int K = segs(blk.seg0 + s).K;
    const int first = ((slice - kb) % nslices + nslices) % nslices;
    if (first < K) total += (K - first + nslices - 1) / nslices;
    kb += K;
  }
  if (total == 0) return;
  RowCursor<SEGS> lc{segs, blk.seg0, blk.nseg, 0, 0};
  lc.sg = segs(blk.seg0);
  lc.skip(slice);  // also steps over empty leading segments
  RowCursor<SEGS> cc = lc;
  TA a[kRowsDepth];  // as loaded; widened to FP64 where consumed
#pragma unroll
  for (int d = 0; d < kRowsDepth; ++d) {
    if (d < total) {
      a[d] = load_or_zero<TA>(at<const TA>(Abase, lc.sg.a_off) + ((int64_t)row * lc.sg.sam +
                                                                  a_koff<PK>(lc.k, lc.sg.sak)),
                              valid);
      lc.skip(nslices);
    }
  }
  for (int base = 0; base < total; base += kRowsDepth) {
#pragma unroll
    for (int d = 0; d < kRowsDepth; ++d) {
      const int c = base + d;
      if (c < total) {
        gptr<const TB> bk = at<const TB>(Bbase, cc.sg.b_off) +
                            ((PK & kPackB) ? (int64_t)(cc.k >> 2) * cc.sg.sbk + (cc.k & 3)
                                 : (int64_t)cc.k * cc.sg.sbk);
#pragma unroll
        for (int n = 0; n < kRowsN; ++n)
          if (n < N) acc[n] += (double)a[d] * (double)bk[(int64_t)n * cc.sg.sbn];
        cc.skip(nslices);
        if (c + kRowsDepth < total) {
          a[d] = load_or_zero<TA>(
              at<const TA>(Abase, lc.sg.a_off) +
                  ((int64_t)row * lc.sg.sam + a_koff<PK>(lc.k, lc.sg.sak)),
              valid);
          lc.skip(nslices);
        }
      }
    }
  }
  drain_loads();
}

template <typename TA, typename TB, typename TC, int PK, typename SEGS>
__device__ __forceinline__ void wave_rows(const char *__restrict__ Abase,
                                          const char *__restrict__ Bbase,
                                          char *__restrict__ Cbase, const GBlk &blk, SEGS segs,
                                          int m0, int lane) {
  const int row = m0 + lane;
  const bool valid = row < blk.M;
  double acc[kRowsN];
  rows_accumulate<TA, TB, PK>(acc, Abase, Bbase, blk, segs, row, valid, 0, 1);
  if (!valid) return;
  gptr<TC> Cp = at<TC>(Cbase, blk.c_off) + (int64_t)row * blk.ldc;
#pragma unroll
  for (int n = 0; n < kRowsN; ++n)
    if (n < blk.N) store_c(Cp + n, acc[n], blk);
}

// Split-K row tile: the four waves take interleaved k values of the same 64 rows.
template <typename TA, typename TB, typename TC, int PK, typename SEGS>
__device__ __forceinline__ void wg_rows(const char *__restrict__ Abase,
                                        const char *__restrict__ Bbase, char *__restrict__ Cbase,
                                        const GBlk &blk, SEGS segs, int m0, int lane, int wave,
                                        double (*red)[16][64]) {
  const int row = m0 + lane;
  const bool valid = row < blk.M;
  double acc[kRowsN];
  rows_accumulate<TA, TB, PK>(acc, Abase, Bbase, blk, segs, row, valid, wave, kWaves);
#pragma unroll
  for (int n = 0; n < kRowsN; ++n) red[wave][n][lane] = acc[n];
  __syncthreads();
  if (wave == 0 && valid) {
    gptr<TC> Cp = at<TC>(Cbase, blk.c_off) + (int64_t)row * blk.ldc;
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
// conj: bit 0 conjugates A, bit 1 conjugates B as they are loaded (the adjoint contractions).
template <typename CT, int MI, int NI, typename SEGS>
__device__ __forceinline__ void wave_tile_cplx(const char *__restrict__ Abase,
                                               const char *__restrict__ Bbase,
                                               char *__restrict__ Cbase, const GBlk &blk,
                                               SEGS segs, int m0, int n0, int lane, int conj) {
  constexpr bool F64 = std::is_same<CT, cd>::value;
  using acc_t = typename std::conditional<F64, f64x4, f32x4>::type;
  using op_t = typename std::conditional<F64, double, float>::type;
  // complex elements move as two-wide vectors (x = real, y = imaginary; same size and
  // alignment as CT), which load and store through global-address-space pointers
  typedef op_t c2 __attribute__((ext_vector_type(2)));
  static_assert(sizeof(c2) == sizeof(CT), "complex element layout");
  const op_t sa = (conj & 1) ? op_t(-1) : op_t(1), sb = (conj & 2) ? op_t(-1) : op_t(1);
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
    gptr<const c2> Ap = at<const c2>(Abase, sg.a_off) + (int64_t)m0 * sg.sam;
    gptr<const c2> Bp = at<const c2>(Bbase, sg.b_off) + (int64_t)n0 * sg.sbn;
    const int K = sg.K;
    for (int k0 = 0; k0 < K; k0 += 16) {
      c2 va[MI][4], vb[NI][4];  // every load of the chunk issued before the first use
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int row = 16 * i + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + 4 * g + j;
          va[i][j] = load_or_zero<c2>(Ap + ((int64_t)row * sg.sam + (int64_t)k * sg.sak),
                                      row < mt && k < K);
        }
      }
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int col = 16 * i + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + 4 * g + j;
          vb[i][j] = load_or_zero<c2>(Bp + ((int64_t)k * sg.sbk + (int64_t)col * sg.sbn),
                                      col < nt && k < K);
        }
      }
      op_t ar[MI][4], ai[MI][4], br[NI][4], bi[NI][4];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ar[i][j] = va[i][j].x;
          ai[i][j] = sa * va[i][j].y;
        }
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          br[i][j] = vb[i][j].x;
          bi[i][j] = sb * vb[i][j].y;
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
  gptr<c2> Cp = at<c2>(Cbase, blk.c_off);
  // complex alpha and beta (Gemm_Batch passes Scalars); contractions use alpha = 1, beta = 0
  const double ar_ = blk.alpha, ai_ = blk.alpha_im, br_ = blk.beta, bi_ = blk.beta_im;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int jj = 0; jj < NI; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + acc_row(F64, g, v);
        const int col = 16 * jj + r;
        if (row < mt && col < nt) {
          const double xr = (double)accr[i][jj][v], xi = (double)acci[i][jj][v];
          double re = ar_ * xr - ai_ * xi, im = ar_ * xi + ai_ * xr;
          gptr<c2> dst = Cp + (int64_t)(m0 + row) * blk.ldc + (n0 + col);
          if (br_ != 0.0 || bi_ != 0.0) {
            const c2 o = *dst;
            re += br_ * (double)o.x - bi_ * (double)o.y;
            im += br_ * (double)o.y + bi_ * (double)o.x;
          }
          c2 w;
          w.x = (op_t)re;
          w.y = (op_t)im;
          *dst = w;
        }
      }
}

// Tile classes: each class is its own kernel instantiation so that it carries only its own
// register budget (occupancy matters most for the memory-bound row tiles).
enum TileClass { TC_MIXED = 0, TC_ROWS = 1, TC_BIG = 2 };

template <int PREC, typename TA, typename TB, typename TC, bool CPLX, int CLS, int PK,
          typename SEGS>
__device__ __forceinline__ void run_tile(int kind, const char *A, const char *B, char *C,
                                         const GBlk &blk, SEGS segs, int m0, int n0, int lane,
                                         int conj, const double *dotx = nullptr,
                                         double *dslot = nullptr, int64_t nslots = 0) {
  if constexpr (CPLX) {
    switch (kind) {
      case TK_1x1: wave_tile_cplx<TA, 1, 1>(A, B, C, blk, segs, m0, n0, lane, conj); break;
      case TK_2x1: wave_tile_cplx<TA, 2, 1>(A, B, C, blk, segs, m0, n0, lane, conj); break;
      case TK_1x2: wave_tile_cplx<TA, 1, 2>(A, B, C, blk, segs, m0, n0, lane, conj); break;
      default: wave_tile_cplx<TA, 2, 2>(A, B, C, blk, segs, m0, n0, lane, conj); break;
    }
  } else if constexpr (CLS == TC_ROWS) {
    wave_rows<TA, TB, TC, PK>(A, B, C, blk, segs, m0, lane);
  } else if constexpr (CLS == TC_BIG) {
    wave_tile_real<PREC, TA, TB, TC, 4, 2, PK>(A, B, C, blk, segs, m0, n0, lane, dotx, dslot,
                                               nslots);
  } else {
    switch (kind) {
      case TK_1x1:
        wave_tile_real<PREC, TA, TB, TC, 1, 1, PK>(A, B, C, blk, segs, m0, n0, lane, dotx, dslot,
                                                   nslots);
        break;
      case TK_2x1:
        wave_tile_real<PREC, TA, TB, TC, 2, 1, PK>(A, B, C, blk, segs, m0, n0, lane, dotx, dslot,
                                                   nslots);
        break;
      case TK_1x2:
        wave_tile_real<PREC, TA, TB, TC, 1, 2, PK>(A, B, C, blk, segs, m0, n0, lane, dotx, dslot,
                                                   nslots);
        break;
      default:
        wave_tile_real<PREC, TA, TB, TC, 2, 2, PK>(A, B, C, blk, segs, m0, n0, lane, dotx, dslot,
                                                   nslots);
        break;
    }
  }
}

// grouped launch: tiles, blocks and segments come from device arrays; one wave per tile.
// Workgroup w runs on XCD w % kXcds and takes tiles from that XCD's range of the list.
template <int PREC, typename TA, typename TB, typename TC, bool CPLX, int CLS, int PK>
__global__ __launch_bounds__(64 * kWaves) void k_grouped(const GTile *__restrict__ tiles,
                                                        int per_xcd, const GBlk *__restrict__ blks,
                                                        const GSeg *__restrict__ segs,
                                                        const char *A, const char *B, char *C,
                                                        int conj, const double *dotx,
                                                        double *dotp) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = blockIdx.x % kXcds;
  const GTile *range = tiles + (int64_t)x * per_xcd;
  const int stride = (gridDim.x / kXcds) * kWaves;
  for (int t = (blockIdx.x / kXcds) * kWaves + wave; t < per_xcd; t += stride) {
    const GTile tl = range[t];
    double *dslot = dotp ? dotp + (int64_t)x * per_xcd + t : nullptr;
    if (tl.blk < 0) {  // padding
      if (dslot && lane == 0) dslot[0] = dslot[(int64_t)kXcds * per_xcd] = 0.0;
      continue;
    }
    const GBlk blk = blks[tl.blk];
    run_tile<PREC, TA, TB, TC, CPLX, CLS, PK>(tl.kind, A, B, C, blk, SegArray{segs}, tl.m0,
                                               tl.n0, lane, conj, dotx, dslot,
                                               (int64_t)kXcds * per_xcd);
  }
}

// grouped launch with workgroup split-K: one workgroup per tile (real dtypes; the mixed MFMA
// class or the row class)
template <int PREC, typename TA, typename TB, typename TC, int CLS, int WAVES, int PK>
__global__ __launch_bounds__(64 * WAVES) void k_grouped_split(
    const GTile *__restrict__ tiles, int per_xcd, const GBlk *__restrict__ blks,
    const GSeg *__restrict__ segs, const char *A, const char *B, char *C,
    const double *dotx, double *dotp) {
  static_assert(CLS != TC_ROWS || WAVES == kWaves, "row tiles split over four waves");
  __shared__ double red[WAVES][16][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = blockIdx.x % kXcds;
  const GTile *range = tiles + (int64_t)x * per_xcd;
  const int stride = gridDim.x / kXcds;
  for (int t = blockIdx.x / kXcds; t < per_xcd; t += stride) {
    const GTile tl = range[t];
    double *dslot = dotp ? dotp + (int64_t)x * per_xcd + t : nullptr;
    if (tl.blk < 0) {  // padding (uniform over the workgroup)
      if (dslot && threadIdx.x == 0) dslot[0] = dslot[(int64_t)kXcds * per_xcd] = 0.0;
      continue;
    }
    const GBlk blk = blks[tl.blk];
    const SegArray sa{segs};
    if constexpr (CLS == TC_ROWS) {
      wg_rows<TA, TB, TC, PK>(A, B, C, blk, sa, tl.m0, lane, wave, red);
    } else {
      switch (tl.kind) {
        case TK_1x1:
          wg_tile_real<PREC, TA, TB, TC, 1, 1, WAVES, PK>(A, B, C, blk, sa, tl.m0, tl.n0, lane,
                                                          wave,
                                                     red, dotx, dslot,
                                                     (int64_t)kXcds * per_xcd);
          break;
        case TK_2x1:
          wg_tile_real<PREC, TA, TB, TC, 2, 1, WAVES, PK>(A, B, C, blk, sa, tl.m0, tl.n0, lane,
                                                          wave,
                                                     red, dotx, dslot,
                                                     (int64_t)kXcds * per_xcd);
          break;
        case TK_1x2:
          wg_tile_real<PREC, TA, TB, TC, 1, 2, WAVES, PK>(A, B, C, blk, sa, tl.m0, tl.n0, lane,
                                                          wave,
                                                     red, dotx, dslot,
                                                     (int64_t)kXcds * per_xcd);
          break;
        default:
          wg_tile_real<PREC, TA, TB, TC, 2, 2, WAVES, PK>(A, B, C, blk, sa, tl.m0, tl.n0, lane,
                                                          wave,
                                                     red, dotx, dslot,
                                                     (int64_t)kXcds * per_xcd);
          break;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// mix launches: one workgroup per tile of an item, of kMixSlices waves (wave s computes slice s:
// the longest dependent load chain of a tile is one slice's K range) or of 4 waves (wave w
// computes slices w and w + 4: no idle waves on items with up to 4 slices); the slice tiles meet
// in LDS, and the waves then write the item's outputs, each a fixed-order weighted sum of the
// slice tiles.
// ---------------------------------------------------------------------------------------
struct SegStride {
  const GSeg *p;
  int base, stride;
  __device__ __forceinline__ GSeg operator()(int i) const { return p[base + i * stride]; }
};

// Segment i of `segs` (nseg <= NACC) accumulates into acc[i]; one load pipeline runs across
// all of them (as in mfma_accumulate).
template <int PREC, typename TA, typename TB, int MI, int NI, int NACC, int PK, typename SEGS>
__device__ __forceinline__ void mfma_accumulate_multi(RealAcc<PREC, MI, NI> (&acc)[NACC],
                                                      const char *__restrict__ Abase,
                                                      const char *__restrict__ Bbase, SEGS segs,
                                                      int nseg, int M, int N, int m0, int n0,
                                                      int lane) {
  constexpr int D = pipe_depth<PREC, MI, NI>();
  const int r = lane & 15, g = lane >> 4;
  const int mt = min(16 * MI, M - m0);
  const int nt = min(16 * NI, N - n0);
  int nc[NACC];
  int total = 0;
#pragma unroll
  for (int q = 0; q < NACC; ++q) {
    nc[q] = q < nseg ? (segs(q).K + 15) >> 4 : 0;
    total += nc[q];
  }
  if (total == 0) return;
  ChunkCursor<TA, TB, SEGS> cur{segs, 0, nseg, 1, m0, n0, 0, 0};
  cur.enter(Abase, Bbase);
  cur.skip(0, Abase, Bbase);
  int cs = 0, left = nc[0];
#pragma unroll
  for (int q = 1; q < NACC; ++q)
    if (left == 0 && cs == q - 1) {
      cs = q;
      left = nc[q];
    }
  Frag<TA, TB, MI, NI> f[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (d < total) {
      cur.template load<MI, NI, PK>(f[d], mt, nt, r, g);
      cur.skip(1, Abase, Bbase);
    }
  }
  // chunk -> accumulator bookkeeping
  auto consume = [&](const Frag<TA, TB, MI, NI> &fr) {
#pragma unroll
    for (int q = 0; q < NACC; ++q)
      if (q == cs) mma_frag<PREC, TA, TB, MI, NI>(acc[q], fr);
    if (--left == 0) {
#pragma unroll
      for (int q = 1; q < NACC; ++q)
        if (left == 0 && cs == q - 1) {
          cs = q;
          left = nc[q];
        }
    }
  };
  for (int base = 0; base < total; base += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int c = base + d;
      if (c < total) {
        consume(f[d]);
        if (c + D < total) {
          cur.template load<MI, NI, PK>(f[d], mt, nt, r, g);
          cur.skip(1, Abase, Bbase);
        }
      }
    }
  }
  drain_loads();
}

template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, int WAVES, int PK>
__global__ __launch_bounds__(64 * WAVES) void k_mix(const GTile *__restrict__ tiles,
                                                   int per_xcd, const GBlk *__restrict__ items,
                                                   const GSeg *__restrict__ slices,
                                                   const MixOut *__restrict__ outs,
                                                   const char *A, const char *B, char *C) {
  using acc_t = typename std::conditional<PREC == CXA_PREC_FP32, float, double>::type;
  constexpr int E = MI * NI * 4;              // accumulator values per lane and slice
  constexpr int NACC = kMixSlices / WAVES;    // slices per wave
  constexpr bool F64 = RealAcc<PREC, MI, NI>::F64;
  __shared__ acc_t xs[kMixSlices][E][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const GTile *range = tiles + (int64_t)(blockIdx.x % kXcds) * per_xcd;
  const int stride = gridDim.x / kXcds;
  for (int t = blockIdx.x / kXcds; t < per_xcd; t += stride) {
    const GTile tl = range[t];
    if (tl.blk < 0) continue;  // padding (uniform over the workgroup)
    const GBlk it = items[tl.blk];
    const int mine = it.nseg > wave ? (it.nseg - wave + WAVES - 1) / WAVES : 0;
    // this wave's first output record, loaded before the K loop so that its latency overlaps
    // the loop instead of following it
    MixOut o_first{};
    if (wave < it.nout) o_first = outs[it.out0 + wave];
    if (mine > 0) {
      RealAcc<PREC, MI, NI> acc[NACC];
#pragma unroll
      for (int q = 0; q < NACC; ++q) acc[q].zero();
      mfma_accumulate_multi<PREC, TA, TB, MI, NI, NACC, PK>(
          acc, A, B, SegStride{slices, it.seg0 + wave, WAVES}, mine, it.M, it.N, tl.m0, tl.n0,
          lane);
#pragma unroll
      for (int q = 0; q < NACC; ++q) {
        const int sl = wave + q * WAVES;
        if (sl < it.nseg) {
#pragma unroll
          for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int jj = 0; jj < NI; ++jj)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                xs[sl][(i * NI + jj) * 4 + v][lane] = (acc_t)acc[q].get(i, jj, v);
        }
      }
    }
    __syncthreads();
    // Each lane writes rows 16i + 4g .. 16i + 4g + 3 of its column: the slice tiles are in LDS,
    // so the values come from whichever lane and register held those rows (the FP64 MFMA
    // layout spreads four consecutive rows over the four lane groups), and a packed output
    // gets them as one k-quad store where the quad is aligned and whole.
    const int mt = min(16 * MI, it.M - tl.m0), nt = min(16 * NI, it.N - tl.n0);
    for (int j = wave; j < it.nout; j += WAVES) {
      const MixOut o = j == wave ? o_first : outs[it.out0 + j];
      gptr<TC> Cp = at<TC>(C, o.c_off);
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int jj = 0; jj < NI; ++jj) {
          const int col = 16 * jj + r;
          const int row0 = 16 * i + 4 * g;
          TC vals[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            // row row0 + v sits in lane group gs, register slot vs (acc_row(F64, gs, vs))
            const int gs = F64 ? v : g, vs = F64 ? g : v;
            const int e = (i * NI + jj) * 4 + vs;
            const int src = r + 16 * gs;
            double val = 0.0;
#pragma unroll
            for (int q = 0; q < kMixSlices; ++q)
              if (q < it.nseg) val += o.coef[q] * (double)xs[q][e][src];
            vals[v] = (TC)val;
          }
          if (col < nt) {
            if constexpr ((PK & kPackC) != 0) {  // c_off: region base, ldc: first k
              const int64_t kk = o.ldc + tl.m0 + row0;
              const int64_t cpos = (int64_t)(tl.n0 + col) * 4;
              const int64_t qs = 4 * (int64_t)it.N;
              if ((kk & 3) == 0 && row0 + 3 < mt) {
                typedef TC tc4 __attribute__((ext_vector_type(4)));
                tc4 w;
                w.x = vals[0];
                w.y = vals[1];
                w.z = vals[2];
                w.w = vals[3];
                *(gptr<tc4>)(Cp + ((kk >> 2) * qs + cpos)) = w;
              } else {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  if (row0 + v < mt)
                    Cp[((kk + v) >> 2) * qs + cpos + ((kk + v) & 3)] = vals[v];
              }
            } else {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (row0 + v < mt)
                  Cp[(int64_t)(tl.m0 + row0 + v) * o.ldc + (tl.n0 + col)] = vals[v];
            }
          }
        }
    }
    __syncthreads();  // the next tile reuses xs
  }
}

// Packed (k4) B operands are built for the fixed operands of a chain: FP32 copies in FP32 mode,
// the FP64 tensors otherwise; only those kernel variants exist.
template <int PREC, typename TB>
constexpr bool k4_ok() {
  return PREC == CXA_PREC_FP32 ? std::is_same<TB, float>::value
                               : std::is_same<TB, double>::value;
}

template <int PREC, typename TA, typename TB, typename TC, int MI, int NI, int PK>
void launch_mix_kind_t(const GemmLaunch &g, int grid, hipStream_t s) {
  const char *A = (const char *)g.A, *B = (const char *)g.B;
  char *C = (char *)g.C;
  if (g.splitk)  // four waves, two slices each
    hipLaunchKernelGGL((k_mix<PREC, TA, TB, TC, MI, NI, 4, PK>), dim3(grid), dim3(64 * 4), 0,
                       s, g.tiles, g.per_xcd, g.blks, g.segs, g.outs, A, B, C);
  else  // one wave per slice
    hipLaunchKernelGGL((k_mix<PREC, TA, TB, TC, MI, NI, kMixSlices, PK>), dim3(grid),
                       dim3(64 * kMixSlices), 0, s, g.tiles, g.per_xcd, g.blks, g.segs, g.outs,
                       A, B, C);
}

// A packed (k4) A operand of a mix launch is the FP32 running vector of an FP32 chain, read
// with packed B and written to packed outputs; only that variant exists.
template <int PREC, typename TA, typename TB, typename TC>
constexpr bool mix_a4_ok() {
  return PREC == CXA_PREC_FP32 && std::is_same<TA, float>::value && k4_ok<PREC, TB>();
}

// mix layouts: plain, packed B, packed B with packed outputs, packed A and B with packed
// outputs (launch_mix_t checks the flags)
template <int PREC, typename TA, typename TB, typename TC, int MI, int NI>
void launch_mix_kind(const GemmLaunch &g, int grid, hipStream_t s) {
  if constexpr (mix_a4_ok<PREC, TA, TB, TC>()) {
    if (g.a_k4) {
      launch_mix_kind_t<PREC, TA, TB, TC, MI, NI, kPackA | kPackB | kPackC>(g, grid, s);
      return;
    }
  }
  if constexpr (k4_ok<PREC, TB>()) {
    if (g.b_k4 && g.c_k4) {
      launch_mix_kind_t<PREC, TA, TB, TC, MI, NI, kPackB | kPackC>(g, grid, s);
      return;
    }
    if (g.b_k4) {
      launch_mix_kind_t<PREC, TA, TB, TC, MI, NI, kPackB>(g, grid, s);
      return;
    }
  }
  launch_mix_kind_t<PREC, TA, TB, TC, MI, NI, 0>(g, grid, s);
}

template <int PREC, typename TA, typename TB, typename TC>
int launch_mix_t(const GemmLaunch &g, hipStream_t s) {
  if (g.per_xcd <= 0) return CXA_OK;
  if (g.ntiles != kXcds * g.per_xcd)
    return fail(CXA_ERR_INVALID, "mix GEMM: tile list of %d is not %d ranges of %d", g.ntiles,
                kXcds, g.per_xcd);
  if ((g.b_k4 && !k4_ok<PREC, TB>()) || (g.c_k4 && !g.b_k4) ||
      (g.a_k4 && !(mix_a4_ok<PREC, TA, TB, TC>() && g.b_k4 && g.c_k4)))
    return fail(CXA_ERR_UNSUPPORTED, "mix GEMM: no variant for these packed layouts / dtypes");
  const int grid = kXcds * std::min(g.per_xcd, 2048);
  switch (g.kind0) {
    case TK_1x1: launch_mix_kind<PREC, TA, TB, TC, 1, 1>(g, grid, s); break;
    case TK_2x1: launch_mix_kind<PREC, TA, TB, TC, 2, 1>(g, grid, s); break;
    case TK_1x2: launch_mix_kind<PREC, TA, TB, TC, 1, 2>(g, grid, s); break;
    case TK_2x2: launch_mix_kind<PREC, TA, TB, TC, 2, 2>(g, grid, s); break;
    default:
      return fail(CXA_ERR_INVALID, "mix GEMM: tile kind %d is not 16x16, 32x16, 16x32 or 32x32",
                  g.kind0);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

// single-block launch: one block with one segment, tiles derived from the launch index, so a
// plain GEMM needs no device-side task list (mixed MFMA class)
template <int PREC, typename TA, typename TB, typename TC, bool CPLX>
__global__ __launch_bounds__(64 * kWaves) void k_single(GBlk blk, GSeg seg, int kind,
                                                       int tiles_n, int ntiles, const char *A,
                                                       const char *B, char *C, int conj) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tr = kind == TK_1x1 || kind == TK_1x2 ? 16 : 32;
  const int tc = kind == TK_1x1 || kind == TK_2x1 ? 16 : 32;
  for (int t = blockIdx.x * kWaves + wave; t < ntiles; t += gridDim.x * kWaves) {
    const int m0 = (t / tiles_n) * tr, n0 = (t % tiles_n) * tc;
    run_tile<PREC, TA, TB, TC, CPLX, TC_MIXED, 0>(kind, A, B, C, blk, SegOne{seg}, m0, n0,
                                                      lane,
                                               conj);
  }
}

// the real-dtype grouped launches of one class / split mode
template <int PREC, typename TA, typename TB, typename TC, int PK>
void launch_real(const GemmLaunch &g, int cls, int wg_grid, int wave_grid, hipStream_t s) {
  const int P = g.per_xcd;
  const char *A = (const char *)g.A, *B = (const char *)g.B;
  char *C = (char *)g.C;
  const double *dx = g.dot_x;
  double *dp = g.dot_part;
  if (g.splitk && cls == TC_ROWS)
    hipLaunchKernelGGL((k_grouped_split<PREC, TA, TB, TC, TC_ROWS, kWaves, PK>), dim3(wg_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, nullptr,
                       nullptr);
  else if (g.splitk == 8 && cls == TC_MIXED)
    hipLaunchKernelGGL((k_grouped_split<PREC, TA, TB, TC, TC_MIXED, 8, PK>), dim3(wg_grid),
                       dim3(64 * 8), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, dx, dp);
  else if (g.splitk && cls == TC_MIXED)
    hipLaunchKernelGGL((k_grouped_split<PREC, TA, TB, TC, TC_MIXED, kWaves, PK>), dim3(wg_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, dx, dp);
  else if (cls == TC_ROWS)
    hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_ROWS, PK>), dim3(wave_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, g.conj,
                       nullptr, nullptr);
  else if (cls == TC_BIG)
    hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_BIG, PK>), dim3(wave_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, g.conj, dx,
                       dp);
  else
    hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, false, TC_MIXED, PK>), dim3(wave_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs, A, B, C, g.conj, dx,
                       dp);
}

template <int PREC, typename TA, typename TB, typename TC, bool CPLX>
int launch_t(const GemmLaunch &g, hipStream_t s) {
  if (g.dot_part) {  // the dot epilogue: real MFMA tile classes with an FP64 C only
    bool ok = !CPLX && !g.single && g.tile_class != TC_ROWS && g.dot_x != nullptr;
    if constexpr (!std::is_same<TC, double>::value) ok = false;
    if (!ok)
      return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: no dot epilogue for this launch");
  }
  if (g.single) {
    if (g.a_k4 || g.b_k4 || g.c_k4)
      return fail(CXA_ERR_INVALID, "single-block GEMM: packed layouts are not supported");
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
                       (char *)g.C, g.conj);
    CXA_HIP(hipGetLastError());
    return CXA_OK;
  }
  if (g.per_xcd <= 0) return CXA_OK;
  if (g.ntiles != kXcds * g.per_xcd)
    return fail(CXA_ERR_INVALID, "grouped GEMM: tile list of %d is not %d ranges of %d",
                g.ntiles, kXcds, g.per_xcd);
  if (g.grid % kXcds)
    return fail(CXA_ERR_INVALID, "grouped GEMM: grid %d is not a multiple of %d", g.grid, kXcds);
  const int cls = g.tile_class;
  if (CPLX && cls != TC_MIXED)
    return fail(CXA_ERR_UNSUPPORTED, "complex GEMM tiles must be MFMA tiles up to 32x32");
  const int P = g.per_xcd;
  const int wg_grid = g.grid > 0 ? g.grid : kXcds * std::min(P, 2048);
  const int wave_grid = g.grid > 0 ? g.grid : kXcds * std::min((P + kWaves - 1) / kWaves, 1024);
  if constexpr (!CPLX) {
    if (g.c_k4 || (g.a_k4 && !g.b_k4))
      return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: no variant for these packed layouts");
    if (g.a_k4) {
      if constexpr (k4_ok<PREC, TA>() && k4_ok<PREC, TB>())
        launch_real<PREC, TA, TB, TC, kPackA | kPackB>(g, cls, wg_grid, wave_grid, s);
      else
        return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: no packed-A variant for these dtypes");
    } else if (g.b_k4) {
      if constexpr (k4_ok<PREC, TB>())
        launch_real<PREC, TA, TB, TC, kPackB>(g, cls, wg_grid, wave_grid, s);
      else
        return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: no packed-B variant for these dtypes");
    } else {
      launch_real<PREC, TA, TB, TC, 0>(g, cls, wg_grid, wave_grid, s);
    }
  } else {
    if (g.a_k4 || g.b_k4 || g.c_k4)
      return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: packed layouts are real-only");
    hipLaunchKernelGGL((k_grouped<PREC, TA, TB, TC, true, TC_MIXED, 0>), dim3(wave_grid),
                       dim3(64 * kWaves), 0, s, g.tiles, P, g.blks, g.segs,
                       (const char *)g.A, (const char *)g.B, (char *)g.C, g.conj, nullptr,
                       nullptr);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

}  // namespace

// precision-family launchers (one translation unit each)
int launch_gemm_fp64(const GemmLaunch &g, hipStream_t s);
int launch_gemm_fp32(const GemmLaunch &g, hipStream_t s);
int launch_gemm_fp32_a64(const GemmLaunch &g, hipStream_t s);
int launch_gemm_fp32x3(const GemmLaunch &g, hipStream_t s);
int launch_mix(const GemmLaunch &g, hipStream_t s);

}  // namespace cxa
