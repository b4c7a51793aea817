// Physical permute (Tensor::contiguous_) for gfx950.
//
// Host side reduces the permutation (drops unit axes, merges old axes that stay adjacent and in
// order), then picks one of three kernels:
//   * a plain device copy when the reduced permutation is the identity;
//   * a row copy when the innermost axis stays innermost (rows of the inner extent move as
//     16-byte vectors when the row length allows);
//   * an LDS-tiled transpose otherwise: a tile spans a run of the input's trailing axes ("A",
//     contiguous in the input) by a run of the output's trailing axes ("B", contiguous in the
//     output).  Reads stream along A, writes stream along B, the tile in LDS turns one into the
//     other; per-tile offsets of the A and B indices are decoded once into LDS.  For elements of
//     8 bytes or less each workgroup walks several tiles with the next tile's loads issued
//     before the current tile's stores (software pipelining); 16-byte elements go one tile at
//     a time.
// The semantics are those of the reference's MoveMemoryCpu / MoveMemoryGpu
// (src/backend/utils_internal_cpu/Movemem_cpu.cpp:18-211,
//  src/backend/utils_internal_gpu/cuMovemem_gpu.cu:60-304): out = transpose(in, mapper).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "contract.hpp"
#include "cxa_common.hpp"

namespace cxa {
namespace {

// A 16-byte element (and the row copies' 16-byte vector): two 64-bit lanes of a vector type, so
// that the kernels' register staging arrays of it stay in registers (an array of a 16-byte struct
// was placed in scratch memory, which tripled the tiled kernels' memory traffic).
typedef uint64_t U128 __attribute__((ext_vector_type(2)));

constexpr int kMaxOuter = 14;
constexpr int kMaxGroup = 8;
constexpr int kThreads = 256;
constexpr int kMaxTile = 256;  // tile extent bound along A and along B
constexpr int kMinTile = 1024;  // elements a tile should span when the axes allow

struct TileArgs {
  int na, nb, no;
  int w, h;  // tile width along A, height along B
  int64_t a_ext[kMaxGroup], a_ostr[kMaxGroup];
  int64_t b_ext[kMaxGroup], b_istr[kMaxGroup];
  int64_t o_ext[kMaxOuter], o_istr[kMaxOuter], o_ostr[kMaxOuter];
  int64_t prodA, prodB, nca, ncb, ntiles;
};

struct GatherArgs {
  int rank;  // reduced rank
  int64_t total;
  int64_t n_ext[16];   // output (new) extents
  int64_t n_istr[16];  // input stride of each new axis
};

template <typename T, int MAXE>
__global__ __launch_bounds__(kThreads) void permute_tiled(const T *__restrict__ src,
                                                          T *__restrict__ dst, TileArgs p) {
  __shared__ T tile[MAXE + 64];
  __shared__ int64_t aoff[1024];
  __shared__ int64_t boff[1024];
  const int w = p.w, h = p.h, ws = w + 1;
  for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    int64_t rem = t;
    const int64_t ca = rem % p.nca;
    rem /= p.nca;
    const int64_t cb = rem % p.ncb;
    rem /= p.ncb;
    int64_t oin = 0, oout = 0;
    for (int i = p.no - 1; i >= 0; --i) {
      const int64_t e = p.o_ext[i];
      const int64_t x = rem % e;
      rem /= e;
      oin += x * p.o_istr[i];
      oout += x * p.o_ostr[i];
    }
    const int64_t a0 = ca * w, b0 = cb * h;
    const int tw = (int)min((int64_t)w, p.prodA - a0);
    const int th = (int)min((int64_t)h, p.prodB - b0);
    __syncthreads();  // the previous tile is fully written out
    for (int i = threadIdx.x; i < w + h; i += kThreads) {
      if (i < w) {
        int64_t a = a0 + i, off = 0;
        for (int k = p.na - 1; k >= 0; --k) {
          off += (a % p.a_ext[k]) * p.a_ostr[k];
          a /= p.a_ext[k];
        }
        aoff[i] = off;
      } else {
        int64_t b = b0 + (i - w), off = 0;
        for (int k = p.nb - 1; k >= 0; --k) {
          off += (b % p.b_ext[k]) * p.b_istr[k];
          b /= p.b_ext[k];
        }
        boff[i - w] = off;
      }
    }
    __syncthreads();
    // read: consecutive threads walk A, which is contiguous in the input
    const T *s = src + oin + a0;
    for (int i = threadIdx.x; i < w * h; i += kThreads) {
      const int a = i % w, b = i / w;
      if (a < tw && b < th) tile[b * ws + a] = s[boff[b] + a];
    }
    __syncthreads();
    // write: consecutive threads walk B, which is contiguous in the output
    T *d = dst + oout + b0;
    for (int i = threadIdx.x; i < w * h; i += kThreads) {
      const int b = i % h, a = i / h;
      if (a < tw && b < th) d[aoff[a] + b] = tile[b * ws + a];
    }
  }
}

// Row copy / generic gather: each thread moves V consecutive output elements that share one
// input row (V divides the inner extent), decoding the output index once per vector.
template <typename T, int V>
__global__ __launch_bounds__(kThreads) void permute_gather(const T *__restrict__ src,
                                                           T *__restrict__ dst, GatherArgs p) {
  const int64_t nvec = p.total / V;
  for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < nvec;
       v += (int64_t)gridDim.x * kThreads) {
    int64_t o = v * V, rem = o, in = 0;
    for (int k = p.rank - 1; k >= 0; --k) {
      in += (rem % p.n_ext[k]) * p.n_istr[k];
      rem /= p.n_ext[k];
    }
    if constexpr (V == 1) {
      dst[o] = src[in];
    } else {
      // inner axis unchanged: the V elements are contiguous in the input as well
#pragma unroll
      for (int j = 0; j < V; ++j) dst[o + j] = src[in + j];
    }
  }
}

// ---------------------------------------------------------------------------------------
// 32-bit forms (tensors below 2^31 elements, the common case): every index decode is a
// multiply-high, an add and a shift instead of 64-bit division, and each thread stages several
// elements in registers so that its loads are in flight together.
// ---------------------------------------------------------------------------------------

// Division of n < 2^31 by a run-time invariant d in [1, 2^31] (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, section 4):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(n, m) + n) >> l.
struct FastDiv {
  uint32_t d, m, l;
  static FastDiv make(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.l = 0;
    while ((uint64_t(1) << f.l) < d) ++f.l;
    f.m = (uint32_t)((((uint64_t(1) << 32) * ((uint64_t(1) << f.l) - d)) / d) + 1);
    return f;
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    return (__umulhi(n, m) + n) >> l;
  }
};

// A tile is w (along A) x h (along B) x c: c > 1 batches consecutive indices of the innermost
// remaining axis (outer axis no - 1, extent cext, strides c_istr / c_ostr) into one tile, for
// permutations whose A and B runs are short and whole (a batch of small transposes, e.g. the last
// two axes swapped): the tile then spans c * w * h elements, contiguous in the input and in the
// output when that axis comes right before the A and B axes in both.
struct TileArgs32 {
  int na, nb, no, w, h, c;
  uint32_t prodA, prodB, ntiles, cext, c_istr, c_ostr;
  FastDiv dnca, dncb, dw, dh, dwh;
  FastDiv a_div[kMaxGroup];  // A axes (input-contiguous), outermost first
  uint32_t a_ostr[kMaxGroup];
  FastDiv b_div[kMaxGroup];  // B axes (output-contiguous), outermost first
  uint32_t b_istr[kMaxGroup];
  FastDiv o_div[kMaxOuter];  // remaining axes
  uint32_t o_istr[kMaxOuter], o_ostr[kMaxOuter];
};

// decode x = sum_k idx_k * stride_k for a mixed-radix index (outermost digit first)
__device__ __forceinline__ uint32_t decode32(uint32_t x, int n, const FastDiv *div,
                                             const uint32_t *stride) {
  uint32_t off = 0;
  for (int k = n - 1; k >= 0; --k) {
    const uint32_t q = div[k].div(x);
    off += (x - q * div[k].d) * stride[k];
    x = q;
  }
  return off;
}

struct TileAt {
  uint32_t oin, oout, a0, b0, tw, th, tc;
};

__device__ __forceinline__ TileAt tile_at(const TileArgs32 &p, uint32_t t) {
  TileAt x;
  uint32_t rem = t, q = p.dnca.div(rem);
  const uint32_t ca = rem - q * p.dnca.d;
  rem = q;
  q = p.dncb.div(rem);
  const uint32_t cb = rem - q * p.dncb.d;
  rem = q;
  x.oin = 0;
  x.oout = 0;
  x.tc = 1;
  for (int i = p.no - 1; i >= 0; --i) {
    q = p.o_div[i].div(rem);
    const uint32_t o = rem - q * p.o_div[i].d;
    rem = q;
    // with batching, the innermost outer axis's digit counts chunks of c
    const uint32_t o0 = (i == p.no - 1) ? o * (uint32_t)p.c : o;
    if (i == p.no - 1) x.tc = min((uint32_t)p.c, p.cext - o0);
    x.oin += o0 * p.o_istr[i];
    x.oout += o0 * p.o_ostr[i];
  }
  x.a0 = ca * p.w;
  x.b0 = cb * p.h;
  x.tw = min((uint32_t)p.w, p.prodA - x.a0);
  x.th = min((uint32_t)p.h, p.prodB - x.b0);
  return x;
}

// Element i of a tile in read order (a fastest, then b, then the batch index cc) and in write
// order (b fastest, then a, then cc).
struct TileIdx {
  uint32_t cc, a, b;
};
// BATCH false (c = 1): no batch digit to split off.
template <bool BATCH>
__device__ __forceinline__ TileIdx read_idx(const TileArgs32 &p, uint32_t i) {
  TileIdx r;
  r.cc = BATCH ? p.dwh.div(i) : 0;
  const uint32_t e = BATCH ? i - r.cc * p.dwh.d : i;
  r.b = p.dw.div(e);
  r.a = e - r.b * p.w;
  return r;
}
template <bool BATCH>
__device__ __forceinline__ TileIdx write_idx(const TileArgs32 &p, uint32_t i) {
  TileIdx r;
  r.cc = BATCH ? p.dwh.div(i) : 0;
  const uint32_t e = BATCH ? i - r.cc * p.dwh.d : i;
  r.a = p.dh.div(e);
  r.b = e - r.a * p.h;
  return r;
}
template <bool BATCH>
__device__ __forceinline__ bool in_tile(const TileIdx &r, const TileAt &x) {
  return r.a < x.tw && r.b < x.th && (!BATCH || r.cc < x.tc);
}

// One tile at a time: offsets, loads, LDS, stores (16-byte elements, where the pipelined form's
// extra registers and offset tables cost more occupancy than the overlap gains).
template <typename T, int MAXE, bool BATCH>
__global__ __launch_bounds__(kThreads) void permute_tiled32_once(const T *__restrict__ src,
                                                                 T *__restrict__ dst,
                                                                 TileArgs32 p) {
  constexpr int PER = MAXE / kThreads;
  __shared__ T tile[MAXE + 64];
  __shared__ uint32_t aoff[kMaxTile];
  __shared__ uint32_t boff[kMaxTile];
  const int w = p.w, h = p.h, ws = w + 1;
  const uint32_t n = (uint32_t)(w * h * p.c);
  for (uint32_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    const TileAt x = tile_at(p, t);
    __syncthreads();  // the previous tile is fully written out
    for (int i = threadIdx.x; i < w + h; i += kThreads) {
      if (i < w)
        aoff[i] = decode32(x.a0 + i, p.na, p.a_div, p.a_ostr);
      else
        boff[i - w] = decode32(x.b0 + (i - w), p.nb, p.b_div, p.b_istr);
    }
    __syncthreads();
    const T *s = src + x.oin + x.a0;
    T r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = read_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, x)) r[k] = s[e.cc * p.c_istr + boff[e.b] + e.a];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = read_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, x)) tile[(e.cc * h + e.b) * ws + e.a] = r[k];
    }
    __syncthreads();
    T *d = dst + x.oout + x.b0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = write_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, x)) r[k] = tile[(e.cc * h + e.b) * ws + e.a];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = write_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, x)) d[e.cc * p.c_ostr + aoff[e.a] + e.b] = r[k];
    }
  }
}

// Software-pipelined over the workgroup's tiles (t, t + grid, ...): the loads of tile t+grid
// are issued right after tile t is staged in LDS, so they are in flight while tile t is written
// out -- reads and writes overlap within each workgroup instead of alternating in lockstep
// across the whole launch.  The per-tile offset tables are double-buffered for that.
template <typename T, int MAXE, bool BATCH>
__global__ __launch_bounds__(kThreads) void permute_tiled32(const T *__restrict__ src,
                                                            T *__restrict__ dst, TileArgs32 p) {
  constexpr int PER = MAXE / kThreads;
  __shared__ T tile[MAXE + 64];
  __shared__ uint32_t aoff[2][kMaxTile];
  __shared__ uint32_t boff[2][kMaxTile];
  const int w = p.w, h = p.h, ws = w + 1;
  const uint32_t n = (uint32_t)(w * h * p.c);
  uint32_t t = blockIdx.x;
  if (t >= p.ntiles) return;  // uniform over the workgroup
  auto decode_offsets = [&](int buf, const TileAt &x) {
    for (int i = threadIdx.x; i < w + h; i += kThreads) {
      if (i < w)
        aoff[buf][i] = decode32(x.a0 + i, p.na, p.a_div, p.a_ostr);
      else
        boff[buf][i - w] = decode32(x.b0 + (i - w), p.nb, p.b_div, p.b_istr);
    }
  };
  // read: consecutive threads walk A, which is contiguous in the input
  auto load = [&](int buf, const TileAt &x, T (&r)[PER]) {
    const T *s = src + x.oin + x.a0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = read_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, x)) r[k] = s[e.cc * p.c_istr + boff[buf][e.b] + e.a];
    }
  };
  TileAt cur = tile_at(p, t);
  decode_offsets(0, cur);
  __syncthreads();
  T r[PER];
  load(0, cur, r);
  int buf = 0;
  for (;;) {
    const uint32_t tn = t + gridDim.x;
    const bool more = tn < p.ntiles;  // uniform
    TileAt nxt = cur;
    if (more) {
      nxt = tile_at(p, tn);
      decode_offsets(buf ^ 1, nxt);  // that buffer was last read before the previous barrier
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = read_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, cur)) tile[(e.cc * h + e.b) * ws + e.a] = r[k];
    }
    __syncthreads();  // tile staged; next offsets decoded
    if (more) load(buf ^ 1, nxt, r);  // next tile's loads in flight during the writes below
    // write: consecutive threads walk B, which is contiguous in the output
    T *d = dst + cur.oout + cur.b0;
    T o[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = write_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, cur)) o[k] = tile[(e.cc * h + e.b) * ws + e.a];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * kThreads;
      const TileIdx e = write_idx<BATCH>(p, i);
      if (i < n && in_tile<BATCH>(e, cur)) d[e.cc * p.c_ostr + aoff[buf][e.a] + e.b] = o[k];
    }
    __syncthreads();  // tile and this buffer's offsets consumed
    if (!more) break;
    t = tn;
    cur = nxt;
    buf ^= 1;
  }
}

struct RowArgs32 {
  int nd;              // row axes (the output's axes except the kept innermost one)
  uint32_t nvec;       // total / V
  FastDiv dvpr;        // vectors per row
  FastDiv n_div[16];   // row axes in output order, outermost first
  uint32_t n_istr[16];  // their input strides (in elements)
};

// Innermost axis kept: output vector v (V elements) is vector v % vpr of output row v / vpr;
// the row's input offset is decoded once per vector with FastDiv.  U vectors per thread: all U
// offsets are decoded first and the U loads then issued back to back with no branch between
// them (a load placed between two decode loops is waited for at the next loop's header, which
// serialises the loads); a vector past the end reads vector 0 and is not stored.
template <typename T, typename VT, int U>
__global__ __launch_bounds__(kThreads) void permute_rows32(const T *__restrict__ src,
                                                           T *__restrict__ dst, RowArgs32 p) {
  constexpr int V = sizeof(VT) / sizeof(T);
  const uint32_t G = gridDim.x * kThreads;
  for (uint32_t v0 = blockIdx.x * kThreads + threadIdx.x; v0 < p.nvec; v0 += U * G) {
    uint32_t in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t v = v0 + u * G < p.nvec ? v0 + u * G : 0;
      const uint32_t row = p.dvpr.div(v), j = v - row * p.dvpr.d;
      in[u] = decode32(row, p.nd, p.n_div, p.n_istr) + j * V;
    }
    VT r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = *reinterpret_cast<const VT *>(src + in[u]);
    // a use of every loaded value here keeps the loads ahead of the conditional stores (the
    // compiler would otherwise sink each load into its store's branch, one after another)
#pragma unroll
    for (int u = 0; u < U; ++u) __asm__ volatile("" ::"v"(r[u].x), "v"(r[u].y));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t v = v0 + u * G;
      if (v < p.nvec) *reinterpret_cast<VT *>(dst + (size_t)v * V) = r[u];
    }
  }
}

// Identity permute: a streaming copy in 16-byte vectors.  U vectors per thread, the U loads
// issued back to back ahead of the (guarded) stores, non-temporal on both sides (the data is
// touched once).
template <int U>
__global__ __launch_bounds__(kThreads) void copy16(const U128 *__restrict__ src,
                                                   U128 *__restrict__ dst, uint64_t nvec) {
  const uint64_t G = (uint64_t)gridDim.x * kThreads;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kThreads + threadIdx.x; v0 < nvec; v0 += U * G) {
    U128 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = v0 + u * G < nvec ? v0 + u * G : v0;
      r[u] = __builtin_nontemporal_load(src + v);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) __asm__ volatile("" ::"v"(r[u].x), "v"(r[u].y));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = v0 + u * G;
      if (v < nvec) __builtin_nontemporal_store(r[u], dst + v);
    }
  }
}

struct Reduced {
  int rank = 0;
  int64_t shape[16];  // old (input) extents, row-major
  int perm[16];       // new axis j = old axis perm[j]
  int64_t total = 1;
};

Reduced reduce_perm(int rank, const uint64_t *old_shape, const uint64_t *mapper) {
  Reduced r;
  // keep non-unit axes; renumber them in old order
  int newid[16];
  int kept = 0;
  for (int i = 0; i < rank; ++i) {
    r.total *= (int64_t)old_shape[i];
    newid[i] = old_shape[i] == 1 ? -1 : kept++;
  }
  std::vector<int> order;  // new order of kept old axes (compacted ids)
  std::vector<int64_t> ext(kept);
  for (int i = 0; i < rank; ++i)
    if (newid[i] >= 0) ext[newid[i]] = (int64_t)old_shape[i];
  for (int j = 0; j < rank; ++j)
    if (newid[mapper[j]] >= 0) order.push_back(newid[mapper[j]]);
  // merge runs of consecutive old axes that stay consecutive
  std::vector<std::pair<int, int>> groups;  // [first, last] old axes, in new order
  for (size_t j = 0; j < order.size(); ++j) {
    if (!groups.empty() && order[j] == groups.back().second + 1)
      groups.back().second = order[j];
    else
      groups.push_back({order[j], order[j]});
  }
  std::vector<int> by_old(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) by_old[g] = (int)g;
  std::sort(by_old.begin(), by_old.end(),
            [&](int x, int y) { return groups[x].first < groups[y].first; });
  std::vector<int> old_pos(groups.size());
  for (size_t i = 0; i < by_old.size(); ++i) old_pos[by_old[i]] = (int)i;
  r.rank = (int)groups.size();
  for (int i = 0; i < r.rank; ++i) {
    const auto &g = groups[by_old[i]];
    int64_t e = 1;
    for (int a = g.first; a <= g.second; ++a) e *= ext[a];
    r.shape[i] = e;
  }
  for (int j = 0; j < r.rank; ++j) r.perm[j] = old_pos[j];
  return r;
}

int grid_for(int64_t work_items, int per_block) {
  int64_t g = (work_items + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 256 * 64));
}

template <typename T>
int launch_tiled(const Reduced &r, const void *src, void *dst, hipStream_t s) {
  const int es = (int)sizeof(T);
  const int r_ = r.rank;
  int64_t istr[16], ostr[16], nshape[16];
  int64_t acc = 1;
  for (int i = r_ - 1; i >= 0; --i) {
    istr[i] = acc;
    acc *= r.shape[i];
  }
  for (int j = 0; j < r_; ++j) nshape[j] = r.shape[r.perm[j]];
  int64_t ostr_new[16];
  acc = 1;
  for (int j = r_ - 1; j >= 0; --j) {
    ostr_new[j] = acc;
    acc *= nshape[j];
  }
  int inv[16];
  for (int j = 0; j < r_; ++j) inv[r.perm[j]] = j;
  for (int i = 0; i < r_; ++i) ostr[i] = ostr_new[inv[i]];  // output stride of old axis i

  const int64_t target = std::max(8, 256 / es);  // elements per contiguous run
  const int q = r.perm[r_ - 1];                  // old axis that is fastest in the output
  bool inA[16] = {false};
  int A[kMaxGroup], na = 0;  // old axes, innermost first
  int64_t prodA = 1;
  for (int i = r_ - 1; i >= 0 && na < kMaxGroup; --i) {
    if (i == q || prodA >= target) break;
    A[na++] = i;
    inA[i] = true;
    prodA *= r.shape[i];
  }
  int B[kMaxGroup], nb = 0;  // old axes in output order, innermost first
  int64_t prodB = 1;
  bool inB[16] = {false};
  // 8-byte and smaller elements: B grows past the target while A is short (A stops at the
  // output-fastest axis), so that a tile still spans kMinTile elements
  const bool pipelined = es <= 8;
  for (int j = r_ - 1; j >= 0 && nb < kMaxGroup; --j) {
    const int x = r.perm[j];
    if (inA[x] || (prodB >= target && (!pipelined || prodA * prodB >= kMinTile))) break;
    B[nb++] = x;
    inB[x] = true;
    prodB *= r.shape[x];
  }
  TileArgs p{};
  p.na = na;
  p.nb = nb;
  // store groups outermost first
  for (int k = 0; k < na; ++k) {
    const int ax = A[na - 1 - k];
    p.a_ext[k] = r.shape[ax];
    p.a_ostr[k] = ostr[ax];
  }
  for (int k = 0; k < nb; ++k) {
    const int ax = B[nb - 1 - k];
    p.b_ext[k] = r.shape[ax];
    p.b_istr[k] = istr[ax];
  }
  int no = 0;
  for (int i = 0; i < r_; ++i) {
    if (inA[i] || inB[i]) continue;
    if (no >= kMaxOuter) return -1;
    p.o_ext[no] = r.shape[i];
    p.o_istr[no] = istr[i];
    p.o_ostr[no] = ostr[i];
    ++no;
  }
  p.no = no;
  constexpr int MAXE = (sizeof(T) >= 16) ? 1024 : 2048;
  // Tile shape and grid, measured on the rank-6 sweep (16 MiB / 32 MiB tensors, 32 seeded
  // permutations): 8-byte and smaller elements run the pipelined kernel on tiles 128 bytes wide
  // with at most 1024 workgroups (several tiles each, so loads and stores overlap); 16-byte
  // elements run one tile at a time, 256 bytes wide and 32 rows high.
  const int64_t wmax = std::max<int64_t>(1, (pipelined ? 128 : 256) / es);
  const int64_t hmax = pipelined ? kMaxTile : 32;
  p.w = (int)std::min<int64_t>({prodA, wmax, (int64_t)kMaxTile});
  p.h = (int)std::min<int64_t>({prodB, (int64_t)(MAXE / (p.w + 1)), (int64_t)kMaxTile, hmax});
  // A short output run (B stopped at an axis already in A) leaves tiles of a few hundred bytes,
  // whose offset tables and barriers then cost more than their data: widen such a tile along A
  // (the height stays, so tile rows keep their alignment).  Measured on the rank-6 sweep: below
  // 1024 elements for the pipelined kernel, below 256 for 16-byte elements (whose 16 x 32 tiles
  // are faster than wider ones).
  const int64_t small = pipelined ? kMinTile : kMinTile / 4;
  if ((int64_t)p.w * p.h < small && p.w < prodA) {
    int64_t w2 = std::min<int64_t>({prodA, (int64_t)kMaxTile,
                                    std::max<int64_t>(p.w, kMinTile / p.h),
                                    (int64_t)(MAXE / p.h) - 1});
    if (w2 < prodA) {  // a partial width: a power of two, so tile rows stay aligned
      int64_t pw = 1;
      while (pw * 2 <= w2) pw *= 2;
      w2 = pw;
    }
    if (w2 > p.w) p.w = (int)w2;
  }
  p.prodA = prodA;
  p.prodB = prodB;
  p.nca = (prodA + p.w - 1) / p.w;
  p.ncb = (prodB + p.h - 1) / p.h;
  int64_t outer = 1;
  for (int k = 0; k < no; ++k) outer *= p.o_ext[k];
  p.ntiles = outer * p.nca * p.ncb;
  // pipelined: about two tiles per workgroup, at least 1024 workgroups (a single round when the
  // tensor is small), at most 16384
  const int64_t grid_cap = pipelined
                               ? std::min<int64_t>(256 * 64, std::max<int64_t>(1024, p.ntiles / 2))
                               : 4096;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(p.ntiles, grid_cap));
  if (r.total < (int64_t(1) << 31)) {  // every offset and index fits in 32 bits
    TileArgs32 q{};
    q.na = p.na;
    q.nb = p.nb;
    q.no = p.no;
    q.w = p.w;
    q.h = p.h;
    q.c = 1;
    q.prodA = (uint32_t)p.prodA;
    q.prodB = (uint32_t)p.prodB;
    // Whole A and B runs in a tile still smaller than `small` (a batch of small transposes):
    // batch c consecutive indices of the innermost outer axis into the tile when that axis is
    // contiguous with the tile on one side (it directly precedes A and B in the input or in the
    // output), so that a tile moves about kMinTile elements.
    int64_t cext = no > 0 ? p.o_ext[no - 1] : 1;
    if (no > 0 && p.w == prodA && p.h == prodB && (int64_t)p.w * p.h < small &&
        (p.o_istr[no - 1] == prodA * prodB || p.o_ostr[no - 1] == prodA * prodB)) {
      const int64_t cmax = std::min<int64_t>((int64_t)MAXE / (p.w * p.h),
                                             (int64_t)(MAXE + 64) / (p.h * (p.w + 1)));
      q.c = (int)std::max<int64_t>(1, std::min<int64_t>(cext, cmax));
    }
    q.cext = (uint32_t)cext;
    q.c_istr = q.c > 1 ? (uint32_t)p.o_istr[no - 1] : 0;
    q.c_ostr = q.c > 1 ? (uint32_t)p.o_ostr[no - 1] : 0;
    const int64_t nchunks = (cext + q.c - 1) / q.c;
    q.ntiles = (uint32_t)(p.ntiles / cext * nchunks);
    q.dnca = FastDiv::make((uint32_t)p.nca);
    q.dncb = FastDiv::make((uint32_t)p.ncb);
    q.dw = FastDiv::make((uint32_t)p.w);
    q.dh = FastDiv::make((uint32_t)p.h);
    q.dwh = FastDiv::make((uint32_t)(p.w * p.h));
    for (int k = 0; k < p.na; ++k) {
      q.a_div[k] = FastDiv::make((uint32_t)p.a_ext[k]);
      q.a_ostr[k] = (uint32_t)p.a_ostr[k];
    }
    for (int k = 0; k < p.nb; ++k) {
      q.b_div[k] = FastDiv::make((uint32_t)p.b_ext[k]);
      q.b_istr[k] = (uint32_t)p.b_istr[k];
    }
    for (int k = 0; k < p.no; ++k) {
      q.o_div[k] = FastDiv::make((uint32_t)(k == p.no - 1 ? nchunks : p.o_ext[k]));
      q.o_istr[k] = (uint32_t)p.o_istr[k];
      q.o_ostr[k] = (uint32_t)p.o_ostr[k];
    }
    const int grid32 =
        (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)q.ntiles,
                                                    pipelined ? std::min<int64_t>(
                                                                    256 * 64,
                                                                    std::max<int64_t>(
                                                                        1024, q.ntiles / 2))
                                                              : 4096));
    if (pipelined && q.c > 1)
      hipLaunchKernelGGL((permute_tiled32<T, MAXE, true>), dim3(grid32), dim3(kThreads), 0, s,
                         (const T *)src, (T *)dst, q);
    else if (pipelined)
      hipLaunchKernelGGL((permute_tiled32<T, MAXE, false>), dim3(grid32), dim3(kThreads), 0, s,
                         (const T *)src, (T *)dst, q);
    else if (q.c > 1)
      hipLaunchKernelGGL((permute_tiled32_once<T, MAXE, true>), dim3(grid32), dim3(kThreads), 0,
                         s, (const T *)src, (T *)dst, q);
    else
      hipLaunchKernelGGL((permute_tiled32_once<T, MAXE, false>), dim3(grid32), dim3(kThreads), 0,
                         s, (const T *)src, (T *)dst, q);
    return 0;
  }
  hipLaunchKernelGGL((permute_tiled<T, MAXE>), dim3(grid), dim3(kThreads), 0, s,
                     (const T *)src, (T *)dst, p);
  return 0;
}

template <typename T>
int launch_gather(const Reduced &r, const void *src, void *dst, hipStream_t s) {
  GatherArgs g{};
  g.rank = r.rank;
  g.total = r.total;
  int64_t istr[16], acc = 1;
  for (int i = r.rank - 1; i >= 0; --i) {
    istr[i] = acc;
    acc *= r.shape[i];
  }
  for (int j = 0; j < r.rank; ++j) {
    g.n_ext[j] = r.shape[r.perm[j]];
    g.n_istr[j] = istr[r.perm[j]];
  }
  const int64_t inner = g.n_ext[r.rank - 1];
  const bool inner_kept = r.perm[r.rank - 1] == r.rank - 1;
  constexpr int V = sizeof(T) >= 16 ? 1 : 16 / (int)sizeof(T);
  const bool aligned = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  if (inner_kept && inner % V == 0 && aligned && r.total < (int64_t(1) << 31)) {
    RowArgs32 q{};
    q.nd = r.rank - 1;
    q.nvec = (uint32_t)(r.total / V);
    q.dvpr = FastDiv::make((uint32_t)(inner / V));
    for (int j = 0; j + 1 < r.rank; ++j) {
      q.n_div[j] = FastDiv::make((uint32_t)g.n_ext[j]);
      q.n_istr[j] = (uint32_t)g.n_istr[j];
    }
    constexpr int U = 4;
    const int grid = grid_for((q.nvec + U - 1) / U, kThreads);
    hipLaunchKernelGGL((permute_rows32<T, U128, U>), dim3(grid), dim3(kThreads), 0, s,
                       (const T *)src, (T *)dst, q);
    return 0;
  }
  if (inner_kept && V > 1 && inner % V == 0 && aligned) {
    hipLaunchKernelGGL((permute_gather<T, V>), dim3(grid_for(r.total / V, kThreads)),
                       dim3(kThreads), 0, s, (const T *)src, (T *)dst, g);
  } else {
    hipLaunchKernelGGL((permute_gather<T, 1>), dim3(grid_for(r.total, kThreads)),
                       dim3(kThreads), 0, s, (const T *)src, (T *)dst, g);
  }
  return 0;
}

template <typename T>
int run_permute(const Reduced &r, const void *src, void *dst, hipStream_t s) {
  if (r.rank <= 1) {
    const uint64_t bytes = (uint64_t)r.total * sizeof(T);
    if (bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0) {
      constexpr int U = 4;
      const uint64_t nvec = bytes / 16;
      const int grid = grid_for((int64_t)((nvec + U - 1) / U), kThreads);
      hipLaunchKernelGGL((copy16<U>), dim3(grid), dim3(kThreads), 0, s, (const U128 *)src,
                         (U128 *)dst, nvec);
      CXA_HIP(hipGetLastError());
    } else {
      CXA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, s));
    }
    return CXA_OK;
  }
  int rc;
  if (r.perm[r.rank - 1] == r.rank - 1)
    rc = launch_gather<T>(r, src, dst, s);
  else
    rc = launch_tiled<T>(r, src, dst, s);
  if (rc != 0) rc = launch_gather<T>(r, src, dst, s);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

// Batched small-block permute: one workgroup per task; the job's reduced extents and source
// strides sit in LDS, every thread decodes its elements' coordinates (32-bit: a job holds at
// most kBatchCopyMaxElems elements) and gathers them; writes are contiguous.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_batch_copy(const BatchCopyJob *__restrict__ jobs,
                                                         const BatchCopyTask *__restrict__ tasks,
                                                         const T *__restrict__ src,
                                                         T *__restrict__ dst) {
  __shared__ int32_t ext[16], sstr[16];
  const BatchCopyTask tk = tasks[blockIdx.x];
  const BatchCopyJob *j = jobs + tk.job;
  const int rank = j->rank;
  if (threadIdx.x < 16) {
    ext[threadIdx.x] = threadIdx.x < (unsigned)rank ? j->ext[threadIdx.x] : 1;
    sstr[threadIdx.x] = threadIdx.x < (unsigned)rank ? j->sstr[threadIdx.x] : 0;
  }
  __syncthreads();
  const T *s = src + j->src_off;
  T *d = dst + j->dst_off;
  const int end = tk.start + tk.count;
  for (int e = tk.start + (int)threadIdx.x; e < end; e += kThreads) {
    int rem = e, off = 0;
    for (int a = rank - 1; a > 0; --a) {
      const int q = rem / ext[a];
      off += (rem - q * ext[a]) * sstr[a];
      rem = q;
    }
    off += rem * sstr[0];
    d[e] = s[off];
  }
}

}  // namespace

bool batch_copy_job(int rank, const uint64_t *shape, const uint64_t *mapper, int64_t src_off,
                    int64_t dst_off, BatchCopyJob *job) {
  if (rank < 0 || rank > 16) return false;
  const Reduced r = reduce_perm(rank, shape, mapper);
  if (r.total > kBatchCopyMaxElems) return false;
  *job = BatchCopyJob{};
  job->src_off = src_off;
  job->dst_off = dst_off;
  job->n = (int32_t)r.total;
  job->rank = std::max(1, r.rank);
  // row-major strides of the reduced old axes, then the new axes' extents and source strides
  int64_t ostr[16];
  int64_t acc = 1;
  for (int i = r.rank - 1; i >= 0; --i) {
    ostr[i] = acc;
    acc *= r.shape[i];
  }
  if (r.rank == 0) {
    job->ext[0] = (int32_t)r.total;
    job->sstr[0] = 1;
    return true;
  }
  for (int d = 0; d < r.rank; ++d) {
    job->ext[d] = (int32_t)r.shape[r.perm[d]];
    job->sstr[d] = (int32_t)ostr[r.perm[d]];
  }
  return true;
}

int launch_batch_copy(int elem_size, const BatchCopyJob *jobs, const BatchCopyTask *tasks,
                      int ntasks, const void *src, void *dst, hipStream_t s) {
  if (ntasks <= 0) return CXA_OK;
  switch (elem_size) {
    case 4:
      hipLaunchKernelGGL((k_batch_copy<uint32_t>), dim3(ntasks), dim3(kThreads), 0, s, jobs,
                         tasks, (const uint32_t *)src, (uint32_t *)dst);
      break;
    case 8:
      hipLaunchKernelGGL((k_batch_copy<uint64_t>), dim3(ntasks), dim3(kThreads), 0, s, jobs,
                         tasks, (const uint64_t *)src, (uint64_t *)dst);
      break;
    case 16:
      hipLaunchKernelGGL((k_batch_copy<U128>), dim3(ntasks), dim3(kThreads), 0, s, jobs, tasks,
                         (const U128 *)src, (U128 *)dst);
      break;
    default:
      return fail(CXA_ERR_UNSUPPORTED, "batched copy: element size %d", elem_size);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

}  // namespace cxa

extern "C" int cxa_permute(int dtype, int rank, const uint64_t *old_shape,
                           const uint64_t *mapper, const void *src, void *dst,
                           cxa_stream_t stream) {
  using namespace cxa;
  const int es = elem_size(dtype);
  CXA_CHECK(es > 0, CXA_ERR_UNSUPPORTED, "cxa_permute: unknown dtype %d", dtype);
  CXA_CHECK(rank >= 0 && rank <= 16, CXA_ERR_UNSUPPORTED, "cxa_permute: rank %d > 16", rank);
  CXA_CHECK(rank == 0 || (old_shape && mapper), CXA_ERR_INVALID, "cxa_permute: null metadata");
  bool seen[16] = {false};
  for (int i = 0; i < rank; ++i) {
    CXA_CHECK(mapper[i] < (uint64_t)rank, CXA_ERR_INVALID,
              "cxa_permute: invalid rank index %llu", (unsigned long long)mapper[i]);
    CXA_CHECK(!seen[mapper[i]], CXA_ERR_INVALID, "cxa_permute: duplicated index %llu",
              (unsigned long long)mapper[i]);
    seen[mapper[i]] = true;
  }
  Reduced r = reduce_perm(rank, old_shape, mapper);
  if (r.total == 0) return CXA_OK;
  CXA_CHECK(src && dst, CXA_ERR_INVALID, "cxa_permute: null data pointer");
  hipStream_t s = (hipStream_t)stream;
  switch (es) {
    case 1: return run_permute<uint8_t>(r, src, dst, s);
    case 2: return run_permute<uint16_t>(r, src, dst, s);
    case 4: return run_permute<uint32_t>(r, src, dst, s);
    case 8: return run_permute<uint64_t>(r, src, dst, s);
    default: return run_permute<U128>(r, src, dst, s);
  }
}

// ------------------------------------------------------------------------------------------
// Block permute plans (cxa_block_permute_*).
// ------------------------------------------------------------------------------------------
struct cxa_bp_plan {
  int dtype = 0, es = 0, rank = 0;
  uint64_t mapper[16] = {};
  struct Big {
    uint64_t shape[16];
    int64_t src_off, dst_off;
  };
  std::vector<Big> big;  // blocks moved by cxa_permute one by one
  int ntasks = 0;
  void *dev = nullptr;  // batched jobs, then tasks (256-byte aligned)
  const cxa::BatchCopyJob *jobs = nullptr;
  const cxa::BatchCopyTask *tasks = nullptr;
  ~cxa_bp_plan() {
    if (dev) (void)hipFree(dev);
  }
};

extern "C" int cxa_block_permute_plan(int dtype, int rank, int64_t nblocks,
                                      const uint64_t *shapes, const int64_t *src_offsets,
                                      const int64_t *dst_offsets, const uint64_t *mapper,
                                      cxa_bp_plan **out) {
  using namespace cxa;
  CXA_CHECK(out, CXA_ERR_INVALID, "cxa_block_permute_plan: null output handle");
  *out = nullptr;
  const int es = elem_size(dtype);
  CXA_CHECK(es > 0, CXA_ERR_UNSUPPORTED, "cxa_block_permute_plan: unknown dtype %d", dtype);
  CXA_CHECK(rank >= 0 && rank <= 16, CXA_ERR_UNSUPPORTED, "cxa_block_permute_plan: rank %d > 16",
            rank);
  CXA_CHECK(nblocks >= 0 && (nblocks == 0 || (src_offsets && dst_offsets)) &&
                (rank == 0 || nblocks == 0 || (shapes && mapper)),
            CXA_ERR_INVALID, "cxa_block_permute_plan: null metadata");
  bool seen[16] = {false};
  for (int i = 0; i < rank; ++i) {
    CXA_CHECK(mapper[i] < (uint64_t)rank && !seen[mapper[i]], CXA_ERR_INVALID,
              "cxa_block_permute_plan: mapper is not a permutation");
    seen[mapper[i]] = true;
  }
  std::unique_ptr<cxa_bp_plan> p(new cxa_bp_plan());
  p->dtype = dtype;
  p->es = es;
  p->rank = rank;
  for (int i = 0; i < rank; ++i) p->mapper[i] = mapper[i];
  std::vector<BatchCopyJob> jobs;
  std::vector<BatchCopyTask> tasks;
  std::vector<int64_t> small_ids;  // blocks in the batched launch (two or more of them)
  const bool batched = es == 4 || es == 8 || es == 16;
  for (int64_t b = 0; b < nblocks; ++b) {
    CXA_CHECK(src_offsets[b] >= 0 && dst_offsets[b] >= 0, CXA_ERR_INVALID,
              "cxa_block_permute_plan: negative offset of block %lld", (long long)b);
    BatchCopyJob j;
    if (batched &&
        batch_copy_job(rank, shapes + b * rank, mapper, src_offsets[b], dst_offsets[b], &j)) {
      small_ids.push_back(b);
      jobs.push_back(j);
      continue;
    }
    cxa_bp_plan::Big g{};
    for (int i = 0; i < rank; ++i) g.shape[i] = shapes[b * rank + i];
    g.src_off = src_offsets[b];
    g.dst_off = dst_offsets[b];
    p->big.push_back(g);
  }
  if (jobs.size() < 2) {  // a lone small block: the tiled permute moves it faster
    for (int64_t b : small_ids) {
      cxa_bp_plan::Big g{};
      for (int i = 0; i < rank; ++i) g.shape[i] = shapes[b * rank + i];
      g.src_off = src_offsets[b];
      g.dst_off = dst_offsets[b];
      p->big.push_back(g);
    }
    jobs.clear();
  }
  for (size_t id = 0; id < jobs.size(); ++id)
    for (int32_t at = 0; at < jobs[id].n; at += kBatchCopyTaskElems)
      tasks.push_back({(int32_t)id, at, std::min(kBatchCopyTaskElems, jobs[id].n - at), 0});
  if (!tasks.empty()) {
    const size_t nj = (jobs.size() * sizeof(BatchCopyJob) + 255) / 256 * 256;
    const size_t nt = tasks.size() * sizeof(BatchCopyTask);
    std::vector<char> host(nj + nt, 0);
    std::memcpy(host.data(), jobs.data(), jobs.size() * sizeof(BatchCopyJob));
    std::memcpy(host.data() + nj, tasks.data(), nt);
    CXA_HIP(hipMalloc(&p->dev, host.size()));
    const int rc = upload_now(p->dev, host.data(), host.size());
    if (rc) return rc;
    p->jobs = (const BatchCopyJob *)p->dev;
    p->tasks = (const BatchCopyTask *)((char *)p->dev + nj);
    p->ntasks = (int)tasks.size();
  }
  *out = p.release();
  return CXA_OK;
}

extern "C" int cxa_block_permute_exec(const cxa_bp_plan *p, const void *src, void *dst,
                                      cxa_stream_t stream) {
  using namespace cxa;
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_permute_exec: null plan");
  if (p->ntasks == 0 && p->big.empty()) return CXA_OK;
  CXA_CHECK(src && dst, CXA_ERR_INVALID, "cxa_block_permute_exec: null data pointer");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_batch_copy(p->es, p->jobs, p->tasks, p->ntasks, src, dst, s);
  if (rc) return rc;
  for (const cxa_bp_plan::Big &g : p->big) {
    rc = cxa_permute(p->dtype, p->rank, g.shape, p->mapper, (const char *)src + g.src_off * p->es,
                     (char *)dst + g.dst_off * p->es, stream);
    if (rc) return rc;
  }
  return CXA_OK;
}

extern "C" int cxa_block_permute_destroy(cxa_bp_plan *p) {
  delete p;
  return CXA_OK;
}
