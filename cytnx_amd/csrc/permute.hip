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
//     other; per-tile offsets of the A and B indices are decoded once into LDS.
// The semantics are those of the reference's MoveMemoryCpu / MoveMemoryGpu
// (src/backend/utils_internal_cpu/Movemem_cpu.cpp:18-211,
//  src/backend/utils_internal_gpu/cuMovemem_gpu.cu:60-304): out = transpose(in, mapper).
#include <algorithm>
#include <vector>

#include "cxa_common.hpp"

namespace cxa {
namespace {

struct alignas(16) U128 {
  uint64_t lo, hi;
};

constexpr int kMaxOuter = 14;
constexpr int kMaxGroup = 8;
constexpr int kThreads = 256;

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
  for (int j = r_ - 1; j >= 0 && nb < kMaxGroup; --j) {
    const int x = r.perm[j];
    if (inA[x] || prodB >= target) break;
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
  const int64_t wmax = std::max<int64_t>(1, 256 / es);
  p.w = (int)std::min<int64_t>(prodA, wmax);
  p.h = (int)std::min<int64_t>({prodB, (int64_t)(MAXE / (p.w + 1)), (int64_t)1024});
  p.prodA = prodA;
  p.prodB = prodB;
  p.nca = (prodA + p.w - 1) / p.w;
  p.ncb = (prodB + p.h - 1) / p.h;
  int64_t outer = 1;
  for (int k = 0; k < no; ++k) outer *= p.o_ext[k];
  p.ntiles = outer * p.nca * p.ncb;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(p.ntiles, 256 * 64));
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
    CXA_HIP(hipMemcpyAsync(dst, src, (size_t)r.total * sizeof(T), hipMemcpyDeviceToDevice, s));
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

}  // namespace
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
