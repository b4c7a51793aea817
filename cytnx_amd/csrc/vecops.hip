// Vector kernels for the Lanczos driver (FP64): dot products, norms, the three-term update,
// re-orthogonalisation and Ritz-vector assembly.  Scalars live in device memory and are indexed
// through a device-side iteration counter, so an iteration enqueues without any host round trip
// and the same launch sequence can be captured once and replayed.  Reductions are two-stage
// with a fixed number of partial sums, so results are deterministic run to run.
#include <algorithm>

#include "cxa_common.hpp"
#include "vecops.hpp"
#include "contract.hpp"

namespace cxa {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kT / 64; ++i) s += red[i];
  return s;  // valid in thread 0
}

// Reductions finish in a separate one-block launch rather than in the last-arriving block of
// the partials launch: the L2s of the XCDs are not coherent with each other, so a device-scope
// fence in every block writes back and invalidates its L2, which costs far more than the
// kernel boundary that orders the partials for the finishing launch.

__global__ __launch_bounds__(kT) void k_dot_partials(const double *__restrict__ x,
                                                     const double *__restrict__ y, int64_t n,
                                                     double *__restrict__ part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    s += x[i] * y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[counter + offset] = sum(part[0..P)) (mode 0) or its square root (mode 1)
__global__ __launch_bounds__(kT) void k_finalize(const double *__restrict__ part, int P,
                                                 double *out, const int *counter, int offset,
                                                 int mode) {
  double s = 0.0;
  for (int i = threadIdx.x; i < P; i += kT) s += part[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    const int idx = (counter ? *counter : 0) + offset;
    out[idx] = mode == 1 ? sqrt(s) : s;
  }
}

// Sum of the P partials a previous launch left in part, in the fixed order of k_finalize, by
// every block of the consuming launch (so no separate finishing launch is needed: the kernel
// boundary orders the partials); the value is returned to all threads.
__device__ __forceinline__ double all_sum_partials(const double *__restrict__ part, int P) {
  __shared__ double bc;
  double s = 0.0;
  for (int i = threadIdx.x; i < P; i += kT) s += part[i];
  s = block_sum(s);
  if (threadIdx.x == 0) bc = s;
  __syncthreads();
  return bc;
}

// One Lanczos iteration is two launches after the matvec y = H u.  The matvec input u is the
// previous iteration's residual, not yet divided by its norm beta (it = 0: u = V[0], beta = 1);
// the division happens in the first launch, so no launch of its own normalises the next vector.
//
// First: beta = sqrt(sum of part_b) (recorded in Bs[it - 1] by block 0), V[it] = u / beta, and
// partials of (y / beta) . V[it] into part_a.
__global__ __launch_bounds__(kT) void k_dot_basis(const double *__restrict__ y,
                                                  const double *__restrict__ u,
                                                  double *__restrict__ V, int64_t n,
                                                  const double *__restrict__ part_b,
                                                  double *__restrict__ Bs, IterIdx iter,
                                                  double *__restrict__ part_a) {
  const int it = iter.get();
  const double b = it > 0 ? sqrt(all_sum_partials(part_b, kPartials)) : 1.0;
  if (it > 0 && blockIdx.x == 0 && threadIdx.x == 0) Bs[it - 1] = b;
  double *v = V + (int64_t)it * n;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double vi = u[i] / b;
    if (it > 0) v[i] = vi;
    s += (y[i] / b) * vi;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part_a[blockIdx.x] = s;
}

// Second: alpha = sum of part_a (recorded in As[it] by block 0); the residual
// u = y / beta - (alpha V[it] + beta V[it - 1]) (no V[-1] term at it = 0), the next matvec input;
// partials of |u|^2 into part_b.
__global__ __launch_bounds__(kT) void k_residual(const double *__restrict__ y,
                                                 double *__restrict__ u,
                                                 const double *__restrict__ V, int64_t n,
                                                 const double *__restrict__ part_a,
                                                 double *__restrict__ As,
                                                 const double *__restrict__ Bs, IterIdx iter,
                                                 double *__restrict__ part_b) {
  const int it = iter.get();
  const double a = all_sum_partials(part_a, kPartials);
  if (blockIdx.x == 0 && threadIdx.x == 0) As[it] = a;
  const double b = it > 0 ? Bs[it - 1] : 1.0;
  const double bp = it > 0 ? b : 0.0;
  const double *cur = V + (int64_t)it * n;
  const double *prev = it > 0 ? V + (int64_t)(it - 1) * n : cur;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double t = y[i] / b - (a * cur[i] + bp * prev[i]);
    u[i] = t;
    s += t * t;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part_b[blockIdx.x] = s;
}

// The residual launch when the matvec's last launch has already produced, per tile slot, the
// partials of (H u) . u and of u . u (the chain's dot-product epilogue, dotp[0..S) and
// dotp[S..2S)): beta = |u| (1 at it = 0), alpha = (Hu . u) / beta^2 (recorded in As[it], beta
// in Bs[it - 1] by block 0), V[it] = u / beta, and the residual u = Hu / beta - alpha V[it] -
// beta V[it - 1] with its |u|^2 partials into part_b.  This replaces k_dot_basis + k_residual:
// one launch and one read of u fewer per iteration.  With p32 non-null the residual also goes,
// rounded to FP32, to its packed position pmap[i] (skipped where pmap[i] < 0).
template <bool PK>
__global__ __launch_bounds__(kT) void k_residual_fused(const double *__restrict__ y,
                                                       double *__restrict__ u,
                                                       double *__restrict__ V, int64_t n,
                                                       const double *__restrict__ dotp, int S,
                                                       double *__restrict__ As,
                                                       double *__restrict__ Bs, IterIdx iter,
                                                       double *__restrict__ part_b,
                                                       const int *__restrict__ pmap,
                                                       float *__restrict__ p32) {
  const int it = iter.get();
  double *cur = V + (int64_t)it * n;
  const double *prev = it > 0 ? V + (int64_t)(it - 1) * n : cur;
  // The first partials are loaded first, then the thread's first vector elements, so that
  // both rounds of loads are in flight together (the vector does not depend on alpha and
  // beta) and the partial sums wait only for their own loads (loads retire in issue order).
  // The loads are unconditional, at clamped indices, with the values selected afterwards, so
  // that no branch splits the sequence and the wait counts stay exact.
  constexpr int PP = 4, R = 2;
  double pa[PP], pb[PP];
#pragma unroll
  for (int q = 0; q < PP; ++q) {
    const int i = threadIdx.x + q * kT;
    const int ic = i < S ? i : S - 1;
    const double va = dotp[ic], vb = dotp[S + ic];
    pa[q] = i < S ? va : 0.0;
    pb[q] = i < S ? vb : 0.0;
  }
  const int64_t i0 = (int64_t)blockIdx.x * kT + threadIdx.x, step = (int64_t)gridDim.x * kT;
  double yr[R], ur[R], pr[R];
  int dr[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int64_t i = i0 + q * step;
    const bool in = i < n;
    const int64_t ic = in ? i : n - 1;
    const double yv = y[ic], uv = u[ic], pv = prev[ic];  // prev is V[0] at it = 0
    yr[q] = in ? yv : 0.0;
    ur[q] = in ? uv : 0.0;
    pr[q] = in && it > 0 ? pv : 0.0;
    if constexpr (PK) {
      const int dv = pmap[ic];
      dr[q] = in ? dv : -1;
    } else {
      dr[q] = -1;
    }
  }
  // both sums in one pass over the partials (then two block reductions)
  __shared__ double bc[2];
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int q = 0; q < PP; ++q) {
    a0 += pa[q];
    a1 += pb[q];
  }
  for (int i = threadIdx.x + PP * kT; i < S; i += kT) {
    a0 += dotp[i];
    a1 += dotp[S + i];
  }
  a0 = block_sum(a0);
  a1 = block_sum(a1);
  if (threadIdx.x == 0) {
    bc[0] = a0;
    bc[1] = a1;
  }
  __syncthreads();
  const double wu = bc[0], uu = bc[1];
  const double b = it > 0 ? sqrt(uu) : 1.0;
  const double a = (wu / b) / b;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    As[it] = a;
    if (it > 0) Bs[it - 1] = b;
  }
  const double bp = it > 0 ? b : 0.0;
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int64_t i = i0 + q * step;
    if (i < n) {
      const double vi = ur[q] / b;
      if (it > 0) cur[i] = vi;
      const double t = yr[q] / b - (a * vi + bp * pr[q]);
      u[i] = t;
      if (PK && dr[q] >= 0) p32[dr[q]] = (float)t;
      s += t * t;
    }
  }
  for (int64_t i = i0 + R * step; i < n; i += step) {
    const double vi = u[i] / b;
    if (it > 0) cur[i] = vi;
    const double t = y[i] / b - (a * vi + bp * prev[i]);
    u[i] = t;
    if (PK) {
      const int d = pmap[i];
      if (d >= 0) p32[d] = (float)t;
    }
    s += t * t;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part_b[blockIdx.x] = s;
}

// h_j partial sums: part[j * P + block] = sum over the block's slice of V_j * w, j <= it.
// Block (x, y) takes the basis vectors in groups of four, groups y, y + gridDim.y, ...: with the
// iteration known at launch there is one row of blocks per group; a captured launch (iteration
// read on the device) has a fixed number of rows that walk the groups.  Each partial sums the
// same elements in the same order whatever the grid.
// CPLX: the n doubles are n / 2 complex values (re, im) and h_j = conj(V_j) . w is complex: the
// real parts are the real dot of the doubles, the imaginary parts sum vr wi - vi wr and go to
// part_im[j * P + block].
template <bool CPLX>
__global__ __launch_bounds__(kT) void k_proj_partials(const double *__restrict__ V,
                                                      const double *__restrict__ w, int64_t n,
                                                      IterIdx iter, double *__restrict__ part,
                                                      double *__restrict__ part_im) {
  const int it = iter.get();
  const int64_t P = gridDim.x;
  const int64_t ne = CPLX ? n / 2 : n;
  for (int j0 = 4 * blockIdx.y; j0 <= it; j0 += 4 * gridDim.y) {
    const int nj = min(4, it + 1 - j0);
    double s[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne; i += P * kT) {
      if constexpr (CPLX) {
        const double wr = w[2 * i], wi = w[2 * i + 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (q < nj) {
            const double *v = V + (int64_t)(j0 + q) * n + 2 * i;
            s[q] += v[0] * wr + v[1] * wi;
            si[q] += v[0] * wi - v[1] * wr;
          }
        }
      } else {
        const double wi = w[i];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < nj) s[q] += V[(int64_t)(j0 + q) * n + i] * wi;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < nj) {  // uniform over the block
        const double t = block_sum(s[q]);
        if (threadIdx.x == 0) part[(int64_t)(j0 + q) * P + blockIdx.x] = t;
        if constexpr (CPLX) {
          const double ti = block_sum(si[q]);
          if (threadIdx.x == 0) part_im[(int64_t)(j0 + q) * P + blockIdx.x] = ti;
        }
      }
    }
  }
}

// h[j] (CPLX: h[2j], h[2j + 1] = re, im) from the partials, j <= it
template <bool CPLX>
__global__ __launch_bounds__(kT) void k_proj_finalize(const double *__restrict__ part,
                                                      const double *__restrict__ part_im, int P,
                                                      IterIdx iter, double *__restrict__ h) {
  const int j = blockIdx.x;
  if (j > iter.get()) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < P; i += kT) s += part[(int64_t)j * P + i];
  s = block_sum(s);
  if constexpr (CPLX) {
    double si = 0.0;
    for (int i = threadIdx.x; i < P; i += kT) si += part_im[(int64_t)j * P + i];
    si = block_sum(si);
    if (threadIdx.x == 0) {
      h[2 * j] = s;
      h[2 * j + 1] = si;
    }
  } else {
    if (threadIdx.x == 0) h[j] = s;
  }
}

// w -= sum_{j <= it} h_j V_j (CPLX: complex h_j times complex V_j), then partial sums of |w|^2
template <bool CPLX>
__global__ __launch_bounds__(kT) void k_proj_apply(double *__restrict__ w,
                                                   const double *__restrict__ V, int64_t n,
                                                   const double *__restrict__ h, IterIdx iter,
                                                   double *__restrict__ part) {
  const int it = iter.get();
  const int64_t ne = CPLX ? n / 2 : n;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne;
       i += (int64_t)gridDim.x * kT) {
    if constexpr (CPLX) {
      double vr = w[2 * i], vi = w[2 * i + 1];
      for (int j = 0; j <= it; ++j) {
        const double *b = V + (int64_t)j * n + 2 * i;
        const double hr = h[2 * j], hi = h[2 * j + 1];
        vr -= hr * b[0] - hi * b[1];
        vi -= hr * b[1] + hi * b[0];
      }
      w[2 * i] = vr;
      w[2 * i + 1] = vi;
      s += vr * vr + vi * vi;
    } else {
      double v = w[i];
      for (int j = 0; j <= it; ++j) v -= h[j] * V[(int64_t)j * n + i];
      w[i] = v;
      s += v * v;
    }
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_counter_set(int *counter, int v) { *counter = v; }
__global__ void k_counter_add(int *counter, int d) { *counter += d; }

// cur = V[0] = x / nrm[0]
__global__ __launch_bounds__(kT) void k_normalize_init(double *__restrict__ cur,
                                                       double *__restrict__ V,
                                                       const double *__restrict__ x, int64_t n,
                                                       const double *__restrict__ nrm) {
  const double s = nrm[0];
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double v = x[i] / s;
    cur[i] = v;
    V[i] = v;
  }
}

// out = sum_j coef[j] V_j, accumulated in the order j = 0, 1, ...  MODE 0: real coefficients
// (any vector of doubles); MODE 1: complex coefficients (re, im pairs) and complex vectors of
// n / 2 values; MODE 2: complex coefficients and real vectors of n values, out complex (2n
// doubles: real parts from the coefficients' real parts, imaginary from their imaginary parts).
template <int MODE>
__global__ __launch_bounds__(kT) void k_combine(double *__restrict__ out,
                                                const double *__restrict__ V,
                                                const double *__restrict__ coef, int k,
                                                int64_t n) {
  const int64_t ne = MODE == 1 ? n / 2 : n;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne;
       i += (int64_t)gridDim.x * kT) {
    if constexpr (MODE == 0) {
      double v = coef[0] * V[i];
      for (int j = 1; j < k; ++j) v += coef[j] * V[(int64_t)j * n + i];
      out[i] = v;
    } else if constexpr (MODE == 1) {
      double vr = 0.0, vi = 0.0;
      for (int j = 0; j < k; ++j) {
        const double *b = V + (int64_t)j * n + 2 * i;
        const double cr = coef[2 * j], ci = coef[2 * j + 1];
        vr += cr * b[0] - ci * b[1];
        vi += cr * b[1] + ci * b[0];
      }
      out[2 * i] = vr;
      out[2 * i + 1] = vi;
    } else {
      double vr = 0.0, vi = 0.0;
      for (int j = 0; j < k; ++j) {
        const double b = V[(int64_t)j * n + i];
        vr += coef[2 * j] * b;
        vi += coef[2 * j + 1] * b;
      }
      out[2 * i] = vr;
      out[2 * i + 1] = vi;
    }
  }
}

__global__ __launch_bounds__(kT) void k_cast_f64_f32(const double *__restrict__ x,
                                                     float *__restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    y[i] = (float)x[i];
}

// ---- dtype-generic dot product / 2-norm (the Vectordot / Norm drop-ins) -------------------
// Products are formed and summed in FP64 whatever the element type; the P partial sums (two
// rows of P for complex: real parts, then imaginary parts) are reduced by one block in a fixed
// order and the result is written in the element type (the norm in its real type).
template <typename T>
struct ElemIO {  // real element: value as FP64
  static constexpr bool kCplx = false;
  __device__ static double re(const T &v) { return (double)v; }
  __device__ static double im(const T &) { return 0.0; }
};
struct cplx64 {
  double re, im;
};
struct cplx32 {
  float re, im;
};
template <>
struct ElemIO<cplx64> {
  static constexpr bool kCplx = true;
  __device__ static double re(const cplx64 &v) { return v.re; }
  __device__ static double im(const cplx64 &v) { return v.im; }
};
template <>
struct ElemIO<cplx32> {
  static constexpr bool kCplx = true;
  __device__ static double re(const cplx32 &v) { return (double)v.re; }
  __device__ static double im(const cplx32 &v) { return (double)v.im; }
};

// mode 0: x.y, mode 1: conj(x).y, mode 2: |x|^2 (y unused)
template <typename T, int MODE>
__global__ __launch_bounds__(kT) void k_vdot_partials(const T *__restrict__ x,
                                                      const T *__restrict__ y, int64_t n,
                                                      double *__restrict__ part) {
  using IO = ElemIO<T>;
  double sr = 0.0, si = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double xr = IO::re(x[i]), xi = IO::im(x[i]);
    if constexpr (MODE == 2) {
      sr += xr * xr + xi * xi;
    } else {
      const double yr = IO::re(y[i]), yi = IO::im(y[i]);
      if constexpr (MODE == 1) {
        sr += xr * yr + xi * yi;
        si += xr * yi - xi * yr;
      } else {
        sr += xr * yr - xi * yi;
        si += xr * yi + xi * yr;
      }
    }
  }
  sr = block_sum(sr);
  if (threadIdx.x == 0) part[blockIdx.x] = sr;
  if constexpr (IO::kCplx && MODE != 2) {
    si = block_sum(si);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = si;
  }
}

// result[0] (and result[1] for a complex dot) in the element type RT; sqrt for the norm
template <typename RT, bool CPLX, bool SQRT>
__global__ __launch_bounds__(kT) void k_vdot_finalize(const double *__restrict__ part, int P,
                                                      RT *__restrict__ result) {
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < P; i += kT) sr += part[i];
  sr = block_sum(sr);
  if constexpr (CPLX) {
    for (int i = threadIdx.x; i < P; i += kT) si += part[P + i];
    si = block_sum(si);
  }
  if (threadIdx.x == 0) {
    result[0] = (RT)(SQRT ? sqrt(sr) : sr);
    if constexpr (CPLX) result[1] = (RT)si;
  }
}

template <typename T, typename RT, int MODE>
int launch_vdot(const void *x, const void *y, int64_t n, double *part, void *result,
                hipStream_t s) {
  constexpr bool cplx = ElemIO<T>::kCplx && MODE != 2;
  hipLaunchKernelGGL((k_vdot_partials<T, MODE>), dim3(kPartials), dim3(kT), 0, s,
                     (const T *)x, (const T *)y, n, part);
  hipLaunchKernelGGL((k_vdot_finalize<RT, cplx, MODE == 2>), dim3(1), dim3(kT), 0, s,
                     (const double *)part, kPartials, (RT *)result);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int grid_of(int64_t n) {
  int64_t g = (n + kT - 1) / kT;
  if (g > kPartials) g = kPartials;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

// ---- internal launchers (declared in vecops.hpp) -------------------------------------------
int vec_dot(const double *x, const double *y, int64_t n, double *part, double *out,
            const int *counter, int offset, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partials, dim3(kPartials), dim3(kT), 0, s, x, y, n, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kT), 0, s, part, kPartials, out, counter, offset,
                     mode);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_dot_partials(const double *x, const double *y, int64_t n, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partials, dim3(kPartials), dim3(kT), 0, s, x, y, n, part);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_dot_basis(const double *y, const double *u, double *V, int64_t n, const double *part_b,
                  double *Bs, IterIdx iter, double *part_a, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_basis, dim3(kPartials), dim3(kT), 0, s, y, u, V, n, part_b, Bs, iter,
                     part_a);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_residual(const double *y, double *u, const double *V, int64_t n, const double *part_a,
                 double *As, const double *Bs, IterIdx iter, double *part_b, hipStream_t s) {
  hipLaunchKernelGGL(k_residual, dim3(kPartials), dim3(kT), 0, s, y, u, V, n, part_a, As, Bs, iter,
                     part_b);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_residual_fused(const double *y, double *u, double *V, int64_t n, const double *dotp,
                       int S, double *As, double *Bs, IterIdx iter, double *part_b,
                       const int *pmap, float *p32, hipStream_t s) {
  if (p32)
    hipLaunchKernelGGL(k_residual_fused<true>, dim3(kPartials), dim3(kT), 0, s, y, u, V, n, dotp,
                       S, As, Bs, iter, part_b, pmap, p32);
  else
    hipLaunchKernelGGL(k_residual_fused<false>, dim3(kPartials), dim3(kT), 0, s, y, u, V, n,
                       dotp, S, As, Bs, iter, part_b, pmap, p32);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_reorth(double *w, const double *V, int64_t n, int max_j, IterIdx iter, double *part,
               double *h, double *part_out, bool cplx, hipStream_t s) {
  // with the iteration known on the host only the live projections are launched
  const int nj = iter.p ? max_j : iter.v + 1;
  const int groups = iter.p ? std::min(8, (max_j + 3) / 4) : (nj + 3) / 4;
  double *part_im = part + (int64_t)kPartials * max_j;
  if (cplx) {
    hipLaunchKernelGGL(k_proj_partials<true>, dim3(kPartials, groups), dim3(kT), 0, s, V, w, n,
                       iter, part, part_im);
    hipLaunchKernelGGL(k_proj_finalize<true>, dim3(nj), dim3(kT), 0, s, part, part_im,
                       kPartials, iter, h);
    hipLaunchKernelGGL(k_proj_apply<true>, dim3(kPartials), dim3(kT), 0, s, w, V, n, h, iter,
                       part_out);
  } else {
    hipLaunchKernelGGL(k_proj_partials<false>, dim3(kPartials, groups), dim3(kT), 0, s, V, w, n,
                       iter, part, part_im);
    hipLaunchKernelGGL(k_proj_finalize<false>, dim3(nj), dim3(kT), 0, s, part, part_im,
                       kPartials, iter, h);
    hipLaunchKernelGGL(k_proj_apply<false>, dim3(kPartials), dim3(kT), 0, s, w, V, n, h, iter,
                       part_out);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_project(const double *V, const double *x, int64_t n, int k, double *part, double *h,
                bool cplx, hipStream_t s) {
  const IterIdx iter{nullptr, k - 1};
  double *part_im = part + (int64_t)kPartials * k;
  if (cplx) {
    hipLaunchKernelGGL(k_proj_partials<true>, dim3(kPartials, (k + 3) / 4), dim3(kT), 0, s, V, x,
                       n, iter, part, part_im);
    hipLaunchKernelGGL(k_proj_finalize<true>, dim3(k), dim3(kT), 0, s, part, part_im, kPartials,
                       iter, h);
  } else {
    hipLaunchKernelGGL(k_proj_partials<false>, dim3(kPartials, (k + 3) / 4), dim3(kT), 0, s, V,
                       x, n, iter, part, part_im);
    hipLaunchKernelGGL(k_proj_finalize<false>, dim3(k), dim3(kT), 0, s, part, part_im, kPartials,
                       iter, h);
  }
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_finalize(const double *part, double *out, const int *counter, int offset, int mode,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kT), 0, s, part, kPartials, out, counter, offset,
                     mode);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_counter_add(int *counter, int d, hipStream_t s) {
  hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(1), 0, s, counter, d);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_counter_set(int *counter, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(1), 0, s, counter, v);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_normalize_init(double *cur, double *V, const double *x, int64_t n, const double *nrm,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_normalize_init, dim3(grid_of(n)), dim3(kT), 0, s, cur, V, x, n, nrm);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_combine(double *out, const double *V, const double *coef, int k, int64_t n, int mode,
                hipStream_t s) {
  const int grid = grid_of(mode == 1 ? n / 2 : n);
  if (mode == 1)
    hipLaunchKernelGGL(k_combine<1>, dim3(grid), dim3(kT), 0, s, out, V, coef, k, n);
  else if (mode == 2)
    hipLaunchKernelGGL(k_combine<2>, dim3(grid), dim3(kT), 0, s, out, V, coef, k, n);
  else
    hipLaunchKernelGGL(k_combine<0>, dim3(grid), dim3(kT), 0, s, out, V, coef, k, n);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_cast_f64_f32(const double *x, float *y, int64_t n, hipStream_t s) {
  if (n <= 0) return CXA_OK;
  hipLaunchKernelGGL(k_cast_f64_f32, dim3(grid_of(n)), dim3(kT), 0, s, x, y, n);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

// One workgroup per region (grid-stride over regions): packed element e of a region holds
// source element (k, x) with k = 4 * (e / 4X) + e % 4, x = (e % 4X) / 4, or zero for k >= K.
// Consecutive threads write consecutive packed elements.
// The region loops of the packing launches: block (x, y) takes regions x, x + gridDim.x, ...
// and, within each, the packed elements y * kT + threadIdx.x, stepping gridDim.y * kT (a few
// large regions still spread over the whole grid).  Element e of a region is (k, x) with
// k = 4 (e / 4X) + (e & 3), x = (e mod 4X) / 4.
template <typename F>
__device__ __forceinline__ void for_region_elems(const PackRegion *__restrict__ regions,
                                                 int nregions, F f) {
  for (int r = blockIdx.x; r < nregions; r += gridDim.x) {
    const PackRegion R = regions[r];
    const int64_t row = 4 * (int64_t)R.X;
    const int64_t nel = (int64_t)((R.K + 3) / 4) * row;
    const int64_t step = (int64_t)gridDim.y * kT;
    if (nel < (int64_t)1 << 31) {  // 32-bit index arithmetic
      const uint32_t row32 = (uint32_t)row;
      for (int64_t e = (int64_t)blockIdx.y * kT + threadIdx.x; e < nel; e += step) {
        const uint32_t e32 = (uint32_t)e;
        const uint32_t q = e32 / row32;
        const uint32_t rem = e32 - q * row32;
        f(R, e, (int64_t)(4 * q + (rem & 3)), (int64_t)(rem >> 2));
      }
    } else {
      for (int64_t e = (int64_t)blockIdx.y * kT + threadIdx.x; e < nel; e += step) {
        const int64_t q = e / row;
        const int64_t rem = e - q * row;
        f(R, e, 4 * q + (rem & 3), rem >> 2);
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kT) void k_pack_b(const double *__restrict__ src, T *__restrict__ dst,
                                               const PackRegion *__restrict__ regions,
                                               int nregions) {
  for_region_elems(regions, nregions, [&](const PackRegion &R, int64_t e, int64_t k, int64_t x) {
    dst[R.dst_off + e] = k < R.K ? (T)src[R.src_off + k * R.sk + x * R.sx] : (T)0;
  });
}

__global__ __launch_bounds__(kT) void k_pack_map(const PackRegion *__restrict__ regions,
                                                int nregions, int64_t n, int *__restrict__ pmap,
                                                int *__restrict__ conflict) {
  for_region_elems(regions, nregions, [&](const PackRegion &R, int64_t e, int64_t k, int64_t x) {
    if (k >= R.K) return;  // padding: no source element
    const int64_t src = R.src_off + k * R.sk + x * R.sx;
    const int dst = (int)(R.dst_off + e);
    if (src < 0 || src >= n) {
      atomicOr(conflict, 1);
      return;
    }
    const int prev = atomicCAS(pmap + src, -1, dst);
    if (prev != -1 && prev != dst) atomicOr(conflict, 1);
  });
}

// grid of the packing launches: regions over x, 16 element slices over y
static dim3 pack_grid(int nregions) { return dim3(nregions < 4096 ? nregions : 4096, 16); }

int vec_pack_map(const PackRegion *regions, int nregions, int64_t n, int *pmap, int *conflict,
                 hipStream_t s) {
  if (nregions <= 0) return CXA_OK;
  hipLaunchKernelGGL(k_pack_map, pack_grid(nregions), dim3(kT), 0, s, regions, nregions, n, pmap,
                     conflict);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

int vec_pack_b(const double *src, void *dst, bool dst_f32, const PackRegion *regions,
               int nregions, hipStream_t s) {
  if (nregions <= 0) return CXA_OK;
  if (dst_f32)
    hipLaunchKernelGGL(k_pack_b<float>, pack_grid(nregions), dim3(kT), 0, s, src, (float *)dst,
                       regions, nregions);
  else
    hipLaunchKernelGGL(k_pack_b<double>, pack_grid(nregions), dim3(kT), 0, s, src,
                       (double *)dst, regions, nregions);
  CXA_HIP(hipGetLastError());
  return CXA_OK;
}

}  // namespace cxa

// ---- public entry points ---------------------------------------------------------------------
// Replaces cuVectordot_internal_{d,f,cd,cf} (src/backend/linalg_internal_gpu/
// cuVectordot_internal.cu:10-180, type Vectordotfunc_oii): result = sum_i x_i y_i, or with
// is_conj sum_i conj(x_i) y_i (the zdotc / cdotc path), written to device memory in the dtype.
extern "C" int cxa_dot(int dtype, int64_t n, const void *x, const void *y, int is_conj,
                       void *result, void *workspace, cxa_stream_t stream) {
  CXA_CHECK(n >= 0 && result && workspace, CXA_ERR_INVALID, "cxa_dot: bad arguments");
  CXA_CHECK(n == 0 || (x && y), CXA_ERR_INVALID, "cxa_dot: null vector");
  hipStream_t s = (hipStream_t)stream;
  double *part = (double *)workspace;
  switch (dtype) {
    case CXA_F64: return cxa::launch_vdot<double, double, 0>(x, y, n, part, result, s);
    case CXA_F32: return cxa::launch_vdot<float, float, 0>(x, y, n, part, result, s);
    case CXA_C128:
      return is_conj ? cxa::launch_vdot<cxa::cplx64, double, 1>(x, y, n, part, result, s)
                     : cxa::launch_vdot<cxa::cplx64, double, 0>(x, y, n, part, result, s);
    case CXA_C64:
      return is_conj ? cxa::launch_vdot<cxa::cplx32, float, 1>(x, y, n, part, result, s)
                     : cxa::launch_vdot<cxa::cplx32, float, 0>(x, y, n, part, result, s);
    default:
      return cxa::fail(CXA_ERR_UNSUPPORTED, "cxa_dot: dtype %d is not F64, F32, C128 or C64", dtype);
  }
}

// Replaces cuNorm_internal_{d,f,cd,cf} (src/backend/linalg_internal_gpu/cuNorm_internal.cu
// :14-60, type Normfunc_oii): result[0] = sqrt(sum_i |x_i|^2) in the dtype's real type (double
// for F64 / C128, float for F32 / C64), written to device memory.
extern "C" int cxa_nrm2(int dtype, int64_t n, const void *x, void *result, void *workspace,
                        cxa_stream_t stream) {
  CXA_CHECK(n >= 0 && result && workspace, CXA_ERR_INVALID, "cxa_nrm2: bad arguments");
  CXA_CHECK(n == 0 || x, CXA_ERR_INVALID, "cxa_nrm2: null vector");
  hipStream_t s = (hipStream_t)stream;
  double *part = (double *)workspace;
  switch (dtype) {
    case CXA_F64: return cxa::launch_vdot<double, double, 2>(x, x, n, part, result, s);
    case CXA_F32: return cxa::launch_vdot<float, float, 2>(x, x, n, part, result, s);
    case CXA_C128: return cxa::launch_vdot<cxa::cplx64, double, 2>(x, x, n, part, result, s);
    case CXA_C64: return cxa::launch_vdot<cxa::cplx32, float, 2>(x, x, n, part, result, s);
    default:
      return cxa::fail(CXA_ERR_UNSUPPORTED, "cxa_nrm2: dtype %d is not F64, F32, C128 or C64", dtype);
  }
}

// partial sums of a complex dot: two rows of kPartials doubles
extern "C" int cxa_vec_workspace_bytes(void) { return 2 * cxa::kPartials * (int)sizeof(double); }
