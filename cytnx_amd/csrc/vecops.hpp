// Internal launchers of the Lanczos vector kernels (vecops.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cxa {

constexpr int kPartials = 512;  // partial sums per reduction (fixed: deterministic order)

// out[(*counter) + offset] = x.y (mode 0) or sqrt(x.y) (mode 1); counter may be null (0)
int vec_dot(const double *x, const double *y, int64_t n, double *part, double *out,
            const int *counter, int offset, int mode, hipStream_t s);
// The Lanczos iteration index: a device counter plus an offset fixed at capture (launches
// captured once and replayed: the counter holds a captured batch's first iteration and
// advances once per batch) or passed by value (launches enqueued per iteration).
struct IterIdx {
  const int *p;
  int v;
  __device__ __forceinline__ int get() const { return p ? *p + v : v; }
};

// Lanczos iteration kernels.  Each reduction leaves kPartials partial sums that the next launch
// sums itself in a fixed order (no separate finishing launch).
// part = kPartials partials of x.y
int vec_dot_partials(const double *x, const double *y, int64_t n, double *part, hipStream_t s);
// After y = H u (u: the previous residual, unnormalised; it = 0: u = V[0]):
// beta = sqrt(sum(part_b)) -> Bs[it-1]; V[it] = u / beta; part_a = partials of (y/beta).V[it]
int vec_dot_basis(const double *y, const double *u, double *V, int64_t n, const double *part_b,
                  double *Bs, IterIdx iter, double *part_a, hipStream_t s);
// As[it] = sum(part_a); u = y/beta - As[it] V[it] - beta V[it-1]; part_b = partials of |u|^2
int vec_residual(const double *y, double *u, const double *V, int64_t n, const double *part_a,
                 double *As, const double *Bs, IterIdx iter, double *part_b, hipStream_t s);
// The two launches above in one, when the matvec produced per-slot partials of Hu . u and
// u . u (dotp[0..S), dotp[S..2S)): beta, alpha, V[it] and the residual with its |u|^2 partials.
// With p32 non-null the residual is also written, rounded to FP32, to p32[pmap[i]] (elements
// with pmap[i] < 0 are skipped): the next matvec's packed input.
int vec_residual_fused(const double *y, double *u, double *V, int64_t n, const double *dotp,
                       int S, double *As, double *Bs, IterIdx iter, double *part_b,
                       const int *pmap, float *p32, hipStream_t s);
// w -= sum_{j <= it} (V_j . w) V_j (classical Gram-Schmidt against the stored basis);
// part_out = partials of the new |w|^2; part: 2 * kPartials * max_j scratch.  cplx: the n
// doubles are n / 2 complex values, the coefficients conj(V_j) . w are complex (h: 2 doubles
// each).
int vec_reorth(double *w, const double *V, int64_t n, int max_j, IterIdx iter, double *part,
               double *h, double *part_out, bool cplx, hipStream_t s);
// h[j] = V_j . x (cplx: conj(V_j) . x, re / im pairs) for j < k (no update of x); part:
// 2 * kPartials * k scratch
int vec_project(const double *V, const double *x, int64_t n, int k, double *part, double *h,
                bool cplx, hipStream_t s);
int vec_finalize(const double *part, double *out, const int *counter, int offset, int mode,
                 hipStream_t s);
int vec_counter_add(int *counter, int d, hipStream_t s);
int vec_counter_set(int *counter, int v, hipStream_t s);
int vec_normalize_init(double *cur, double *V, const double *x, int64_t n, const double *nrm,
                       hipStream_t s);
// out = sum_{j < k} coef_j V_j; mode 0: real coefficients; 1: complex coefficients (re, im
// pairs) and complex vectors (n doubles = n / 2 values); 2: complex coefficients and real
// vectors of n doubles, out complex (2n doubles)
int vec_combine(double *out, const double *V, const double *coef, int k, int64_t n, int mode,
                hipStream_t s);
int vec_cast_f64_f32(const double *x, float *y, int64_t n, hipStream_t s);

// Fill the k4-packed copy of a B operand (regions of contract.hpp's PackRegion, nregions of
// them, on the device): dst = FP64 or FP32 (dst_f32), src = FP64.
struct PackRegion;
int vec_pack_b(const double *src, void *dst, bool dst_f32, const PackRegion *regions,
               int nregions, hipStream_t s);
// The map from source elements to packed positions of the same regions: pmap[src] = dst for
// every element a region reads (pmap must hold -1 in all n entries beforehand); *conflict
// (device, 0 beforehand) becomes nonzero if an element has two packed positions or a region
// reads outside [0, n).
int vec_pack_map(const PackRegion *regions, int nregions, int64_t n, int *pmap, int *conflict,
                 hipStream_t s);

}  // namespace cxa
