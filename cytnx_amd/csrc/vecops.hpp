// Internal launchers of the Lanczos vector kernels (vecops.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cxa {

constexpr int kPartials = 512;  // partial sums per reduction (fixed: deterministic order)

// out[(*counter) + offset] = x.y (mode 0) or sqrt(x.y) (mode 1); counter may be null (0)
int vec_dot(const double *x, const double *y, int64_t n, double *part, double *out,
            const int *counter, int offset, int mode, hipStream_t s);
int vec_three_term(double *w, const double *cur, const double *V, int64_t n, const double *As,
                   const double *Bs, const int *counter, double *part, hipStream_t s);
int vec_reorth(double *w, const double *V, int64_t n, int max_j, const int *counter, double *part,
               double *h, hipStream_t s);
int vec_finalize(const double *part, double *out, const int *counter, int offset, int mode,
                 hipStream_t s);
int vec_next(double *cur, double *V, const double *w, int64_t n, const double *Bs,
             const int *counter, hipStream_t s);
int vec_counter_add(int *counter, int d, hipStream_t s);
int vec_counter_set(int *counter, int v, hipStream_t s);
int vec_normalize_init(double *cur, double *V, const double *x, int64_t n, const double *nrm,
                       hipStream_t s);
int vec_combine(double *out, const double *V, const double *coef, int k, int64_t n,
                hipStream_t s);
int vec_cast_f64_f32(const double *x, float *y, int64_t n, hipStream_t s);

}  // namespace cxa
