// Mix launches (grouped GEMM with a linear-combination epilogue), every precision.
#define CXA_GEMM_TU mix
#include "gemm_kernels.hpp"

namespace cxa {

int launch_mix(const GemmLaunch &g, hipStream_t s) {
  const int p = g.prec, a = g.a_dt, b = g.b_dt, c = g.c_dt;
  if (p == CXA_PREC_FP64 && a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
    return launch_mix_t<CXA_PREC_FP64, double, double, double>(g, s);
  if (p == CXA_PREC_FP32X3 && a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
    return launch_mix_t<CXA_PREC_FP32X3, double, double, double>(g, s);
  if (p == CXA_PREC_FP32) {
    if (a == CXA_F32 && b == CXA_F64 && c == CXA_F32)
      return launch_mix_t<CXA_PREC_FP32, float, double, float>(g, s);
    if (a == CXA_F64 && b == CXA_F32 && c == CXA_F32)  // psi as the left operand
      return launch_mix_t<CXA_PREC_FP32, double, float, float>(g, s);
    if (a == CXA_F32 && b == CXA_F32 && c == CXA_F32)
      return launch_mix_t<CXA_PREC_FP32, float, float, float>(g, s);
    if (a == CXA_F32 && b == CXA_F64 && c == CXA_F64)
      return launch_mix_t<CXA_PREC_FP32, float, double, double>(g, s);
    if (a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
      return launch_mix_t<CXA_PREC_FP32, double, double, double>(g, s);
  }
  return fail(CXA_ERR_UNSUPPORTED, "mix GEMM: unsupported precision %d with dtypes (%d,%d,%d)",
              p, a, b, c);
}

}  // namespace cxa
