// Grouped GEMM launches in FP64 (real and complex128).
#define CXA_GEMM_TU fp64
#include "gemm_kernels.hpp"

namespace cxa {

int launch_gemm_fp64(const GemmLaunch &g, hipStream_t s) {
  const int a = g.a_dt, b = g.b_dt, c = g.c_dt;
  if (a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
    return launch_t<CXA_PREC_FP64, double, double, double, false>(g, s);
  if (a == CXA_C128 && b == CXA_C128 && c == CXA_C128)
    return launch_t<CXA_PREC_FP64, cd, cd, cd, true>(g, s);
  return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: FP64 with dtypes (%d,%d,%d)", a, b, c);
}

}  // namespace cxa
