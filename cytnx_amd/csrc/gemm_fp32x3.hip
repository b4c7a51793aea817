// Grouped GEMM launches in FP32X3 (FP64 operands split hi + lo on FP32 MFMA).
#define CXA_GEMM_TU fp32x3
#include "gemm_kernels.hpp"

namespace cxa {

int launch_gemm_fp32x3(const GemmLaunch &g, hipStream_t s) {
  const int a = g.a_dt, b = g.b_dt, c = g.c_dt;
  if (a == CXA_F64 && b == CXA_F64 && c == CXA_F64)
    return launch_t<CXA_PREC_FP32X3, double, double, double, false>(g, s);
  return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: FP32X3 with dtypes (%d,%d,%d)", a, b, c);
}

}  // namespace cxa
