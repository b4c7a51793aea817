// Grouped GEMM launches in FP32 with an FP32 A operand, and complex64.
#define CXA_GEMM_TU fp32
#include "gemm_kernels.hpp"

namespace cxa {

int launch_gemm_fp32(const GemmLaunch &g, hipStream_t s) {
  const int a = g.a_dt, b = g.b_dt, c = g.c_dt;
#define CXA_F32_CASE(TA_, TB_, TC_, IA, IB, IC) \
  if (a == IA && b == IB && c == IC) return launch_t<CXA_PREC_FP32, TA_, TB_, TC_, false>(g, s);
  CXA_F32_CASE(float, float, float, CXA_F32, CXA_F32, CXA_F32)
  CXA_F32_CASE(float, float, double, CXA_F32, CXA_F32, CXA_F64)
  CXA_F32_CASE(float, double, float, CXA_F32, CXA_F64, CXA_F32)
  CXA_F32_CASE(float, double, double, CXA_F32, CXA_F64, CXA_F64)
#undef CXA_F32_CASE
  if (a == CXA_C64 && b == CXA_C64 && c == CXA_C64)
    return launch_t<CXA_PREC_FP32, cf, cf, cf, true>(g, s);
  if (a == CXA_F64) return launch_gemm_fp32_a64(g, s);
  return fail(CXA_ERR_UNSUPPORTED, "grouped GEMM: FP32 with dtypes (%d,%d,%d)", a, b, c);
}

}  // namespace cxa
