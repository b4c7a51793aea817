// Internal interface of the block-contraction plans (contract.cpp) for other library parts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cytnx_amd.h"

namespace cxa {

// conj (complex plans): bit 0 conjugates L, bit 1 conjugates R as they are read
int bc_exec_typed(const cxa_bc_plan *p, const void *L, const void *R, void *C, int a_dt,
                  int b_dt, int c_dt, int prec, void *workspace, hipStream_t s, int conj = 0);
bool bc_is_mix(const cxa_bc_plan *p);
int64_t bc_workspace_elems(const cxa_bc_plan *p);
double bc_flops(const cxa_bc_plan *p);
int bc_dtype(const cxa_bc_plan *p);

}  // namespace cxa
