// Internal interface of the block-contraction plans (contract.cpp) for other library parts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/cytnx_amd.h"
#include "gemm.hpp"

namespace cxa {

// conj (complex plans): bit 0 conjugates L, bit 1 conjugates R as they are read.  segs / b_k4:
// run the plan's tiles with another segment array whose B views are k4-packed (bc_pack_b).
int bc_exec_typed(const cxa_bc_plan *p, const void *L, const void *R, void *C, int a_dt,
                  int b_dt, int c_dt, int prec, void *workspace, hipStream_t s, int conj = 0,
                  const GSeg *segs = nullptr, bool b_k4 = false);

// One distinct B view of a plan and where its packed copy goes: source element (k, n) at
// src_off + k * sbk + n * sbn, packed element (k, n) at dst_off + (k >> 2) * 4N + 4n + (k & 3)
// for k < 4 * ceil(K / 4) (zero for k >= K).
struct PackRegion {
  int64_t dst_off, src_off, sbk, sbn;
  int32_t K, N;
};

// The k4-packed form of a real plan's B operand (read directly from the caller's buffer, no
// permute copy): the plan's segments with their B views remapped into the packed buffer
// (b_off = region start, sbk = 4N, sbn = 4), the regions to fill (one per distinct view, each
// a multiple of 4 elements), and the packed buffer's size.  Views are checked against the B
// buffer's b_elems elements.  CXA_ERR_UNSUPPORTED for plans that cannot be packed.
int bc_pack_b(const cxa_bc_plan *p, int64_t b_elems, std::vector<GSeg> &segs,
              std::vector<PackRegion> &regions, int64_t *packed_elems);
bool bc_is_mix(const cxa_bc_plan *p);
int64_t bc_workspace_elems(const cxa_bc_plan *p);
double bc_flops(const cxa_bc_plan *p);
int bc_dtype(const cxa_bc_plan *p);

}  // namespace cxa
