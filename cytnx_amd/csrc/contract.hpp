// Internal interface of the block-contraction plans (contract.cpp) for other library parts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/cytnx_amd.h"
#include "gemm.hpp"

namespace cxa {

// Operand layouts of an execution other than the plan's own (null: the plan's): segments with
// k4-packed views (GemmLaunch::a_k4 / b_k4), mix outputs written k4-packed (c_k4).
struct ExecLayout {
  const GSeg *segs = nullptr;
  const MixOut *outs = nullptr;
  bool a_k4 = false, b_k4 = false, c_k4 = false;
  // dot-product epilogue (GemmLaunch::dot_x / dot_part): dot_part gets bc_dot_slots entries
  const double *dot_x = nullptr;
  double *dot_part = nullptr;
};

// Batched permute of many small blocks in one launch (permute.hip).  Job j writes its
// n elements contiguously: dst[dst_off + e] = src[src_off + sum_d idx_d(e) * sstr[d]], idx(e)
// the row-major coordinates of e over ext[0 .. rank); a task is a contiguous element range of
// one job, one workgroup each.
constexpr int kBatchCopyMaxElems = 1 << 15;  // larger blocks go through cxa_permute one by one
constexpr int kBatchCopyTaskElems = 4096;
struct BatchCopyJob {
  int64_t src_off, dst_off;  // elements
  int32_t n, rank;
  int32_t ext[16];
  int32_t sstr[16];
};
struct BatchCopyTask {
  int32_t job, start, count, pad;
};
// The job that permutes one row-major block of extents shape[rank] by mapper (new axis d = old
// axis mapper[d]); false when the block is too large for the batched launch.
bool batch_copy_job(int rank, const uint64_t *shape, const uint64_t *mapper, int64_t src_off,
                    int64_t dst_off, BatchCopyJob *job);
// One launch over ntasks tasks (device arrays) for elements of elem_size bytes (4, 8 or 16).
int launch_batch_copy(int elem_size, const BatchCopyJob *jobs, const BatchCopyTask *tasks,
                      int ntasks, const void *src, void *dst, hipStream_t s);

// conj (complex plans): bit 0 conjugates L, bit 1 conjugates R as they are read.
int bc_exec_typed(const cxa_bc_plan *p, const void *L, const void *R, void *C, int a_dt,
                  int b_dt, int c_dt, int prec, void *workspace, hipStream_t s, int conj = 0,
                  const ExecLayout *lay = nullptr);

// One distinct operand view of a plan and where its packed copy goes: source element (k, x) at
// src_off + k * sk + x * sx (x = n of a B view, m of an A view), packed element (k, x) at
// dst_off + (k >> 2) * 4X + 4x + (k & 3) for k < 4 * ceil(K / 4) (zero for k >= K).
struct PackRegion {
  int64_t dst_off, src_off, sk, sx;
  int32_t K, X;
};

// The k4-packed form of a real plan's B (bc_pack_b) or A (bc_pack_a) operand, read directly
// from the caller's buffer (no permute copy): `segs` (in: the segments to remap, empty for the
// plan's own; out: those views remapped into the packed buffer, offset = region start, stride
// along k = 4X, along the free index 4), the regions (one per distinct view, each a multiple
// of 4 elements) and the packed buffer's size.  Views are checked against the operand buffer's
// `elems` elements.  CXA_ERR_UNSUPPORTED for plans that cannot be packed.
int bc_pack_b(const cxa_bc_plan *p, int64_t elems, std::vector<GSeg> &segs,
              std::vector<PackRegion> &regions, int64_t *packed_elems);
int bc_pack_a(const cxa_bc_plan *p, int64_t elems, std::vector<GSeg> &segs,
              std::vector<PackRegion> &regions, int64_t *packed_elems);

// A mix plan's outputs redirected into the packed A regions of the launch that reads them
// (regions from bc_pack_a on that launch's plan): every output must be rows k0.. k0+M-1 of
// one region read with unit stride along x (c_off = region start + k0 * sk, ldc = sk,
// N = X).  On success outs holds the packed outputs (c_off = region's dst_off, ldc = k0).
int bc_mix_pack_outs(const cxa_bc_plan *mix, const std::vector<PackRegion> &regions,
                     std::vector<MixOut> &outs);
bool bc_is_mix(const cxa_bc_plan *p);
// Make a plan's tile list for one precision if it is not made yet (the plan makes only its own
// precision at creation); refused while `s` is capturing a graph.  bc_prepare makes all three.
int bc_ensure(const cxa_bc_plan *p, int prec, hipStream_t s);
int bc_prepare(const cxa_bc_plan *p);
// Tile slots a dot-product epilogue of an execution at precision prec writes (the launch's tile
// list length), or 0 when that launch cannot carry one (mix, complex, row-streaming tiles).
int bc_dot_slots(const cxa_bc_plan *p, int prec);
int64_t bc_workspace_elems(const cxa_bc_plan *p);
double bc_flops(const cxa_bc_plan *p);
int bc_dtype(const cxa_bc_plan *p);

}  // namespace cxa
