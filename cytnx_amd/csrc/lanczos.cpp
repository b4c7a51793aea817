// Matvec chains (the effective-Hamiltonian program) and the ground-state Lanczos driver.
//
// A chain is a fixed sequence of block-contraction plans: step 1 contracts the input vector
// with a fixed operand (an environment), each later step contracts the running intermediate
// with the next fixed operand (MPO tensors, the other environment).  Intermediates and FP32
// copies of the fixed operands are allocated once, so applying a chain allocates nothing and can
// be captured into a graph.
//
// The Lanczos driver restates Lanczos_Gnd_Ut (src/linalg/Lanczos_Gnd_Ut.cpp:29-162): plain
// three-term recurrence (optionally with FP64 re-orthogonalisation), tridiagonal eigensolve,
// stop at the first iteration with |E - E_prev| < CvgCrit or beta == 0, Ritz vector from the
// stored Krylov basis.  Device scalars and a device iteration counter let a batch of iterations
// run without host round trips; the host reads the alpha/beta history once per batch and
// applies the reference's stopping rule iteration by iteration, so the result is the one a
// per-iteration check would give.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "contract.hpp"
#include "cxa_common.hpp"
#include "gemm.hpp"
#include "vecops.hpp"

using namespace cxa;

// ------------------------------------------------------------------------------------------
// Ground state of a symmetric tridiagonal matrix (diag a[0..n), off-diagonal b[0..n-1)):
// smallest eigenvalue by Sturm-sequence bisection, eigenvector by inverse iteration with a
// partially pivoted tridiagonal LU (the LINPACK/LAPACK gtsv elimination).  Only the lowest
// pair is needed by the ground-state Lanczos, where the reference computes all pairs with
// LAPACKE_dstev (src/backend/linalg_internal_cpu/Tridiag_internal.cpp:9-42) and uses [0].
// ------------------------------------------------------------------------------------------
namespace {

int sturm_count(int n, const double *a, const double *b2, double x, double pivmin) {
  int cnt = 0;
  double q = a[0] - x;
  if (std::fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0;
  for (int i = 1; i < n; ++i) {
    q = (a[i] - x) - b2[i - 1] / q;
    if (std::fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0;
  }
  return cnt;
}

// Sturm counts at three shifts in one pass: three independent recurrences, so the divisions of
// one overlap those of the others instead of waiting on each other.
void sturm_count3(int n, const double *a, const double *b2, const double *x, double pivmin,
                  int *cnt) {
  double q[3];
  int c[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    q[k] = a[0] - x[k];
    if (std::fabs(q[k]) < pivmin) q[k] = -pivmin;
    c[k] += q[k] < 0;
  }
  for (int i = 1; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      q[k] = (a[i] - x[k]) - b2[i - 1] / q[k];
      if (std::fabs(q[k]) < pivmin) q[k] = -pivmin;
      c[k] += q[k] < 0;
    }
  }
  for (int k = 0; k < 3; ++k) cnt[k] = c[k];
}

// Smallest eigenvalue inside the bracket [lo, hi] (count(lo) == 0 <= 1 <= count(hi)) by
// quadrisection: each pass keeps the quarter that holds the eigenvalue, until the bracket is
// two ulps wide relative to its ends.
double lowest_in(int n, const double *a, const double *b2, double lo, double hi, double pivmin) {
  const double eps = 2.220446049250313e-16;
  for (int it = 0; it < 200; ++it) {
    if (hi - lo <= 2.0 * eps * std::fmax(std::fabs(lo), std::fabs(hi))) break;
    const double x[3] = {lo + 0.25 * (hi - lo), 0.5 * (lo + hi), hi - 0.25 * (hi - lo)};
    if (!(x[0] > lo && x[2] < hi)) break;  // no representable interior points left
    int c[3];
    sturm_count3(n, a, b2, x, pivmin, c);
    if (c[0] >= 1) {
      hi = x[0];
    } else if (c[1] >= 1) {
      lo = x[0];
      hi = x[1];
    } else if (c[2] >= 1) {
      lo = x[1];
      hi = x[2];
    } else {
      lo = x[2];
    }
  }
  return 0.5 * (lo + hi);
}

// solve (T - lam I) x = r in place (r -> x) with partial pivoting
void gt_solve(int n, const double *a, const double *b, double lam, double tiny,
              std::vector<double> &r) {
  std::vector<double> dl(n > 1 ? n - 1 : 1), d(n), du(n > 1 ? n - 1 : 1), du2(n > 2 ? n - 2 : 1, 0.0);
  std::vector<char> swp(n > 1 ? n - 1 : 1, 0);
  for (int i = 0; i < n; ++i) d[i] = a[i] - lam;
  for (int i = 0; i + 1 < n; ++i) dl[i] = du[i] = b[i];
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (d[i] == 0.0) d[i] = tiny;
      const double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
      if (i + 2 < n) du2[i] = 0.0;
    } else {
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      const double t = d[i + 1];
      d[i + 1] = du[i] - f * t;
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du2[i];
      }
      du[i] = t;
      dl[i] = f;
      swp[i] = 1;
    }
  }
  if (d[n - 1] == 0.0) d[n - 1] = tiny;
  for (int i = 0; i + 1 < n; ++i) {
    if (swp[i]) std::swap(r[i], r[i + 1]);
    r[i + 1] -= dl[i] * r[i];
  }
  r[n - 1] /= d[n - 1];
  if (n > 1) r[n - 2] = (r[n - 2] - du[n - 2] * r[n - 1]) / d[n - 2];
  for (int i = n - 3; i >= 0; --i) r[i] = (r[i] - du[i] * r[i + 1] - du2[i] * r[i + 2]) / d[i];
}

// Lowest eigenvalue of the symmetric tridiagonal matrix with diagonal a and off-diagonal b.
// Without `upper` the bracket is the Gershgorin interval.  With `upper` = the lowest eigenvalue of
// the leading (n-1) x (n-1) block (what the previous Lanczos step computed) the bracket is
// narrowed: by Cauchy interlacing the eigenvalue lies below `upper`, and by Weyl's inequality
// it lies above min(upper, a[n-1]) - |b[n-2]|; a first guess of `upper - step` (step: a few times
// the previous change) is tried before that bound.  Every end is confirmed by a Sturm count and
// the Gershgorin end is kept where a confirmation fails, so the result does not depend on the hint
// beyond rounding.
int tridiag_lowest(int n, const double *a, const double *b, const double *upper, double step,
                   double *E, double *nrm_out) {
  std::vector<double> b2(n - 1);
  double lo = a[0], hi = a[0], nrm = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(b[i]) : 0.0);
    lo = std::fmin(lo, a[i] - r);
    hi = std::fmax(hi, a[i] + r);
    nrm = std::fmax(nrm, std::fabs(a[i]) + r);
  }
  for (int i = 0; i < n; ++i)
    CXA_CHECK(std::isfinite(a[i]) && (i + 1 == n || std::isfinite(b[i])), CXA_ERR_NUMERIC,
              "cxa_tridiag_ground: non-finite entries");
  for (int i = 0; i + 1 < n; ++i) b2[i] = b[i] * b[i];
  const double eps = 2.220446049250313e-16;
  // smallest pivot magnitude allowed in the Sturm recurrence (keeps b^2/q finite)
  const double pivmin = 1e-290 * std::fmax(1.0, nrm * nrm);
  lo -= eps * nrm + pivmin;
  hi += eps * nrm + pivmin;
  if (upper && std::isfinite(*upper)) {
    const double u = *upper;
    const double slack = 8.0 * eps * (nrm + std::fabs(u)) + pivmin;
    const double h = u + slack;
    if (h < hi && sturm_count(n, a, b2.data(), h, pivmin) >= 1) hi = h;
    const double guess = hi - std::fmax(4.0 * step, 64.0 * eps * std::fmax(nrm, 1.0));
    const double weyl = std::fmin(u, a[n - 1]) - std::fabs(b[n - 2]) - slack;
    if (guess > lo && guess < hi && sturm_count(n, a, b2.data(), guess, pivmin) == 0)
      lo = guess;
    else if (weyl > lo && weyl < hi && sturm_count(n, a, b2.data(), weyl, pivmin) == 0)
      lo = weyl;
  }
  *E = lowest_in(n, a, b2.data(), lo, hi, pivmin);
  if (nrm_out) *nrm_out = nrm;
  return CXA_OK;
}

}  // namespace

extern "C" int cxa_tridiag_ground(int n, const double *a, const double *b, double *E,
                                  double *z) {
  CXA_CHECK(n >= 1 && a && E, CXA_ERR_INVALID, "cxa_tridiag_ground: bad arguments");
  CXA_CHECK(n == 1 || b, CXA_ERR_INVALID, "cxa_tridiag_ground: null off-diagonal");
  if (n == 1) {
    *E = a[0];
    if (z) z[0] = 1.0;
    return CXA_OK;
  }
  const double eps = 2.220446049250313e-16;
  double lam = 0.0, nrm = 0.0;
  int rc = tridiag_lowest(n, a, b, nullptr, 0.0, &lam, &nrm);
  if (rc) return rc;
  *E = lam;
  if (!z) return CXA_OK;
  std::vector<double> r(n, 1.0 / std::sqrt((double)n));
  const double tiny = eps * std::fmax(nrm, 1e-300);
  for (int k = 0; k < 3; ++k) {
    gt_solve(n, a, b, lam, tiny, r);
    double s = 0.0;
    for (double v : r) s += v * v;
    s = std::sqrt(s);
    CXA_CHECK(s > 0 && std::isfinite(s), CXA_ERR_NUMERIC,
              "cxa_tridiag_ground: inverse iteration failed");
    for (double &v : r) v /= s;
  }
  std::memcpy(z, r.data(), sizeof(double) * n);
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// chains
// ------------------------------------------------------------------------------------------
struct cxa_chain {
  int device = 0;
  int nsteps = 0;
  int dtype = CXA_F64;  // every plan's dtype: F64, or C128 (complex chains: FP64, unpacked)
  std::vector<const cxa_bc_plan *> plans;
  std::vector<const void *> fixed;
  std::vector<int> fixed_left;
  std::vector<int64_t> fixed_elems, out_elems;
  std::vector<void *> mid64, mid32, fixed32;
  // Steps whose fixed operand is B read it k4-packed (one vector load per lane and k-quad):
  // the plan's segments remapped into the packed buffer (device), the packed FP64 copy (made
  // at creation) and FP32 copy (made by cxa_chain_prepare for FP32), null where not packed.
  std::vector<void *> packed_segs, packed64, packed32;
  std::vector<int64_t> packed_elems;
  std::vector<std::vector<cxa::PackRegion>> pack_regions;
  std::vector<void *> pack_regions_dev;  // the same on the device (uploaded at the first fill)
  // A mix step followed by a step that reads its output as A: the output is written k4-packed
  // (packed_outs: the mix plan's outputs redirected, device) and read as a packed A.
  std::vector<void *> packed_outs;
  std::vector<int> a_packed, c_packed;
  std::vector<int64_t> mid_elems;  // elements of each intermediate buffer (packed: padded)
  // The running vector as a k4-packed FP32 A operand of a first mix step (FP32 ground-state
  // Lanczos): pk_segs0 holds that step's segments with B (and here A) remapped, pk_regions its
  // A regions; cxa_chain_prepare(FP32) uploads them and builds pk_map (vector element ->
  // packed position, -1 where no view reads it) and the zero-padded packed copy pk_psi32,
  // which the fused residual launch rewrites every iteration.  Empty / null where the first
  // step is not such a launch or a vector element has two packed positions.
  std::vector<cxa::GSeg> pk_segs0;
  std::vector<cxa::PackRegion> pk_regions;
  int64_t pk_n = 0, pk_elems = 0;
  void *pk_segs = nullptr, *pk_reg_dev = nullptr;
  int *pk_map = nullptr;
  float *pk_psi32 = nullptr;
  void *ws = nullptr;  // permute workspace (max over steps), 16 bytes per element
  int prec = CXA_PREC_FP64;
  double flops = 0;
  ~cxa_chain() {
    for (auto *v : {&mid64, &mid32, &fixed32, &packed_segs, &packed64, &packed32, &packed_outs,
                    &pack_regions_dev})
      for (void *p : *v)
        if (p) cxa::pool_release(p);
    for (void *p : {pk_segs, pk_reg_dev, (void *)pk_map, (void *)pk_psi32})
      if (p) cxa::pool_release(p);
    cxa::pool_release(ws);
  }
};

namespace {

// Fill a packed copy of step k's fixed operand (FP64 source) and wait for it: on the library's
// upload stream (the fixed operands were synchronised when the chain was created), with the
// region table uploaded once and kept with the chain.
int fill_packed(cxa_chain *c, int k, void *dst, bool f32) {
  const std::vector<cxa::PackRegion> &reg = c->pack_regions[k];
  if (reg.empty()) return CXA_OK;
  hipStream_t s;
  if (const int rc = cxa::upload_stream(&s)) return rc;
  if (!c->pack_regions_dev[k]) {
    CXA_RC(cxa::pool_alloc((void **)&c->pack_regions_dev[k],
                           reg.size() * sizeof(cxa::PackRegion)));
    const int rc = cxa::upload_now(c->pack_regions_dev[k], reg.data(),
                                   reg.size() * sizeof(cxa::PackRegion));
    if (rc) return rc;
  }
  const int rc = cxa::vec_pack_b((const double *)c->fixed[k], dst, f32,
                                 (const cxa::PackRegion *)c->pack_regions_dev[k],
                                 (int)reg.size(), s);
  if (rc) return rc;
  CXA_HIP(hipStreamSynchronize(s));
  return CXA_OK;
}

}  // namespace

extern "C" int cxa_chain_create(int nsteps, cxa_bc_plan *const *plans, const void *const *fixed,
                                const int *fixed_is_left, const int64_t *fixed_elems,
                                const int64_t *out_elems, cxa_chain **out) {
  CXA_CHECK(out && nsteps >= 1 && plans && fixed && fixed_is_left && fixed_elems && out_elems,
            CXA_ERR_INVALID, "cxa_chain_create: bad arguments");
  *out = nullptr;
  std::unique_ptr<cxa_chain> c(new cxa_chain());
  CXA_HIP(hipGetDevice(&c->device));
  c->nsteps = nsteps;
  int64_t wsmax = 0;
  CXA_CHECK(plans[0], CXA_ERR_INVALID, "cxa_chain_create: null plan at step 0");
  c->dtype = bc_dtype(plans[0]);
  CXA_CHECK(c->dtype == CXA_F64 || c->dtype == CXA_C128, CXA_ERR_UNSUPPORTED,
            "cxa_chain_create: chains run F64 or C128 plans");
  const bool cplx = c->dtype == CXA_C128;
  const size_t es = cplx ? 16 : 8;
  for (int k = 0; k < nsteps; ++k) {
    CXA_CHECK(plans[k], CXA_ERR_INVALID, "cxa_chain_create: null plan at step %d", k);
    CXA_CHECK(bc_dtype(plans[k]) == c->dtype, CXA_ERR_UNSUPPORTED,
              "cxa_chain_create: every step's plan must have the first step's dtype (step %d)",
              k);
    CXA_CHECK(!cplx || !bc_is_mix(plans[k]), CXA_ERR_UNSUPPORTED,
              "cxa_chain_create: mix plans are real (step %d)", k);
    c->plans.push_back(plans[k]);
    c->fixed.push_back(fixed[k]);
    c->fixed_left.push_back(fixed_is_left[k] ? 1 : 0);
    c->fixed_elems.push_back(fixed_elems[k]);
    c->out_elems.push_back(out_elems[k]);
    c->flops += bc_flops(plans[k]);
    wsmax = std::max(wsmax, bc_workspace_elems(plans[k]));
  }
  c->mid64.assign(nsteps > 1 ? nsteps - 1 : 0, nullptr);
  for (int k = 0; k + 1 < nsteps; ++k)
    CXA_RC(cxa::pool_alloc((void **)&c->mid64[k], std::max<int64_t>(1, out_elems[k]) * es));
  if (wsmax > 0) CXA_RC(cxa::pool_alloc((void **)&c->ws, wsmax * 16));
  // k4-packed copies of the fixed B operands (read once here: the chain keeps using these
  // copies, so a fixed operand changed after creation needs a new chain)
  c->packed_segs.assign(nsteps, nullptr);
  c->packed64.assign(nsteps, nullptr);
  c->packed32.assign(nsteps, nullptr);
  c->packed_elems.assign(nsteps, 0);
  c->pack_regions.assign(nsteps, {});
  c->pack_regions_dev.assign(nsteps, nullptr);
  c->packed_outs.assign(nsteps, nullptr);
  c->a_packed.assign(nsteps, 0);
  c->c_packed.assign(nsteps, 0);
  c->mid_elems.assign(c->mid64.size(), 0);
  for (size_t k = 0; k < c->mid64.size(); ++k) c->mid_elems[k] = std::max<int64_t>(1, out_elems[k]);
  std::vector<std::vector<cxa::GSeg>> segs_host(nsteps);
  bool synced = false;
  for (int k = 0; k < nsteps && !cplx; ++k) {
    if (c->fixed_left[k]) continue;
    std::vector<cxa::GSeg> &segs = segs_host[k];
    int64_t elems = 0;
    if (cxa::bc_pack_b(plans[k], fixed_elems[k], segs, c->pack_regions[k], &elems) != CXA_OK) {
      segs.clear();
      continue;  // this plan reads B through a workspace: unpacked
    }
    if (!synced) {  // the fixed operands may come from any stream
      CXA_HIP(hipDeviceSynchronize());
      synced = true;
    }
    c->packed_elems[k] = elems;
    CXA_RC(cxa::pool_alloc((void **)&c->packed_segs[k],
                           std::max<size_t>(1, segs.size()) * sizeof(cxa::GSeg)));
    int rc = cxa::upload_now(c->packed_segs[k], segs.data(), segs.size() * sizeof(cxa::GSeg));
    if (rc) return rc;
    CXA_RC(cxa::pool_alloc((void **)&c->packed64[k], std::max<int64_t>(1, elems) * sizeof(double)));
    rc = fill_packed(c.get(), k, c->packed64[k], false);
    if (rc) return rc;
  }
  // a mix step whose output the next step reads as its (running) A operand: write it packed
  for (int k = 0; k + 1 < nsteps && !cplx; ++k) {
    if (!cxa::bc_is_mix(plans[k]) || !c->packed_segs[k] || c->fixed_left[k + 1] ||
        !c->packed_segs[k + 1])
      continue;
    std::vector<cxa::GSeg> segs = segs_host[k + 1];  // B already remapped
    std::vector<cxa::PackRegion> regions;
    std::vector<cxa::MixOut> outs;
    int64_t elems = 0;
    if (cxa::bc_pack_a(plans[k + 1], out_elems[k], segs, regions, &elems) != CXA_OK ||
        cxa::bc_mix_pack_outs(plans[k], regions, outs) != CXA_OK)
      continue;
    int rc = cxa::upload_now(c->packed_segs[k + 1], segs.data(),
                             segs.size() * sizeof(cxa::GSeg));
    if (rc) return rc;
    CXA_RC(cxa::pool_alloc((void **)&c->packed_outs[k],
                           std::max<size_t>(1, outs.size()) * sizeof(cxa::MixOut)));
    rc = cxa::upload_now(c->packed_outs[k], outs.data(), outs.size() * sizeof(cxa::MixOut));
    if (rc) return rc;
    // the packed intermediate: quads past K are never written and stay zero
    c->mid_elems[k] = std::max<int64_t>(1, elems);
    cxa::pool_release(c->mid64[k]);
    c->mid64[k] = nullptr;
    CXA_RC(cxa::pool_alloc((void **)&c->mid64[k], c->mid_elems[k] * sizeof(double)));
    rc = cxa::zero_now(c->mid64[k], c->mid_elems[k] * sizeof(double));
    if (rc) return rc;
    c->a_packed[k + 1] = 1;
    c->c_packed[k] = 1;
  }
  // the running vector packed for a first mix step that writes packed outputs (the vector
  // maps to itself: its length is the last step's output)
  if (!cplx && nsteps > 1 && c->c_packed[0] && !c->fixed_left[0]) {
    std::vector<cxa::GSeg> segs = segs_host[0];
    int64_t elems = 0;
    if (cxa::bc_pack_a(plans[0], out_elems[nsteps - 1], segs, c->pk_regions, &elems) == CXA_OK &&
        elems <= INT32_MAX) {
      c->pk_segs0 = std::move(segs);
      c->pk_n = out_elems[nsteps - 1];
      c->pk_elems = elems;
    } else {
      c->pk_regions.clear();
    }
  }
  *out = c.release();
  return CXA_OK;
}

namespace {

// The packed running vector's map and buffers (FP32 set-up), built on the device: every
// vector element that a view reads must get exactly one packed position, or the chain keeps
// reading the FP64 vector (the map and buffers are released again).
int build_packed_vector(cxa_chain *c) {
  if (c->pk_regions.empty() || c->pk_n <= 0) return CXA_OK;
  // set-up work with no dependence on the caller's queued work: the library's upload stream
  hipStream_t s;
  int rc = upload_stream(&s);
  if (rc) return rc;
  const size_t nreg = c->pk_regions.size();
  CXA_RC(cxa::pool_alloc((void **)&c->pk_reg_dev, nreg * sizeof(cxa::PackRegion)));
  CXA_HIP(hipMemcpyAsync(c->pk_reg_dev, c->pk_regions.data(), nreg * sizeof(cxa::PackRegion),
                         hipMemcpyHostToDevice, s));
  // + the conflict flag
  CXA_RC(cxa::pool_alloc((void **)&c->pk_map, sizeof(int32_t) * (c->pk_n + 1)));
  int *flag = c->pk_map + c->pk_n;
  CXA_HIP(hipMemsetAsync(c->pk_map, 0xff, sizeof(int32_t) * c->pk_n, s));
  CXA_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
  rc = vec_pack_map((const cxa::PackRegion *)c->pk_reg_dev, (int)nreg, c->pk_n, c->pk_map, flag,
                    s);
  if (rc) return rc;
  int conflict = 0;
  CXA_HIP(hipMemcpyAsync(&conflict, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  CXA_HIP(hipStreamSynchronize(s));
  if (conflict) {  // no packed vector for this chain
    cxa::pool_release(c->pk_map);
    c->pk_map = nullptr;
    cxa::pool_release(c->pk_reg_dev);
    c->pk_reg_dev = nullptr;
    return CXA_OK;
  }
  CXA_RC(cxa::pool_alloc((void **)&c->pk_segs,
                         std::max<size_t>(1, c->pk_segs0.size()) * sizeof(cxa::GSeg)));
  if (!c->pk_segs0.empty())
    CXA_HIP(hipMemcpyAsync(c->pk_segs, c->pk_segs0.data(),
                           c->pk_segs0.size() * sizeof(cxa::GSeg), hipMemcpyHostToDevice, s));
  CXA_RC(cxa::pool_alloc((void **)&c->pk_psi32,
                         sizeof(float) * std::max<int64_t>(1, c->pk_elems)));
  CXA_HIP(hipMemsetAsync(c->pk_psi32, 0, sizeof(float) * std::max<int64_t>(1, c->pk_elems), s));
  CXA_HIP(hipStreamSynchronize(s));  // ready for runs on any stream
  return CXA_OK;
}

}  // namespace

// Allocate and fill the FP32 copies a precision mode needs (call before graph capture).
extern "C" int cxa_chain_prepare(cxa_chain *c, int prec, cxa_stream_t stream) {
  CXA_CHECK(c, CXA_ERR_INVALID, "cxa_chain_prepare: null chain");
  CXA_CHECK(prec >= CXA_PREC_FP64 && prec <= CXA_PREC_FP32X3, CXA_ERR_INVALID,
            "cxa_chain_prepare: unknown precision %d", prec);
  CXA_CHECK(c->dtype == CXA_F64 || prec == CXA_PREC_FP64, CXA_ERR_UNSUPPORTED,
            "cxa_chain_prepare: complex chains run in FP64 only");
  c->prec = prec;
  for (int k = 0; k < c->nsteps; ++k) {  // tile lists of this precision, before any capture
    const int rc = bc_ensure(c->plans[k], prec, nullptr);
    if (rc) return rc;
  }
  if (prec != CXA_PREC_FP32 || !c->fixed32.empty()) return CXA_OK;
  hipStream_t s = (hipStream_t)stream;
  c->fixed32.assign(c->nsteps, nullptr);
  c->mid32.assign(c->mid64.size(), nullptr);
  for (int k = 0; k < c->nsteps; ++k) {
    if (c->packed_segs[k]) {  // packed steps keep only the packed FP32 copy
      CXA_RC(cxa::pool_alloc((void **)&c->packed32[k],
                             std::max<int64_t>(1, c->packed_elems[k]) * sizeof(float)));
      int rc = fill_packed(c, k, c->packed32[k], true);
      if (rc) return rc;
      continue;
    }
    CXA_RC(cxa::pool_alloc((void **)&c->fixed32[k],
                           std::max<int64_t>(1, c->fixed_elems[k]) * sizeof(float)));
    int rc = vec_cast_f64_f32((const double *)c->fixed[k], (float *)c->fixed32[k],
                              c->fixed_elems[k], s);
    if (rc) return rc;
  }
  for (size_t k = 0; k < c->mid32.size(); ++k) {
    CXA_RC(cxa::pool_alloc((void **)&c->mid32[k], c->mid_elems[k] * sizeof(float)));
    if (c->c_packed[k]) {  // the padding of a packed intermediate must read as zero
      const int rc = zero_now(c->mid32[k], c->mid_elems[k] * sizeof(float));
      if (rc) return rc;
    }
  }
  return build_packed_vector(c);
}

namespace {

// dotx / dotp (last step of a real chain only): the launch's dot-product epilogue, with dotp
// holding 2 * chain_dot_slots doubles.
// pk: step 0 reads the packed FP32 running vector (pk_psi32, kept equal to x by the caller)
// instead of x.
int chain_step(cxa_chain *c, int k, const void *x, void *y, hipStream_t s,
               const double *dotx = nullptr, double *dotp = nullptr, bool pk = false) {
  if (c->dtype != CXA_F64) {  // complex chain: FP64, every operand as the plan lays it out
    const void *in = k == 0 ? x : c->mid64[k - 1];
    void *outp = k == c->nsteps - 1 ? y : c->mid64[k];
    const int dt = c->dtype;
    return c->fixed_left[k]
               ? bc_exec_typed(c->plans[k], c->fixed[k], in, outp, dt, dt, dt, c->prec, c->ws, s)
               : bc_exec_typed(c->plans[k], in, c->fixed[k], outp, dt, dt, dt, c->prec, c->ws, s);
  }
  const bool f32 = c->prec == CXA_PREC_FP32;
  const bool last = k == c->nsteps - 1;
  const void *in = k == 0 ? x : (f32 ? c->mid32[k - 1] : c->mid64[k - 1]);
  const int in_dt = (k == 0 || !f32) ? CXA_F64 : CXA_F32;
  void *outp = last ? y : (f32 ? c->mid32[k] : c->mid64[k]);
  const int out_dt = (last || !f32) ? CXA_F64 : CXA_F32;
  const int fx_dt = f32 ? CXA_F32 : CXA_F64;
  cxa::ExecLayout lay;
  lay.dot_x = dotx;
  lay.dot_part = dotp;
  if (c->fixed_left[k])
    return bc_exec_typed(c->plans[k], f32 ? c->fixed32[k] : c->fixed[k], in, outp, fx_dt, in_dt,
                         out_dt, c->prec, c->ws, s, 0, &lay);
  if (c->packed_segs[k]) {
    lay.segs = (const cxa::GSeg *)c->packed_segs[k];
    lay.b_k4 = true;
    lay.a_k4 = c->a_packed[k] != 0;
    if (c->c_packed[k]) {
      lay.outs = (const cxa::MixOut *)c->packed_outs[k];
      lay.c_k4 = true;
    }
    if (pk && k == 0 && f32) {
      lay.segs = (const cxa::GSeg *)c->pk_segs;
      lay.a_k4 = true;
      return bc_exec_typed(c->plans[0], c->pk_psi32, c->packed32[0], outp, CXA_F32, fx_dt,
                           out_dt, c->prec, c->ws, s, 0, &lay);
    }
    return bc_exec_typed(c->plans[k], in, f32 ? c->packed32[k] : c->packed64[k], outp, in_dt,
                         fx_dt, out_dt, c->prec, c->ws, s, 0, &lay);
  }
  return bc_exec_typed(c->plans[k], in, f32 ? c->fixed32[k] : c->fixed[k], outp, in_dt, fx_dt,
                       out_dt, c->prec, c->ws, s, 0, &lay);
}

int chain_apply(cxa_chain *c, const void *x, void *y, hipStream_t s) {
  for (int k = 0; k < c->nsteps; ++k) {
    int rc = chain_step(c, k, x, y, s);
    if (rc) return rc;
  }
  return CXA_OK;
}

// Tile slots of the dot-product epilogue of the chain's last launch at its current precision,
// or 0 when that launch cannot carry it (complex chain, mix or row-streaming launch).
int chain_dot_slots(const cxa_chain *c) {
  if (c->dtype != CXA_F64) return 0;
  return bc_dot_slots(c->plans[c->nsteps - 1], c->prec);
}

// y = H x with the last launch also writing the per-slot partials of y . x and x . x into dotp
// (x and y share the vector layout, as the Lanczos vectors do); pk as in chain_step.
int chain_apply_dot(cxa_chain *c, const void *x, void *y, double *dotp, hipStream_t s,
                    bool pk = false) {
  for (int k = 0; k < c->nsteps; ++k) {
    const bool last = k == c->nsteps - 1;
    int rc = chain_step(c, k, x, y, s, last ? (const double *)x : nullptr,
                        last ? dotp : nullptr, pk);
    if (rc) return rc;
  }
  return CXA_OK;
}

}  // namespace

// The chain as a matvec callback: ctx is the chain, x/y are FP64 device vectors.
extern "C" int cxa_chain_matvec(void *ctx, const void *x, void *y, cxa_stream_t stream) {
  cxa_chain *c = (cxa_chain *)ctx;
  CXA_CHECK(c && x && y, CXA_ERR_INVALID, "cxa_chain_matvec: bad arguments");
  CXA_CHECK(c->prec != CXA_PREC_FP32 || !c->fixed32.empty(), CXA_ERR_INVALID,
            "cxa_chain_matvec: FP32 copies missing; call cxa_chain_prepare first");
  return chain_apply(c, x, y, (hipStream_t)stream);
}

// One step of the chain (for per-kernel timing): step 0 reads x, the last step writes y, the
// others use the chain's intermediates left by the previous step.
extern "C" int cxa_chain_step(void *ctx, int step, const void *x, void *y, cxa_stream_t stream) {
  cxa_chain *c = (cxa_chain *)ctx;
  CXA_CHECK(c && step >= 0 && step < c->nsteps, CXA_ERR_INVALID, "cxa_chain_step: bad step");
  CXA_CHECK((step > 0 || x) && (step < c->nsteps - 1 || y), CXA_ERR_INVALID,
            "cxa_chain_step: missing input or output vector");
  CXA_CHECK(c->prec != CXA_PREC_FP32 || !c->fixed32.empty(), CXA_ERR_INVALID,
            "cxa_chain_step: FP32 copies missing; call cxa_chain_prepare first");
  return chain_step(c, step, x, y, (hipStream_t)stream);
}

extern "C" int cxa_chain_layout(const cxa_chain *c, int step, int *flags) {
  CXA_CHECK(c && flags && step >= 0 && step < c->nsteps, CXA_ERR_INVALID,
            "cxa_chain_layout: bad arguments");
  *flags = (c->packed_segs[step] ? 1 : 0) | (c->a_packed[step] ? 2 : 0) |
           (c->c_packed[step] ? 4 : 0) | (step == 0 && c->pk_psi32 ? 8 : 0);
  return CXA_OK;
}

extern "C" int cxa_chain_flops(const cxa_chain *c, double *flops) {
  CXA_CHECK(c && flops, CXA_ERR_INVALID, "cxa_chain_flops: bad arguments");
  *flops = c->flops;
  return CXA_OK;
}

extern "C" int cxa_chain_destroy(cxa_chain *c) {
  delete c;
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Tile-shape autotuner: for every step of the chain, at the chain's current precision, time
// each candidate tiling (tile shape x workgroup split-K) on the chain's own buffers and keep the
// fastest.  The choice is stored in the step's plan for that precision only.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_chain_tune(cxa_chain *c, const void *x, void *y, int reps, int *kinds,
                              int *splits, double *ms, cxa_stream_t stream) {
  CXA_CHECK(c && x && y && reps >= 1, CXA_ERR_INVALID, "cxa_chain_tune: bad arguments");
  CXA_CHECK(c->prec != CXA_PREC_FP32 || !c->fixed32.empty(), CXA_ERR_INVALID,
            "cxa_chain_tune: FP32 copies missing; call cxa_chain_prepare first");
  hipStream_t s = (hipStream_t)stream;
  static const int cand[][2] = {{-1, -1},    {-1, 0},     {-1, 1},     {-1, 2},
                                {TK_1x1, 0}, {TK_1x1, 1}, {TK_1x1, 2}, {TK_2x1, 0},
                                {TK_2x1, 1}, {TK_2x1, 2}, {TK_1x2, 0}, {TK_1x2, 1},
                                {TK_1x2, 2}, {TK_2x2, 0}, {TK_2x2, 1}, {TK_2x2, 2},
                                {TK_4x2, 0}, {TK_ROWS, 0}, {TK_ROWS, 1}};
  const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
  hipEvent_t e0, e1;
  CXA_HIP(hipEventCreate(&e0));
  CXA_HIP(hipEventCreate(&e1));
  int rc = chain_apply(c, x, y, s);  // every intermediate in place
  for (int k = 0; k < c->nsteps && rc == CXA_OK; ++k) {
    cxa_bc_plan *plan = const_cast<cxa_bc_plan *>(c->plans[k]);
    // FP32 chains with a packed running vector: step 0 runs in two forms, reading the packed
    // FP32 copy (FP32 ground-state Lanczos runs) or the FP64 vector (every other matvec); a
    // tiling is scored by the sum of both times and reported by the packed form's
    const bool pk = k == 0 && c->prec == CXA_PREC_FP32 && c->pk_psi32;
    double best = 1e30, best_ms = 0.0;
    int bk = -1, bs = -1;
    // mean time of `reps` launches of step k in one form (after two untimed ones)
    auto time_form = [&](bool packed, double *per) -> int {
      for (int w = 0; w < 2; ++w) {
        const int r2 = chain_step(c, k, x, y, s, nullptr, nullptr, packed);
        if (r2) return r2;
      }
      CXA_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) {
        const int r2 = chain_step(c, k, x, y, s, nullptr, nullptr, packed);
        if (r2) return r2;
      }
      float t = 0;
      CXA_HIP(hipEventRecord(e1, s));
      CXA_HIP(hipEventSynchronize(e1));
      CXA_HIP(hipEventElapsedTime(&t, e0, e1));
      *per = (double)t / reps;
      return CXA_OK;
    };
    for (int ci = 0; ci < ncand && rc == CXA_OK; ++ci) {
      // mix plans: explicit 16x16, 32x16, 16x32 or 32x32 tiles; split 0 / 1 = eight / four
      // waves
      if (bc_is_mix(plan) && (cand[ci][0] < 0 || cand[ci][0] == TK_4x2 ||
                              cand[ci][0] == TK_ROWS || cand[ci][1] == 2))
        continue;
      // the big-tile and row classes are real-only (complex plans fall back to the default)
      if (c->dtype != CXA_F64 && (cand[ci][0] == TK_4x2 || cand[ci][0] == TK_ROWS)) continue;
      if ((rc = cxa_block_contract_set_config(plan, c->prec, cand[ci][0], cand[ci][1]))) break;
      double per = 0.0, per_plain = 0.0;
      if ((rc = time_form(pk, &per))) break;
      if (pk && (rc = time_form(false, &per_plain))) break;
      if (per + per_plain < best) {
        best = per + per_plain;
        best_ms = per;
        bk = cand[ci][0];
        bs = cand[ci][1];
      }
    }
    if (rc == CXA_OK) rc = cxa_block_contract_set_config(plan, c->prec, bk, bs);
    if (kinds) kinds[k] = bk;
    if (splits) splits[k] = bs;
    if (ms) ms[k] = best_ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// ------------------------------------------------------------------------------------------
// A chain application captured as a graph for fixed x / y (repeats applications back to back),
// replayed with one launch.
// ------------------------------------------------------------------------------------------
struct cxa_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  ~cxa_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

extern "C" int cxa_chain_graph_create(cxa_chain *c, const void *x, void *y, int repeats,
                                      cxa_graph **out) {
  CXA_CHECK(c && x && y && out && repeats >= 1, CXA_ERR_INVALID,
            "cxa_chain_graph_create: bad arguments");
  CXA_CHECK(c->prec != CXA_PREC_FP32 || !c->fixed32.empty(), CXA_ERR_INVALID,
            "cxa_chain_graph_create: FP32 copies missing; call cxa_chain_prepare first");
  *out = nullptr;
  std::unique_ptr<cxa_graph> g(new cxa_graph());
  CXA_HIP(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
  CXA_HIP(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
  int rc = CXA_OK;
  for (int r = 0; r < repeats && rc == CXA_OK; ++r) rc = chain_apply(c, x, y, g->cap);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(g->cap, &graph);
  g->graph = graph;
  if (rc) return rc;
  CXA_HIP(e);
  CXA_HIP(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
  *out = g.release();
  return CXA_OK;
}

extern "C" int cxa_graph_launch(cxa_graph *g, cxa_stream_t stream) {
  CXA_CHECK(g && g->exec, CXA_ERR_INVALID, "cxa_graph_launch: null graph");
  CXA_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return CXA_OK;
}

extern "C" int cxa_graph_destroy(cxa_graph *g) {
  delete g;
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Lanczos
// ------------------------------------------------------------------------------------------
struct cxa_lanczos {
  int device = 0;
  int64_t n = 0;
  int maxiter = 0;
  // what the buffers were allocated for (cxa_lanczos_reshape stays within it)
  int64_t cap_n = 0, cap_basis = 0;
  int cap_maxiter = 0;
  double *V = nullptr, *w = nullptr, *cur = nullptr, *part = nullptr, *sc = nullptr,
         *coef = nullptr;
  int *counter = nullptr;
  double *hbuf = nullptr;  // pinned: alpha, beta history and Ritz coefficients
  // A run whose matvec is a real chain with a dot-capable last launch takes the fused form:
  // that launch writes the partials of Hu . u and u . u (2 * fused doubles in dotp) and one
  // residual launch follows, instead of two vector launches.
  double *dotp = nullptr;
  int dot_cap = 0, fused = 0;
  // FP32 ground-state runs of a chain with a packed running vector: the residual launch also
  // writes the next matvec input, rounded to FP32, into the chain's packed copy
  bool pk = false;
  hipStream_t own = nullptr;
  // captured iterations: one (gexec1) and `gbatch` back to back (gexecB)
  hipGraphExec_t gexec1 = nullptr, gexecB = nullptr;
  int gbatch = 0;
  void drop_graphs() {
    if (gexec1) (void)hipGraphExecDestroy(gexec1);
    if (gexecB) (void)hipGraphExecDestroy(gexecB);
    gexec1 = gexecB = nullptr;
    gbatch = 0;
  }
  ~cxa_lanczos() {
    drop_graphs();
    for (double *p : {V, w, cur, part, sc, coef, dotp})
      if (p) (void)hipFree(p);
    if (counter) (void)hipFree(counter);
    if (hbuf) (void)hipHostFree(hbuf);
    if (own) (void)hipStreamDestroy(own);
  }
  // device scalars: alphas, betas, projection coefficients (complex: two doubles each), norm
  double *As() { return sc; }
  double *Bs() { return sc + (maxiter + 1); }
  double *h() { return sc + 2 * (maxiter + 1); }
  double *nrm() { return sc + 4 * (maxiter + 1); }
};

extern "C" int cxa_lanczos_create(int64_t n, int maxiter, cxa_lanczos **out) {
  CXA_CHECK(out && n > 0 && maxiter >= 2, CXA_ERR_INVALID,
            "cxa_lanczos_create: need n > 0 and Maxiter > 1");
  *out = nullptr;
  std::unique_ptr<cxa_lanczos> L(new cxa_lanczos());
  CXA_HIP(hipGetDevice(&L->device));
  L->n = L->cap_n = n;
  L->maxiter = L->cap_maxiter = maxiter;
  const int64_t m1 = maxiter + 1;
  L->cap_basis = n * m1;
  CXA_HIP(hipMalloc(&L->V, sizeof(double) * n * m1));
  CXA_HIP(hipMalloc(&L->w, sizeof(double) * n));
  CXA_HIP(hipMalloc(&L->cur, sizeof(double) * n));
  // partial sums: two rows of m1 (the imaginary parts of complex projections in the second)
  CXA_HIP(hipMalloc(&L->part, sizeof(double) * kPartials * 2 * m1));
  CXA_HIP(hipMalloc(&L->sc, sizeof(double) * (4 * m1 + 8)));
  CXA_HIP(hipMalloc(&L->coef, sizeof(double) * 2 * m1));
  CXA_HIP(hipMalloc(&L->counter, sizeof(int)));
  // pinned: alphas, betas, then up to 2 * m1 doubles of staged coefficients
  CXA_HIP(hipHostMalloc((void **)&L->hbuf, sizeof(double) * 4 * m1, hipHostMallocDefault));
  *out = L.release();
  return CXA_OK;
}

extern "C" int cxa_lanczos_destroy(cxa_lanczos *L) {
  delete L;
  return CXA_OK;
}

// The context's buffers reused for vectors of n doubles and Maxiter iterations, within what
// they were allocated for: n <= the creation n, Maxiter <= the creation Maxiter and
// n * (Maxiter + 1) basis doubles <= the creation count.  Captured iteration graphs (made for
// the old sizes) are dropped.
extern "C" int cxa_lanczos_reshape(cxa_lanczos *L, int64_t n, int maxiter) {
  CXA_CHECK(L && n > 0 && maxiter >= 2, CXA_ERR_INVALID,
            "cxa_lanczos_reshape: need a context, n > 0 and Maxiter > 1");
  CXA_CHECK(n <= L->cap_n && maxiter <= L->cap_maxiter && n * (maxiter + 1) <= L->cap_basis,
            CXA_ERR_INVALID,
            "cxa_lanczos_reshape: %lld x %d exceeds the context's %lld x %d", (long long)n,
            maxiter, (long long)L->cap_n, L->cap_maxiter);
  L->drop_graphs();
  L->n = n;
  L->maxiter = maxiter;
  return CXA_OK;
}

namespace {

// One iteration.  it >= 0: the iteration index is passed to the kernels by value; it < 0: they
// read the device counter plus `off` (the captured form; the capture advances the counter once
// per captured batch).
int lanczos_body(cxa_lanczos *L, cxa_matvec_fn fn, void *ctx, int reorth, bool cplx, int it,
                 hipStream_t s, int off = 0) {
  const IterIdx iter{it < 0 ? L->counter : nullptr, it < 0 ? off : it};
  double *part_a = L->part, *part_b = L->part + kPartials;
  int rc;
  if (L->fused > 0) {  // the chain's last launch leaves the Hu . u and u . u partials
    cxa_chain *c = (cxa_chain *)ctx;
    if ((rc = chain_apply_dot(c, L->cur, L->w, L->dotp, s, L->pk))) return rc;
    if ((rc = vec_residual_fused(L->w, L->cur, L->V, L->n, L->dotp, L->fused, L->As(), L->Bs(),
                                 iter, part_b, L->pk ? c->pk_map : nullptr,
                                 L->pk ? c->pk_psi32 : nullptr, s)))
      return rc;
    for (int pass = 0; pass < reorth; ++pass)
      if ((rc = vec_reorth(L->cur, L->V, L->n, L->maxiter + 1, iter, L->part, L->h(), part_b,
                           cplx, s)))
        return rc;
    return CXA_OK;
  }
  // cur holds the previous residual u (V[0] at it = 0); w = H u
  if ((rc = fn(ctx, L->cur, L->w, (cxa_stream_t)s))) return rc;
  // beta of u and V[it] = u / beta with the alpha partials; alpha with the residual (into cur);
  // (re-orthogonalisation of the residual against V[0..it])
  if ((rc = vec_dot_basis(L->w, L->cur, L->V, L->n, part_b, L->Bs(), iter, part_a, s))) return rc;
  if ((rc = vec_residual(L->w, L->cur, L->V, L->n, part_a, L->As(), L->Bs(), iter, part_b, s)))
    return rc;
  for (int pass = 0; pass < reorth; ++pass)  // classical Gram-Schmidt passes
    if ((rc = vec_reorth(L->cur, L->V, L->n, L->maxiter + 1, iter, L->part, L->h(), part_b, cplx,
                         s)))
      return rc;
  return CXA_OK;
}

// `iters` iterations (device-counter form) captured on stream ws as one executable graph
int capture_iterations(cxa_lanczos *L, cxa_matvec_fn fn, void *ctx, int reorth, bool cplx,
                       int iters, hipStream_t ws, hipGraphExec_t *out) {
  CXA_HIP(hipStreamBeginCapture(ws, hipStreamCaptureModeThreadLocal));
  int rc = CXA_OK;
  for (int i = 0; i < iters && rc == CXA_OK; ++i)
    rc = lanczos_body(L, fn, ctx, reorth, cplx, -1, ws, i);
  if (rc == CXA_OK) rc = vec_counter_add(L->counter, iters, ws);  // one launch per batch
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ws, &g);
  if (rc == CXA_OK && e == hipSuccess) {
    if (hipGraphInstantiate(out, g, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(CXA_ERR_HIP, "Lanczos: graph instantiation failed");
  } else if (rc == CXA_OK) {
    rc = fail(CXA_ERR_HIP, "Lanczos: graph capture failed (%s)", hipGetErrorString(e));
  }
  if (g) (void)hipGraphDestroy(g);
  return rc;
}

}  // namespace

namespace {

// The stream a run enqueues on: the caller's, or (graphs on the null stream, which cannot be
// captured) the context's own, joined to the caller's with an event at the end.
struct RunStream {
  hipStream_t s = nullptr, ws = nullptr;
  hipEvent_t ev = nullptr;
};

int run_begin(cxa_lanczos *L, bool use_graph, hipStream_t s, RunStream &r) {
  r.s = r.ws = s;
  if (use_graph && s == nullptr) {
    if (!L->own) CXA_HIP(hipStreamCreateWithFlags(&L->own, hipStreamNonBlocking));
    r.ws = L->own;
    CXA_HIP(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
    CXA_HIP(hipEventRecord(r.ev, s));
    CXA_HIP(hipStreamWaitEvent(r.ws, r.ev, 0));
  }
  return CXA_OK;
}

int run_end(RunStream &r) {
  if (r.ws != r.s) {
    CXA_HIP(hipEventRecord(r.ev, r.ws));
    CXA_HIP(hipStreamWaitEvent(r.s, r.ev, 0));
    CXA_HIP(hipEventDestroy(r.ev));
    r.ev = nullptr;
  } else {
    CXA_HIP(hipStreamSynchronize(r.ws));  // the pinned coefficients are reused by the next run
  }
  return CXA_OK;
}

// psi_0 = Tin / |Tin| into V[0] and cur, then (graphs) the iterations captured afresh: the matvec
// behind (fn, ctx) may have changed since the last run (a chain's precision, its buffers).  One
// iteration, and a whole batch of them so that a full batch is one host launch.
int run_prologue(cxa_lanczos *L, cxa_matvec_fn matvec, void *ctx, const double *Tin, int reorth,
                 bool cplx, bool use_graph, int batch, int maxiter, hipStream_t ws,
                 bool allow_pk = false) {
  int rc;
  // the fused iteration for real chains with a dot-capable last launch (decided per run: the
  // chain's precision, hence its launch, may have changed since the last one)
  L->fused = (!cplx && matvec == cxa_chain_matvec) ? chain_dot_slots((const cxa_chain *)ctx) : 0;
  // the packed FP32 running vector: only where nothing but the residual launch changes the
  // vector between matvecs (no re-orthogonalisation)
  const cxa_chain *pc = L->fused > 0 ? (const cxa_chain *)ctx : nullptr;
  L->pk = allow_pk && pc && reorth == 0 && pc->prec == CXA_PREC_FP32 && pc->pk_psi32 &&
          pc->pk_n == L->n;
  if (L->fused > L->dot_cap) {
    CXA_HIP(hipDeviceSynchronize());  // an earlier run's launches may still read the old buffer
    if (L->dotp) CXA_HIP(hipFree(L->dotp));
    L->dotp = nullptr;
    L->dot_cap = 0;
    CXA_HIP(hipMalloc(&L->dotp, sizeof(double) * 2 * L->fused));
    L->dot_cap = L->fused;
  }
  if ((rc = vec_counter_set(L->counter, 0, ws))) return rc;
  if ((rc = vec_dot(Tin, Tin, L->n, L->part, L->nrm(), nullptr, 0, 1, ws))) return rc;
  if ((rc = vec_normalize_init(L->cur, L->V, Tin, L->n, L->nrm(), ws))) return rc;
  if (L->pk && (rc = vec_pack_b(L->cur, pc->pk_psi32, true,
                                (const cxa::PackRegion *)pc->pk_reg_dev,
                                (int)pc->pk_regions.size(), ws)))
    return rc;
  if (use_graph) {
    L->drop_graphs();
    if ((rc = capture_iterations(L, matvec, ctx, reorth, cplx, 1, ws, &L->gexec1))) return rc;
    if (batch > 1 && maxiter >= batch) {
      if ((rc = capture_iterations(L, matvec, ctx, reorth, cplx, batch, ws, &L->gexecB)))
        return rc;
      L->gbatch = batch;
    }
  }
  return CXA_OK;
}

// Iterations [it, it + nb), then alphas / betas [0, it + nb) copied into the pinned host buffers
// hA / hB; returns after they have arrived.
int run_batch(cxa_lanczos *L, cxa_matvec_fn matvec, void *ctx, int reorth, bool cplx,
              bool use_graph, int it, int nb, hipStream_t ws) {
  int rc;
  if (use_graph && nb == L->gbatch) {
    CXA_HIP(hipGraphLaunch(L->gexecB, ws));
  } else {
    for (int b = 0; b < nb; ++b) {
      if (use_graph) {
        CXA_HIP(hipGraphLaunch(L->gexec1, ws));
      } else if ((rc = lanczos_body(L, matvec, ctx, reorth, cplx, it + b, ws))) {
        return rc;
      }
    }
  }
  const int m1 = L->maxiter + 1, upto = it + nb;
  // beta of the batch's last residual (the next iteration's first launch would record it)
  if ((rc = vec_finalize(L->part + kPartials, L->Bs(), use_graph ? L->counter : nullptr,
                         use_graph ? -1 : upto - 1, 1, ws)))
    return rc;
  CXA_HIP(hipMemcpyAsync(L->hbuf, L->As(), sizeof(double) * upto, hipMemcpyDeviceToHost, ws));
  CXA_HIP(hipMemcpyAsync(L->hbuf + m1, L->Bs(), sizeof(double) * upto, hipMemcpyDeviceToHost,
                         ws));
  CXA_HIP(hipStreamSynchronize(ws));
  return CXA_OK;
}

// out = sum_{j < k} coef[j] V_j with coef given on the host (staged through the pinned buffer);
// mode as vec_combine (1, 2: k complex coefficients, 2k doubles).
int combine_host_coef(cxa_lanczos *L, const double *coef, int k, double *out, int mode,
                      hipStream_t ws) {
  double *hz = L->hbuf + 2 * (L->maxiter + 1);
  const int nd = mode ? 2 * k : k;
  CXA_HIP(hipStreamSynchronize(ws));  // the staging buffer may still feed an earlier upload
  std::memcpy(hz, coef, sizeof(double) * nd);
  CXA_HIP(hipMemcpyAsync(L->coef, hz, sizeof(double) * nd, hipMemcpyHostToDevice, ws));
  return vec_combine(out, L->V, L->coef, k, L->n, mode, ws);
}

// Eigenvalues of the symmetric tridiagonal matrix (diagonal a[n], off-diagonal b[n-1]) by the
// implicit symmetric QR iteration with Wilkinson shifts (Golub & Van Loan, Matrix Computations,
// 4th ed., Algorithms 8.3.2 and 8.3.3): each step chases the bulge of one shifted QR step down
// the unreduced block with plane rotations, and off-diagonals below eps (|d_i| + |d_i+1|) are
// set to zero.  w = eigenvalues ascending.  Of the orthogonal matrix of eigenvectors Z (column k
// belongs to w[k]) only the rows listed in `rows` are accumulated, into zr (rows.size() x n,
// row-major): the whole Z costs O(n^3), a few rows O(n^2).
void tridiag_qr(int n, const double *a, const double *b, std::vector<double> &w,
                const std::vector<int> &rows, std::vector<double> &zr) {
  const double eps = 2.220446049250313e-16;
  std::vector<double> d(a, a + n), e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = b[i];
  const int nr = (int)rows.size();
  zr.assign((size_t)nr * n, 0.0);
  for (int r = 0; r < nr; ++r) zr[(size_t)r * n + rows[r]] = 1.0;
  auto rotate = [&](int k, double c, double s) {  // Z <- Z G on columns k, k + 1
    for (int r = 0; r < nr; ++r) {
      double *z = &zr[(size_t)r * n];
      const double zk = z[k], zk1 = z[k + 1];
      z[k] = c * zk - s * zk1;
      z[k + 1] = s * zk + c * zk1;
    }
  };
  int h = n - 1, guard = 0;
  while (h > 0 && guard++ < 60 * n + 60) {
    if (std::fabs(e[h - 1]) <= eps * (std::fabs(d[h - 1]) + std::fabs(d[h]))) {
      e[h - 1] = 0.0;
      --h;
      continue;
    }
    int l = h - 1;
    while (l > 0 && std::fabs(e[l - 1]) > eps * (std::fabs(d[l - 1]) + std::fabs(d[l]))) --l;
    if (l > 0) e[l - 1] = 0.0;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to d[h]
    const double delta = 0.5 * (d[h - 1] - d[h]);
    const double rad = std::hypot(delta, e[h - 1]);
    const double mu = d[h] - e[h - 1] * (e[h - 1] / (delta + (delta >= 0 ? rad : -rad)));
    double x = d[l] - mu, z = e[l], bulge = 0.0;
    for (int k = l; k < h; ++k) {
      // rotation in the (k, k+1) plane with G^T [x; z] = [r; 0]
      const double r = std::hypot(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = -z / r;
      }
      if (k > l) e[k - 1] = r;  // x, z were (T[k-1][k], bulge T[k-1][k+1])
      const double ak = d[k], bk = e[k], ck = d[k + 1];
      d[k] = c * c * ak - 2.0 * c * s * bk + s * s * ck;
      d[k + 1] = s * s * ak + 2.0 * c * s * bk + c * c * ck;
      e[k] = (ak - ck) * c * s + bk * (c * c - s * s);
      if (k + 1 < h) {
        bulge = -s * e[k + 1];
        e[k + 1] *= c;
      }
      rotate(k, c, s);
      x = e[k];
      z = bulge;
    }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  std::vector<double> tmp(n);
  for (int k = 0; k < n; ++k) w[k] = d[ord[k]];
  for (int r = 0; r < nr; ++r) {
    double *z = &zr[(size_t)r * n];
    for (int k = 0; k < n; ++k) tmp[k] = z[ord[k]];
    std::copy(tmp.begin(), tmp.end(), z);
  }
}

// every row of Z (the full eigenvector matrix, row-major, Z[i * n + k] = component i of
// eigenvector k)
void tridiag_qr_full(int n, const double *a, const double *b, std::vector<double> &w,
                     std::vector<double> &Z) {
  std::vector<int> rows(n);
  for (int i = 0; i < n; ++i) rows[i] = i;
  tridiag_qr(n, a, b, w, rows, Z);
}

}  // namespace

// flags: bit 0 = FP64 full re-orthogonalisation, bit 1 = replay captured iteration graphs,
// bit 2 = complex Hermitian operator on C128 vectors of n / 2 values (alpha and beta are real
// dots of the n doubles: Re <v|Hv> and |u|, as Lanczos_Gnd_Ut.cpp:96-99 takes them; the
// re-orthogonalisation projects with complex coefficients)
extern "C" int cxa_lanczos_run(cxa_lanczos *L, cxa_matvec_fn matvec, void *ctx, const double *Tin,
                               double CvgCrit, int maxiter, int batch, int flags, double *E_out,
                               double *evec, int *niter, double *alphas, double *betas,
                               cxa_stream_t stream) {
  CXA_CHECK(L && matvec && Tin && E_out, CXA_ERR_INVALID, "cxa_lanczos_run: bad arguments");
  CXA_CHECK(CvgCrit > 0, CXA_ERR_INVALID, "[ERROR][Lanczos] converge criteria must >0");
  CXA_CHECK(maxiter >= 2, CXA_ERR_INVALID, "[ERROR][Lanczos] Maxiter must >1");
  CXA_CHECK(maxiter <= L->maxiter, CXA_ERR_INVALID,
            "cxa_lanczos_run: Maxiter %d above the context's %d", maxiter, L->maxiter);
  if (batch < 1) batch = 1;
  const int reorth = flags & 1;
  const bool use_graph = (flags >> 1) & 1;
  const bool cplx = (flags >> 2) & 1;
  CXA_CHECK(!cplx || L->n % 2 == 0, CXA_ERR_INVALID,
            "cxa_lanczos_run: complex vectors need an even number of doubles (got %lld)",
            (long long)L->n);
  RunStream r;
  int rc;
  if ((rc = run_begin(L, use_graph, (hipStream_t)stream, r))) return rc;
  const hipStream_t ws = r.ws;
  if ((rc = run_prologue(L, matvec, ctx, Tin, reorth, cplx, use_graph, batch, maxiter, ws,
                         true)))
    return rc;

  const int m1 = L->maxiter + 1;
  const double *hA = L->hbuf, *hB = L->hbuf + m1;
  std::vector<double> z(m1);
  double E = 0.0, eig = 0.0, dE = 0.0;
  std::vector<double> dEs;  // |E_j - E_{j-1}| per checked iteration j >= 1
  int done = -1, it = 0;
  while (done < 0 && it < maxiter) {
    // The next batch stops near the predicted convergence point: |dE| has fallen geometrically
    // over the last four iterations at rate r, so about log(CvgCrit / |dE|) / log(r) more are
    // needed.  A batch that is too short costs one more host check; one that runs past the stop
    // costs whole iterations.  Results do not depend on the batch sizes.
    int nb = std::min(batch, maxiter - it);
    const size_t nd = dEs.size();
    if (nb > 2 && nd >= 5 && dEs[nd - 1] > 0 && dEs[nd - 5] > dEs[nd - 1]) {
      const double r = std::pow(dEs[nd - 1] / dEs[nd - 5], 0.25);
      const double need = std::log(CvgCrit / dEs[nd - 1]) / std::log(r);
      if (need < nb - 1) nb = std::max(2, (int)std::ceil(need) + 1);
    }
    if ((rc = run_batch(L, matvec, ctx, reorth, cplx, use_graph, it, nb, ws))) return rc;
    const int upto = it + nb;  // alphas/betas [0, upto) are valid after the batch
    for (int j = it; j < upto && done < 0; ++j) {
      if (j == 0) {
        E = hA[0];
        eig = hA[0];
        if (hB[0] == 0.0) done = 0;  // Tin is an eigenvector
        continue;
      }
      // E (the lowest eigenvalue of the previous tridiagonal) bounds this one from above
      rc = tridiag_lowest(j + 1, hA, hB, &E, dE, &eig, nullptr);
      if (rc) {  // the reference stops on a failed tridiagonal solve
        done = j - 1;
        break;
      }
      if (hB[j] == 0.0) {
        done = j;
        break;
      }
      if (std::fabs(E - eig) < CvgCrit) {
        done = j;
        break;
      }
      dE = std::fabs(E - eig);
      dEs.push_back(dE);
      E = eig;
    }
    it = upto;
  }
  if (done < 0) done = maxiter - 1;
  const int k = done + 1;  // tridiagonal size at the stop
  if ((rc = cxa_tridiag_ground(k, hA, hB, &eig, z.data()))) return rc;
  *E_out = eig;
  if (niter) *niter = k;
  if (alphas) std::memcpy(alphas, hA, sizeof(double) * k);
  if (betas) std::memcpy(betas, hB, sizeof(double) * k);
  // real coefficients: the same combination of the doubles for real and complex vectors
  if (evec && (rc = combine_host_coef(L, z.data(), k, evec, 0, ws))) return rc;
  return run_end(r);
}

extern "C" int cxa_tridiag_eigh(int n, const double *a, const double *b, double *w, double *z) {
  CXA_CHECK(n >= 1 && a && w && (n == 1 || b), CXA_ERR_INVALID,
            "cxa_tridiag_eigh: bad arguments");
  for (int i = 0; i < n; ++i)
    CXA_CHECK(std::isfinite(a[i]) && (i + 1 == n || std::isfinite(b[i])), CXA_ERR_NUMERIC,
              "cxa_tridiag_eigh: non-finite entries");
  std::vector<double> ev, Z;
  tridiag_qr_full(n, a, b, ev, Z);
  std::memcpy(w, ev.data(), sizeof(double) * n);
  if (z)  // row k = the eigenvector of w[k]
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) z[(size_t)k * n + i] = Z[(size_t)i * n + k];
  return CXA_OK;
}

// Lanczos_Exp (src/linalg/Lanczos_Exp.cpp:266-390): the plain three-term recurrence from
// Tin / |Tin|; at step i (i >= 1) the step stops before v_i when beta_i is at the round-off
// level 100 eps max(1, |alpha|, |alpha| + beta_i + beta_{i-1}, ...) sqrt(i), otherwise after v_i
// when |exp(tau T_{i+1})[i, 0]| < CvgCrit or i = min(Maxiter, n) - 1; then
// out = V_k exp(tau T_k) (V_k^H Tin).  tau = tau_re + i tau_im.  flags bit 2: complex Hermitian
// operator on C128 vectors of n / 2 values (out: n doubles).  Real vectors with tau_im != 0:
// out is complex, 2n doubles (n C128 values).  T is the real symmetric tridiagonal either way
// (alpha = Re <v|Hv>), and exp(tau T) = Z diag(exp(tau w)) Z^T from its eigenpairs.
extern "C" int cxa_lanczos_exp(cxa_lanczos *L, cxa_matvec_fn matvec, void *ctx, const double *Tin,
                               double tau_re, double tau_im, double CvgCrit, int Maxiter,
                               int flags, double *out, int *niter, cxa_stream_t stream) {
  CXA_CHECK(L && matvec && Tin && out, CXA_ERR_INVALID, "cxa_lanczos_exp: bad arguments");
  CXA_CHECK(CvgCrit > 0, CXA_ERR_INVALID, "[ERROR][Lanczos_Exp] converge criteria must >0");
  CXA_CHECK(Maxiter >= 2, CXA_ERR_INVALID, "[ERROR][Lanczos_Exp] Maxiter must >1");
  CXA_CHECK(std::isfinite(tau_re) && std::isfinite(tau_im), CXA_ERR_INVALID,
            "[ERROR][Lanczos_Exp] tau must be finite");
  const bool cplx = (flags >> 2) & 1;
  CXA_CHECK(!cplx || L->n % 2 == 0, CXA_ERR_INVALID,
            "cxa_lanczos_exp: complex vectors need an even number of doubles (got %lld)",
            (long long)L->n);
  const int64_t ne = cplx ? L->n / 2 : L->n;  // vector length in elements
  const int imp = (int)std::min<int64_t>(Maxiter, ne);  // Krylov vectors at most
  CXA_CHECK(imp <= L->maxiter, CXA_ERR_INVALID,
            "cxa_lanczos_exp: %d Krylov vectors above the context's %d", imp, L->maxiter);
  const double eps = 2.220446049250313e-16;
  RunStream r;
  int rc;
  if ((rc = run_begin(L, false, (hipStream_t)stream, r))) return rc;
  const hipStream_t ws = r.ws;
  if ((rc = run_prologue(L, matvec, ctx, Tin, 0, cplx, false, 1, imp, ws))) return rc;
  const int m1 = L->maxiter + 1;
  const double *hA = L->hbuf, *hB = L->hbuf + m1;
  // exp(tau w) for the eigenvalues w of T (complex tau)
  auto expw = [&](double w, double *re, double *im) {
    const double mag = std::exp(tau_re * w);
    *re = mag * std::cos(tau_im * w);
    *im = tau_im == 0.0 ? 0.0 : mag * std::sin(tau_im * w);
  };
  // |exp(tau T_m)[m-1][0]| from rows 0 and m-1 of the eigenvectors of the tridiagonal T_m
  auto corner = [&](int m) {
    std::vector<double> w, zr;
    tridiag_qr(m, hA, hB, w, {0, m - 1}, zr);
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < m; ++k) {
      double er, ei;
      expw(w[k], &er, &ei);
      const double z = zr[(size_t)m + k] * zr[k];
      ar += z * er;
      ai += z * ei;
    }
    return std::hypot(ar, ai);
  };
  int size = -1, it = 0;
  double scale = 1.0;
  const int batch = 8;
  while (size < 0) {
    // one more iteration than the host checks: alpha_i of step i comes from iteration i
    const int nb = std::min(batch, imp - it);
    if (nb <= 0) break;
    if ((rc = run_batch(L, matvec, ctx, 0, cplx, false, it, nb, ws))) return rc;
    const int upto = it + nb;
    for (int i = std::max(it, 1); i < upto && size < 0; ++i) {
      if (i == 1) scale = std::max(std::fabs(hA[0]), 1.0);
      const double beta = hB[i - 1], beta_prev = i >= 2 ? hB[i - 2] : 0.0;
      const double local = std::fabs(hA[i - 1]) + std::fabs(beta) + std::fabs(beta_prev);
      scale = std::max(scale, std::max(local, 1.0));
      if (!(beta > 100.0 * eps * scale * std::sqrt((double)i))) {
        size = i;  // breakdown: v_0 .. v_{i-1}
        break;
      }
      if (corner(i + 1) < CvgCrit || i == imp - 1) size = i + 1;
    }
    it = upto;
  }
  if (size < 0) size = std::max(1, std::min(it, imp));
  // c = V^H Tin on the device (complex for complex vectors), then y = exp(tau T) c on the host,
  // out = V y
  if ((rc = vec_project(L->V, Tin, L->n, size, L->part, L->h(), cplx, ws))) return rc;
  double *hc = L->hbuf + 2 * m1;
  CXA_HIP(hipMemcpyAsync(hc, L->h(), sizeof(double) * size * (cplx ? 2 : 1),
                         hipMemcpyDeviceToHost, ws));
  CXA_HIP(hipStreamSynchronize(ws));
  std::vector<double> w, Z;
  tridiag_qr_full(size, hA, hB, w, Z);
  const bool cout = cplx || tau_im != 0.0;  // complex coefficients
  std::vector<double> cr(size), ci(size, 0.0), tr(size, 0.0), ti(size, 0.0);
  for (int j = 0; j < size; ++j) {
    cr[j] = cplx ? hc[2 * j] : hc[j];
    if (cplx) ci[j] = hc[2 * j + 1];
  }
  for (int k = 0; k < size; ++k) {  // t = diag(exp(tau w)) Z^T c
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < size; ++j) {
      sr += Z[(size_t)j * size + k] * cr[j];
      si += Z[(size_t)j * size + k] * ci[j];
    }
    double er, ei;
    expw(w[k], &er, &ei);
    tr[k] = er * sr - ei * si;
    ti[k] = er * si + ei * sr;
  }
  std::vector<double> y(cout ? 2 * size : size, 0.0);  // y = Z t
  for (int i = 0; i < size; ++i) {
    double yr = 0.0, yi = 0.0;
    for (int k = 0; k < size; ++k) {
      yr += Z[(size_t)i * size + k] * tr[k];
      yi += Z[(size_t)i * size + k] * ti[k];
    }
    if (cout) {
      y[2 * i] = yr;
      y[2 * i + 1] = yi;
    } else {
      y[i] = yr;
    }
  }
  if ((rc = combine_host_coef(L, y.data(), size, out, cplx ? 1 : (cout ? 2 : 0), ws))) return rc;
  if (niter) *niter = size;
  return run_end(r);
}

namespace {

// Eigenpairs of a dense symmetric m x m matrix A (row-major, overwritten) by the cyclic Jacobi
// method (Golub & Van Loan, Matrix Computations, 4th ed., Algorithm 8.5.3: 2 x 2 symmetric Schur
// rotations in row order, sweeps until the off-diagonal mass is below eps times the norm).
// w ascending; Z row-major with Z[i * m + k] = component i of the eigenvector of w[k].  Used for
// the projected matrix of a restarted Lanczos (a tridiagonal with an arrow), of order ncv.
void sym_eig_jacobi(int m, std::vector<double> &A, std::vector<double> &w,
                    std::vector<double> &Z) {
  const double eps = 2.220446049250313e-16;
  std::vector<double> V((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) V[(size_t)i * m + i] = 1.0;
  double fro = 0.0;
  for (double x : A) fro += x * x;
  fro = std::sqrt(fro);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += 2.0 * A[(size_t)p * m + q] * A[(size_t)p * m + q];
    if (std::sqrt(off) <= eps * fro) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[(size_t)p * m + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        for (int r = 0; r < m; ++r) {  // A <- A J (columns p, q)
          const double arp = A[(size_t)r * m + p], arq = A[(size_t)r * m + q];
          A[(size_t)r * m + p] = c * arp - sn * arq;
          A[(size_t)r * m + q] = sn * arp + c * arq;
        }
        for (int r = 0; r < m; ++r) {  // A <- J^T A (rows p, q)
          const double apr = A[(size_t)p * m + r], aqr = A[(size_t)q * m + r];
          A[(size_t)p * m + r] = c * apr - sn * aqr;
          A[(size_t)q * m + r] = sn * apr + c * aqr;
        }
        A[(size_t)p * m + q] = A[(size_t)q * m + p] = 0.0;
        for (int r = 0; r < m; ++r) {  // V <- V J
          const double vrp = V[(size_t)r * m + p], vrq = V[(size_t)r * m + q];
          V[(size_t)r * m + p] = c * vrp - sn * vrq;
          V[(size_t)r * m + q] = sn * vrp + c * vrq;
        }
      }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return A[(size_t)x * m + x] < A[(size_t)y * m + y]; });
  w.resize(m);
  Z.assign((size_t)m * m, 0.0);
  for (int k = 0; k < m; ++k) {
    w[k] = A[(size_t)ord[k] * m + ord[k]];
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + k] = V[(size_t)i * m + ord[k]];
  }
}

}  // namespace

// k extreme eigenpairs of a symmetric (complex Hermitian with flags bit 2) operator: which 0 =
// "SA" (smallest algebraic), 1 = "LA" (largest algebraic), 2 = "LM" (largest magnitude), the
// contract of linalg::Lanczos(Hop, Tin, which, maxiter, cvg_crit, k, is_V, ncv)
// (src/linalg/Lanczos.cpp:249-402), which runs ARPACK's implicitly restarted Lanczos with ncv
// basis vectors (default min(n, 2k + 10)) for at most maxiter restart cycles.  Here: a thick-
// restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000) 602-616) with full FP64
// re-orthogonalisation.  A cycle extends the basis to ncv vectors; the projected matrix (a
// tridiagonal, with an arrow after a restart) is diagonalised on the host, and the Ritz pairs
// are accepted with ARPACK's dsconv test |beta_m y_i[m - 1]| <= tol max(eps^(2/3), |theta_i|)
// (tol <= 0: machine epsilon).  Otherwise the basis restarts from the k + (ncv - k) / 2 most
// wanted Ritz vectors (one GEMM over the basis) and the residual; their couplings to it,
// beta_m y_i[m - 1], are the arrow.  A start vector in an invariant subspace of at least k
// dimensions ends the run early.  Eigenvalues in the reference's order (SA ascending, LA
// descending, LM by descending magnitude); evecs (device, may be NULL): the k Ritz vectors as
// rows; niter = matvecs applied; nconv = pairs within the bound.  The context must hold
// 2 * ncv basis vectors (ncv for the basis, the rest for the restart's product).
extern "C" int cxa_lanczos_eigsh(cxa_lanczos *L, cxa_matvec_fn matvec, void *ctx,
                                 const double *Tin, int which, int k, double tol, int Maxiter,
                                 int ncv, int flags, double *evals, double *evecs, int *niter,
                                 int *nconv, cxa_stream_t stream) {
  CXA_CHECK(L && matvec && Tin && evals, CXA_ERR_INVALID, "cxa_lanczos_eigsh: bad arguments");
  CXA_CHECK(which >= 0 && which <= 2, CXA_ERR_INVALID,
            "[ERROR][Lanczos] 'which' should be 'LM', 'LA', 'SA'");
  CXA_CHECK(k >= 1, CXA_ERR_INVALID, "[ERROR][Lanczos] k should be >0");
  const bool cplx = (flags >> 2) & 1;
  CXA_CHECK(!cplx || L->n % 2 == 0, CXA_ERR_INVALID,
            "cxa_lanczos_eigsh: complex vectors need an even number of doubles (got %lld)",
            (long long)L->n);
  const int64_t ne = cplx ? L->n / 2 : L->n;
  CXA_CHECK(k <= ne, CXA_ERR_INVALID,
            "[ERROR][Lanczos] k can only be up to total dimension of input vector D");
  CXA_CHECK(Maxiter >= 1, CXA_ERR_INVALID, "cxa_lanczos_eigsh: Maxiter must be positive");
  CXA_CHECK(ncv == 0 || (ncv >= k + 2 && ncv <= ne) || (ncv == ne && ne < k + 2),
            CXA_ERR_INVALID, "[ERROR][Lanczos] ncv should be 2+k<=ncv<=nx");
  const int m = (int)std::min<int64_t>(ne, ncv > 0 ? ncv : 2 * (int64_t)k + 10);
  const int full = m == ne;  // the basis spans the space: one cycle, no restart
  CXA_CHECK(2 * m <= L->maxiter + 1, CXA_ERR_INVALID,
            "cxa_lanczos_eigsh: the context holds %d vectors, a basis of %d needs %d",
            L->maxiter + 1, m, 2 * m);
  const double eps = 2.220446049250313e-16, eps23 = std::cbrt(eps * eps);
  if (!(tol > 0)) tol = eps;
  RunStream r;
  int rc;
  if ((rc = run_begin(L, false, (hipStream_t)stream, r))) return rc;
  const hipStream_t ws = r.ws;
  if ((rc = run_prologue(L, matvec, ctx, Tin, 1, cplx, false, 1, m, ws))) return rc;
  const int m1 = L->maxiter + 1;
  const double *hA = L->hbuf, *hB = L->hbuf + m1;
  // the wanted Ritz values in the reference's output order (first `cnt` of the m)
  auto select = [&](const std::vector<double> &w, int mm, int cnt) {
    std::vector<int> idx(mm);
    for (int i = 0; i < mm; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
      if (which == 0) return w[x] < w[y];
      if (which == 1) return w[x] > w[y];
      return std::fabs(w[x]) > std::fabs(w[y]);
    });
    idx.resize(std::min(cnt, mm));
    return idx;
  };
  std::vector<double> theta, svec;  // kept Ritz values and their arrow couplings
  std::vector<double> w, Z, T;
  std::vector<int> pick;
  double *dY = nullptr;  // the restart's coefficient matrix (kept x m) on the device
  int start = 0, mm = m, conv = 0, matvecs = 0, cycles = 0;
  const int batch = 16;
  for (;;) {
    // extend the basis from `start` to m vectors (the first vector after a restart gets a
    // second Gram-Schmidt pass: its residual sheds the arrow's components)
    bool invariant = false;
    mm = m;
    for (int it = start; it < m;) {
      const int nb = std::min(batch, m - it);
      for (int b = 0; b < nb && rc == CXA_OK; ++b)
        rc = lanczos_body(L, matvec, ctx, (it + b == start && start > 0) ? 2 : 1, cplx, it + b,
                          ws);
      if (rc) break;
      matvecs += nb;
      it += nb;
      rc = vec_finalize(L->part + kPartials, L->Bs(), nullptr, it - 1, 1, ws);
      if (rc) break;
      CXA_HIP(hipMemcpyAsync(L->hbuf, L->As(), sizeof(double) * it, hipMemcpyDeviceToHost, ws));
      CXA_HIP(hipMemcpyAsync(L->hbuf + m1, L->Bs(), sizeof(double) * it, hipMemcpyDeviceToHost,
                             ws));
      CXA_HIP(hipStreamSynchronize(ws));
      // an invariant subspace ends the basis early: beta_j at round-off level
      double tnorm = 0.0;
      for (int j = 0; j < (int)theta.size(); ++j)
        tnorm = std::max(tnorm, std::fabs(theta[j]) + std::fabs(svec[j]));
      for (int j = start; j < it; ++j)
        tnorm = std::max(tnorm, std::fabs(hA[j]) + std::fabs(hB[j]) +
                                    (j > start ? std::fabs(hB[j - 1]) : 0.0));
      for (int j = start; j < it; ++j)
        if (!(hB[j] > 100.0 * eps * std::max(tnorm, 1.0) * std::sqrt((double)(j + 1)))) {
          mm = j + 1;
          invariant = true;
          break;
        }
      if (invariant) break;
    }
    if (rc) break;
    // the projected matrix: kept Ritz values with their arrow, then the new tridiagonal part
    const int ks = start;
    T.assign((size_t)mm * mm, 0.0);
    for (int i = 0; i < ks; ++i) {
      T[(size_t)i * mm + i] = theta[i];
      if (ks < mm) T[(size_t)i * mm + ks] = T[(size_t)ks * mm + i] = svec[i];
    }
    for (int j = ks; j < mm; ++j) {
      T[(size_t)j * mm + j] = hA[j];
      if (j + 1 < mm) T[(size_t)j * mm + j + 1] = T[(size_t)(j + 1) * mm + j] = hB[j];
    }
    sym_eig_jacobi(mm, T, w, Z);
    pick = select(w, mm, k);
    const double bm = invariant ? 0.0 : hB[mm - 1];
    conv = 0;
    for (int i : pick)
      if (std::fabs(bm * Z[(size_t)(mm - 1) * mm + i]) <= tol * std::max(eps23, std::fabs(w[i])))
        ++conv;
    ++cycles;
    if ((int)pick.size() == k && conv == k) break;
    if (invariant) {
      if (mm < k) rc = fail(CXA_ERR_NUMERIC, "[ERROR][Lanczos] the start vector spans an "
                                             "invariant subspace of dimension %d < k = %d",
                            mm, k);
      break;
    }
    if (full || cycles >= Maxiter) break;
    // thick restart: U = V Y for the `kept` most wanted Ritz vectors, written over V[0..kept);
    // the residual (cur, its norm partials) continues as the next basis vector
    const int kept = std::min(m - 1, k + (m - k) / 2);
    std::vector<int> keep = select(w, mm, kept);
    std::vector<double> Yt((size_t)kept * mm);
    theta.assign(kept, 0.0);
    svec.assign(kept, 0.0);
    for (int i = 0; i < kept; ++i) {
      theta[i] = w[keep[i]];
      svec[i] = bm * Z[(size_t)(mm - 1) * mm + keep[i]];
      for (int j = 0; j < mm; ++j) Yt[(size_t)i * mm + j] = Z[(size_t)j * mm + keep[i]];
    }
    if (!dY) {
      if (hipMalloc(&dY, sizeof(double) * (size_t)m * m) != hipSuccess) {
        rc = fail(CXA_ERR_NOMEM, "cxa_lanczos_eigsh: restart coefficients");
        break;
      }
    }
    double *tmp = L->V + (size_t)m * L->n;  // rows m .. m + kept - 1 of the context
    if (hipMemcpyAsync(dY, Yt.data(), sizeof(double) * Yt.size(), hipMemcpyHostToDevice, ws) !=
        hipSuccess) {
      rc = fail(CXA_ERR_HIP, "cxa_lanczos_eigsh: coefficient upload failed");
      break;
    }
    // tmp (kept x n) = Yt (kept x mm) . V (mm x n), every basis vector a row of n doubles
    if ((rc = cxa_matmul(CXA_F64, tmp, dY, L->V, kept, mm, L->n, (cxa_stream_t)ws))) break;
    if (hipMemcpyAsync(L->V, tmp, sizeof(double) * (size_t)kept * L->n, hipMemcpyDeviceToDevice,
                       ws) != hipSuccess ||
        hipStreamSynchronize(ws) != hipSuccess) {  // Yt is read by the upload above
      rc = fail(CXA_ERR_HIP, "cxa_lanczos_eigsh: restart copy failed");
      break;
    }
    start = kept;
  }
  if (dY) (void)hipFree(dY);
  if (rc) return rc;
  for (int i = 0; i < k && i < (int)pick.size(); ++i) evals[i] = w[pick[i]];
  if (evecs) {
    std::vector<double> y(mm);
    for (int i = 0; i < k && i < (int)pick.size(); ++i) {
      for (int j = 0; j < mm; ++j) y[j] = Z[(size_t)j * mm + pick[i]];
      if ((rc = combine_host_coef(L, y.data(), mm, evecs + (size_t)i * L->n, 0, ws))) return rc;
    }
  }
  if (niter) *niter = matvecs;
  if (nconv) *nconv = conv;
  return run_end(r);
}
