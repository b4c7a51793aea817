// Contraction entry points: dense matmul, batched GEMM, tensordot, and the block-sparse
// contraction planner that turns every charge-matched block pair of a BlockUniTensor
// contraction into segments of one grouped GEMM launch.
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "cxa_common.hpp"
#include "contract.hpp"
#include "gemm.hpp"

namespace cxa {
namespace {

bool is_complex(int dt) { return dt == CXA_C128 || dt == CXA_C64; }

// Can the axes `axes` (in this order) of a row-major array with extents `ext` be addressed
// as one index with a single stride?  Unit axes are ignored.  On success *stride is the
// element stride of the combined index (0 when every axis is a unit axis).
bool merged_stride(int rank, const uint64_t *ext, const int *axes, int n, int64_t *stride) {
  int64_t st[16];
  int64_t acc = 1;
  for (int i = rank - 1; i >= 0; --i) {
    st[i] = acc;
    acc *= (int64_t)ext[i];
  }
  int prev = -1;
  for (int i = 0; i < n; ++i) {
    const int ax = axes[i];
    if (ext[ax] == 1) continue;
    if (prev >= 0 && st[prev] != st[ax] * (int64_t)ext[ax]) return false;
    prev = ax;
  }
  *stride = prev >= 0 ? st[prev] : 0;
  return true;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
int upload(DevBuf &d, const std::vector<T> &v) {
  if (v.empty()) return CXA_OK;
  CXA_HIP(hipMalloc(&d.p, v.size() * sizeof(T)));
  return upload_now(d.p, v.data(), v.size() * sizeof(T));
}

struct CopyJob {
  int64_t src_off;  // element offset of the source block
  int64_t dst_off;  // element offset in the workspace
  int rank;
  uint64_t shape[16];
  uint64_t mapper[16];
};

}  // namespace
}  // namespace cxa

// One tiling of a plan's blocks: the tile list on the device plus how it is launched.  A plan
// keeps one per precision mode, so switching precision never re-uploads (launches stay
// capturable) and the tuner can pick different shapes for FP64 and FP32.
struct PlanVariant {
  bool built = false;   // tile list made and uploaded (done on first use, see bc_ensure)
  int kind = -1;        // tile kind override, -1 = per-block default
  int split_mode = -1;  // -1 automatic, 0 off, 1 on (four waves per tile), 2 eight waves
  cxa::DevBuf d_tiles;  // kXcds ranges of per_xcd tiles (made in a reused host buffer)
  int ntiles = 0, nlive = 0;  // tiles in the list, and those that are not padding
  int per_xcd = 0;
  int splitk = 0;  // waves per tile of the split-K launch, 0 = none (mix plans: 1 = four-wave)
  int tile_class = 0;
  int launch_kind = -1;  // mix plans: the one tile kind of the launch
};

struct cxa_bc_plan {
  int dtype = 0;
  int es = 0;
  int device = 0;
  std::vector<cxa::GBlk> blks;
  std::vector<cxa::GSeg> segs;
  cxa::DevBuf d_meta;  // blks, segs and (mix plans) outs in one device allocation
  const cxa::GBlk *dev_blks = nullptr;
  const cxa::GSeg *dev_segs = nullptr;
  const cxa::MixOut *dev_outs = nullptr;
  PlanVariant var[3];  // indexed by CXA_PREC_*
  std::mutex mu;       // guards building a variant on first use
  bool copyA = false, copyB = false;
  std::vector<cxa::CopyJob> jobsA, jobsB;  // blocks too large for the batched copy
  // the other operand blocks, permuted into the workspace by one batched launch per operand
  std::vector<cxa::BatchCopyJob> bjobsA, bjobsB;
  std::vector<cxa::BatchCopyTask> btasksA, btasksB;
  const cxa::BatchCopyJob *dev_bjobsA = nullptr, *dev_bjobsB = nullptr;
  const cxa::BatchCopyTask *dev_btasksA = nullptr, *dev_btasksB = nullptr;
  int64_t wsA = 0, wsB = 0;  // workspace elements for A and B copies
  double flops = 0;
  int64_t npairs = 0;
  // mix plans (cxa_mix_contract_plan): blks are items, segs their slices, outs the outputs
  bool mix = false;
  std::vector<cxa::MixOut> outs;
};

namespace cxa {
namespace {

// tile kinds a mix launch can run (every slice of an item stays in the wave's registers)
bool mix_kind_ok(int kind) {
  return kind == TK_1x1 || kind == TK_2x1 || kind == TK_1x2 || kind == TK_2x2;
}

int build_variant(cxa_bc_plan *p, int prec) {
  const bool real = p->dtype == CXA_F64 || p->dtype == CXA_F32;
  PlanVariant &v = p->var[prec];
  v.built = false;
  // the host tile list lives only until it is uploaded: one buffer per thread keeps its pages
  thread_local std::vector<GTile> tiles;
  if (p->mix) {
    v.launch_kind = mix_kind_ok(v.kind) ? v.kind : TK_1x1;
    v.tile_class = make_tiles(p->blks, p->segs, tiles, &v.per_xcd, v.launch_kind, false);
    v.splitk = v.split_mode == 1 ? 1 : 0;  // mix plans: four-wave workgroups (two slices each)
  } else {
    v.tile_class = make_tiles(p->blks, p->segs, tiles, &v.per_xcd, v.kind, real);
    const bool split = real && v.tile_class != 2 &&
                       (v.split_mode >= 1 ||
                        (v.split_mode < 0 && prefer_splitk(p->blks, p->segs, tiles,
                                                           v.tile_class)));
    // eight-wave split-K only for MFMA tiles (row tiles always split over four waves)
    v.splitk = !split ? 0 : (v.split_mode == 2 && v.tile_class == 0 ? 8 : 4);
  }
  v.ntiles = (int)tiles.size();
  v.nlive = 0;
  for (const GTile &t : tiles) v.nlive += t.blk >= 0;
  if (v.d_tiles.p) {
    CXA_HIP(hipFree(v.d_tiles.p));
    v.d_tiles.p = nullptr;
  }
  const int rc = upload(v.d_tiles, tiles);
  v.built = rc == CXA_OK;
  return rc;
}

// The precision an exec of a plan of this dtype runs when asked for FP64 (FP32 data runs FP32).
int native_prec(int dtype) {
  return dtype == CXA_F32 || dtype == CXA_C64 ? CXA_PREC_FP32 : CXA_PREC_FP64;
}

// The task metadata in one allocation and one copy; only the tile list of the plan's own
// precision is made now, the others on first use (a tile list of a large contraction takes
// about a millisecond of host time, and most plans only ever run one precision).
int upload_plan(cxa_bc_plan *p) {
  // blks, segs, outs and the batched copy jobs / tasks of both operands, each 256-byte aligned
  struct Part {
    const void *src;
    size_t bytes, at;
  };
  Part parts[7] = {{p->blks.data(), p->blks.size() * sizeof(GBlk), 0},
                   {p->segs.data(), p->segs.size() * sizeof(GSeg), 0},
                   {p->outs.data(), p->mix ? p->outs.size() * sizeof(MixOut) : 0, 0},
                   {p->bjobsA.data(), p->bjobsA.size() * sizeof(BatchCopyJob), 0},
                   {p->btasksA.data(), p->btasksA.size() * sizeof(BatchCopyTask), 0},
                   {p->bjobsB.data(), p->bjobsB.size() * sizeof(BatchCopyJob), 0},
                   {p->btasksB.data(), p->btasksB.size() * sizeof(BatchCopyTask), 0}};
  size_t total = 0;
  for (Part &q : parts) {
    q.at = total;
    total += (q.bytes + 255) / 256 * 256;
  }
  if (total) {
    thread_local std::vector<char> host;  // keeps its pages between plans
    host.assign(total, 0);
    for (const Part &q : parts)
      if (q.bytes) std::memcpy(host.data() + q.at, q.src, q.bytes);
    CXA_HIP(hipMalloc(&p->d_meta.p, host.size()));
    const int rc = upload_now(p->d_meta.p, host.data(), host.size());
    if (rc) return rc;
    char *base = (char *)p->d_meta.p;
    auto at = [&](int i) -> const void * { return parts[i].bytes ? base + parts[i].at : nullptr; };
    p->dev_blks = (const GBlk *)at(0);
    p->dev_segs = (const GSeg *)at(1);
    p->dev_outs = (const MixOut *)at(2);
    p->dev_bjobsA = (const BatchCopyJob *)at(3);
    p->dev_btasksA = (const BatchCopyTask *)at(4);
    p->dev_bjobsB = (const BatchCopyJob *)at(5);
    p->dev_btasksB = (const BatchCopyTask *)at(6);
  }
  return build_variant(p, native_prec(p->dtype));
}

// Split an operand's permute-to-workspace jobs: small blocks into one batched launch (jobs +
// tasks), the rest (large blocks, or a lone small block, which the tiled permute moves faster)
// onto the one-by-one list.
void split_copies(const std::vector<CopyJob> &all, std::vector<CopyJob> &big,
                  std::vector<BatchCopyJob> &bj, std::vector<BatchCopyTask> &bt) {
  std::vector<BatchCopyJob> small;
  std::vector<const CopyJob *> from;
  for (const CopyJob &j : all) {
    BatchCopyJob b;
    if (batch_copy_job(j.rank, j.shape, j.mapper, j.src_off, j.dst_off, &b)) {
      small.push_back(b);
      from.push_back(&j);
    } else {
      big.push_back(j);
    }
  }
  if (small.size() < 2) {
    for (const CopyJob *j : from) big.push_back(*j);
    return;
  }
  for (const BatchCopyJob &b : small) {
    const int id = (int)bj.size();
    bj.push_back(b);
    for (int32_t at = 0; at < b.n; at += kBatchCopyTaskElems)
      bt.push_back({id, at, std::min(kBatchCopyTaskElems, b.n - at), 0});
  }
}

}  // namespace

// Make (once) the tile list of one precision.  A host-side copy runs, so an exec inside a graph
// capture refuses a precision that was never prepared.
int bc_ensure(const cxa_bc_plan *cp, int prec, hipStream_t s) {
  cxa_bc_plan *p = const_cast<cxa_bc_plan *>(cp);
  std::lock_guard<std::mutex> lk(p->mu);
  if (p->var[prec].built) return CXA_OK;
  if (s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    CXA_HIP(hipStreamIsCapturing(s, &st));
    CXA_CHECK(st == hipStreamCaptureStatusNone, CXA_ERR_INVALID,
              "block contract: the precision %d tile list of this plan is made on first use, "
              "which cannot happen during a graph capture (run or configure it once first)",
              prec);
  }
  return build_variant(p, prec);
}

int bc_prepare(const cxa_bc_plan *p) {
  for (int q = 0; q < 3; ++q) {
    const int rc = bc_ensure(p, q, nullptr);
    if (rc) return rc;
  }
  return CXA_OK;
}

}  // namespace cxa

using namespace cxa;

// ------------------------------------------------------------------------------------------
// Block-sparse contraction planner.  Pairing follows BlockUniTensor::contract
// (src/BlockUniTensor.cpp:948-1228): left blocks in order, right blocks whose common-qindex
// tuple matches (in order), target = the output block with itoi (L free ++ R free).  Each pair
// becomes one GEMM segment of its target block: M = prod(L free dims), K = prod(common dims),
// N = prod(R free dims).  Operands are read through strides when the permuted block is still a
// strided matrix; otherwise the plan copies the operand blocks to the workspace first.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_block_contract_plan(int dtype, int rankL, int nbL, const uint64_t *itoiL,
                                       const uint64_t *shapeL, const int64_t *offL, int rankR,
                                       int nbR, const uint64_t *itoiR, const uint64_t *shapeR,
                                       const int64_t *offR, int ncomm, const uint64_t *comm1,
                                       const uint64_t *comm2, int nbC, const uint64_t *itoiC,
                                       const uint64_t *shapeC, const int64_t *offC,
                                       cxa_bc_plan **out) {
  CXA_CHECK(out, CXA_ERR_INVALID, "cxa_block_contract_plan: null output handle");
  *out = nullptr;
  const int es = elem_size(dtype);
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "block contract: unsupported dtype %d", dtype);
  CXA_CHECK(rankL >= 0 && rankL <= 16 && rankR >= 0 && rankR <= 16, CXA_ERR_UNSUPPORTED,
            "block contract: rank above 16");
  CXA_CHECK(ncomm >= 0 && ncomm <= rankL && ncomm <= rankR, CXA_ERR_INVALID,
            "block contract: %d common legs for ranks %d and %d", ncomm, rankL, rankR);
  bool isc1[16] = {false}, isc2[16] = {false};
  for (int c = 0; c < ncomm; ++c) {
    CXA_CHECK(comm1[c] < (uint64_t)rankL && comm2[c] < (uint64_t)rankR && !isc1[comm1[c]] &&
                  !isc2[comm2[c]],
              CXA_ERR_INVALID, "block contract: invalid common index list");
    isc1[comm1[c]] = true;
    isc2[comm2[c]] = true;
  }
  int nc1[16], nc2[16], n1 = 0, n2 = 0;
  for (int i = 0; i < rankL; ++i)
    if (!isc1[i]) nc1[n1++] = i;
  for (int i = 0; i < rankR; ++i)
    if (!isc2[i]) nc2[n2++] = i;
  const int rankC = n1 + n2;

  std::unique_ptr<cxa_bc_plan> p(new cxa_bc_plan());
  p->dtype = dtype;
  p->es = es;
  CXA_HIP(hipGetDevice(&p->device));

  // right blocks by common-qindex tuple (ascending block order), output blocks by itoi
  std::map<std::vector<uint64_t>, std::vector<int>> mp;
  for (int b = 0; b < nbR; ++b) {
    std::vector<uint64_t> key(ncomm);
    for (int c = 0; c < ncomm; ++c) key[c] = itoiR[(int64_t)b * rankR + comm2[c]];
    mp[key].push_back(b);
  }
  std::map<std::vector<uint64_t>, int> mpC;
  for (int t = 0; t < nbC; ++t)
    mpC[std::vector<uint64_t>(itoiC + (int64_t)t * rankC, itoiC + (int64_t)t * rankC + rankC)] =
        t;

  struct Pair {
    int a, b, t;
    int64_t M, K, N;
  };
  std::vector<Pair> pairs;
  for (int a = 0; a < nbL; ++a) {
    std::vector<uint64_t> key(ncomm);
    for (int c = 0; c < ncomm; ++c) key[c] = itoiL[(int64_t)a * rankL + comm1[c]];
    auto it = mp.find(key);
    if (it == mp.end()) continue;
    const uint64_t *sa = shapeL + (int64_t)a * rankL;
    int64_t M = 1, K = 1;
    for (int i = 0; i < n1; ++i) M *= (int64_t)sa[nc1[i]];
    for (int c = 0; c < ncomm; ++c) K *= (int64_t)sa[comm1[c]];
    for (int b : it->second) {
      const uint64_t *sb = shapeR + (int64_t)b * rankR;
      int64_t N = 1, Kb = 1;
      for (int i = 0; i < n2; ++i) N *= (int64_t)sb[nc2[i]];
      for (int c = 0; c < ncomm; ++c) {
        CXA_CHECK(sb[comm2[c]] == sa[comm1[c]], CXA_ERR_INVALID,
                  "block contract: contracted dimension mismatch (%llu vs %llu)",
                  (unsigned long long)sa[comm1[c]], (unsigned long long)sb[comm2[c]]);
        Kb *= (int64_t)sb[comm2[c]];
      }
      std::vector<uint64_t> tk(rankC);
      for (int i = 0; i < n1; ++i) tk[i] = itoiL[(int64_t)a * rankL + nc1[i]];
      for (int i = 0; i < n2; ++i) tk[n1 + i] = itoiR[(int64_t)b * rankR + nc2[i]];
      auto ct = mpC.find(tk);
      CXA_CHECK(ct != mpC.end(), CXA_ERR_INVALID,
                "block contract: pair (%d,%d) has no output block (flux rule violated)", a, b);
      pairs.push_back({a, b, ct->second, M, K, N});
      (void)Kb;
    }
  }
  p->npairs = (int64_t)pairs.size();

  // can every operand be read in place?
  int axA[32], axB[32];
  for (int i = 0; i < n1; ++i) axA[i] = nc1[i];
  for (int c = 0; c < ncomm; ++c) axA[n1 + c] = (int)comm1[c];
  for (int c = 0; c < ncomm; ++c) axB[c] = (int)comm2[c];
  for (int i = 0; i < n2; ++i) axB[ncomm + i] = nc2[i];
  std::vector<int64_t> sam(nbL), sak(nbL), sbk(nbR), sbn(nbR);
  std::vector<char> usedA(nbL, 0), usedB(nbR, 0);
  for (const Pair &pr : pairs) {
    usedA[pr.a] = 1;
    usedB[pr.b] = 1;
  }
  for (int a = 0; a < nbL; ++a) {
    if (!usedA[a]) continue;
    const uint64_t *sa = shapeL + (int64_t)a * rankL;
    if (!merged_stride(rankL, sa, axA, n1, &sam[a]) ||
        !merged_stride(rankL, sa, axA + n1, ncomm, &sak[a]))
      p->copyA = true;
  }
  for (int b = 0; b < nbR; ++b) {
    if (!usedB[b]) continue;
    const uint64_t *sb = shapeR + (int64_t)b * rankR;
    if (!merged_stride(rankR, sb, axB, ncomm, &sbk[b]) ||
        !merged_stride(rankR, sb, axB + ncomm, n2, &sbn[b]))
      p->copyB = true;
  }
  std::vector<int64_t> wsoffA(nbL, 0), wsoffB(nbR, 0);
  std::vector<CopyJob> copiesA, copiesB;
  if (p->copyA) {
    for (int a = 0; a < nbL; ++a) {
      if (!usedA[a]) continue;
      const uint64_t *sa = shapeL + (int64_t)a * rankL;
      CopyJob j{};
      j.src_off = offL[a];
      j.dst_off = p->wsA;
      j.rank = rankL;
      int64_t M = 1, K = 1, n = 1;
      for (int i = 0; i < rankL; ++i) {
        j.shape[i] = sa[i];
        j.mapper[i] = (uint64_t)axA[i];
        n *= (int64_t)sa[i];
      }
      for (int i = 0; i < n1; ++i) M *= (int64_t)sa[nc1[i]];
      for (int c = 0; c < ncomm; ++c) K *= (int64_t)sa[comm1[c]];
      wsoffA[a] = p->wsA;
      sam[a] = K;
      sak[a] = 1;
      p->wsA += (n + 15) / 16 * 16;
      copiesA.push_back(j);
      (void)M;
    }
  }
  if (p->copyB) {
    for (int b = 0; b < nbR; ++b) {
      if (!usedB[b]) continue;
      const uint64_t *sb = shapeR + (int64_t)b * rankR;
      CopyJob j{};
      j.src_off = offR[b];
      j.dst_off = p->wsB;
      j.rank = rankR;
      int64_t N = 1, n = 1;
      for (int i = 0; i < rankR; ++i) {
        j.shape[i] = sb[i];
        j.mapper[i] = (uint64_t)axB[i];
        n *= (int64_t)sb[i];
      }
      for (int i = 0; i < n2; ++i) N *= (int64_t)sb[nc2[i]];
      wsoffB[b] = p->wsB;
      sbk[b] = N;
      sbn[b] = 1;
      p->wsB += (n + 15) / 16 * 16;
      copiesB.push_back(j);
    }
  }

  split_copies(copiesA, p->jobsA, p->bjobsA, p->btasksA);
  split_copies(copiesB, p->jobsB, p->bjobsB, p->btasksB);

  // one GBlk per output block, segments grouped by target in pair order
  std::vector<std::vector<int>> by_target(nbC);
  for (size_t i = 0; i < pairs.size(); ++i) by_target[pairs[i].t].push_back((int)i);
  p->blks.resize(nbC);
  const double cplx = is_complex(dtype) ? 4.0 : 1.0;
  for (int t = 0; t < nbC; ++t) {
    GBlk &g = p->blks[t];
    const uint64_t *sc = shapeC + (int64_t)t * rankC;
    int64_t M = 1, N = 1;
    for (int i = 0; i < n1; ++i) M *= (int64_t)sc[i];
    for (int i = 0; i < n2; ++i) N *= (int64_t)sc[n1 + i];
    CXA_CHECK(M <= INT32_MAX && N <= INT32_MAX, CXA_ERR_UNSUPPORTED,
              "block contract: output block too large");
    g.c_off = offC[t];
    g.ldc = N;
    g.alpha = 1.0;
    g.beta = 0.0;
    g.M = (int32_t)M;
    g.N = (int32_t)N;
    g.seg0 = (int32_t)p->segs.size();
    g.nseg = (int32_t)by_target[t].size();
    for (int pi : by_target[t]) {
      const Pair &pr = pairs[pi];
      CXA_CHECK(pr.M == M && pr.N == N, CXA_ERR_INVALID,
                "block contract: output block %d shape does not match its pair (%lld x %lld)", t,
                (long long)pr.M, (long long)pr.N);
      CXA_CHECK(pr.K <= INT32_MAX, CXA_ERR_UNSUPPORTED, "block contract: K too large");
      GSeg s{};
      s.a_off = p->copyA ? wsoffA[pr.a] : offL[pr.a];
      s.b_off = p->copyB ? wsoffB[pr.b] : offR[pr.b];
      s.sam = sam[pr.a];
      s.sak = sak[pr.a];
      s.sbk = sbk[pr.b];
      s.sbn = sbn[pr.b];
      s.K = (int32_t)pr.K;
      p->segs.push_back(s);
      p->flops += 2.0 * cplx * (double)pr.M * (double)pr.N * (double)pr.K;
    }
  }
  int rc = upload_plan(p.get());
  if (rc) return rc;
  *out = p.release();
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Mix plan: items of one M x N shape, each the weighted sums of up to CXA_MIX_SLICES products
// (see cytnx_amd.h).  Runs through the block-contract exec / config / destroy entry points.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_mix_contract_plan(int dtype, int nitems, const int64_t *items,
                                     int64_t nslices, const int64_t *slices, int64_t nouts,
                                     const int64_t *outs, const double *coef, cxa_bc_plan **out) {
  CXA_CHECK(out, CXA_ERR_INVALID, "cxa_mix_contract_plan: null output handle");
  *out = nullptr;
  CXA_CHECK(dtype == CXA_F64, CXA_ERR_UNSUPPORTED,
            "cxa_mix_contract_plan: mix plans run real FP64 operands (dtype %d)", dtype);
  CXA_CHECK(nitems >= 0 && nslices >= 0 && nouts >= 0 && (nitems == 0 || items) &&
                (nslices == 0 || slices) && (nouts == 0 || (outs && coef)),
            CXA_ERR_INVALID, "cxa_mix_contract_plan: bad arguments");
  CXA_CHECK(CXA_MIX_SLICES == kMixSlices, CXA_ERR_INVALID, "mix slice bound mismatch");
  std::unique_ptr<cxa_bc_plan> p(new cxa_bc_plan());
  p->dtype = dtype;
  p->es = elem_size(dtype);
  p->mix = true;
  CXA_HIP(hipGetDevice(&p->device));
  int64_t s_at = 0, o_at = 0;
  for (int i = 0; i < nitems; ++i) {
    const int64_t M = items[4 * i], N = items[4 * i + 1];
    const int64_t ns = items[4 * i + 2], no = items[4 * i + 3];
    CXA_CHECK(M >= 0 && N >= 0 && M <= INT32_MAX && N <= INT32_MAX, CXA_ERR_INVALID,
              "mix plan: item %d has shape %lld x %lld", i, (long long)M, (long long)N);
    CXA_CHECK(ns >= 0 && ns <= kMixSlices, CXA_ERR_UNSUPPORTED,
              "mix plan: item %d has %lld slices (at most %d)", i, (long long)ns, kMixSlices);
    CXA_CHECK(no >= 0 && s_at + ns <= nslices && o_at + no <= nouts, CXA_ERR_INVALID,
              "mix plan: item %d runs past the slice or output list", i);
    GBlk b{};
    b.M = (int32_t)M;
    b.N = (int32_t)N;
    b.alpha = 1.0;
    b.seg0 = (int32_t)s_at;
    b.nseg = (int32_t)ns;
    b.out0 = (int32_t)o_at;
    b.nout = (int32_t)no;
    p->blks.push_back(b);
    for (int64_t s = s_at; s < s_at + ns; ++s) {
      const int64_t *d = slices + 7 * s;
      CXA_CHECK(d[0] >= 0 && d[1] >= 0 && d[6] >= 0 && d[6] <= INT32_MAX, CXA_ERR_INVALID,
                "mix plan: slice %lld has a negative offset or bad K", (long long)s);
      GSeg sg{};
      sg.a_off = d[0];
      sg.b_off = d[1];
      sg.sam = d[2];
      sg.sak = d[3];
      sg.sbk = d[4];
      sg.sbn = d[5];
      sg.K = (int32_t)d[6];
      p->segs.push_back(sg);
      p->flops += 2.0 * (double)M * (double)N * (double)sg.K;
    }
    for (int64_t o = o_at; o < o_at + no; ++o) {
      CXA_CHECK(outs[2 * o] >= 0 && outs[2 * o + 1] >= N, CXA_ERR_INVALID,
                "mix plan: output %lld has a negative offset or a row stride below N",
                (long long)o);
      MixOut mo{};
      mo.c_off = outs[2 * o];
      mo.ldc = outs[2 * o + 1];
      for (int s = 0; s < kMixSlices; ++s) mo.coef[s] = s < ns ? coef[kMixSlices * o + s] : 0.0;
      p->outs.push_back(mo);
    }
    s_at += ns;
    o_at += no;
  }
  CXA_CHECK(s_at == nslices && o_at == nouts, CXA_ERR_INVALID,
            "mix plan: items cover %lld slices / %lld outputs, lists hold %lld / %lld",
            (long long)s_at, (long long)o_at, (long long)nslices, (long long)nouts);
  p->npairs = nslices;
  int rc = upload_plan(p.get());
  if (rc) return rc;
  *out = p.release();
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Strided-batch plan: nbatch GEMMs of one shape whose operands are strided matrix views of
// three buffers (see cytnx_amd.h).  One GBlk with one GSeg per batch entry; no copies.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_strided_batch_plan(int dtype, int64_t nbatch, int64_t M, int64_t N, int64_t K,
                                      int64_t a_bstride, int64_t sam, int64_t sak,
                                      int64_t b_bstride, int64_t sbk, int64_t sbn,
                                      int64_t c_bstride, int64_t ldc, cxa_bc_plan **out) {
  CXA_CHECK(out, CXA_ERR_INVALID, "cxa_strided_batch_plan: null output handle");
  *out = nullptr;
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "strided batch plan: unsupported dtype %d", dtype);
  CXA_CHECK(nbatch >= 0 && nbatch <= INT32_MAX && M >= 0 && N >= 0 && K >= 0 &&
                M <= INT32_MAX && N <= INT32_MAX && K <= INT32_MAX,
            CXA_ERR_INVALID, "strided batch plan: bad sizes (%lld x %lld x %lld, %lld)",
            (long long)M, (long long)N, (long long)K, (long long)nbatch);
  CXA_CHECK(a_bstride >= 0 && sam >= 0 && sak >= 0 && b_bstride >= 0 && sbk >= 0 && sbn >= 0 &&
                c_bstride >= 0 && ldc >= N,
            CXA_ERR_INVALID, "strided batch plan: negative stride or ldc below N");
  std::unique_ptr<cxa_bc_plan> p(new cxa_bc_plan());
  p->dtype = dtype;
  p->es = elem_size(dtype);
  CXA_HIP(hipGetDevice(&p->device));
  const double cplx = is_complex(dtype) ? 4.0 : 1.0;
  for (int64_t b = 0; b < nbatch; ++b) {
    GBlk g{};
    g.c_off = b * c_bstride;
    g.ldc = ldc;
    g.alpha = 1.0;
    g.beta = 0.0;
    g.M = (int32_t)M;
    g.N = (int32_t)N;
    g.seg0 = (int32_t)b;
    g.nseg = 1;
    p->blks.push_back(g);
    GSeg sg{};
    sg.a_off = b * a_bstride;
    sg.b_off = b * b_bstride;
    sg.sam = sam;
    sg.sak = sak;
    sg.sbk = sbk;
    sg.sbn = sbn;
    sg.K = (int32_t)K;
    p->segs.push_back(sg);
    p->flops += 2.0 * cplx * (double)M * (double)N * (double)K;
  }
  p->npairs = nbatch;
  int rc = upload_plan(p.get());
  if (rc) return rc;
  *out = p.release();
  return CXA_OK;
}

extern "C" int cxa_block_contract_info(const cxa_bc_plan *p, double *flops,
                                       int64_t *workspace_bytes, int64_t *npairs) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_info: null plan");
  if (flops) *flops = p->flops;
  if (workspace_bytes) *workspace_bytes = (p->wsA + p->wsB) * p->es;
  if (npairs) *npairs = p->npairs;
  return CXA_OK;
}

extern "C" int cxa_block_contract_set_config(cxa_bc_plan *p, int prec, int kind, int split) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_set_config: null plan");
  CXA_CHECK(prec >= -1 && prec <= CXA_PREC_FP32X3, CXA_ERR_INVALID, "unknown precision %d", prec);
  CXA_CHECK(kind >= -1 && kind < TK_COUNT, CXA_ERR_INVALID, "unknown tile kind %d", kind);
  CXA_CHECK(split >= -1 && split <= 2, CXA_ERR_INVALID, "split mode must be -1, 0, 1 or 2");
  CXA_CHECK(!p->mix || kind == -1 || mix_kind_ok(kind), CXA_ERR_UNSUPPORTED,
            "mix plans run 16x16, 32x16, 16x32 or 32x32 tiles (kind %d)", kind);
  CXA_CHECK(!p->mix || split <= 1, CXA_ERR_UNSUPPORTED,
            "mix plans run eight- or four-wave workgroups (split 0 or 1)");
  std::lock_guard<std::mutex> lk(p->mu);
  for (int q = 0; q < 3; ++q) {
    if (prec >= 0 && q != prec) continue;
    PlanVariant &v = p->var[q];
    if (v.kind == kind && v.split_mode == split && v.built) continue;
    v.kind = kind;
    v.split_mode = split;
    if (!v.built && q != prec) continue;  // made with this configuration on first use
    int rc = build_variant(p, q);
    if (rc) return rc;
  }
  return CXA_OK;
}

extern "C" int cxa_block_contract_get_config(const cxa_bc_plan *p, int prec, int *kind,
                                             int *split, int *ntiles) {
  CXA_CHECK(p && prec >= 0 && prec <= CXA_PREC_FP32X3, CXA_ERR_INVALID,
            "cxa_block_contract_get_config: bad arguments");
  int rc = bc_ensure(p, prec, nullptr);
  if (rc) return rc;
  const PlanVariant &v = p->var[prec];
  if (kind) *kind = v.kind;
  if (split) *split = v.splitk == 8 ? 2 : (v.splitk ? 1 : 0);
  if (ntiles) *ntiles = v.nlive;
  return CXA_OK;
}

extern "C" int cxa_block_contract_set_tile(cxa_bc_plan *p, int kind) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_set_tile: null plan");
  CXA_CHECK(kind >= -1 && kind < TK_COUNT, CXA_ERR_INVALID, "unknown tile kind %d", kind);
  std::lock_guard<std::mutex> lk(p->mu);
  for (int q = 0; q < 3; ++q) {
    p->var[q].kind = kind;
    if (!p->var[q].built) continue;  // made with this kind on first use
    int rc = build_variant(p, q);
    if (rc) return rc;
  }
  return CXA_OK;
}

namespace cxa {

// Execute a plan with explicit buffer dtypes (a_dt / b_dt / c_dt may differ from the plan's
// dtype: FP32 copies of the same layout, see the chain).  Offsets are in elements, so the same
// task list addresses either copy.  Workspace offsets are in elements of the copied operand.
int bc_exec_typed(const cxa_bc_plan *p, const void *L, const void *R, void *C, int a_dt,
                  int b_dt, int c_dt, int prec, void *workspace, hipStream_t s, int conj,
                  const ExecLayout *lay) {
  const ExecLayout none;
  if (!lay) lay = &none;
  CXA_CHECK(!(lay->a_k4 || lay->b_k4) || (lay->segs && !is_complex(p->dtype)),
            CXA_ERR_INVALID, "block contract exec: packed operands need remapped segments of a "
                             "real plan");
  CXA_CHECK(!(lay->a_k4 && p->copyA) && !(lay->b_k4 && p->copyB), CXA_ERR_INVALID,
            "block contract exec: a packed operand cannot also be permuted into the workspace");
  CXA_CHECK(!lay->c_k4 || (p->mix && lay->outs), CXA_ERR_INVALID,
            "block contract exec: packed outputs need a mix plan and its packed outputs");
  CXA_CHECK(p, CXA_ERR_INVALID, "block contract exec: null plan");
  CXA_CHECK(!(p->copyA || p->copyB) || workspace, CXA_ERR_INVALID,
            "block contract exec: this plan needs a workspace of %lld bytes",
            (long long)((p->wsA + p->wsB) * 16));
  const int esA = elem_size(a_dt), esB = elem_size(b_dt);
  char *ws = (char *)workspace;
  for (const CopyJob &j : p->jobsA) {
    int rc = cxa_permute(a_dt, j.rank, j.shape, j.mapper, (const char *)L + j.src_off * esA,
                         ws + j.dst_off * esA, s);
    if (rc) return rc;
  }
  if (!p->btasksA.empty()) {
    int rc = launch_batch_copy(esA, p->dev_bjobsA, p->dev_btasksA, (int)p->btasksA.size(), L, ws,
                               s);
    if (rc) return rc;
  }
  char *wsB = ws ? ws + p->wsA * esA : nullptr;
  for (const CopyJob &j : p->jobsB) {
    int rc = cxa_permute(b_dt, j.rank, j.shape, j.mapper, (const char *)R + j.src_off * esB,
                         wsB + j.dst_off * esB, s);
    if (rc) return rc;
  }
  if (!p->btasksB.empty()) {
    int rc = launch_batch_copy(esB, p->dev_bjobsB, p->dev_btasksB, (int)p->btasksB.size(), R,
                               wsB, s);
    if (rc) return rc;
  }
  int prec_eff = prec;
  if (is_complex(p->dtype)) prec_eff = p->dtype == CXA_C128 ? CXA_PREC_FP64 : CXA_PREC_FP32;
  if ((a_dt == CXA_F32 || b_dt == CXA_F32) && prec == CXA_PREC_FP64) prec_eff = CXA_PREC_FP32;
  GemmLaunch g;
  g.prec = prec_eff;
  g.a_dt = a_dt;
  g.b_dt = b_dt;
  g.c_dt = c_dt;
  g.A = p->copyA ? (const void *)ws : L;
  g.B = p->copyB ? (const void *)wsB : R;
  g.C = C;
  int rc = bc_ensure(p, prec_eff, s);
  if (rc) return rc;
  const PlanVariant &v = p->var[prec_eff];
  g.tiles = (const GTile *)v.d_tiles.p;
  g.blks = p->dev_blks;
  g.segs = lay->segs ? lay->segs : p->dev_segs;
  g.a_k4 = lay->a_k4;
  g.b_k4 = lay->b_k4;
  g.c_k4 = lay->c_k4;
  g.ntiles = v.ntiles;
  g.per_xcd = v.per_xcd;
  g.splitk = v.splitk;
  g.tile_class = v.tile_class;
  g.conj = is_complex(p->dtype) ? conj : 0;
  g.dot_x = lay->dot_x;
  g.dot_part = lay->dot_part;
  if (p->mix) {
    g.outs = lay->outs ? lay->outs : p->dev_outs;
    g.kind0 = v.launch_kind;
  }
  return launch_grouped_gemm(g, s);
}

namespace {

// Shared body of bc_pack_a / bc_pack_b: `which` 0 packs the A views (free index m, extent the
// block's M), 1 the B views (free index n, extent N).
int pack_views(const cxa_bc_plan *p, int which, int64_t elems, std::vector<GSeg> &segs,
               std::vector<PackRegion> &regions, int64_t *packed_elems) {
  CXA_CHECK(p && packed_elems, CXA_ERR_INVALID, "pack views: bad arguments");
  if (segs.empty()) segs = p->segs;
  CXA_CHECK(segs.size() == p->segs.size(), CXA_ERR_INVALID, "pack views: segment count");
  if ((which == 0 ? p->copyA : p->copyB) || is_complex(p->dtype))
    return fail(CXA_ERR_UNSUPPORTED, "pack views: operand permuted into a workspace or complex");
  regions.clear();
  std::map<std::array<int64_t, 5>, int64_t> at;  // view -> packed offset
  int64_t total = 0;
  for (const GBlk &b : p->blks) {
    for (int s = b.seg0; s < b.seg0 + b.nseg; ++s) {
      GSeg &sg = segs[s];
      int64_t &off0 = which == 0 ? sg.a_off : sg.b_off;
      int64_t &sk = which == 0 ? sg.sak : sg.sbk;
      int64_t &sx = which == 0 ? sg.sam : sg.sbn;
      const int64_t K = sg.K, X = which == 0 ? b.M : b.N;
      if (K > 0 && X > 0) {
        CXA_CHECK(off0 >= 0 && sk >= 0 && sx >= 0 && off0 + (K - 1) * sk + (X - 1) * sx < elems,
                  CXA_ERR_INVALID, "pack views: segment %d reads outside its operand", s);
      }
      const std::array<int64_t, 5> key{off0, sk, sx, K, X};
      auto it = at.find(key);
      int64_t off;
      if (it != at.end()) {
        off = it->second;
      } else {
        off = total;
        at.emplace(key, off);
        CXA_CHECK(K <= INT32_MAX && X <= INT32_MAX, CXA_ERR_UNSUPPORTED,
                  "pack views: view too large");
        regions.push_back(PackRegion{off, off0, sk, sx, (int32_t)K, (int32_t)X});
        total += (K + 3) / 4 * X * 4;
      }
      off0 = off;
      sk = 4 * X;
      sx = 4;
    }
  }
  *packed_elems = total;
  return CXA_OK;
}

}  // namespace

int bc_pack_b(const cxa_bc_plan *p, int64_t elems, std::vector<GSeg> &segs,
              std::vector<PackRegion> &regions, int64_t *packed_elems) {
  return pack_views(p, 1, elems, segs, regions, packed_elems);
}

int bc_pack_a(const cxa_bc_plan *p, int64_t elems, std::vector<GSeg> &segs,
              std::vector<PackRegion> &regions, int64_t *packed_elems) {
  return pack_views(p, 0, elems, segs, regions, packed_elems);
}

int bc_mix_pack_outs(const cxa_bc_plan *mix, const std::vector<PackRegion> &regions,
                     std::vector<MixOut> &outs) {
  CXA_CHECK(mix && mix->mix, CXA_ERR_INVALID, "bc_mix_pack_outs: not a mix plan");
  // regions by source start (views of distinct blocks do not overlap)
  std::vector<const PackRegion *> by(regions.size());
  for (size_t i = 0; i < regions.size(); ++i) by[i] = &regions[i];
  std::sort(by.begin(), by.end(),
            [](const PackRegion *x, const PackRegion *y) { return x->src_off < y->src_off; });
  outs = mix->outs;
  for (const GBlk &it : mix->blks) {
    for (int j = it.out0; j < it.out0 + it.nout; ++j) {
      MixOut &o = outs[j];
      auto up = std::upper_bound(by.begin(), by.end(), o.c_off,
                                 [](int64_t v, const PackRegion *r) { return v < r->src_off; });
      if (up == by.begin())
        return fail(CXA_ERR_UNSUPPORTED, "mix output %d lies before every region", j);
      const PackRegion &R = **(up - 1);
      const int64_t d = o.c_off - R.src_off;
      // strides of extent-1 indices are arbitrary (the planner reports 0 for them)
      int64_t k0 = 0;
      bool ok = it.N == R.X && (R.X == 1 || R.sx == 1);
      if (ok && R.K == 1) {
        ok = d == 0 && it.M == 1;
      } else if (ok) {
        ok = R.sk > 0 && d >= 0 && d % R.sk == 0 && (it.M == 1 || o.ldc == R.sk);
        if (ok) k0 = d / R.sk;
      }
      if (!ok || k0 + it.M > R.K)
        return fail(CXA_ERR_UNSUPPORTED,
                    "mix output %d (c_off %lld, ldc %lld, item %d x %d) does not map onto the "
                    "packed region at %lld (K %d, X %d, sk %lld, sx %lld)",
                    j, (long long)o.c_off, (long long)o.ldc, it.M, it.N, (long long)R.src_off,
                    R.K, R.X, (long long)R.sk, (long long)R.sx);
      o.c_off = R.dst_off;
      o.ldc = k0;
    }
  }
  return CXA_OK;
}

bool bc_is_mix(const cxa_bc_plan *p) { return p->mix; }
int64_t bc_workspace_elems(const cxa_bc_plan *p) { return p->wsA + p->wsB; }

int bc_dot_slots(const cxa_bc_plan *p, int prec) {
  if (!p || p->mix || is_complex(p->dtype) || prec < 0 || prec > CXA_PREC_FP32X3) return 0;
  if (bc_ensure(p, prec, nullptr)) return 0;
  const PlanVariant &v = p->var[prec];
  if (v.tile_class == 1 || v.ntiles == 0) return 0;  // make_tiles' row-streaming class
  return v.ntiles;
}
double bc_flops(const cxa_bc_plan *p) { return p->flops; }
int bc_dtype(const cxa_bc_plan *p) { return p->dtype; }

}  // namespace cxa

extern "C" int cxa_block_contract_set_split(cxa_bc_plan *p, int mode) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_set_split: null plan");
  CXA_CHECK(mode >= -1 && mode <= 2, CXA_ERR_INVALID, "split mode must be -1, 0, 1 or 2");
  std::lock_guard<std::mutex> lk(p->mu);
  for (int q = 0; q < 3; ++q) {
    p->var[q].split_mode = mode;
    if (!p->var[q].built) continue;  // made with this mode on first use
    int rc = build_variant(p, q);
    if (rc) return rc;
  }
  return CXA_OK;
}

extern "C" int cxa_block_contract_exec(const cxa_bc_plan *p, const void *L, const void *R,
                                       void *C, void *workspace, int prec, cxa_stream_t stream) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_exec: null plan");
  return bc_exec_typed(p, L, R, C, p->dtype, p->dtype, p->dtype, prec, workspace,
                       (hipStream_t)stream);
}

extern "C" int cxa_block_contract_exec_ex(const cxa_bc_plan *p, const void *L, const void *R,
                                          void *C, void *workspace, int prec, int flags,
                                          cxa_stream_t stream) {
  CXA_CHECK(p, CXA_ERR_INVALID, "cxa_block_contract_exec_ex: null plan");
  CXA_CHECK(flags >= 0 && flags <= 3, CXA_ERR_INVALID, "cxa_block_contract_exec_ex: flags %d",
            flags);
  return bc_exec_typed(p, L, R, C, p->dtype, p->dtype, p->dtype, prec, workspace,
                       (hipStream_t)stream, flags);
}

extern "C" int cxa_block_contract_destroy(cxa_bc_plan *p) {
  delete p;
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Dense row-major matmul: out(Ml x Nr) = inl(Ml x Comm) . inr(Comm x Nr).
// ------------------------------------------------------------------------------------------
extern "C" int cxa_matmul(int dtype, void *out, const void *inl, const void *inr, int64_t Ml,
                          int64_t Comm, int64_t Nr, cxa_stream_t stream) {
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "cxa_matmul: unsupported dtype %d", dtype);
  CXA_CHECK(Ml >= 0 && Comm >= 0 && Nr >= 0, CXA_ERR_INVALID, "cxa_matmul: negative size");
  CXA_CHECK(Ml <= INT32_MAX && Nr <= INT32_MAX && Comm <= INT32_MAX, CXA_ERR_UNSUPPORTED,
            "cxa_matmul: dimension above 2^31");
  if (Ml == 0 || Nr == 0) return CXA_OK;
  GemmLaunch g;
  g.single = true;
  g.prec = (dtype == CXA_F64 || dtype == CXA_C128) ? CXA_PREC_FP64 : CXA_PREC_FP32;
  g.a_dt = g.b_dt = g.c_dt = dtype;
  g.A = inl;
  g.B = inr;
  g.C = out;
  g.blk0 = GBlk{0, Nr, 1.0, 0.0, (int32_t)Ml, (int32_t)Nr, 0, Comm > 0 ? 1 : 0};
  g.seg0 = GSeg{0, 0, Comm, 1, Nr, 1, (int32_t)Comm, 0};
  return launch_grouped_gemm(g, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------------
// Batched GEMM, BLAS column-major convention (the reference's Gemm_Batchfunc_oii):
// for group g and each of its group_size[g] matrices,
//   C = alpha[g] * op(A) * op(B) + beta[g] * C   (column-major, leading dims lda/ldb/ldc)
// Mapped to row-major by computing C^T = op(B)^T op(A)^T.  One grouped launch; the task list is
// uploaded and the call returns after the launch has been enqueued and the upload completed.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_gemm_batch(int dtype, const char *transa, const char *transb,
                              const int64_t *m, const int64_t *n, const int64_t *k,
                              const double *alpha, const void *const *a, const int64_t *lda,
                              const void *const *b, const int64_t *ldb, const double *beta,
                              void *const *c, const int64_t *ldc, int64_t group_count,
                              const int64_t *group_size, cxa_stream_t stream) {
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "cxa_gemm_batch: unsupported dtype %d", dtype);
  CXA_CHECK(group_count >= 0, CXA_ERR_INVALID, "cxa_gemm_batch: negative group count");
  std::vector<GBlk> blks;
  std::vector<GSeg> segs;
  const int es = elem_size(dtype);
  const bool cplx = dtype == CXA_C128 || dtype == CXA_C64;
  int64_t idx = 0;
  for (int64_t gi = 0; gi < group_count; ++gi) {
    const char ta = transa[gi], tb = transb[gi];
    CXA_CHECK((ta == 'N' || ta == 'T') && (tb == 'N' || tb == 'T'), CXA_ERR_UNSUPPORTED,
              "cxa_gemm_batch: transpose flags must be 'N' or 'T' (got %c,%c)", ta, tb);
    CXA_CHECK(m[gi] >= 0 && n[gi] >= 0 && k[gi] >= 0 && group_size[gi] >= 0, CXA_ERR_INVALID,
              "cxa_gemm_batch: negative size in group %lld", (long long)gi);
    CXA_CHECK(m[gi] <= INT32_MAX && n[gi] <= INT32_MAX && k[gi] <= INT32_MAX,
              CXA_ERR_UNSUPPORTED, "cxa_gemm_batch: dimension above 2^31");
    for (int64_t j = 0; j < group_size[gi]; ++j, ++idx) {
      // row-major C'(n x m) = A'(n x k) . B'(k x m) with A' = op(B)^T, B' = op(A)^T
      CXA_CHECK((uintptr_t)c[idx] % es == 0 && (uintptr_t)a[idx] % es == 0 &&
                    (uintptr_t)b[idx] % es == 0,
                CXA_ERR_INVALID, "cxa_gemm_batch: matrix %lld is not element-aligned",
                (long long)idx);
      GBlk g{};
      g.c_off = (int64_t)((uintptr_t)c[idx] / es);
      g.ldc = ldc[gi];
      if (cplx) {  // (re, im) pairs per group
        g.alpha = alpha[2 * gi];
        g.alpha_im = alpha[2 * gi + 1];
        g.beta = beta[2 * gi];
        g.beta_im = beta[2 * gi + 1];
      } else {
        g.alpha = alpha[gi];
        g.beta = beta[gi];
      }
      g.M = (int32_t)n[gi];
      g.N = (int32_t)m[gi];
      g.seg0 = (int32_t)segs.size();
      g.nseg = k[gi] > 0 ? 1 : 0;
      blks.push_back(g);
      if (g.nseg) {
        GSeg s{};
        s.a_off = (int64_t)((uintptr_t)b[idx] / es);
        s.b_off = (int64_t)((uintptr_t)a[idx] / es);
        if (tb == 'N') {  // op(B)[kk][nn] = B[kk + nn*ldb]
          s.sam = ldb[gi];
          s.sak = 1;
        } else {  // op(B)[kk][nn] = B[nn + kk*ldb]
          s.sam = 1;
          s.sak = ldb[gi];
        }
        if (ta == 'N') {  // op(A)[mm][kk] = A[mm + kk*lda]
          s.sbk = lda[gi];
          s.sbn = 1;
        } else {  // op(A)[mm][kk] = A[kk + mm*lda]
          s.sbk = 1;
          s.sbn = lda[gi];
        }
        s.K = (int32_t)k[gi];
        segs.push_back(s);
      }
    }
  }
  std::vector<GTile> tiles;
  int per_xcd = 0;
  const int cls = make_tiles(blks, segs, tiles, &per_xcd, -1, dtype == CXA_F64 || dtype == CXA_F32);
  if (tiles.empty()) return CXA_OK;
  hipStream_t s = (hipStream_t)stream;
  DevBuf d_blks, d_segs, d_tiles;
  int rc;
  if ((rc = upload(d_blks, blks)) || (rc = upload(d_segs, segs)) || (rc = upload(d_tiles, tiles)))
    return rc;
  GemmLaunch g;
  g.prec = (dtype == CXA_F64 || dtype == CXA_C128) ? CXA_PREC_FP64 : CXA_PREC_FP32;
  g.a_dt = g.b_dt = g.c_dt = dtype;
  g.A = nullptr;  // absolute addresses (in elements) are carried in the offsets
  g.B = nullptr;
  g.C = nullptr;
  g.tiles = (const GTile *)d_tiles.p;
  g.blks = (const GBlk *)d_blks.p;
  g.segs = (const GSeg *)d_segs.p;
  g.ntiles = (int)tiles.size();
  g.per_xcd = per_xcd;
  g.tile_class = cls;
  rc = launch_grouped_gemm(g, s);
  if (rc) return rc;
  // the task arrays are freed on return: wait for the launch that reads them
  CXA_HIP(hipStreamSynchronize(s));
  return CXA_OK;
}

// ------------------------------------------------------------------------------------------
// Dense tensordot: out = sum over idxl/idxr of L * R, output axes = L free then R free
// (the decomposition of _Tensordot_generic, src/linalg/Tensordot.cpp:29-110).  Operands that
// stay strided matrices under the permutation are read in place (the permute is fused into
// the GEMM loads); others are permuted into the caller's workspace first.
// ------------------------------------------------------------------------------------------
extern "C" int cxa_tensordot_workspace(int dtype, int rankL, const uint64_t *shapeL, int rankR,
                                       const uint64_t *shapeR, int ncomm, const uint64_t *idxl,
                                       const uint64_t *idxr, int64_t *bytes) {
  const int es = elem_size(dtype);
  CXA_CHECK(es > 0 && bytes, CXA_ERR_INVALID, "cxa_tensordot_workspace: bad arguments");
  CXA_CHECK(rankL <= 16 && rankR <= 16 && ncomm <= rankL && ncomm <= rankR, CXA_ERR_INVALID,
            "cxa_tensordot_workspace: bad ranks");
  bool isc1[16] = {false}, isc2[16] = {false};
  for (int c = 0; c < ncomm; ++c) {
    CXA_CHECK(idxl[c] < (uint64_t)rankL && idxr[c] < (uint64_t)rankR, CXA_ERR_INVALID,
              "cxa_tensordot_workspace: axis out of range");
    isc1[idxl[c]] = isc2[idxr[c]] = true;
  }
  int axA[32], axB[32], n1 = 0, n2 = 0;
  for (int i = 0; i < rankL; ++i)
    if (!isc1[i]) axA[n1++] = i;
  for (int c = 0; c < ncomm; ++c) axA[n1 + c] = (int)idxl[c];
  for (int c = 0; c < ncomm; ++c) axB[c] = (int)idxr[c];
  for (int i = 0; i < rankR; ++i)
    if (!isc2[i]) axB[ncomm + n2++] = i;
  int64_t s0, s1, need = 0, nL = 1, nR = 1;
  for (int i = 0; i < rankL; ++i) nL *= (int64_t)shapeL[i];
  for (int i = 0; i < rankR; ++i) nR *= (int64_t)shapeR[i];
  if (!merged_stride(rankL, shapeL, axA, n1, &s0) ||
      !merged_stride(rankL, shapeL, axA + n1, ncomm, &s1))
    need += (nL + 15) / 16 * 16;
  if (!merged_stride(rankR, shapeR, axB, ncomm, &s0) ||
      !merged_stride(rankR, shapeR, axB + ncomm, n2, &s1))
    need += (nR + 15) / 16 * 16;
  *bytes = need * es;
  return CXA_OK;
}

extern "C" int cxa_tensordot(int dtype, void *out, const void *L, int rankL,
                             const uint64_t *shapeL, const void *R, int rankR,
                             const uint64_t *shapeR, int ncomm, const uint64_t *idxl,
                             const uint64_t *idxr, void *workspace, cxa_stream_t stream) {
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "cxa_tensordot: unsupported dtype %d", dtype);
  CXA_CHECK(rankL <= 16 && rankR <= 16, CXA_ERR_UNSUPPORTED, "cxa_tensordot: rank above 16");
  CXA_CHECK(ncomm >= 1 && ncomm <= rankL && ncomm <= rankR, CXA_ERR_INVALID,
            "cxa_tensordot: index lists must be non-empty and fit the ranks");
  const int es = elem_size(dtype);
  bool isc1[16] = {false}, isc2[16] = {false};
  for (int c = 0; c < ncomm; ++c) {
    CXA_CHECK(idxl[c] < (uint64_t)rankL && idxr[c] < (uint64_t)rankR, CXA_ERR_INVALID,
              "cxa_tensordot: axis out of range");
    CXA_CHECK(!isc1[idxl[c]] && !isc2[idxr[c]], CXA_ERR_INVALID,
              "cxa_tensordot: duplicated axis");
    CXA_CHECK(shapeL[idxl[c]] == shapeR[idxr[c]], CXA_ERR_INVALID,
              "cxa_tensordot: contracted dimensions differ (%llu vs %llu)",
              (unsigned long long)shapeL[idxl[c]], (unsigned long long)shapeR[idxr[c]]);
    isc1[idxl[c]] = isc2[idxr[c]] = true;
  }
  int axA[32], axB[32], n1 = 0, n2 = 0;
  int64_t M = 1, K = 1, N = 1, nL = 1, nR = 1;
  for (int i = 0; i < rankL; ++i) {
    nL *= (int64_t)shapeL[i];
    if (!isc1[i]) {
      axA[n1++] = i;
      M *= (int64_t)shapeL[i];
    }
  }
  for (int c = 0; c < ncomm; ++c) {
    axA[n1 + c] = (int)idxl[c];
    K *= (int64_t)shapeL[idxl[c]];
  }
  for (int c = 0; c < ncomm; ++c) axB[c] = (int)idxr[c];
  for (int i = 0; i < rankR; ++i) {
    nR *= (int64_t)shapeR[i];
    if (!isc2[i]) {
      axB[ncomm + n2++] = i;
      N *= (int64_t)shapeR[i];
    }
  }
  CXA_CHECK(M <= INT32_MAX && N <= INT32_MAX && K <= INT32_MAX, CXA_ERR_UNSUPPORTED,
            "cxa_tensordot: matricized dimension above 2^31");
  if (M == 0 || N == 0) return CXA_OK;
  int64_t sam = 0, sak = 0, sbk = 0, sbn = 0;
  const void *A = L, *B = R;
  char *ws = (char *)workspace;
  int64_t used = 0;
  if (!merged_stride(rankL, shapeL, axA, n1, &sam) ||
      !merged_stride(rankL, shapeL, axA + n1, ncomm, &sak)) {
    CXA_CHECK(ws, CXA_ERR_INVALID, "cxa_tensordot: operand L needs a workspace");
    uint64_t mp[16];
    for (int i = 0; i < rankL; ++i) mp[i] = (uint64_t)axA[i];
    int rc = cxa_permute(dtype, rankL, shapeL, mp, L, ws, stream);
    if (rc) return rc;
    A = ws;
    sam = K;
    sak = 1;
    used = (nL + 15) / 16 * 16;
  }
  if (!merged_stride(rankR, shapeR, axB, ncomm, &sbk) ||
      !merged_stride(rankR, shapeR, axB + ncomm, n2, &sbn)) {
    CXA_CHECK(ws, CXA_ERR_INVALID, "cxa_tensordot: operand R needs a workspace");
    uint64_t mp[16];
    for (int i = 0; i < rankR; ++i) mp[i] = (uint64_t)axB[i];
    int rc = cxa_permute(dtype, rankR, shapeR, mp, R, ws + used * es, stream);
    if (rc) return rc;
    B = ws + used * es;
    sbk = N;
    sbn = 1;
  }
  GemmLaunch g;
  g.single = true;
  g.prec = (dtype == CXA_F64 || dtype == CXA_C128) ? CXA_PREC_FP64 : CXA_PREC_FP32;
  g.a_dt = g.b_dt = g.c_dt = dtype;
  g.A = A;
  g.B = B;
  g.C = out;
  g.blk0 = GBlk{0, N, 1.0, 0.0, (int32_t)M, (int32_t)N, 0, K > 0 ? 1 : 0};
  g.seg0 = GSeg{0, 0, sam, sak, sbk, sbn, (int32_t)K, 0};
  return launch_grouped_gemm(g, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------------
// Tensordot that allocates nothing: the operand copies that cxa_tensordot permutes into the
// caller's workspace go to a fixed device array of this library instead (`scratch`, 32 MiB on
// every GPU the library runs on, held until the process ends).  Every call on one GPU shares
// that GPU's array: calls that overlap on one GPU (different streams, or threads) race.  The
// copies overwrite every scratch element the GEMM reads, so nothing needs clearing.
// ------------------------------------------------------------------------------------------
namespace {

constexpr int64_t kScratchBytes = CXA_TENSORDOT_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device ordinal

// The address of `scratch` on the current device, looked up once per device (host code must
// not pass the array's own name to a kernel: on the host it is not a device address).
int scratch_on_current_device(void **out) {
  int dev = -1;
  CXA_HIP(hipGetDevice(&dev));
  CXA_CHECK(dev >= 0 && dev < kMaxDevices, CXA_ERR_UNSUPPORTED,
            "cxa_tensordot_noalloc: device ordinal %d above %d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    CXA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CXA_CHECK(p != nullptr, CXA_ERR_HIP, "cxa_tensordot_noalloc: no device address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CXA_OK;
}

}  // namespace

extern "C" int cxa_tensordot_noalloc(int dtype, void *out, const void *L, int rankL,
                                     const uint64_t *shapeL, const void *R, int rankR,
                                     const uint64_t *shapeR, int ncomm, const uint64_t *idxl,
                                     const uint64_t *idxr, cxa_stream_t stream) {
  CXA_CHECK(dtype == CXA_F64 || dtype == CXA_F32 || dtype == CXA_C128 || dtype == CXA_C64,
            CXA_ERR_UNSUPPORTED, "cxa_tensordot_noalloc: unsupported dtype %d", dtype);
  CXA_CHECK(rankL >= 0 && rankR >= 0 && ncomm >= 1 && idxl && idxr && shapeL && shapeR,
            CXA_ERR_INVALID, "cxa_tensordot_noalloc: bad arguments");
  int64_t need = 0;
  int rc = cxa_tensordot_workspace(dtype, rankL, shapeL, rankR, shapeR, ncomm, idxl, idxr, &need);
  if (rc) return rc;
  CXA_CHECK(need <= kScratchBytes, CXA_ERR_INVALID,
            "cxa_tensordot_noalloc: the permuted operands need %lld bytes, above the %lld-byte "
            "static scratch; use cxa_tensordot with a workspace",
            (long long)need, (long long)kScratchBytes);
  void *ws = nullptr;
  if (need > 0 && (rc = scratch_on_current_device(&ws))) return rc;
  return cxa_tensordot(dtype, out, L, rankL, shapeL, R, rankR, shapeR, ncomm, idxl, idxr, ws,
                       stream);
}
