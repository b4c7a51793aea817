// U(1) block bookkeeping on the host: zero-flux block enumeration.
#include <vector>

#include "cxa_common.hpp"

// Enumerates the qindex tuples whose total U(1) charge is zero, in the reference's odometer
// order (last bond fastest), as BlockUniTensor::Init does (src/BlockUniTensor.cpp:190-230)
// with the flux rule of include/UniTensor.hpp:1139-1166: a KET bond adds its qnum, a BRA bond
// adds the reversed (negated, src/Symmetry.cpp:33-50) qnum.
// The walk prunes on partial sums: once the first j bonds are fixed, the remaining bonds can
// only reach charges in [min_rest, max_rest], so prefixes that cannot close to zero are skipped.
extern "C" int cxa_u1_enumerate_blocks(int rank, const int *btype, const int64_t *nq,
                                       const int64_t *qnums, int64_t max_blocks, uint64_t *itoi,
                                       int64_t *nblocks) {
  CXA_CHECK(rank >= 0 && rank <= 16, CXA_ERR_UNSUPPORTED, "enumerate: rank %d above 16", rank);
  CXA_CHECK(nblocks, CXA_ERR_INVALID, "enumerate: null count pointer");
  *nblocks = 0;
  if (rank == 0) return CXA_OK;
  std::vector<std::vector<int64_t>> q(rank);  // signed charge contribution per qindex
  int64_t base = 0;
  for (int i = 0; i < rank; ++i) {
    CXA_CHECK(btype[i] == -1 || btype[i] == 1, CXA_ERR_INVALID,
              "enumerate: bond %d is neither KET nor BRA (regular bonds carry no charges)", i);
    CXA_CHECK(nq[i] >= 1, CXA_ERR_INVALID, "enumerate: bond %d has no qnums", i);
    q[i].resize(nq[i]);
    for (int64_t k = 0; k < nq[i]; ++k) q[i][k] = btype[i] == -1 ? qnums[base + k] : -qnums[base + k];
    base += nq[i];
  }
  // reachable range of the suffix sums
  std::vector<int64_t> lo(rank + 1, 0), hi(rank + 1, 0);
  for (int i = rank - 1; i >= 0; --i) {
    int64_t mn = q[i][0], mx = q[i][0];
    for (int64_t v : q[i]) {
      mn = v < mn ? v : mn;
      mx = v > mx ? v : mx;
    }
    lo[i] = lo[i + 1] + mn;
    hi[i] = hi[i + 1] + mx;
  }
  std::vector<int64_t> loc(rank, 0), part(rank + 1, 0);
  int depth = 0;
  int64_t count = 0;
  // iterative depth-first walk in lexicographic order
  loc[0] = 0;
  while (depth >= 0) {
    if (loc[depth] >= nq[depth]) {
      --depth;
      if (depth >= 0) ++loc[depth];
      continue;
    }
    part[depth + 1] = part[depth] + q[depth][loc[depth]];
    const int64_t need = -part[depth + 1];
    if (need < lo[depth + 1] || need > hi[depth + 1]) {
      ++loc[depth];
      continue;
    }
    if (depth == rank - 1) {
      if (part[rank] == 0) {
        if (count < max_blocks && itoi)
          for (int i = 0; i < rank; ++i) itoi[count * rank + i] = (uint64_t)loc[i];
        ++count;
      }
      ++loc[depth];
    } else {
      ++depth;
      loc[depth] = 0;
    }
  }
  *nblocks = count;
  return CXA_OK;
}
