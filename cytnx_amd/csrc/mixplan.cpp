// Host planner of the two-site effective Hamiltonian's mix launch: from the block metadata of
// the environment L (permuted to (l, w, l')), the two-site state psi (l, s1, s2, r), the
// output layout Y (l', s1', s2', w'', r in any order, r or l' contiguous) and the dense two-site
// MPO W12 (w, s1, s1', w'', s2, s2'), the task arrays of cxa_mix_contract_plan.  One item per
// (l', r) sector pair of Y; its slices are the products X = L[:, w, :]^T psi[:, s1, s2, :] of
// every (w, s1, s2) element that reaches one of the item's outputs, its outputs the Y
// sub-matrices (s1', s2', w''), each the W12-weighted sum of the slices.  This is the
// contraction order of the reference's two-site network (example/DMRG/dmrg_two_sites_U1.py:
// 18-37), L.psi then the two MPO tensors, with L.psi never stored.
#include <algorithm>
#include <map>
#include <utility>
#include <vector>

#include "cxa_common.hpp"

namespace {

constexpr int kSlices = CXA_MIX_SLICES;

struct Block {
  const int64_t *q, *shape;
  int64_t off;
  int rank;
};

void row_strides(const int64_t *shape, int rank, int64_t *st) {
  st[rank - 1] = 1;
  for (int d = rank - 2; d >= 0; --d) st[d] = st[d + 1] * shape[d + 1];
}

struct Cand {
  int64_t desc[7];
  int64_t w, s1, s2;  // MPO and physical indices of the slice (dense, ungrouped)
  int64_t M, N;
};

struct Out {
  int64_t c_off, ldc;  // within block yi
  int64_t s1p, w2, s2p;
  int yi;
};

struct Record {
  int64_t key;
  int64_t M, N;
  std::vector<const Cand *> cand;
  std::vector<Out> outs;
  std::vector<double> rows;  // nout x kSlices
};

}  // namespace

extern "C" int cxa_heff2_mix_plan(
    int nbL, const int64_t *itoiL, const int64_t *shapeL, const int64_t *offL, const int *posL,
    int nbP, const int64_t *itoiP, const int64_t *shapeP, const int64_t *offP, const int *posP,
    int nbY, const int64_t *itoiY, const int64_t *shapeY, const int64_t *offY, const int *posY,
    const int64_t *sec_off, const int64_t *sec_n, const double *W12, const int64_t *wdims,
    const unsigned char *read, int64_t cap_items, int64_t cap_slices, int64_t cap_outs,
    int64_t *items, int64_t *slices, int64_t *outs, double *coef, unsigned char *live,
    int64_t *counts) {
  CXA_CHECK(nbL >= 0 && nbP >= 0 && nbY >= 0 && posL && posP && posY && sec_off && sec_n &&
                W12 && wdims && items && slices && outs && coef && counts,
            CXA_ERR_INVALID, "cxa_heff2_mix_plan: bad arguments");
  for (int i = 0; i < 6; ++i)
    CXA_CHECK(wdims[i] > 0, CXA_ERR_INVALID, "cxa_heff2_mix_plan: bad W12 dimensions");
  // sector offsets of the six bonds: w (of L), s1, s2 (of psi), s1', s2', w'' (of Y)
  const int64_t *soff[6];
  int64_t sn[6];
  {
    int64_t at = 0;
    for (int b = 0; b < 6; ++b) {
      soff[b] = sec_off + at;
      sn[b] = sec_n[b];
      at += sn[b];
    }
  }
  auto sec = [&](int b, int64_t q) -> int64_t {
    return (q >= 0 && q < sn[b]) ? soff[b][q] : -1;
  };
  const int Ll = posL[0], Lw = posL[1], Llp = posL[2];
  const int Pl = posP[0], Ps1 = posP[1], Ps2 = posP[2], Pr = posP[3];
  const int Ylp = posY[0], Ys1p = posY[1], Ys2p = posY[2], Yw2 = posY[3], Yr = posY[4];
  const bool psi_left = Ylp > Yr;  // l' contiguous: transposed tiles, psi as the left operand
  const int col = psi_left ? Ylp : Yr, row = psi_left ? Yr : Ylp;

  std::map<std::pair<int64_t, int64_t>, std::vector<int>> L_by;
  for (int b = 0; b < nbL; ++b) L_by[{itoiL[3 * b + Ll], itoiL[3 * b + Llp]}].push_back(b);
  std::map<int64_t, std::vector<int>> P_by;
  for (int b = 0; b < nbP; ++b) P_by[itoiP[4 * b + Pr]].push_back(b);
  // Y's (l', r) sector pairs in order of first appearance
  std::map<std::pair<int64_t, int64_t>, int> Y_at;
  std::vector<std::vector<int>> Y_groups;
  std::vector<std::pair<int64_t, int64_t>> Y_keys;
  for (int b = 0; b < nbY; ++b) {
    const std::pair<int64_t, int64_t> key{itoiY[5 * b + Ylp], itoiY[5 * b + Yr]};
    auto it = Y_at.find(key);
    if (it == Y_at.end()) {
      Y_at[key] = (int)Y_groups.size();
      Y_groups.push_back({b});
      Y_keys.push_back(key);
    } else {
      Y_groups[it->second].push_back(b);
    }
  }
  auto wv = [&](int64_t w, int64_t s1, int64_t s1p, int64_t w2, int64_t s2, int64_t s2p) {
    if (w < 0 || s1 < 0 || s1p < 0 || w2 < 0 || s2 < 0 || s2p < 0 || w >= wdims[0] ||
        s1 >= wdims[1] || s1p >= wdims[2] || w2 >= wdims[3] || s2 >= wdims[4] ||
        s2p >= wdims[5])
      return 0.0;
    return W12[((((w * wdims[1] + s1) * wdims[2] + s1p) * wdims[3] + w2) * wdims[4] + s2) *
                   wdims[5] +
               s2p];
  };

  std::vector<unsigned char> keep_blocks(nbY, 0), reached_any(nbY, 0);
  std::vector<Record> records;
  std::vector<std::vector<Cand>> cand_store(Y_groups.size());
  for (size_t gi = 0; gi < Y_groups.size(); ++gi) {
    const int64_t qap = Y_keys[gi].first, qc = Y_keys[gi].second;
    std::vector<Cand> &cand = cand_store[gi];
    auto pit = P_by.find(qc);
    if (pit != P_by.end()) {
      for (int pi : pit->second) {
        const int64_t *tp = itoiP + 4 * pi, *shp = shapeP + 4 * pi;
        int64_t stp[4];
        row_strides(shp, 4, stp);
        const int64_t qa = tp[Pl];
        auto lit = L_by.find({qa, qap});
        if (lit == L_by.end()) continue;
        for (int li : lit->second) {
          const int64_t *tl = itoiL + 3 * li, *shl = shapeL + 3 * li;
          int64_t stl[3];
          row_strides(shl, 3, stl);
          const int64_t K = shl[Ll];
          CXA_CHECK(K == shp[Pl], CXA_ERR_INVALID,
                    "Heff2 mix plan: L and psi disagree on a sector of l");
          for (int64_t iw = 0; iw < shl[Lw]; ++iw)
            for (int64_t i1 = 0; i1 < shp[Ps1]; ++i1)
              for (int64_t i2 = 0; i2 < shp[Ps2]; ++i2) {
                Cand c;
                const int64_t a_l = offL[li] + iw * stl[Lw];
                const int64_t a_p = offP[pi] + i1 * stp[Ps1] + i2 * stp[Ps2];
                if (psi_left) {  // X[r, l'] = psi^T . Lp
                  const int64_t d[7] = {a_p, a_l, stp[Pr], stp[Pl], stl[Ll], stl[Llp], K};
                  std::copy(d, d + 7, c.desc);
                  c.M = shp[Pr];
                  c.N = shl[Llp];
                } else {  // X[l', r] = Lp^T . psi
                  const int64_t d[7] = {a_l, a_p, stl[Llp], stl[Ll], stp[Pl], stp[Pr], K};
                  std::copy(d, d + 7, c.desc);
                  c.M = shl[Llp];
                  c.N = shp[Pr];
                }
                c.w = sec(0, tl[Lw]) + iw;
                c.s1 = sec(1, tp[Ps1]) + i1;
                c.s2 = sec(2, tp[Ps2]) + i2;
                cand.push_back(c);
              }
        }
      }
    }
    std::vector<Out> olist;
    int64_t M = -1, N = -1;
    for (int yi : Y_groups[gi]) {
      const int64_t *ty = itoiY + 5 * yi, *shy = shapeY + 5 * yi;
      int64_t sty[5];
      row_strides(shy, 5, sty);
      CXA_CHECK(sty[col] == 1, CXA_ERR_INVALID,
                "Heff2 mix plan: the %s leg must be the contiguous axis of Y",
                psi_left ? "l'" : "r");
      M = shy[row];
      N = shy[col];
      for (int64_t i1p = 0; i1p < shy[Ys1p]; ++i1p)
        for (int64_t i2p = 0; i2p < shy[Ys2p]; ++i2p)
          for (int64_t j = 0; j < shy[Yw2]; ++j) {
            Out o;
            o.c_off = i1p * sty[Ys1p] + i2p * sty[Ys2p] + j * sty[Yw2];
            o.ldc = sty[row];
            o.s1p = sec(3, ty[Ys1p]) + i1p;
            o.w2 = sec(5, ty[Yw2]) + j;
            o.s2p = sec(4, ty[Ys2p]) + i2p;
            o.yi = yi;
            olist.push_back(o);
          }
    }
    const size_t no = olist.size(), nc = cand.size();
    std::vector<double> C(no * nc);
    for (size_t o = 0; o < no; ++o)
      for (size_t s = 0; s < nc; ++s)
        C[o * nc + s] = wv(cand[s].w, cand[s].s1, olist[o].s1p, olist[o].w2, cand[s].s2,
                           olist[o].s2p);
    std::vector<unsigned char> reached(no, 0);
    for (size_t o = 0; o < no; ++o)
      for (size_t s = 0; s < nc && !reached[o]; ++s) reached[o] = C[o * nc + s] != 0.0;
    // a block is reached when any of its outputs is
    std::map<int, bool> block_reached;
    for (size_t o = 0; o < no; ++o)
      if (reached[o]) block_reached[olist[o].yi] = true;
    std::vector<size_t> kept;
    if (read) {
      for (size_t o = 0; o < no; ++o)
        if (block_reached.count(olist[o].yi) && read[olist[o].yi]) kept.push_back(o);
      for (auto &kv : block_reached)
        if (read[kv.first]) keep_blocks[kv.first] = 1;
      if (kept.empty()) continue;
    } else {
      for (auto &kv : block_reached) reached_any[kv.first] = 1;
      for (size_t o = 0; o < no; ++o) kept.push_back(o);
    }
    // slices that some kept output uses
    std::vector<size_t> used;
    for (size_t s = 0; s < nc; ++s) {
      bool any = false;
      for (size_t o : kept) any = any || C[o * nc + s] != 0.0;
      if (any) used.push_back(s);
    }
    CXA_CHECK((int)used.size() <= kSlices, CXA_ERR_UNSUPPORTED,
              "an item of the two-site operator has %d slices (at most %d)", (int)used.size(),
              kSlices);
    Record rec;
    rec.M = M;
    rec.N = N;
    int64_t key = -1;
    for (size_t s : used) {
      CXA_CHECK(cand[s].M == M && cand[s].N == N, CXA_ERR_INVALID,
                "Heff2 mix plan: slice shape differs from its outputs'");
      rec.cand.push_back(&cand[s]);
    }
    for (size_t o : kept) {
      rec.outs.push_back(olist[o]);
      for (int q = 0; q < kSlices; ++q)
        rec.rows.push_back(q < (int)used.size() ? C[o * nc + used[q]] : 0.0);
      key = key < 0 ? olist[o].yi : std::min<int64_t>(key, olist[o].yi);
    }
    rec.key = kept.empty() ? (int64_t)records.size() : key;
    records.push_back(std::move(rec));
  }
  // with a read mask: items ordered by the first block of the pruned layout they write
  if (read)
    std::stable_sort(records.begin(), records.end(),
                     [](const Record &a, const Record &b) { return a.key < b.key; });
  // element offsets of the layout the outputs address
  std::vector<int64_t> base(nbY, 0);
  if (read) {
    int64_t acc = 0;
    for (int b = 0; b < nbY; ++b) {
      base[b] = acc;
      if (keep_blocks[b]) {
        int64_t sz = 1;
        for (int d = 0; d < 5; ++d) sz *= shapeY[5 * b + d];
        acc += sz;
      }
    }
  } else {
    CXA_CHECK(offY != nullptr, CXA_ERR_INVALID, "cxa_heff2_mix_plan: Y offsets missing");
    for (int b = 0; b < nbY; ++b) base[b] = offY[b];
  }
  int64_t ni = 0, ns = 0, nout = 0;
  for (const Record &r : records) {
    CXA_CHECK(ni < cap_items && ns + (int64_t)r.cand.size() <= cap_slices &&
                  nout + (int64_t)r.outs.size() <= cap_outs,
              CXA_ERR_INVALID, "cxa_heff2_mix_plan: output arrays too small");
    int64_t *it = items + 4 * ni;
    it[0] = r.M;
    it[1] = r.N;
    it[2] = (int64_t)r.cand.size();
    it[3] = (int64_t)r.outs.size();
    ++ni;
    for (const Cand *c : r.cand) {
      std::copy(c->desc, c->desc + 7, slices + 7 * ns);
      ++ns;
    }
    for (size_t o = 0; o < r.outs.size(); ++o) {
      outs[2 * nout] = base[r.outs[o].yi] + r.outs[o].c_off;
      outs[2 * nout + 1] = r.outs[o].ldc;
      std::copy(r.rows.begin() + o * kSlices, r.rows.begin() + (o + 1) * kSlices,
                coef + kSlices * nout);
      ++nout;
    }
  }
  if (live)
    for (int b = 0; b < nbY; ++b) live[b] = read ? keep_blocks[b] : reached_any[b];
  counts[0] = ni;
  counts[1] = ns;
  counts[2] = nout;
  return CXA_OK;
}
