// Pivots (zero-, one- and two-site): combined local basis, block-structure
// precalculation, per-term expansion and the staged plan; see pivot.hpp.
#include "pivot.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

#include "errors.hpp"
#include "par.hpp"

namespace vmps {

// ---------------------------------------------------------------------------
// TwoSiteBasis: Qbasis::pullData(qloc) (Qbasis.h:580-590 with push_back
// :228-254), Qbasis::combine (:686-798), outer_num (:378-388), inner_num
// (:335-350) and leftOffset (:440-462).
// ---------------------------------------------------------------------------
namespace {
struct PlainEntry {
  Qn q;
  int64_t num;
  int64_t plain;
};

void pull_data(const std::vector<Qn>& qs, std::vector<PlainEntry>& entries, std::vector<int64_t>& inner) {
  entries.clear();
  int64_t curr = 0;
  for (const Qn& q : qs) {
    auto it = std::find_if(entries.begin(), entries.end(), [&](const PlainEntry& e) { return e.q == q; });
    if (it == entries.end()) {
      entries.push_back({q, curr, 1});
    } else {
      // the reference appends to the existing sector without shifting the
      // offsets of later sectors (Qbasis.h:243-251)
      it->plain += 1;
    }
    curr += 1;
  }
  inner.assign(qs.size(), 0);
  for (size_t s = 0; s < qs.size(); ++s) {
    bool found = false;
    for (const auto& e : entries) {
      for (int64_t i = 0; i < e.plain && !found; ++i) {
        if ((int64_t)s == e.num + i) {
          inner[s] = i;
          found = true;
        }
      }
      if (found) break;
    }
  }
}
}  // namespace

void TwoSiteBasis::build(const Sym& sym, const std::vector<Qn>& qloc12, const std::vector<Qn>& qloc34) {
  q12_ = qloc12;
  q34_ = qloc34;
  std::vector<PlainEntry> e12, e34;
  pull_data(qloc12, e12, inner12_);
  pull_data(qloc34, e34, inner34_);
  plain12_.clear();
  plain34_.clear();
  for (auto& e : e12) plain12_[e.q] = e.plain;
  for (auto& e : e34) plain34_[e.q] = e.plain;

  history_.clear();
  for (const auto& a : e12)
    for (const auto& b : e34)
      for (const Qn& q : sym.reduce(a.q, b.q)) {
        auto& h = history_[q];
        bool have = false;
        for (const auto& src : h)
          if (src.q1 == a.q && src.q2 == b.q) have = true;
        if (!have) h.push_back({a.q, b.q, a.plain, b.plain});
      }
  for (auto& kv : history_)
    std::stable_sort(kv.second.begin(), kv.second.end(), [](const Src& x, const Src& y) {
      if (x.q1 != y.q1) return x.q1 < y.q1;
      return x.q2 < y.q2;
    });
  outer_.clear();
  qloc.clear();
  int64_t num = 0;
  for (const auto& kv : history_) {  // std::map iterates in sorted (Qbasis::sort) order
    outer_[kv.first] = num;
    int64_t d = 0;
    for (const auto& src : kv.second) d += src.dim1 * src.dim2;
    for (int64_t i = 0; i < d; ++i) qloc.push_back(kv.first);
    num += d;
  }
}

int64_t TwoSiteBasis::index(int s1, int s3, const Qn& qmerge) const {
  auto ho = history_.find(qmerge);
  if (ho == history_.end()) return -1;
  int64_t out = outer_.at(qmerge);
  const Qn& q1 = q12_[s1];
  const Qn& q2 = q34_[s3];
  for (const auto& src : ho->second) {
    if (!(src.q1 == q1 && src.q2 == q2)) {
      out += src.dim1 * src.dim2;
    } else {
      int64_t p1 = inner12_[s1], p2 = inner34_[s3];
      for (int64_t j = 0; j < src.dim1 * src.dim2; ++j) {
        int64_t t = j / src.dim1;
        if (!(j - src.dim1 * t == p1 && t == p2)) {
          out += 1;
        } else {
          break;
        }
      }
      return out;
    }
  }
  return -1;
}

// ---------------------------------------------------------------------------
size_t Pivot::add_term(const std::vector<Qn>& qop12, const std::vector<Qn>& qop34) {
  PivotTerm t;
  t.qop12 = qop12;
  t.qop34 = qop34;
  if (nsite >= 1)
    t.W12.assign(qloc12.size(), std::vector<std::vector<WBiped>>(qloc12.size(), std::vector<WBiped>(qop12.size())));
  if (nsite == 2)
    t.W34.assign(qloc34.size(), std::vector<std::vector<WBiped>>(qloc34.size(), std::vector<WBiped>(qop34.size())));
  terms.push_back(std::move(t));
  cur = terms.size() - 1;
  return cur;
}

int Pivot::add_V_block(int s, const Qn& qin, const Qn& qout, int64_t rows, int64_t cols) {
  if (s < 0 || s >= (int)V.in.size()) throw VmpsError(VMPS_ERR_ARG, "V block: s out of range");
  if (rows < 0 || cols < 0) throw VmpsError(VMPS_ERR_ARG, "V block: negative dimension");
  Key2 key{qin, qout};
  if (V.dict[s].count(key)) throw VmpsError(VMPS_ERR_LAYOUT, "V block: duplicate (qin,qout) in one s");
  DenseBlock b;
  b.rows = rows;
  b.cols = cols;
  b.off = V.size;
  V.size += rows * cols;
  V.in[s].push_back(qin);
  V.out[s].push_back(qout);
  V.block[s].push_back(b);
  V.dict[s][key] = V.in[s].size() - 1;
  V.nblocks += 1;
  return (int)V.in[s].size() - 1;
}

// ---------------------------------------------------------------------------
// AAWWAA, DmrgIndexGymnastics.h:198-262
// ---------------------------------------------------------------------------
namespace {
struct AawwaaHit {
  Qn qRin, qRout, qRmid, qW;
  size_t q13, q24, qW12, qW34;
};

bool AAWWAA(const Sym& sym, const Qn& Lin, const Qn& Lout, const Qn& Lmid, const Qn& qOp12, const Qn& qOp34,
            const Qn& qmerge13, const Qn& qmerge24, const VLayout& V, size_t s13, size_t s24, const WBiped& W12,
            const WBiped& W34, std::vector<AawwaaHit>& result) {
  result.clear();
  auto qRouts = sym.reduce(Lin, qmerge13);
  auto qRins = sym.reduce(Lout, qmerge24);
  auto qWs = sym.reduce(Lmid, qOp12);
  for (const Qn& qRout : qRouts) {
    auto q13 = V.dict[s13].find(Key2{Lin, qRout});
    if (q13 == V.dict[s13].end()) continue;
    for (const Qn& qRin : qRins) {
      auto q24 = V.dict[s24].find(Key2{Lout, qRin});
      if (q24 == V.dict[s24].end()) continue;
      for (const Qn& qW : qWs) {
        auto w12 = W12.dict.find(Key2{Lmid, qW});
        if (w12 == W12.dict.end()) continue;
        for (const Qn& qRmid : sym.reduce(qW, qOp34)) {
          auto w34 = W34.dict.find(Key2{qW, qRmid});
          if (w34 == W34.dict.end()) continue;
          if (sym.validate3(qRin, qRmid, qRout))
            result.push_back({qRin, qRout, qRmid, qW, q13->second, q24->second, w12->second, w34->second});
        }
      }
    }
  }
  return !result.empty();
}
}  // namespace

namespace {
// LAWA, DmrgIndexGymnastics.h:32-77, on the wavefunction layout (Abra = Aket = V):
// right sectors (Rin, Rout, Rmid) reachable from the left sector (Lin, Lout, Lmid)
// through Vout[s1] (Lin, Rout), Vin[s2] (Lout, Rin) and W[s1][s2][k] (Lmid, Rmid).
struct LawaHit {
  Qn qRin, qRout, qRmid;
  size_t q1, q2, qW;
};

void LAWA_V(const Sym& sym, const Qn& Lin, const Qn& Lout, const Qn& Lmid, const Qn& qloc1, const Qn& qloc2,
            const Qn& qOp, const VLayout& V, size_t s1, size_t s2, const WBiped& W, std::vector<LawaHit>& result) {
  result.clear();
  for (const Qn& Rout : sym.reduce(Lin, qloc1)) {
    auto q1 = V.dict[s1].find(Key2{Lin, Rout});
    if (q1 == V.dict[s1].end()) continue;
    for (const Qn& Rin : sym.reduce(Lout, qloc2)) {
      auto q2 = V.dict[s2].find(Key2{Lout, Rin});
      if (q2 == V.dict[s2].end()) continue;
      for (const Qn& Rmid : sym.reduce(Lmid, qOp)) {
        auto qW = W.dict.find(Key2{Lmid, Rmid});
        if (qW == W.dict.end()) continue;
        if (sym.validate3(Rin, Rmid, Rout)) result.push_back({Rin, Rout, Rmid, q1->second, q2->second, qW->second});
      }
    }
  }
}

// One-site precalc_blockStructure, DmrgIndexGymnastics.h:485-590.
void precalc_one_site(Pivot& P, PivotTerm& T) {
  const Sym& sym = P.sym;
  const bool NA = sym.non_abelian();
  if (P.V.in.size() != P.qloc12.size())
    throw VmpsError(VMPS_ERR_LAYOUT, "PivotVector has " + std::to_string(P.V.in.size()) +
                                         " components, the local basis has " + std::to_string(P.qloc12.size()));
  using Key = std::array<size_t, 2>;
  std::map<Key, std::pair<std::vector<std::array<size_t, 12>>, std::vector<double>>> lookup;
  std::vector<LawaHit> ix;
  for (size_t s1 = 0; s1 < P.qloc12.size(); ++s1)
    for (size_t s2 = 0; s2 < P.qloc12.size(); ++s2)
      for (size_t k = 0; k < T.qop12.size(); ++k) {
        if (!sym.validate3(P.qloc12[s2], T.qop12[k], P.qloc12[s1])) continue;
        for (size_t qL = 0; qL < T.L.dim(); ++qL) {
          LAWA_V(sym, T.L.in(qL), T.L.out(qL), T.L.mid(qL), P.qloc12[s1], P.qloc12[s2], T.qop12[k], P.V, s1, s2,
                 T.W12[s1][s2][k], ix);
          for (const LawaHit& h : ix) {
            auto qR = T.R.dict.find(Key3{h.qRin, h.qRout, h.qRmid});
            if (qR == T.R.dict.end()) continue;
            // skip if no nonzero of W connects a present L block with a present R block
            const SparseBlock& w = T.W12[s1][s2][k].block[h.qW];
            bool any = false;
            for (size_t i = 0; i < w.val.size() && !any; ++i) {
              if (w.row[i] >= (int)T.L.block[qL].size() || w.col[i] >= (int)T.R.block[qR->second].size())
                throw VmpsError(VMPS_ERR_LAYOUT, "MPO index outside the environment's mid sector");
              any = T.L.block[qL][w.row[i]].present() && T.R.block[qR->second][w.col[i]].present();
            }
            if (!any) continue;
            double f = NA ? sym.coeff_HPsi(P.V.in[s2][h.q2], P.qloc12[s2], P.V.out[s2][h.q2], T.L.mid(qL), T.qop12[k],
                                           h.qRmid, P.V.in[s1][h.q1], P.qloc12[s1], P.V.out[s1][h.q1])
                          : 1.0;
            if (std::abs(f) < kCgcEps) continue;
            auto& e = lookup[Key{s1, h.q1}];
            e.first.push_back({s2, h.q2, qL, qR->second, k, h.qW, 0, 0, 0, 0, 0, 0});
            e.second.push_back(f);
          }
        }
      }
  T.qlhs.clear();
  T.qrhs.clear();
  T.factor_cgcs.clear();
  for (auto& kv : lookup) {
    T.qlhs.push_back(kv.first);
    T.qrhs.push_back(std::move(kv.second.first));
    T.factor_cgcs.push_back(std::move(kv.second.second));
  }
  T.has_blockstructure = true;
}
}  // namespace

void precalc_blockStructure(Pivot& P) {
  for (auto& t : P.terms) precalc_blockStructure(P, t);
}

void precalc_blockStructure(Pivot& P, PivotTerm& T) {
  if (P.nsite == 0) {
    T.qlhs.clear();
    T.qrhs.clear();
    T.factor_cgcs.clear();
    T.has_blockstructure = true;
    return;
  }
  if (P.nsite == 1) {
    precalc_one_site(P, T);
    return;
  }
  const Sym& sym = P.sym;
  const bool NA = sym.non_abelian();
  TwoSiteBasis tb;
  tb.build(sym, P.qloc12, P.qloc34);
  if ((int64_t)tb.qloc.size() != (int64_t)P.V.in.size())
    throw VmpsError(VMPS_ERR_LAYOUT, "PivotVector has " + std::to_string(P.V.in.size()) +
                                         " two-site components, the combined local basis has " +
                                         std::to_string(tb.qloc.size()));

  using Key = std::array<size_t, 2>;
  std::map<Key, std::pair<std::vector<std::array<size_t, 12>>, std::vector<double>>> lookup;
  std::vector<AawwaaHit> ixs;
  // Necessary conditions of an AAWWAA match, tested before the call: Vin[s1s3]
  // holds a block with in = L.in, Vin[s2s4] one with in = L.out, and W12 one
  // with in = L.mid.  They only skip qL that would match nothing, so the
  // tables and their order are unchanged.
  // The L sectors are bucketed by their mid quantum number, so a W12 block
  // list selects its candidates directly (visited in ascending qL order).
  std::vector<std::unordered_set<Qn, QnHash>> vin(P.V.in.size());
  for (size_t s = 0; s < P.V.in.size(); ++s) vin[s].insert(P.V.in[s].begin(), P.V.in[s].end());
  std::unordered_map<Qn, std::vector<size_t>, QnHash> by_mid;
  for (size_t qL = 0; qL < T.L.dim(); ++qL) by_mid[T.L.mid(qL)].push_back(qL);
  std::vector<size_t> cand;

  const auto& qop12 = T.qop12;
  const auto& qop34 = T.qop34;
  for (size_t s1 = 0; s1 < P.qloc12.size(); ++s1)
    for (size_t s2 = 0; s2 < P.qloc12.size(); ++s2)
      for (size_t k12 = 0; k12 < qop12.size(); ++k12) {
        if (!sym.validate3(P.qloc12[s2], qop12[k12], P.qloc12[s1])) continue;
        for (size_t s3 = 0; s3 < P.qloc34.size(); ++s3)
          for (size_t s4 = 0; s4 < P.qloc34.size(); ++s4)
            for (size_t k34 = 0; k34 < qop34.size(); ++k34) {
              if (!sym.validate3(P.qloc34[s4], qop34[k34], P.qloc34[s3])) continue;
              auto qOps = sym.reduce(qop12[k12], qop34[k34]);
              for (const Qn& qOp : qOps) {
                auto qmerges13 = sym.reduce(P.qloc12[s1], P.qloc34[s3]);
                auto qmerges24 = sym.reduce(P.qloc12[s2], P.qloc34[s4]);
                for (const Qn& qmerge13 : qmerges13)
                  for (const Qn& qmerge24 : qmerges24) {
                    int64_t s1s3 = tb.index((int)s1, (int)s3, qmerge13);
                    int64_t s2s4 = tb.index((int)s2, (int)s4, qmerge24);
                    if (s1s3 < 0 || s2s4 < 0) continue;
                    double cgc9 = NA ? sym.coeff_tensorProd(P.qloc12[s2], P.qloc34[s4], qmerge24, qop12[k12],
                                                            qop34[k34], qOp, P.qloc12[s1], P.qloc34[s3], qmerge13)
                                     : 1.0;
                    if (std::abs(cgc9) < kCgcEps) continue;
                    const WBiped& w12 = T.W12[s1][s2][k12];
                    cand.clear();
                    for (size_t b = 0; b < w12.in.size(); ++b) {
                      if (std::find(w12.in.begin(), w12.in.begin() + b, w12.in[b]) != w12.in.begin() + b) continue;
                      auto it = by_mid.find(w12.in[b]);
                      if (it != by_mid.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
                    }
                    std::sort(cand.begin(), cand.end());
                    for (size_t qL : cand) {
                      if (!vin[s1s3].count(T.L.in(qL)) || !vin[s2s4].count(T.L.out(qL))) continue;
                      bool found = AAWWAA(sym, T.L.in(qL), T.L.out(qL), T.L.mid(qL), qop12[k12], qop34[k34],
                                          qmerge13, qmerge24, P.V, (size_t)s1s3, (size_t)s2s4, T.W12[s1][s2][k12],
                                          T.W34[s3][s4][k34], ixs);
                      if (!found) continue;
                      for (const auto& ix : ixs) {
                        auto qR = T.R.dict.find(Key3{ix.qRin, ix.qRout, ix.qRmid});
                        double cgc6 = NA ? sym.coeff_Apair(T.L.mid(qL), qop12[k12], ix.qW, qop34[k34], ix.qRmid, qOp)
                                         : 1.0;
                        if (std::abs(cgc6) < kCgcEps) continue;
                        if (qR == T.R.dict.end()) continue;
                        double cgcH =
                            NA ? sym.coeff_HPsi(P.V.in[s2s4][ix.q24], qmerge24, P.V.out[s2s4][ix.q24], T.L.mid(qL), qOp,
                                                T.R.mid(qR->second), P.V.in[s1s3][ix.q13], qmerge13,
                                                P.V.out[s1s3][ix.q13])
                               : 1.0;
                        Key key = {(size_t)s1s3, ix.q13};
                        std::array<size_t, 12> val = {(size_t)s2s4, ix.q24, qL, qR->second, s1, s2, k12, ix.qW12,
                                                      s3, s4, k34, ix.qW34};
                        auto& e = lookup[key];
                        e.first.push_back(val);
                        e.second.push_back(cgc6 * cgc9 * cgcH);
                      }
                    }
                  }
              }
            }
      }
  T.qlhs.clear();
  T.qrhs.clear();
  T.factor_cgcs.clear();
  for (auto& kv : lookup) {
    T.qlhs.push_back(kv.first);
    T.qrhs.push_back(std::move(kv.second.first));
    T.factor_cgcs.push_back(std::move(kv.second.second));
  }
  T.has_blockstructure = true;
}

// ---------------------------------------------------------------------------
// Per-term expansion: the loop nests of OxV for two sites
// (DmrgPivotMatrix2.h:182-230), one site (DmrgPivotMatrix1.h:218-262) and
// zero sites (HxV, DmrgPivotMatrix0.h:31-96), over every term of the sum.
// ---------------------------------------------------------------------------
namespace {
struct TermSink {
  std::vector<Term>& terms;
  std::vector<EnvRef>& lblocks;
  std::vector<EnvRef>& rblocks;
  std::map<EnvRef, int64_t> lid, rid;

  static int64_t id(std::map<EnvRef, int64_t>& m, std::vector<EnvRef>& list, const EnvRef& k) {
    auto it = m.find(k);
    if (it == m.end()) {
      it = m.emplace(k, (int64_t)list.size()).first;
      list.push_back(k);
    }
    return it->second;
  }
  void add(int64_t out_v, const EnvRef& l, int64_t in_v, const EnvRef& r, double c) {
    if (c == 0.0) return;  // exact zero contributes nothing
    terms.push_back({out_v, id(lid, lblocks, l), in_v, id(rid, rblocks, r), c});
  }
};

const DenseBlock& env_block(const Tripod& E, size_t q, int a) {
  if (a < 0 || a >= (int)E.block[q].size())
    throw VmpsError(VMPS_ERR_LAYOUT, "MPO index outside the environment's mid sector");
  return E.block[q][a];
}
}  // namespace

void expand_terms(const Pivot& P, std::vector<Term>& terms, std::vector<EnvRef>& lblocks,
                  std::vector<EnvRef>& rblocks, std::vector<std::pair<int, int>>& vblocks) {
  // global ids of V blocks in packed (s,q) order
  std::vector<std::vector<int64_t>> vid(P.V.in.size());
  vblocks.clear();
  for (size_t s = 0; s < P.V.in.size(); ++s) {
    vid[s].resize(P.V.in[s].size());
    for (size_t q = 0; q < P.V.in[s].size(); ++q) {
      vid[s][q] = (int64_t)vblocks.size();
      vblocks.push_back({(int)s, (int)q});
    }
  }
  lblocks.clear();
  rblocks.clear();
  terms.clear();
  TermSink sink{terms, lblocks, rblocks, {}, {}};
  for (size_t t = 0; t < P.terms.size(); ++t) {
    const PivotTerm& T = P.terms[t];
    if (P.nsite == 0) {
      // Vout[0](qL.in, qL.in) += sum_a L[qL][a] . Vin[0](qL.out, qL.out) . R[{out,in,mid}][a]
      if (P.V.in.size() != 1) throw VmpsError(VMPS_ERR_LAYOUT, "zero-site pivot vectors have one component");
      for (size_t qL = 0; qL < T.L.dim(); ++qL) {
        auto qR = T.R.dict.find(Key3{T.L.out(qL), T.L.in(qL), T.L.mid(qL)});
        if (qR == T.R.dict.end()) continue;
        auto qAin = P.V.dict[0].find(Key2{T.L.out(qL), T.L.out(qL)});
        if (qAin == P.V.dict[0].end()) continue;
        const Qn& qo = T.R.out(qR->second);
        auto qAout = P.V.dict[0].find(Key2{qo, qo});
        if (qAout == P.V.dict[0].end()) continue;
        if (T.L.block[qL].size() != T.R.block[qR->second].size())
          throw VmpsError(VMPS_ERR_LAYOUT, "zero-site pivot: L and R sectors differ in their mid dimension");
        if (!P.V.block[0][qAin->second].present()) continue;
        for (size_t a = 0; a < T.L.block[qL].size(); ++a)
          if (T.L.block[qL][a].present() && T.R.block[qR->second][a].present())
            sink.add(vid[0][qAout->second], {t, qL, a}, vid[0][qAin->second], {t, qR->second, a}, 1.0);
      }
      continue;
    }
    if (!T.has_blockstructure) throw VmpsError(VMPS_ERR_STATE, "pivot has no block structure (call precalc first)");
    for (size_t q = 0; q < T.qlhs.size(); ++q) {
      size_t so = T.qlhs[q][0], qo = T.qlhs[q][1];
      if (so >= P.V.in.size() || qo >= P.V.in[so].size())
        throw VmpsError(VMPS_ERR_LAYOUT, "qlhs entry outside the vector layout");
      for (size_t p = 0; p < T.qrhs[q].size(); ++p) {
        const auto& e = T.qrhs[q][p];
        const double f = T.factor_cgcs[q][p];
        size_t si = e[0], qi = e[1], qL = e[2], qR = e[3];
        if (si >= P.V.in.size() || qi >= P.V.in[si].size() || qL >= T.L.dim() || qR >= T.R.dim())
          throw VmpsError(VMPS_ERR_LAYOUT, "qrhs entry outside the pivot");
        const bool vin_present = P.V.block[si][qi].present();
        if (P.nsite == 1) {
          size_t k = e[4], qW = e[5];
          if (k >= T.W12[so][si].size() || qW >= T.W12[so][si][k].block.size())
            throw VmpsError(VMPS_ERR_LAYOUT, "qrhs entry outside the pivot");
          const SparseBlock& w = T.W12[so][si][k].block[qW];
          for (size_t i = 0; i < w.val.size(); ++i) {
            int a = w.row[i], b = w.col[i];
            if (!(env_block(T.L, qL, a).present() && env_block(T.R, qR, b).present() && vin_present)) continue;
            sink.add(vid[so][qo], {t, qL, (size_t)a}, vid[si][qi], {t, qR, (size_t)b}, f * w.val[i]);
          }
          continue;
        }
        size_t s1 = e[4], s2 = e[5], k12 = e[6], qW12 = e[7], s3 = e[8], s4 = e[9], k34 = e[10], qW34 = e[11];
        if (s1 >= T.W12.size() || s2 >= T.W12[s1].size() || k12 >= T.W12[s1][s2].size() ||
            qW12 >= T.W12[s1][s2][k12].block.size() || s3 >= T.W34.size() || s4 >= T.W34[s3].size() ||
            k34 >= T.W34[s3][s4].size() || qW34 >= T.W34[s3][s4][k34].block.size())
          throw VmpsError(VMPS_ERR_LAYOUT, "qrhs entry outside the pivot");
        const SparseBlock& w12 = T.W12[s1][s2][k12].block[qW12];
        const SparseBlock& w34 = T.W34[s3][s4][k34].block[qW34];
        for (size_t i = 0; i < w12.val.size(); ++i)
          for (size_t j = 0; j < w34.val.size(); ++j) {
            int a = w12.row[i], w = w12.col[i], w2 = w34.row[j], b = w34.col[j];
            if (!(env_block(T.L, qL, a).present() && env_block(T.R, qR, b).present() && vin_present && w == w2))
              continue;
            sink.add(vid[so][qo], {t, qL, (size_t)a}, vid[si][qi], {t, qR, (size_t)b}, f * w12.val[i] * w34.val[j]);
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Every device access of the three stages stays inside its buffer: checked on
// the host before anything is uploaded.
namespace {
void check_gemm(const std::vector<GemmProb>& probs, const std::vector<GemmChunk>& chunks, int64_t a_size,
                int64_t b_size, int64_t c_size, const char* what) {
  for (const GemmProb& p : probs) {
    if (p.M <= 0 || p.N <= 0 || p.c_off < 0 || p.c_off + (int64_t)p.M * p.N > c_size || p.chunk_begin < 0 ||
        p.chunk_end < p.chunk_begin || p.chunk_end > (int32_t)chunks.size())
      throw VmpsError(VMPS_ERR_INTERNAL, std::string(what) + ": output descriptor out of range");
    for (int32_t c = p.chunk_begin; c < p.chunk_end; ++c) {
      const GemmChunk& ch = chunks[c];
      if (ch.K <= 0 || ch.lda < p.M || ch.ldb < ch.K || ch.a_off < 0 || ch.b_off < 0 ||
          ch.a_off + (int64_t)ch.lda * (ch.K - 1) + p.M > a_size ||
          ch.b_off + (int64_t)ch.ldb * (p.N - 1) + ch.K > b_size)
        throw VmpsError(VMPS_ERR_INTERNAL, std::string(what) + ": operand descriptor out of range");
    }
  }
}
}  // namespace

void validate_host_plan(const HostPlan& H) {
  check_gemm(H.probA, H.chunkA, H.L_size, H.ysize, H.T_size, "stage A");
  check_gemm(H.probC, H.chunkC, H.U_size, H.R_size, H.osize, "stage C");
  for (const CombineList& cl : H.comb) {
    if (cl.u_off < 0 || cl.size <= 0 || cl.u_off + cl.size > H.U_size || cl.begin < 0 || cl.end < cl.begin ||
        cl.end > (int32_t)H.comb_t.size() || cl.size > INT32_MAX)
      throw VmpsError(VMPS_ERR_INTERNAL, "stage B: list descriptor out of range");
    for (int32_t j = cl.begin; j < cl.end; ++j) {
      int32_t t = H.comb_t[j];
      if (t < 0 || t >= (int32_t)H.T_off.size() || H.T_off[t] + cl.size > H.T_size)
        throw VmpsError(VMPS_ERR_INTERNAL, "stage B: T reference out of range");
    }
  }
  for (size_t i = 0; i < H.seg_off.size(); ++i)
    if (H.seg_off[i] < 0 || H.seg_off[i] + H.seg_len[i] > H.osize)
      throw VmpsError(VMPS_ERR_INTERNAL, "dot segment out of range");
}

// Flatten the stage tables into per-tile work lists.  Stage B (U = sum c T)
// is folded into stage C: each stage C chunk carries the list of (T, c) that
// forms its U operand; a chunk with more than kMaxTermsPerChunk terms is split
// into several chunks on the same R block (U . R is linear in U).  Tiles are
// ordered by decreasing work so the largest start first.
namespace {
// select: 0 = every problem, > 0 = problems with both dimensions >= select,
// < 0 = problems with a dimension below -select
bool selected(int select, int64_t M, int64_t N) {
  if (select == 0) return true;
  const bool big = M >= std::abs(select) && N >= std::abs(select);
  return select > 0 ? big : !big;
}

void make_tiles(const HostPlan& H, int edge, const std::vector<std::pair<int32_t, int32_t>>& crange,
                std::vector<TileA>& outA, std::vector<TileC>& outC, int select = 0) {
  outA.clear();
  outC.clear();
  std::vector<std::pair<int64_t, size_t>> order;
  for (size_t p = 0; p < H.probA.size(); ++p) {
    const GemmProb& pr = H.probA[p];
    if (!selected(select, pr.M, pr.N)) continue;
    const GemmChunk& ch = H.chunkA[pr.chunk_begin];
    for (int i = 0; i * edge < pr.M; ++i)
      for (int j = 0; j * edge < pr.N; ++j) {
        if (i > 32767 || j > 32767) throw VmpsError(VMPS_ERR_ARG, "block dimension too large for the tile index");
        TileA t;
        t.a_off = ch.a_off;
        t.b_off = ch.b_off;
        t.c_off = pr.c_off;
        t.M = pr.M;
        t.N = pr.N;
        t.K = ch.K;
        t.mt = (int16_t)i;
        t.nt = (int16_t)j;
        outA.push_back(t);
      }
  }
  for (size_t i = 0; i < outA.size(); ++i) order.push_back({-(int64_t)outA[i].K, i});
  std::stable_sort(order.begin(), order.end());
  {
    std::vector<TileA> sorted;
    for (auto& o : order) sorted.push_back(outA[o.second]);
    outA.swap(sorted);
  }
  std::vector<int64_t> cost;
  for (size_t p = 0; p < H.probC.size(); ++p) {
    const GemmProb& pr = H.probC[p];
    if (!selected(select, pr.M, pr.N)) continue;
    int64_t w = 0;
    for (int32_t c = crange[p].first; c < crange[p].second; ++c)
      w += (int64_t)H.chunksC[c].K * (1 + H.chunksC[c].term_end - H.chunksC[c].term_begin);
    for (int i = 0; i * edge < pr.M; ++i)
      for (int j = 0; j * edge < pr.N; ++j) {
        if (i > 32767 || j > 32767) throw VmpsError(VMPS_ERR_ARG, "block dimension too large for the tile index");
        TileC t;
        t.c_off = pr.c_off;
        t.M = pr.M;
        t.N = pr.N;
        t.chunk_begin = crange[p].first;
        t.chunk_end = crange[p].second;
        t.mt = (int16_t)i;
        t.nt = (int16_t)j;
        t.pad = 0;
        outC.push_back(t);
        cost.push_back(w);
      }
  }
  order.clear();
  for (size_t i = 0; i < outC.size(); ++i) order.push_back({-cost[i], i});
  std::stable_sort(order.begin(), order.end());
  {
    std::vector<TileC> sorted;
    for (auto& o : order) sorted.push_back(outC[o.second]);
    outC.swap(sorted);
  }
}

}  // namespace

std::vector<int64_t> xcd_order(const std::vector<int64_t>& group, const std::vector<int64_t>& cost, int per) {
  const size_t n = group.size();
  // supertiles in order of first appearance, with their items in list order
  std::unordered_map<int64_t, size_t> gid;
  std::vector<std::vector<int64_t>> items;
  std::vector<int64_t> gcost;
  for (size_t i = 0; i < n; ++i) {
    auto it = gid.find(group[i]);
    if (it == gid.end()) {
      it = gid.emplace(group[i], items.size()).first;
      items.emplace_back();
      gcost.push_back(0);
    }
    items[it->second].push_back((int64_t)i);
    gcost[it->second] += cost[i];
  }
  std::vector<size_t> gorder(items.size());
  for (size_t g = 0; g < gorder.size(); ++g) gorder[g] = g;
  std::stable_sort(gorder.begin(), gorder.end(), [&](size_t a, size_t b) { return gcost[a] > gcost[b]; });
  // largest supertile first onto the least loaded XCD
  std::vector<std::vector<int64_t>> queue(kXcdCount);
  std::vector<int64_t> load(kXcdCount, 0);
  for (size_t g : gorder) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    queue[x].insert(queue[x].end(), items[g].begin(), items[g].end());
    load[x] += gcost[g];
  }
  size_t len = 0;
  for (auto& q : queue) len = std::max(len, q.size());
  len = (len + per - 1) / per * per;
  std::vector<int64_t> perm;
  perm.reserve(len * kXcdCount);
  for (size_t k = 0; k < len; k += per)
    for (int x = 0; x < kXcdCount; ++x)
      for (int u = 0; u < per; ++u) perm.push_back(k + u < queue[x].size() ? queue[x][k + u] : -1);
  return perm;
}

namespace {
// supertile key of a tile: (problem, 4 x 4 block of tiles)
int64_t supertile_key(std::unordered_map<int64_t, int64_t>& ids, int64_t problem, int mt, int nt) {
  const int64_t key = (problem * 8192 + mt / 4) * 8192 + nt / 4;
  auto it = ids.find(key);
  if (it == ids.end()) it = ids.emplace(key, (int64_t)ids.size()).first;
  return it->second;
}

void order_for_xcds(std::vector<TileA>& tiles) {
  std::unordered_map<int64_t, int64_t> ids, prob;
  std::vector<int64_t> group, cost;
  for (const TileA& t : tiles) {
    const int64_t pkey = t.c_off;  // the T block identifies the (L, V) problem
    group.push_back(supertile_key(ids, pkey, t.mt, t.nt));
    cost.push_back(t.K);
  }
  std::vector<TileA> out;
  for (int64_t i : xcd_order(group, cost)) out.push_back(i >= 0 ? tiles[i] : TileA{0, 0, 0, 0, 0, 0, 0, 0});
  tiles.swap(out);
}

// FP64 / FP32 wave tiles of stage A: 4 consecutive tiles form one workgroup,
// so the XCD queues are interleaved 4 tiles at a time; all tiles of one T
// block go to one XCD, whose L2 then serves the L and V re-reads.
void order_wave_tiles_for_xcds(std::vector<TileA>& tiles) {
  std::unordered_map<int64_t, int64_t> ids;
  std::vector<int64_t> group, cost;
  for (const TileA& t : tiles) {
    group.push_back(supertile_key(ids, t.c_off, t.mt, t.nt));
    cost.push_back(t.K + 8);  // + fixed cost of a tile (descriptor, store)
  }
  std::vector<TileA> out;
  for (int64_t i : xcd_order(group, cost, 4)) out.push_back(i >= 0 ? tiles[i] : TileA{0, 0, 0, 0, 0, 0, 0, 0});
  tiles.swap(out);
}

void order_for_xcds(std::vector<TileC>& tiles, const std::vector<ChunkC>& chunks) {
  std::unordered_map<int64_t, int64_t> ids;
  std::vector<int64_t> group, cost;
  for (const TileC& t : tiles) {
    group.push_back(supertile_key(ids, t.c_off, t.mt, t.nt));
    int64_t w = 0;
    for (int32_t c = t.chunk_begin; c < t.chunk_end; ++c) w += (int64_t)chunks[c].K * (1 + chunks[c].term_end - chunks[c].term_begin);
    cost.push_back(w);
  }
  std::vector<TileC> out;
  for (int64_t i : xcd_order(group, cost)) out.push_back(i >= 0 ? tiles[i] : TileC{0, 0, 0, 0, 0, 0, 0, 0});
  tiles.swap(out);
}

int choose_wave_edge(const HostPlan& H) {
  if (const char* e = std::getenv("VMPS_WAVE_TILE")) {
    const int v = std::atoi(e);
    if (v == 16 || v == 32) return v;
  }
  return H.stats.max_block_dim <= kSmallTileMaxDim ? 16 : 32;
}
}  // namespace

void set_wave_edge(HostPlan& H, int edge, int big_min64) {
  if (edge != 16 && edge != 32) throw VmpsError(VMPS_ERR_ARG, "wave tile edge must be 16 or 32");
  if (big_min64 < 0) throw VmpsError(VMPS_ERR_ARG, "big_min64 must be >= 0");
  if (H.crange.size() != H.probC.size()) throw VmpsError(VMPS_ERR_INTERNAL, "wave tiles: chunk ranges missing");
  H.wave_edge = edge;
  H.big_min64 = big_min64;
  if (big_min64 > 0) {
    make_tiles(H, edge, H.crange, H.tilesA_w, H.tilesC_w, -big_min64);
    make_tiles(H, kBigTile, H.crange, H.tilesA_b, H.tilesC_b, big_min64);
    order_for_xcds(H.tilesA_b);
    order_for_xcds(H.tilesC_b, H.chunksC);
  } else {
    H.tilesA_b.clear();
    H.tilesC_b.clear();
    if (edge == kTile) {
      H.tilesA_w = H.tilesA;
      H.tilesC_w = H.tilesC;
    } else {
      make_tiles(H, edge, H.crange, H.tilesA_w, H.tilesC_w);
    }
  }
  // XCD-local wave tiles (one workgroup per 4 stage A tiles, one per stage C
  // tile): the tiles of a block share one XCD's L2.  Short lists stay in
  // cost order, where the padding of the XCD rounds would cost more.
  if (H.tilesA_w.size() >= kXcdMinTilesA) order_wave_tiles_for_xcds(H.tilesA_w);
  if (H.tilesC_w.size() >= kXcdMinTilesC) order_for_xcds(H.tilesC_w, H.chunksC);
  // bounds of every FP64 / FP32 tile list
  const int64_t src_size = H.explicit_u ? H.U_size : H.T_size;
  for (const auto* list : {&H.tilesC_w, &H.tilesC_b})
    for (const TileC& t : *list)
      for (int32_t c = t.chunk_begin; c < t.chunk_end; ++c) {
        if (c < 0 || c >= (int32_t)H.chunksC.size() || H.chunksC[c].r_off < 0 ||
            H.chunksC[c].r_off + (int64_t)H.chunksC[c].K * t.N > H.R_size)
          throw VmpsError(VMPS_ERR_INTERNAL, "stage C: R block out of range");
        for (int32_t j = H.chunksC[c].term_begin; j < H.chunksC[c].term_end; ++j)
          if (H.termsC[j].t_off + (int64_t)t.M * H.chunksC[c].K > src_size)
            throw VmpsError(VMPS_ERR_INTERNAL, "stage C: T block out of range");
      }
  for (const auto* list : {&H.tilesA_w, &H.tilesA_b})
    for (const TileA& t : *list)
      if (t.a_off < 0 || t.a_off + (int64_t)t.M * t.K > H.L_size || t.b_off < 0 ||
          t.b_off + (int64_t)t.K * t.N > H.ysize || t.c_off < 0 || t.c_off + (int64_t)t.M * t.N > H.T_size)
        throw VmpsError(VMPS_ERR_INTERNAL, "stage A: tile out of range");
}

std::string tile_signature(const HostPlan& H) {
  auto cls = [](int64_t d) { return d <= 16 ? 0 : d <= 32 ? 1 : d <= 64 ? 2 : d <= 128 ? 3 : d <= 256 ? 4 : 5; };
  double share[2][6][6] = {};
  double total = 0;
  for (const GemmProb& p : H.probA) {
    const double f = 2.0 * p.M * p.N * H.chunkA[p.chunk_begin].K;
    share[0][cls(p.M)][cls(p.N)] += f;
    total += f;
  }
  for (const GemmProb& p : H.probC)
    for (int32_t c = p.chunk_begin; c < p.chunk_end; ++c) {
      const double f = 2.0 * p.M * p.N * H.chunkC[c].K;
      share[1][cls(p.M)][cls(p.N)] += f;
      total += f;
    }
  std::string key;
  for (int st = 0; st < 2; ++st)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) {
        const int q = total > 0 ? (int)std::lround(8.0 * share[st][a][b] / total) : 0;
        if (q) key += std::to_string(st) + std::to_string(a) + std::to_string(b) + ":" + std::to_string(q) + ";";
      }
  // the total work decides whether launches or MFMA time dominate
  key += "w" + std::to_string(total > 0 ? (int)std::floor(std::log2(total)) : 0);
  return key;
}

void build_tiles(HostPlan& H) {
  H.chunksC.clear();
  H.termsC.clear();
  H.comb_tasks.clear();
  H.comb_terms.clear();
  if (H.explicit_u) {
    // stage C reads U_u directly: term u = (U offset, 1)
    for (const CombineList& cl : H.comb) H.termsC.push_back({cl.u_off, 1.0});
    for (size_t j = 0; j < H.comb_t.size(); ++j) H.comb_terms.push_back({H.T_off[H.comb_t[j]], H.comb_c[j]});
    for (const CombineList& cl : H.comb)
      for (int64_t e0 = 0; e0 < cl.size; e0 += kCombTaskElems)
        H.comb_tasks.push_back({cl.u_off, e0, (int32_t)std::min<int64_t>(kCombTaskElems, cl.size - e0), cl.begin,
                                cl.end, 0});
  } else {
    for (size_t j = 0; j < H.comb_t.size(); ++j) H.termsC.push_back({H.T_off[H.comb_t[j]], H.comb_c[j]});
  }
  std::vector<std::pair<int32_t, int32_t>>& crange = H.crange;
  crange.assign(H.probC.size(), {0, 0});
  for (size_t p = 0; p < H.probC.size(); ++p) {
    const GemmProb& pr = H.probC[p];
    crange[p].first = (int32_t)H.chunksC.size();
    for (int32_t k = pr.chunk_begin; k < pr.chunk_end; ++k) {
      const GemmChunk& ch = H.chunkC[k];
      const CombineList& cl = H.comb[ch.pad];
      if (H.explicit_u) {
        H.chunksC.push_back({ch.b_off, ch.K, ch.pad, ch.pad + 1, 0});
        continue;
      }
      for (int32_t j0 = cl.begin; j0 < cl.end; j0 += kMaxTermsPerChunk) {
        ChunkC c;
        c.r_off = ch.b_off;
        c.K = ch.K;
        c.term_begin = j0;
        c.term_end = std::min(cl.end, j0 + kMaxTermsPerChunk);
        c.pad = 0;
        H.chunksC.push_back(c);
      }
    }
    crange[p].second = (int32_t)H.chunksC.size();
  }
  make_tiles(H, kTile, crange, H.tilesA, H.tilesC);
  set_wave_edge(H, choose_wave_edge(H), H.big_min64);
  // bounds of stage C: every T (or U) read stays inside T (or U)
  const int64_t src_size = H.explicit_u ? H.U_size : H.T_size;
  for (const CombTask& t : H.comb_tasks) {
    if (t.term_begin < 0 || t.term_end > (int32_t)H.comb_terms.size() || t.term_end < t.term_begin || t.n < 0 ||
        t.u_off + t.e0 + t.n > H.U_size)
      throw VmpsError(VMPS_ERR_INTERNAL, "stage B: task out of range");
    if ((t.u_off + t.e0) % kElemAlign)
      throw VmpsError(VMPS_ERR_INTERNAL, "stage B: U task not aligned");
    for (int32_t j = t.term_begin; j < t.term_end; ++j)
      if (H.comb_terms[j].t_off < 0 || H.comb_terms[j].t_off + t.e0 + t.n > H.T_size ||
          H.comb_terms[j].t_off % kElemAlign)
        throw VmpsError(VMPS_ERR_INTERNAL, "stage B: T block out of range or not aligned");
  }
  for (const ChunkC& c : H.chunksC) {
    if (c.term_end - c.term_begin > kMaxTermsPerChunk || c.term_end < c.term_begin)
      throw VmpsError(VMPS_ERR_INTERNAL, "stage C: chunk term list too long");
    for (int32_t j = c.term_begin; j < c.term_end; ++j)
      if (j < 0 || j >= (int32_t)H.termsC.size() || H.termsC[j].t_off < 0)
        throw VmpsError(VMPS_ERR_INTERNAL, "stage C: term reference out of range");
  }
  for (const auto* list : {&H.tilesC, &H.tilesC_w})
    for (const TileC& t : *list)
      for (int32_t c = t.chunk_begin; c < t.chunk_end; ++c) {
        if (H.chunksC[c].r_off < 0 || H.chunksC[c].r_off + (int64_t)H.chunksC[c].K * t.N > H.R_size)
          throw VmpsError(VMPS_ERR_INTERNAL, "stage C: R block out of range");
        for (int32_t j = H.chunksC[c].term_begin; j < H.chunksC[c].term_end; ++j)
          if (H.termsC[j].t_off + (int64_t)t.M * H.chunksC[c].K > src_size)
            throw VmpsError(VMPS_ERR_INTERNAL, "stage C: T block out of range");
      }
  for (const auto* list : {&H.tilesA, &H.tilesA_w})
    for (const TileA& t : *list)
      if (t.a_off < 0 || t.a_off + (int64_t)t.M * t.K > H.L_size || t.b_off < 0 ||
          t.b_off + (int64_t)t.K * t.N > H.ysize || t.c_off < 0 || t.c_off + (int64_t)t.M * t.N > H.T_size)
        throw VmpsError(VMPS_ERR_INTERNAL, "stage A: tile operand out of range");
}

void triples_from_pivot(const Pivot& P, TripleSet& S) {
  std::vector<EnvRef> lblocks, rblocks;
  std::vector<std::pair<int, int>> vblocks;
  S = TripleSet();
  expand_terms(P, S.terms, lblocks, rblocks, vblocks);
  // gather the referenced environment blocks, packed in first-use order
  // (offsets first, then the copies in parallel)
  // (offsets first, then the copies in parallel; environments resident on
  // the device become device gathers, all of one side or none)
  auto gather = [&](const std::vector<EnvRef>& refs, bool left, std::vector<BlockDim>& dims,
                    std::vector<double>& data, std::vector<Gather>& dev, int64_t& size) {
    int64_t o = 0;
    int ndev = 0;
    for (const auto& r : refs) {
      const Tripod& E = left ? P.terms[r.t].L : P.terms[r.t].R;
      const DenseBlock& b = E.block[r.q][r.a];
      dims.push_back({o, b.rows, b.cols});
      if (E.dev) {
        ++ndev;
        if (b.size()) dev.push_back({E.dev + b.off, o, b.size()});
        if (E.dev_ready && std::find(S.env_waits.begin(), S.env_waits.end(), E.dev_ready) == S.env_waits.end())
          S.env_waits.push_back(E.dev_ready);
      }
      o += b.size();
    }
    size = o;
    if (ndev && ndev != (int)refs.size())
      throw VmpsError(VMPS_ERR_ARG, "the environments of one side must all be on the host or all on the device");
    if (ndev) return;
    data.resize((size_t)o);
    parallel_for((int64_t)refs.size(), o, [&](int64_t i) {
      const auto& r = refs[(size_t)i];
      const Tripod& E = left ? P.terms[r.t].L : P.terms[r.t].R;
      const DenseBlock& b = E.block[r.q][r.a];
      if (b.size()) std::memcpy(data.data() + dims[(size_t)i].off, E.data.data() + b.off, sizeof(double) * b.size());
    });
  };
  gather(lblocks, true, S.Xb, S.Xdata, S.Xgather, S.Xsize);
  gather(rblocks, false, S.Zb, S.Zdata, S.Zgather, S.Zsize);
  for (const auto& vb : vblocks) {
    const DenseBlock& b = P.V.block[vb.first][vb.second];
    S.Yb.push_back({b.off, b.rows, b.cols});
    S.Ow.push_back(P.sym.coeff_dot(P.V.out[vb.first][vb.second]));
  }
  S.Ob = S.Yb;
  S.ysize = S.osize = P.V.size;
}

void build_host_plan(const Pivot& P, HostPlan& H) {
  TripleSet S;
  triples_from_pivot(P, S);
  build_plan(std::move(S), H);
}

namespace {
void build_plan_tables(const TripleSet& S, HostPlan& H);
}

void build_plan(const TripleSet& S, HostPlan& H) {
  H = HostPlan();
  H.Ldata = S.Xdata;
  H.Rdata = S.Zdata;
  H.Lgather = S.Xgather;
  H.Rgather = S.Zgather;
  H.env_waits = S.env_waits;
  H.L_size = S.Xgather.empty() ? (int64_t)S.Xdata.size() : S.Xsize;
  H.R_size = S.Zgather.empty() ? (int64_t)S.Zdata.size() : S.Zsize;
  build_plan_tables(S, H);
}

// the environment data move into the plan; S keeps everything else
void build_plan(TripleSet&& S, HostPlan& H) {
  H = HostPlan();
  H.L_size = S.Xgather.empty() ? (int64_t)S.Xdata.size() : S.Xsize;
  H.R_size = S.Zgather.empty() ? (int64_t)S.Zdata.size() : S.Zsize;
  H.Ldata = std::move(S.Xdata);
  H.Rdata = std::move(S.Zdata);
  H.Lgather = std::move(S.Xgather);
  H.Rgather = std::move(S.Zgather);
  H.env_waits = std::move(S.env_waits);
  build_plan_tables(S, H);
}

namespace {
void build_plan_tables(const TripleSet& S, HostPlan& H) {
  Stats& st = H.stats;
  H.ysize = S.ysize;
  H.osize = S.osize;
  H.vsize = S.osize;
  H.Lblk = S.Xb;
  H.Rblk = S.Zb;
  H.Vblk = S.Yb;
  for (const BlockDim& b : S.Xb) st.max_block_dim = std::max({st.max_block_dim, b.rows, b.cols});
  for (const BlockDim& b : S.Zb) st.max_block_dim = std::max({st.max_block_dim, b.rows, b.cols});
  const std::vector<Term>& terms = S.terms;
  const int64_t nY = (int64_t)S.Yb.size(), nO = (int64_t)S.Ob.size(), nZ = (int64_t)S.Zb.size();

  // shapes: X (m' x m), Y (m x n), Z (n x n'), O (m' x n')
  for (const Term& t : terms) {
    if (t.l_blk < 0 || t.l_blk >= (int64_t)S.Xb.size() || t.r_blk < 0 || t.r_blk >= nZ || t.in_v < 0 ||
        t.in_v >= nY || t.out_v < 0 || t.out_v >= nO)
      throw VmpsError(VMPS_ERR_LAYOUT, "term references a block that does not exist");
    const BlockDim &L = S.Xb[t.l_blk], &R = S.Zb[t.r_blk], &vi = S.Yb[t.in_v], &vo = S.Ob[t.out_v];
    if (L.cols != vi.rows || vi.cols != R.rows || L.rows != vo.rows || R.cols != vo.cols)
      throw VmpsError(VMPS_ERR_LAYOUT, "block shapes of a contributing term do not chain: L " + std::to_string(L.rows) +
                                           "x" + std::to_string(L.cols) + ", Vin " + std::to_string(vi.rows) + "x" +
                                           std::to_string(vi.cols) + ", R " + std::to_string(R.rows) + "x" +
                                           std::to_string(R.cols) + ", Vout " + std::to_string(vo.rows) + "x" +
                                           std::to_string(vo.cols));
    st.flops_ref += 2.0 * L.rows * L.cols * vi.cols + 2.0 * L.rows * vi.cols * R.cols;
  }

  // stage A: distinct (x, y)
  std::unordered_map<int64_t, int32_t> tmap;  // key = x * nY + y
  std::vector<int32_t> term_t(terms.size());
  for (size_t i = 0; i < terms.size(); ++i) {
    const Term& t = terms[i];
    int64_t key = t.l_blk * nY + t.in_v;
    auto it = tmap.find(key);
    if (it == tmap.end()) {
      const BlockDim &L = S.Xb[t.l_blk], &vi = S.Yb[t.in_v];
      int32_t tid = (int32_t)H.probA.size();
      it = tmap.emplace(key, tid).first;
      GemmProb pr;
      pr.c_off = H.T_size;
      pr.M = (int32_t)L.rows;
      pr.N = (int32_t)vi.cols;
      pr.chunk_begin = (int32_t)H.chunkA.size();
      pr.chunk_end = pr.chunk_begin + 1;
      GemmChunk ch;
      ch.a_off = L.off;
      ch.b_off = vi.off;
      ch.K = (int32_t)L.cols;
      ch.lda = (int32_t)L.rows;
      ch.ldb = (int32_t)vi.rows;
      ch.pad = 0;
      H.chunkA.push_back(ch);
      H.probA.push_back(pr);
      H.T_off.push_back(H.T_size);
      H.T_size += align_elems(L.rows * vi.cols);
      st.flops_alg += 2.0 * L.rows * L.cols * vi.cols;
      st.flops_A += 2.0 * L.rows * L.cols * vi.cols;
    }
    term_t[i] = it->second;
  }

  // stage B: distinct (o, z), entries in term order
  std::unordered_map<int64_t, int32_t> umap;  // key = o * nZ + z
  std::vector<std::vector<std::pair<int32_t, double>>> ulist;
  std::vector<std::pair<int64_t, int64_t>> ukey;  // (o, z)
  std::vector<std::vector<int32_t>> out_u(nO);
  for (size_t i = 0; i < terms.size(); ++i) {
    const Term& t = terms[i];
    int64_t key = t.out_v * nZ + t.r_blk;
    auto it = umap.find(key);
    if (it == umap.end()) {
      int32_t uid = (int32_t)ulist.size();
      it = umap.emplace(key, uid).first;
      ulist.emplace_back();
      ukey.push_back({t.out_v, t.r_blk});
      out_u[t.out_v].push_back(uid);
    }
    ulist[it->second].push_back({term_t[i], t.c});
  }
  std::vector<int64_t> uoff(ulist.size());
  for (size_t u = 0; u < ulist.size(); ++u) {
    const BlockDim& vo = S.Ob[ukey[u].first];
    const BlockDim& R = S.Zb[ukey[u].second];
    CombineList cl;
    cl.u_off = H.U_size;
    cl.size = vo.rows * R.rows;
    cl.begin = (int32_t)H.comb_t.size();
    for (auto& e : ulist[u]) {
      H.comb_t.push_back(e.first);
      H.comb_c.push_back(e.second);
    }
    cl.end = (int32_t)H.comb_t.size();
    uoff[u] = H.U_size;
    H.U_size += align_elems(cl.size);
    H.comb.push_back(cl);
    st.flops_alg += 2.0 * vo.rows * R.rows * R.cols;
    st.flops_C += 2.0 * vo.rows * R.rows * R.cols;
  }

  // stage C: one problem per output block (all blocks, so untouched ones are zeroed)
  for (int64_t o = 0; o < nO; ++o) {
    const BlockDim& vo = S.Ob[o];
    if (vo.rows * vo.cols == 0) continue;
    GemmProb pr;
    pr.c_off = vo.off;
    pr.M = (int32_t)vo.rows;
    pr.N = (int32_t)vo.cols;
    pr.chunk_begin = (int32_t)H.chunkC.size();
    for (int32_t u : out_u[o]) {
      const BlockDim& R = S.Zb[ukey[u].second];
      GemmChunk ch;
      ch.a_off = uoff[u];
      ch.b_off = R.off;
      ch.K = (int32_t)R.rows;
      ch.lda = (int32_t)vo.rows;
      ch.ldb = (int32_t)R.rows;
      ch.pad = u;  // combine list that forms this U
      H.chunkC.push_back(ch);
    }
    pr.chunk_end = (int32_t)H.chunkC.size();
    H.probC.push_back(pr);
  }

  // coeff_dot weights per output block (DmrgPivotVector.h:323), split in segments
  for (int64_t o = 0; o < nO; ++o) {
    const BlockDim& b = S.Ob[o];
    const int64_t size = b.rows * b.cols;
    H.vblock_off.push_back(b.off);
    H.vblock_size.push_back(size);
    const double w = S.Ow.empty() ? 1.0 : S.Ow[o];
    for (int64_t e = 0; e < size; e += 4096) {
      H.seg_off.push_back(b.off + e);
      H.seg_len.push_back((int32_t)std::min<int64_t>(4096, size - e));
      H.seg_w.push_back(w);
    }
  }

  validate_host_plan(H);
  // Pivots with blocks of 256 and more start from the explicit stage B and
  // the FP64 / FP32 big tiles, which the autotuner picks on every such pivot
  // measured; the tuner then finds the lists already built for its first
  // candidate instead of building them a second time.
  if (st.max_block_dim >= 2 * kBigTile) {
    H.explicit_u = true;
    if (H.big_min64 <= 0) H.big_min64 = kBigTile;
  }
  build_tiles(H);
  st.n_terms = (int64_t)terms.size();
  st.n_T = (int64_t)H.probA.size();
  st.n_U = (int64_t)H.comb.size();
  st.n_out = (int64_t)H.probC.size();
  st.n_L = (int64_t)S.Xb.size();
  st.n_R = (int64_t)S.Zb.size();
  st.n_V = nO;
  st.L_elems = H.L_size;
  st.R_elems = H.R_size;
  st.V_elems = S.osize;
  st.T_elems = H.T_size;
  st.U_elems = H.U_size;
  st.bytes_alg_fp64 = 8.0 * (double)(st.L_elems + st.R_elems + S.ysize + S.osize);
}
}  // namespace

}  // namespace vmps

namespace vmps {

void build_bf16_layout(const HostPlan& H, Bf16Layout& B) {
  B = Bf16Layout();
  // L^T, row-major M x Kp at 8-aligned offsets
  std::unordered_map<int64_t, int64_t> Lmap, Rmap, Vmap, Tmap, Umap;
  int64_t o = 0;
  for (const BlockDim& b : H.Lblk) {
    Lmap[b.off] = o;
    o += b.rows * pad8(b.cols);
  }
  B.L_size = o;
  for (const BlockDim& b : H.Lblk) {  // conversion tasks of at most 4096 padded elements
    if (b.rows * b.cols == 0) continue;
    const int64_t Kp = pad8(b.cols), base = Lmap.at(b.off);
    const int64_t rows_per = std::max<int64_t>(1, 4096 / Kp);
    for (int64_t i0 = 0; i0 < b.rows; i0 += rows_per)
      B.lconv.push_back({b.off + i0, base + i0 * Kp, (int32_t)b.cols, (int32_t)std::min(rows_per, b.rows - i0),
                         (int32_t)Kp, (int32_t)b.rows});
  }
  // R, column-major Kp x N
  o = 0;
  for (const BlockDim& b : H.Rblk) {
    Rmap[b.off] = o;
    o += pad8(b.rows) * b.cols;
  }
  B.R_size = o;
  for (const BlockDim& b : H.Rblk) {
    if (b.rows * b.cols == 0) continue;
    const int64_t Kp = pad8(b.rows), base = Rmap.at(b.off);
    const int64_t cols_per = std::max<int64_t>(1, 4096 / Kp);
    for (int64_t j0 = 0; j0 < b.cols; j0 += cols_per)
      B.rconv.push_back({b.off + j0 * b.rows, base + j0 * Kp, (int32_t)b.rows, (int32_t)std::min(cols_per, b.cols - j0),
                         (int32_t)Kp, 0});
  }
  // V, column-major Kp x N (filled per call)
  o = 0;
  for (const BlockDim& b : H.Vblk) {
    if (b.rows * b.cols == 0) continue;
    Vmap[b.off] = o;
    // one conversion task per group of columns holding at most 4096 padded elements
    const int64_t Kp = pad8(b.rows);
    const int64_t cols_per = std::max<int64_t>(1, 4096 / Kp);
    for (int64_t j0 = 0; j0 < b.cols; j0 += cols_per) {
      const int64_t nc = std::min(cols_per, b.cols - j0);
      B.vconv.push_back({b.off + j0 * b.rows, o + j0 * Kp, (int32_t)b.rows, (int32_t)nc, (int32_t)Kp, 0});
    }
    o += Kp * b.cols;
  }
  B.V_size = o;
  // T, row-major M x Np
  o = 0;
  for (size_t t = 0; t < H.probA.size(); ++t) {
    Tmap[H.probA[t].c_off] = o;
    o += (int64_t)H.probA[t].M * pad8(H.probA[t].N);
  }
  B.T_size = o;
  // explicit stage B: U blocks in the layout of their T blocks (row-major
  // M x Kp, Kp = K of the stage C chunk padded to 8)
  if (H.explicit_u) {
    o = 0;
    for (const CombineList& cl : H.comb) {
      Umap[cl.u_off] = o;
      const GemmProb& pa = H.probA[H.comb_t[cl.begin]];  // a T block of this U: same M x K
      const int64_t Np = pad8(pa.N);
      for (int64_t e0 = 0; e0 < (int64_t)pa.M * Np; e0 += kCombTaskElems)
        B.comb_tasks.push_back({o, e0, (int32_t)std::min<int64_t>(kCombTaskElems, (int64_t)pa.M * Np - e0),
                                cl.begin, cl.end, 0});
      o += (int64_t)pa.M * Np;
    }
    B.U_size = o;
    for (size_t j = 0; j < H.comb_t.size(); ++j)
      B.comb_terms.push_back({Tmap.at(H.T_off[H.comb_t[j]]), (float)H.comb_c[j], 0});
  }
  const int64_t big_min = H.big_min;
  auto is_big = [big_min](int64_t M, int64_t N) { return big_min > 0 && M >= big_min && N >= big_min; };
  for (const TileA& t : H.tilesA) {
    TileA16 a;
    a.a_off = Lmap.at(t.a_off);
    a.b_off = Vmap.at(t.b_off);
    a.c_off = Tmap.at(t.c_off);
    a.M = t.M;
    a.N = t.N;
    a.Kp = (int32_t)pad8(t.K);
    a.Np = (int32_t)pad8(t.N);
    a.mt = t.mt;
    a.nt = t.nt;
    a.pad = 0;
    if (!is_big(t.M, t.N)) {
      B.tilesA.push_back(a);
    } else if (t.mt % (kBigTile / kTile) == 0 && t.nt % (kBigTile / kTile) == 0) {
      a.mt = (int16_t)(t.mt / (kBigTile / kTile));  // one 128 x 128 tile per 4 x 4 wave tiles
      a.nt = (int16_t)(t.nt / (kBigTile / kTile));
      B.tilesA_big.push_back(a);
    }
  }
  for (const TileC& t : H.tilesC) {
    if (!is_big(t.M, t.N)) {
      B.tilesC_small.push_back(t);
    } else if (t.mt % (kBigTile / kTile) == 0 && t.nt % (kBigTile / kTile) == 0) {
      TileC b = t;
      b.mt = (int16_t)(t.mt / (kBigTile / kTile));
      b.nt = (int16_t)(t.nt / (kBigTile / kTile));
      B.tilesC_big.push_back(b);
    }
  }
  // Split the K iterations of a big stage C tile over several workgroups only
  // when the big tiles alone cannot fill the GPU (the FP32 partials cost a
  // round trip through HBM): aim at ~512 workgroups, at least 8 iterations of
  // kBigK k per split, at most 4 splits.
  const int64_t nbig = (int64_t)B.tilesC_big.size();
  const int64_t want = nbig > 0 ? std::max<int64_t>(1, std::min<int64_t>(4, 512 / nbig)) : 1;
  for (size_t ti = 0; ti < B.tilesC_big.size(); ++ti) {
    const TileC& t = B.tilesC_big[ti];
    int64_t steps = 0;
    for (int32_t c = t.chunk_begin; c < t.chunk_end; ++c) steps += (pad8(H.chunksC[c].K) + kBigK - 1) / kBigK;
    const int64_t S = std::max<int64_t>(1, std::min<int64_t>(want, steps / 8));
    for (int64_t s = 0; s < S; ++s) {
      B.splitC_tile.push_back((int32_t)ti);
      B.splitC_begin.push_back((int32_t)(steps * s / S));
      B.splitC_end.push_back((int32_t)(steps * (s + 1) / S));
      B.splitC_slot.push_back(S == 1 ? -1 : B.partial_slots + s);
    }
    if (S > 1) {
      B.redC_tile.push_back((int32_t)ti);
      B.redC_slot.push_back(B.partial_slots);
      B.redC_nsplit.push_back((int32_t)S);
      B.partial_slots += S;
    }
  }
  // launch order for the XCDs (see xcd_order): big stage A tiles and big
  // stage C work items, grouped by 4 x 4 supertiles of their block
  {
    std::unordered_map<int64_t, int64_t> ids;
    auto key = [&](int64_t problem, int mt, int nt) {
      const int64_t k = (problem * 8192 + mt / 4) * 8192 + nt / 4;
      auto it = ids.find(k);
      if (it == ids.end()) it = ids.emplace(k, (int64_t)ids.size()).first;
      return it->second;
    };
    std::vector<int64_t> group, cost;
    for (const TileA16& a : B.tilesA_big) {
      group.push_back(key(a.c_off, a.mt, a.nt));
      cost.push_back(a.Kp);
    }
    std::vector<TileA16> outA;
    for (int64_t i : xcd_order(group, cost))
      outA.push_back(i >= 0 ? B.tilesA_big[i] : TileA16{0, 0, 0, 0, 0, 0, 0, 0, 0, 0});
    B.tilesA_big.swap(outA);
    group.clear();
    cost.clear();
    ids.clear();
    for (size_t w = 0; w < B.splitC_tile.size(); ++w) {
      const TileC& t = B.tilesC_big[B.splitC_tile[w]];
      group.push_back(key(t.c_off, t.mt, t.nt));
      cost.push_back(B.splitC_end[w] - B.splitC_begin[w]);
    }
    const int32_t empty = (int32_t)B.tilesC_big.size();  // an empty tile for the padding items
    std::vector<int32_t> wt, wb, we;
    std::vector<int64_t> ws;
    bool padded = false;
    for (int64_t i : xcd_order(group, cost)) {
      if (i >= 0) {
        wt.push_back(B.splitC_tile[i]);
        wb.push_back(B.splitC_begin[i]);
        we.push_back(B.splitC_end[i]);
        ws.push_back(B.splitC_slot[i]);
      } else {
        wt.push_back(empty);
        wb.push_back(0);
        we.push_back(0);
        ws.push_back(-1);
        padded = true;
      }
    }
    if (padded) B.tilesC_big.push_back(TileC{0, 0, 0, 0, 0, 0, 0, 0});
    B.splitC_tile.swap(wt);
    B.splitC_begin.swap(wb);
    B.splitC_end.swap(we);
    B.splitC_slot.swap(ws);
  }
  for (const ChunkC& c : H.chunksC) {
    ChunkC16 d;
    d.r_off = Rmap.at(c.r_off);
    d.Kp = (int32_t)pad8(c.K);
    d.term_begin = c.term_begin;
    d.term_end = c.term_end;
    d.pad = 0;
    B.chunksC.push_back(d);
  }
  for (const TermC& t : H.termsC)
    B.termsC.push_back({H.explicit_u ? Umap.at(t.t_off) : Tmap.at(t.t_off), (float)t.c, 0});
  const int64_t src_size = H.explicit_u ? B.U_size : B.T_size;
  // bounds: every padded access stays inside its array
  for (const CombTask& t : B.comb_tasks) {
    if (t.term_begin < 0 || t.term_end > (int32_t)B.comb_terms.size() || t.term_end < t.term_begin ||
        t.u_off + t.e0 + t.n > B.U_size)
      throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage B task out of range");
    if ((t.u_off + t.e0) % kElemAlign) throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage B task not aligned");
    for (int32_t j = t.term_begin; j < t.term_end; ++j)
      if (B.comb_terms[j].t_off + t.e0 + t.n > B.T_size || B.comb_terms[j].t_off % kElemAlign)
        throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage B term out of range or not aligned");
  }
  for (const auto* list : {&B.tilesA, &B.tilesA_big})
    for (const TileA16& a : *list)
      if (a.a_off + (int64_t)a.M * a.Kp > B.L_size || a.b_off + (int64_t)a.Kp * a.N > B.V_size ||
          a.c_off + (int64_t)a.M * a.Np > B.T_size || (a.a_off | a.b_off | a.c_off) % 8)
        throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage A descriptor out of range");
  for (const TileC& t : H.tilesC)
    for (int32_t c = t.chunk_begin; c < t.chunk_end; ++c) {
      const ChunkC16& ch = B.chunksC[c];
      if (ch.r_off % 8 || ch.r_off + (int64_t)ch.Kp * t.N > B.R_size)
        throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage C operand out of range");
      if (ch.term_end - ch.term_begin > kMaxTermsPerChunk || ch.term_end < ch.term_begin ||
          ch.term_end > (int32_t)B.termsC.size())
        throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage C chunk term list out of range");
      for (int32_t j = ch.term_begin; j < ch.term_end; ++j)
        if (B.termsC[j].t_off % 8 || B.termsC[j].t_off + (int64_t)t.M * ch.Kp > src_size)
          throw VmpsError(VMPS_ERR_INTERNAL, "bf16 stage C term out of range");
    }
  for (const VConv& v : B.vconv)
    if (v.dst_off + (int64_t)v.Kp * v.N > B.V_size || v.src_off + (int64_t)v.K * v.N > H.ysize)
      throw VmpsError(VMPS_ERR_INTERNAL, "bf16 conversion task out of range");
}

}  // namespace vmps
