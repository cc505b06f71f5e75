// contract_L / contract_R planners; see env.hpp.
#include "env.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "errors.hpp"

namespace vmps {

namespace {

struct Hit {
  Qn q0, q1, q2;  // (Rin, Rout, Rmid) from LAWA or (Lin, Lout, Lmid) from AWAR
  size_t qAbra, qAket, qW;
};

// LAWA, DmrgIndexGymnastics.h:32-77
void LAWA(const Sym& sym, const Qn& Lin, const Qn& Lout, const Qn& Lmid, const Qn& qloc1, const Qn& qloc2,
          const Qn& qOp, const DBiped& Abra, const DBiped& Aket, const WBiped& W, std::vector<Hit>& out) {
  out.clear();
  for (const Qn& Rout : sym.reduce(Lin, qloc1)) {
    auto ib = Abra.dict.find(Key2{Lin, Rout});
    if (ib == Abra.dict.end()) continue;
    for (const Qn& Rin : sym.reduce(Lout, qloc2)) {
      auto ik = Aket.dict.find(Key2{Lout, Rin});
      if (ik == Aket.dict.end()) continue;
      for (const Qn& Rmid : sym.reduce(Lmid, qOp)) {
        auto iw = W.dict.find(Key2{Lmid, Rmid});
        if (iw == W.dict.end()) continue;
        if (sym.validate3(Rin, Rmid, Rout)) out.push_back({Rin, Rout, Rmid, ib->second, ik->second, iw->second});
      }
    }
  }
}

// AWAR, DmrgIndexGymnastics.h:79-124
void AWAR(const Sym& sym, const Qn& Rin, const Qn& Rout, const Qn& Rmid, const Qn& qloc1, const Qn& qloc2,
          const Qn& qOp, const DBiped& Abra, const DBiped& Aket, const WBiped& W, std::vector<Hit>& out) {
  out.clear();
  for (const Qn& Lin : sym.reduce(Rout, sym.flip(qloc1))) {
    auto ib = Abra.dict.find(Key2{Lin, Rout});
    if (ib == Abra.dict.end()) continue;
    for (const Qn& Lout : sym.reduce(Rin, sym.flip(qloc2))) {
      auto ik = Aket.dict.find(Key2{Lout, Rin});
      if (ik == Aket.dict.end()) continue;
      for (const Qn& Lmid : sym.reduce(Rmid, sym.flip(qOp))) {
        auto iw = W.dict.find(Key2{Lmid, Rmid});
        if (iw == W.dict.end()) continue;
        if (sym.validate3(Lout, Lmid, Lin)) out.push_back({Lin, Lout, Lmid, ib->second, ik->second, iw->second});
      }
    }
  }
}

void pack(const double* src, int64_t rows, int64_t cols, bool transpose, std::vector<double>& dst,
          std::vector<BlockDim>& blocks) {
  const int64_t off = (int64_t)dst.size();
  if (!transpose) {
    dst.insert(dst.end(), src, src + rows * cols);
    blocks.push_back({off, rows, cols});
  } else {
    dst.resize(dst.size() + rows * cols);
    // (cols x rows) col-major, in 32 x 32 tiles so that reads and writes
    // both stay within a few cache lines per row
    double* d = dst.data() + off;
    constexpr int64_t kT = 32;
    for (int64_t j0 = 0; j0 < cols; j0 += kT)
      for (int64_t i0 = 0; i0 < rows; i0 += kT) {
        const int64_t j1 = std::min(cols, j0 + kT), i1 = std::min(rows, i0 + kT);
        for (int64_t j = j0; j < j1; ++j)
          for (int64_t i = i0; i < i1; ++i) d[i * cols + j] = src[j * rows + i];
      }
    blocks.push_back({off, cols, rows});
  }
}

}  // namespace

void plan_env(EnvProblem& P, TripleSet& S) {
  const Sym& sym = P.sym;
  const bool NA = sym.non_abelian();
  const bool L = (P.which == 0);
  S = TripleSet();
  P.new_index.clear();
  P.new_block.clear();
  P.new_adim.clear();

  // Y: every present block of the old environment (offsets into old.data)
  std::vector<std::vector<int64_t>> yid(P.old.dim());
  for (size_t q = 0; q < P.old.dim(); ++q) {
    yid[q].assign(P.old.block[q].size(), -1);
    for (size_t a = 0; a < P.old.block[q].size(); ++a) {
      const DenseBlock& b = P.old.block[q][a];
      if (!b.present()) continue;
      yid[q][a] = (int64_t)S.Yb.size();
      S.Yb.push_back({b.off, b.rows, b.cols});
    }
  }
  S.ysize = P.old.size();

  {  // each A block is packed at most once per operand: reserve that much
    size_t nbra = 0, nket = 0;
    for (const DBiped& A : P.Abra) nbra += A.data.size();
    for (const DBiped& A : P.Aket) nket += A.data.size();
    S.Xdata.reserve(L ? nbra : nket);
    S.Zdata.reserve(L ? nket : nbra);
  }
  std::map<std::pair<size_t, size_t>, int64_t> xmap, zmap;  // (s, q) of A -> packed block id
  std::unordered_map<Key3, size_t, Key3Hash> sec;            // new sector -> index
  std::vector<Hit> hits;
  const size_t nloc = P.qloc.size();
  for (size_t s1 = 0; s1 < nloc; ++s1)
    for (size_t s2 = 0; s2 < nloc; ++s2)
      for (size_t k = 0; k < P.qop.size(); ++k) {
        if (!sym.validate3(P.qloc[s2], P.qop[k], P.qloc[s1])) continue;
        const WBiped& W = P.W[s1][s2][k];
        for (size_t qo = 0; qo < P.old.dim(); ++qo) {
          const Qn &oin = P.old.in(qo), &oout = P.old.out(qo), &omid = P.old.mid(qo);
          if (L)
            LAWA(sym, oin, oout, omid, P.qloc[s1], P.qloc[s2], P.qop[k], P.Abra[s1], P.Aket[s2], W, hits);
          else
            AWAR(sym, oin, oout, omid, P.qloc[s1], P.qloc[s2], P.qop[k], P.Abra[s1], P.Aket[s2], W, hits);
          for (const Hit& h : hits) {
            // new sector: swap of the first two entries of the hit
            const Key3 key{h.q1, h.q0, h.q2};
            const DenseBlock& Bb = P.Abra[s1].block[h.qAbra];
            const DenseBlock& Bk = P.Aket[s2].block[h.qAket];
            if (!Bk.present() || !Bb.present()) continue;
            double f = 1.0;
            if (NA) {
              if (L)
                f = sym.coeff_buildL(P.Aket[s2].in[h.qAket], P.qloc[s2], P.Aket[s2].out[h.qAket], omid, P.qop[k],
                                     h.q2, P.Abra[s1].in[h.qAbra], P.qloc[s1], P.Abra[s1].out[h.qAbra]);
              else
                f = sym.coeff_buildR(P.Aket[s2].in[h.qAket], P.qloc[s2], P.Aket[s2].out[h.qAket], h.q2, P.qop[k],
                                     omid, P.Abra[s1].in[h.qAbra], P.qloc[s1], P.Abra[s1].out[h.qAbra]);
            }
            if (std::abs(f) < kCgcEps) continue;
            const SparseBlock& wb = W.block[h.qW];
            for (size_t n = 0; n < wb.val.size(); ++n) {
              const int a = wb.row[n], b = wb.col[n];
              const int oldidx = L ? a : b, newidx = L ? b : a;
              if (oldidx >= (int)P.old.block[qo].size())
                throw VmpsError(VMPS_ERR_LAYOUT, "MPO index outside the old environment's mid sector");
              if (yid[qo][oldidx] < 0) continue;  // absent old block
              const double c = f * wb.val[n];
              if (c == 0.0) continue;
              // operands
              int64_t x, z;
              const DenseBlock& Xsrc = L ? Bb : Bk;
              const DBiped& Xbip = L ? P.Abra[s1] : P.Aket[s2];
              const DenseBlock& Zsrc = L ? Bk : Bb;
              const DBiped& Zbip = L ? P.Aket[s2] : P.Abra[s1];
              const auto xkey = L ? std::make_pair(s1, h.qAbra) : std::make_pair(s2, h.qAket);
              const auto zkey = L ? std::make_pair(s2, h.qAket) : std::make_pair(s1, h.qAbra);
              auto ix = xmap.find(xkey);
              if (ix == xmap.end()) {
                pack(Xbip.data.data() + Xsrc.off, Xsrc.rows, Xsrc.cols, L, S.Xdata, S.Xb);
                ix = xmap.emplace(xkey, (int64_t)S.Xb.size() - 1).first;
              }
              auto iz = zmap.find(zkey);
              if (iz == zmap.end()) {
                pack(Zbip.data.data() + Zsrc.off, Zsrc.rows, Zsrc.cols, !L, S.Zdata, S.Zb);
                iz = zmap.emplace(zkey, (int64_t)S.Zb.size() - 1).first;
              }
              x = ix->second;
              z = iz->second;
              // output block (created on first touch)
              auto is = sec.find(key);
              if (is == sec.end()) {
                is = sec.emplace(key, P.new_index.size()).first;
                P.new_index.push_back({key.a, key.b, key.c});
                const int adim = L ? wb.cols : wb.rows;
                P.new_adim.push_back(adim);
                P.new_block.emplace_back(adim, -1);
              }
              auto& slot = P.new_block[is->second];
              if (newidx >= (int)slot.size())
                throw VmpsError(VMPS_ERR_LAYOUT, "MPO blocks of one sector disagree on the auxiliary dimension");
              const int64_t rows = S.Xb[x].rows, cols = S.Zb[z].cols;
              if (slot[newidx] < 0) {
                slot[newidx] = (int64_t)S.Ob.size();
                S.Ob.push_back({S.osize, rows, cols});
                S.osize += rows * cols;
              } else if (S.Ob[slot[newidx]].rows != rows || S.Ob[slot[newidx]].cols != cols) {
                throw VmpsError(VMPS_ERR_LAYOUT, "contributions to one environment block differ in shape");
              }
              S.terms.push_back({slot[newidx], x, yid[qo][oldidx], z, c});
            }
          }
        }
      }
}

}  // namespace vmps
