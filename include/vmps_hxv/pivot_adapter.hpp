// pivot_adapter.hpp -- route the reference's HxV / OxV through libvmps_hxv.
//
// Include this header AFTER the reference's pivot headers
// (pivot/DmrgPivotMatrix2.h, pivot/DmrgPivotMatrix1.h, pivot/DmrgPivotMatrix0.h)
// in the translation unit that instantiates the Krylov solvers, and link
// libvmps_hxv.so.  It adds more specialised overloads for Scalar = MpoScalar =
// double:
//
//   OxV(const PivotMatrix2<Sym,double,double>&, const PivotVector<Sym,double>&, PivotVector<Sym,double>&)
//       replaces the generic template of pivot/DmrgPivotMatrix2.h:85-273
//   OxV(const PivotMatrix1<Sym,double,double>&, ...)   replaces DmrgPivotMatrix1.h:131-332
//   HxV(const PivotMatrix0<Sym,double,double>&, ...)   replaces DmrgPivotMatrix0.h:31-96
//   contract_L / contract_R (seven-argument form)      replace DmrgContractions.h:34-311
//
// The reference's HxV (DmrgPivotMatrix2.h:78-83, DmrgPivotMatrix1.h:123-129)
// calls OxV unqualified with dependent arguments, so argument-dependent
// lookup at the point of instantiation finds these overloads and partial
// ordering prefers them over the generic templates.  LanczosSolver and every
// other caller stay unchanged.
//
// The adapter flattens the pivot (environment blocks, MPO nonzeros, the
// PivotVector layout and, for single-term pivots, the caller's own
// qlhs/qrhs/factor_cgcs) into the C ABI of vmps_hxv.h once, keeps the device
// plan, and on every call copies Vin's blocks into a packed buffer, runs the
// device HxV and copies the result into Vout's blocks.
//
// Plan reuse: a plan is cached per calling thread, keyed by the pivot's
// address and a fingerprint of its contents: block shapes, the layout of Vin,
// the MPO nonzeros and block-structure tables, and every value of every
// environment block (a 64-bit hash over all of them, read once per call at
// ~10 GB/s on one core).  A pivot whose values change, in place or by being
// rebuilt at the same address, therefore gets a new plan.  A caller that
// guarantees it never changes environment values behind a pivot's back (the
// reference's solvers build a new PivotMatrix per step) may call
// vmps_adapter::trust_pointers(true): environment blocks then enter the
// fingerprint by address, shape and first/last element only, which costs
// nothing per call; vmps_adapter::invalidate() drops cached plans explicitly.
//
// Errors from the library are thrown as std::runtime_error with the
// library's message (the reference's HxV is void and has no status).
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../vmps_hxv.h"

namespace vmps_adapter {

inline void check(vmps_status st, const char* what) {
  if (st != VMPS_OK)
    throw std::runtime_error(std::string("vmps_hxv adapter: ") + what + ": " + vmps_last_error());
}

// Symmetry::kind() (symmetry/U1.h:60, symmetry/SU2.h, symmetry/S1xS2.h:68):
// KIND::S, KIND::Salt and KIND::T are SU(2) components, the others U(1).
template <typename Symmetry>
std::vector<int> kinds() {
  static_assert(Symmetry::Nq >= 1 && Symmetry::Nq <= 2, "libvmps_hxv handles one or two symmetry components");
  std::vector<int> k;
  for (auto x : Symmetry::kind()) {
    const int v = static_cast<int>(x);
    // enum KIND {S, Salt, T, N, M, Nup, Ndn, Nparity, K} (DmrgTypedefs.h:110)
    if (v == 7 || v == 8) throw std::runtime_error("vmps_hxv adapter: Z_N symmetries are not supported");
    k.push_back(v <= 2 ? VMPS_SYM_SU2 : VMPS_SYM_U1);
  }
  return k;
}

template <typename Q>
void put_q(std::vector<int>& out, const Q& q, size_t nq) {
  for (size_t i = 0; i < nq; ++i) out.push_back(q[i]);
}

template <typename QV>
std::vector<int> flat_qs(const QV& v, size_t nq) {
  std::vector<int> out;
  for (const auto& q : v) put_q(out, q, nq);
  return out;
}

struct PivotHandle {
  vmps_pivot2* p = nullptr;
  ~PivotHandle() {
    if (p) vmps_pivot2_destroy(p);
  }
};

struct PlanEntry {
  uint64_t fingerprint = 0;
  vmps_plan* plan = nullptr;
  int64_t vsize = 0;
  std::vector<double> vin, vout;
  ~PlanEntry() {
    if (plan) vmps_plan_destroy(plan);
  }
};

inline std::unordered_map<const void*, std::unique_ptr<PlanEntry>>& plan_cache() {
  thread_local std::unordered_map<const void*, std::unique_ptr<PlanEntry>> cache;
  return cache;
}

// Fingerprint environments by address (true) or by content (false, default);
// per calling thread.
inline bool& trust_pointers_flag() {
  thread_local bool flag = false;
  return flag;
}
inline void trust_pointers(bool on) { trust_pointers_flag() = on; }

// Drop the cached plans of the calling thread (all, or the one of `pivot`).
inline void invalidate(const void* pivot = nullptr) {
  if (pivot)
    plan_cache().erase(pivot);
  else
    plan_cache().clear();
}

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  }
  template <typename T>
  void add(const T& x) {
    add(&x, sizeof(T));
  }
  // every value of n doubles: four independent multiply-xorshift lanes over
  // the 64-bit patterns, folded into h (order- and position-sensitive)
  void add_values(const double* d, size_t n) {
    uint64_t a[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull, 0x27D4EB2F165667C5ull};
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
      for (int j = 0; j < 4; ++j) {
        uint64_t w;
        std::memcpy(&w, d + i + j, 8);
        a[j] = (a[j] ^ w) * 0xFF51AFD7ED558CCDull;
        a[j] ^= a[j] >> 29;
      }
    for (; i < n; ++i) {
      uint64_t w;
      std::memcpy(&w, d + i, 8);
      a[0] = (a[0] ^ w) * 0xFF51AFD7ED558CCDull;
      a[0] ^= a[0] >> 29;
    }
    add(a);
    add(n);
  }
};

template <typename TripodT>
void fingerprint_env(Fnv& f, const TripodT& E) {
  f.add(E.dim);
  for (size_t q = 0; q < E.dim; ++q) {
    const size_t adim = E.block[q].shape()[0];
    f.add(adim);
    for (size_t a = 0; a < adim; ++a) {
      const auto& M = E.block[q][a][0];
      const int64_t r = M.rows(), c = M.cols();
      f.add(r);
      f.add(c);
      if (M.size() > 0) {
        const double* d = M.data();
        if (trust_pointers_flag()) {
          f.add(d);
          f.add(d[0]);
          f.add(d[M.size() - 1]);
        } else {
          f.add_values(d, static_cast<size_t>(M.size()));
        }
      }
    }
  }
}

// W[s1][s2][k]: shapes and every nonzero (by address only when pointers are trusted)
template <typename WT>
void fingerprint_W(Fnv& f, const WT& W) {
  if (trust_pointers_flag()) {
    f.add(&W);
    return;
  }
  f.add(W.size());
  for (const auto& W1 : W)
    for (const auto& W2 : W1)
      for (const auto& B : W2) {
        f.add(B.dim);
        for (size_t q = 0; q < B.dim; ++q) {
          const auto& M = B.block[q];
          using SM = typename std::decay<decltype(M)>::type;
          f.add(static_cast<int64_t>(M.rows()));
          f.add(static_cast<int64_t>(M.cols()));
          for (int o = 0; o < M.outerSize(); ++o)
            for (typename SM::InnerIterator it(M, o); it; ++it) {
              f.add(static_cast<int64_t>(it.row()));
              f.add(static_cast<int64_t>(it.col()));
              const double v = it.value();
              f.add(v);
            }
        }
      }
}

// qlhs / qrhs / factor_cgcs, A0proj and Epenalty of a pivot
template <typename HT>
void fingerprint_tables(Fnv& f, const HT& H) {
  f.add(H.qlhs.size());
  f.add(H.A0proj.size());
  const double ep = H.Epenalty;
  f.add(ep);
  if (trust_pointers_flag()) return;
  for (size_t i = 0; i < H.qlhs.size(); ++i) {
    f.add(static_cast<uint64_t>(H.qlhs[i][0]));
    f.add(static_cast<uint64_t>(H.qlhs[i][1]));
    f.add(H.qrhs[i].size());
    for (size_t j = 0; j < H.qrhs[i].size(); ++j) {
      for (size_t c = 0; c < H.qrhs[i][j].size(); ++c) f.add(static_cast<uint64_t>(H.qrhs[i][j][c]));
      const double x = H.factor_cgcs[i][j];
      f.add(x);
    }
  }
  for (const auto& P : H.A0proj) {
    f.add(P.size());
    for (const auto& B : P) {
      f.add(B.dim);
      for (size_t r = 0; r < B.dim; ++r) f.add_values(B.block[r].data(), static_cast<size_t>(B.block[r].size()));
    }
  }
}

template <typename PV>
void fingerprint_layout(Fnv& f, const PV& V) {
  f.add(V.data.size());
  for (const auto& B : V.data) {
    f.add(B.dim);
    for (size_t q = 0; q < B.dim; ++q) {
      const int64_t r = B.block[q].rows(), c = B.block[q].cols();
      f.add(r);
      f.add(c);
    }
  }
}

// L (which = 0) or R (which = 1): Tripod index {in, out, mid}, block[q][a][0].
template <typename Symmetry, typename TripodT>
void add_env(vmps_pivot2* p, int which, const TripodT& E) {
  const size_t nq = Symmetry::Nq;
  std::vector<int64_t> rows, cols;
  std::vector<double> data;
  for (size_t q = 0; q < E.dim; ++q) {
    std::vector<int> qi, qo, qm;
    put_q(qi, E.index[q][0], nq);
    put_q(qo, E.index[q][1], nq);
    put_q(qm, E.index[q][2], nq);
    const size_t adim = E.block[q].shape()[0];
    rows.assign(adim, 0);
    cols.assign(adim, 0);
    data.clear();
    for (size_t a = 0; a < adim; ++a) {
      const auto& M = E.block[q][a][0];
      rows[a] = M.rows();
      cols[a] = M.cols();
      data.insert(data.end(), M.data(), M.data() + M.size());
    }
    check(vmps_pivot2_add_env(p, which, qi.data(), qo.data(), qm.data(), static_cast<int>(adim), rows.data(),
                              cols.data(), data.data()),
          "add_env");
  }
}

// W[s1][s2][k]: Biped of column-major sparse blocks, nonzeros in InnerIterator order.
template <typename Symmetry, typename WT>
void add_W(vmps_pivot2* p, int which, const WT& W) {
  const size_t nq = Symmetry::Nq;
  for (size_t s1 = 0; s1 < W.size(); ++s1)
    for (size_t s2 = 0; s2 < W[s1].size(); ++s2)
      for (size_t k = 0; k < W[s1][s2].size(); ++k) {
        const auto& B = W[s1][s2][k];
        for (size_t q = 0; q < B.dim; ++q) {
          const auto& M = B.block[q];
          std::vector<int> row, col, qi, qo;
          std::vector<double> val;
          using SM = typename std::decay<decltype(M)>::type;
          for (int o = 0; o < M.outerSize(); ++o)
            for (typename SM::InnerIterator it(M, o); it; ++it) {
              row.push_back(static_cast<int>(it.row()));
              col.push_back(static_cast<int>(it.col()));
              val.push_back(it.value());
            }
          put_q(qi, B.in[q], nq);
          put_q(qo, B.out[q], nq);
          check(vmps_pivot2_add_W(p, which, static_cast<int>(s1), static_cast<int>(s2), static_cast<int>(k),
                                  qi.data(), qo.data(), static_cast<int>(M.rows()), static_cast<int>(M.cols()),
                                  static_cast<int64_t>(val.size()), row.data(), col.data(), val.data()),
                "add_W");
        }
      }
}

template <typename Symmetry, typename PV>
void add_layout(vmps_pivot2* p, const PV& V) {
  const size_t nq = Symmetry::Nq;
  int ns = 0;
  check(vmps_pivot2_local_basis(p, &ns, nullptr), "local_basis");
  if (static_cast<size_t>(ns) != V.data.size())
    throw std::runtime_error("vmps_hxv adapter: PivotVector has " + std::to_string(V.data.size()) +
                             " components, the pivot's local basis " + std::to_string(ns));
  for (size_t s = 0; s < V.data.size(); ++s) {
    const auto& B = V.data[s];
    for (size_t q = 0; q < B.dim; ++q) {
      std::vector<int> qi, qo;
      put_q(qi, B.in[q], nq);
      put_q(qo, B.out[q], nq);
      int64_t off = 0;
      check(vmps_pivot2_add_V_block(p, static_cast<int>(s), qi.data(), qo.data(), B.block[q].rows(),
                                    B.block[q].cols(), &off),
            "add_V_block");
    }
  }
}

// qlhs / qrhs / factor_cgcs of a single-term pivot (DmrgPivotMatrix2.h:66-70,
// DmrgPivotMatrix1.h:58-62); rows of qrhs are padded to 12 entries.
template <typename HT>
void pass_blockstructure(vmps_pivot2* p, const HT& H) {
  const size_t nlhs = H.qlhs.size();
  std::vector<int64_t> qlhs, begin, qrhs;
  std::vector<double> fac;
  begin.push_back(0);
  for (size_t i = 0; i < nlhs; ++i) {
    qlhs.push_back(static_cast<int64_t>(H.qlhs[i][0]));
    qlhs.push_back(static_cast<int64_t>(H.qlhs[i][1]));
    for (size_t j = 0; j < H.qrhs[i].size(); ++j) {
      const auto& e = H.qrhs[i][j];
      for (size_t c = 0; c < 12; ++c) qrhs.push_back(c < e.size() ? static_cast<int64_t>(e[c]) : 0);
      fac.push_back(H.factor_cgcs[i][j]);
    }
    begin.push_back(static_cast<int64_t>(fac.size()));
  }
  check(vmps_pivot2_set_blockstructure(p, static_cast<int64_t>(nlhs), qlhs.data(), begin.data(), qrhs.data(),
                                       fac.data()),
        "set_blockstructure");
}

// A0proj[n][s] -> packed vectors in Vin's layout (absent blocks are zero).
template <typename HT, typename PV>
void pass_projectors(vmps_pivot2* p, const HT& H, const PV& V, int64_t vsize) {
  if (H.A0proj.empty()) return;
  std::vector<double> A0(H.A0proj.size() * static_cast<size_t>(vsize), 0.0);
  for (size_t n = 0; n < H.A0proj.size(); ++n) {
    int64_t off = 0;
    for (size_t s = 0; s < V.data.size(); ++s)
      for (size_t q = 0; q < V.data[s].dim; ++q) {
        const auto& Vb = V.data[s].block[q];
        if (s < H.A0proj[n].size()) {
          const auto& P = H.A0proj[n][s];
          for (size_t r = 0; r < P.dim; ++r)
            if (P.in[r] == V.data[s].in[q] && P.out[r] == V.data[s].out[q] && P.block[r].size() == Vb.size())
              std::memcpy(&A0[n * vsize + off], P.block[r].data(), sizeof(double) * Vb.size());
        }
        off += Vb.size();
      }
  }
  check(vmps_pivot2_set_projectors(p, H.Epenalty, static_cast<int>(H.A0proj.size()), A0.data()),
        "set_projectors");
}

template <typename PV>
void pack(const PV& V, std::vector<double>& buf) {
  size_t off = 0;
  for (const auto& B : V.data)
    for (size_t q = 0; q < B.dim; ++q) {
      const auto& M = B.block[q];
      std::memcpy(buf.data() + off, M.data(), sizeof(double) * M.size());
      off += M.size();
    }
}

template <typename PV>
void unpack(const std::vector<double>& buf, PV& V) {
  size_t off = 0;
  for (auto& B : V.data)
    for (size_t q = 0; q < B.dim; ++q) {
      auto& M = B.block[q];
      std::memcpy(M.data(), buf.data() + off, sizeof(double) * M.size());
      off += M.size();
    }
}

template <typename Builder, typename PV>
PlanEntry& plan_for(const void* key, uint64_t fp, const PV& Vin, Builder&& build) {
  auto& cache = plan_cache();
  auto it = cache.find(key);
  if (it != cache.end() && it->second->fingerprint == fp) return *it->second;
  auto e = std::make_unique<PlanEntry>();
  e->fingerprint = fp;
  PivotHandle h;
  build(h);
  check(vmps_pivot2_vsize(h.p, &e->vsize), "vsize");
  check(vmps_plan_create(h.p, VMPS_PREC_FP64, &e->plan), "plan_create");
  e->vin.resize(e->vsize);
  e->vout.resize(e->vsize);
  (void)Vin;
  auto& ref = *e;
  cache[key] = std::move(e);
  return ref;
}

template <typename PV>
void run(PlanEntry& e, const PV& Vin, PV& Vout) {
  pack(Vin, e.vin);
  check(vmps_hxv(e.plan, e.vin.data(), e.vout.data()), "hxv");
  unpack(e.vout, Vout);
}

}  // namespace vmps_adapter

// ---- PivotMatrix2 -----------------------------------------------------------
template <typename Symmetry>
void OxV(const PivotMatrix2<Symmetry, double, double>& H, const PivotVector<Symmetry, double>& Vin,
         PivotVector<Symmetry, double>& Vout) {
  namespace A = vmps_adapter;
  const size_t nq = Symmetry::Nq;
  A::Fnv f;
  f.add(H.Terms.size());
  for (const auto& t : H.Terms) {
    A::fingerprint_env(f, t.L);
    A::fingerprint_env(f, t.R);
    A::fingerprint_W(f, t.W12);
    A::fingerprint_W(f, t.W34);
  }
  A::fingerprint_layout(f, Vin);
  A::fingerprint_tables(f, H);
  auto& e = A::plan_for(&H, f.h, Vin, [&](A::PivotHandle& h) {
    const auto& t0 = H.Terms.at(0);
    const auto k = A::kinds<Symmetry>();
    const auto q12 = A::flat_qs(t0.qloc12, nq), q34 = A::flat_qs(t0.qloc34, nq);
    const auto o12 = A::flat_qs(t0.qOp12, nq), o34 = A::flat_qs(t0.qOp34, nq);
    A::check(vmps_pivot2_create(static_cast<int>(nq), k.data(), static_cast<int>(t0.qloc12.size()), q12.data(),
                                static_cast<int>(t0.qloc34.size()), q34.data(), static_cast<int>(t0.qOp12.size()),
                                o12.data(), static_cast<int>(t0.qOp34.size()), o34.data(), &h.p),
             "pivot2_create");
    for (size_t t = 0; t < H.Terms.size(); ++t) {
      const auto& T = H.Terms[t];
      if (t > 0) {
        const auto a = A::flat_qs(T.qOp12, nq), b = A::flat_qs(T.qOp34, nq);
        A::check(vmps_pivot_add_term(h.p, static_cast<int>(T.qOp12.size()), a.data(),
                                     static_cast<int>(T.qOp34.size()), b.data(), nullptr),
                 "add_term");
      }
      A::add_env<Symmetry>(h.p, 0, T.L);
      A::add_env<Symmetry>(h.p, 1, T.R);
      A::add_W<Symmetry>(h.p, 12, T.W12);
      A::add_W<Symmetry>(h.p, 34, T.W34);
    }
    A::add_layout<Symmetry>(h.p, Vin);
    // the reference uses the stored tables only for a single term (DmrgPivotMatrix2.h:157-167)
    if (H.Terms.size() == 1 && !H.qlhs.empty())
      A::pass_blockstructure(h.p, H);
    else
      A::check(vmps_pivot2_precalc(h.p), "precalc");
    int64_t vs = 0;
    A::check(vmps_pivot2_vsize(h.p, &vs), "vsize");
    A::pass_projectors(h.p, H, Vin, vs);
  });
  A::run(e, Vin, Vout);
}

// ---- PivotMatrix1 -----------------------------------------------------------
template <typename Symmetry>
void OxV(const PivotMatrix1<Symmetry, double, double>& H, const PivotVector<Symmetry, double>& Vin,
         PivotVector<Symmetry, double>& Vout) {
  namespace A = vmps_adapter;
  const size_t nq = Symmetry::Nq;
  A::Fnv f;
  f.add(H.Terms.size());
  for (const auto& t : H.Terms) {
    A::fingerprint_env(f, t.L);
    A::fingerprint_env(f, t.R);
    A::fingerprint_W(f, t.W);
  }
  A::fingerprint_layout(f, Vin);
  A::fingerprint_tables(f, H);
  auto& e = A::plan_for(&H, f.h, Vin, [&](A::PivotHandle& h) {
    const auto& t0 = H.Terms.at(0);
    const auto k = A::kinds<Symmetry>();
    const auto ql = A::flat_qs(t0.qloc, nq), qo = A::flat_qs(t0.qOp, nq);
    A::check(vmps_pivot1_create(static_cast<int>(nq), k.data(), static_cast<int>(t0.qloc.size()), ql.data(),
                                static_cast<int>(t0.qOp.size()), qo.data(), &h.p),
             "pivot1_create");
    for (size_t t = 0; t < H.Terms.size(); ++t) {
      const auto& T = H.Terms[t];
      if (t > 0) {
        const auto a = A::flat_qs(T.qOp, nq);
        A::check(vmps_pivot_add_term(h.p, static_cast<int>(T.qOp.size()), a.data(), 0, nullptr, nullptr),
                 "add_term");
      }
      A::add_env<Symmetry>(h.p, 0, T.L);
      A::add_env<Symmetry>(h.p, 1, T.R);
      A::add_W<Symmetry>(h.p, 12, T.W);
    }
    A::add_layout<Symmetry>(h.p, Vin);
    if (H.Terms.size() == 1 && !H.qlhs.empty())
      A::pass_blockstructure(h.p, H);
    else
      A::check(vmps_pivot2_precalc(h.p), "precalc");
    int64_t vs = 0;
    A::check(vmps_pivot2_vsize(h.p, &vs), "vsize");
    A::pass_projectors(h.p, H, Vin, vs);
  });
  A::run(e, Vin, Vout);
}

// ---- PivotMatrix0 -----------------------------------------------------------
template <typename Symmetry>
void HxV(const PivotMatrix0<Symmetry, double, double>& H, const PivotVector<Symmetry, double>& Vin,
         PivotVector<Symmetry, double>& Vout) {
  namespace A = vmps_adapter;
  A::Fnv f;
  A::fingerprint_env(f, H.L);
  A::fingerprint_env(f, H.R);
  A::fingerprint_layout(f, Vin);
  auto& e = A::plan_for(&H, f.h, Vin, [&](A::PivotHandle& h) {
    const auto k = A::kinds<Symmetry>();
    A::check(vmps_pivot0_create(static_cast<int>(Symmetry::Nq), k.data(), &h.p), "pivot0_create");
    A::add_env<Symmetry>(h.p, 0, H.L);
    A::add_env<Symmetry>(h.p, 1, H.R);
    A::add_layout<Symmetry>(h.p, Vin);
    A::check(vmps_pivot2_precalc(h.p), "precalc");
  });
  Vout = Vin;  // Vout takes Vin's block layout (outerResize, DmrgPivotMatrix0.h:34)
  A::run(e, Vin, Vout);
}

// ---- contract_L / contract_R -------------------------------------------------
// The seven-argument calls of DmrgSolver::build_L / build_R
// (solvers/DmrgSolver.h:1942-1973) with RANDOMIZE = false and MODE = FULL
// resolve to these overloads (partial ordering considers only the passed
// arguments); calls that pass RANDOMIZE, MODE or a basis order map keep the
// reference's generic template.
namespace vmps_adapter {

struct EnvHandle {
  vmps_env* e = nullptr;
  ~EnvHandle() {
    if (e) vmps_env_destroy(e);
  }
};

template <typename Symmetry, typename TripodT, typename AV, typename WT, typename QV>
void contract_env(int which, const TripodT& Eold, const AV& Abra, const WT& W, const AV& Aket, const QV& qloc,
                  const QV& qOp, TripodT& Enew) {
  const size_t nq = Symmetry::Nq;
  const auto k = kinds<Symmetry>();
  const auto ql = flat_qs(qloc, nq), qo = flat_qs(qOp, nq);
  EnvHandle h;
  check(vmps_env_create(which, static_cast<int>(nq), k.data(), static_cast<int>(qloc.size()), ql.data(),
                        static_cast<int>(qOp.size()), qo.data(), &h.e),
        "env_create");
  std::vector<int64_t> rows, cols;
  std::vector<double> data;
  for (size_t q = 0; q < Eold.dim; ++q) {
    std::vector<int> qi, qout, qm;
    put_q(qi, Eold.index[q][0], nq);
    put_q(qout, Eold.index[q][1], nq);
    put_q(qm, Eold.index[q][2], nq);
    const size_t adim = Eold.block[q].shape()[0];
    rows.assign(adim, 0);
    cols.assign(adim, 0);
    data.clear();
    for (size_t a = 0; a < adim; ++a) {
      const auto& M = Eold.block[q][a][0];
      rows[a] = M.rows();
      cols[a] = M.cols();
      data.insert(data.end(), M.data(), M.data() + M.size());
    }
    check(vmps_env_add_old(h.e, qi.data(), qout.data(), qm.data(), static_cast<int>(adim), rows.data(),
                           cols.data(), data.data()),
          "env_add_old");
  }
  for (int bra = 1; bra >= 0; --bra) {
    const AV& A = bra ? Abra : Aket;
    for (size_t s = 0; s < A.size(); ++s)
      for (size_t q = 0; q < A[s].dim; ++q) {
        std::vector<int> qi, qout;
        put_q(qi, A[s].in[q], nq);
        put_q(qout, A[s].out[q], nq);
        const auto& M = A[s].block[q];
        check(vmps_env_add_A(h.e, bra, static_cast<int>(s), qi.data(), qout.data(), M.rows(), M.cols(), M.data()),
              "env_add_A");
      }
  }
  for (size_t s1 = 0; s1 < W.size(); ++s1)
    for (size_t s2 = 0; s2 < W[s1].size(); ++s2)
      for (size_t kk = 0; kk < W[s1][s2].size(); ++kk) {
        const auto& B = W[s1][s2][kk];
        for (size_t q = 0; q < B.dim; ++q) {
          const auto& M = B.block[q];
          std::vector<int> row, col, qi, qout;
          std::vector<double> val;
          using SM = typename std::decay<decltype(M)>::type;
          for (int o = 0; o < M.outerSize(); ++o)
            for (typename SM::InnerIterator it(M, o); it; ++it) {
              row.push_back(static_cast<int>(it.row()));
              col.push_back(static_cast<int>(it.col()));
              val.push_back(it.value());
            }
          put_q(qi, B.in[q], nq);
          put_q(qout, B.out[q], nq);
          check(vmps_env_add_W(h.e, static_cast<int>(s1), static_cast<int>(s2), static_cast<int>(kk), qi.data(),
                               qout.data(), static_cast<int>(M.rows()), static_cast<int>(M.cols()),
                               static_cast<int64_t>(val.size()), row.data(), col.data(), val.data()),
                "env_add_W");
        }
      }
  check(vmps_env_contract(h.e, VMPS_PREC_FP64), "env_contract");
  int64_t nsec = 0, nslots = 0, size = 0;
  check(vmps_env_result_size(h.e, &nsec, &nslots, &size), "env_result_size");
  std::vector<int> qi(nsec * nq), qout(nsec * nq), qm(nsec * nq), adim(nsec);
  rows.assign(nslots, 0);
  cols.assign(nslots, 0);
  data.assign(size, 0.0);
  check(vmps_env_result_get(h.e, qi.data(), qout.data(), qm.data(), adim.data(), rows.data(), cols.data(),
                            data.data()),
        "env_result_get");
  // Lnew / Rnew are cleared and rebuilt (DmrgContractions.h:46-47, 191-192);
  // sectors are pushed in the order the reference creates them
  Enew.clear();
  using QT = typename std::decay<decltype(Enew.index[0][0])>::type;
  using MA = typename std::decay<decltype(Enew.block[0])>::type;
  int64_t slot = 0, off = 0;
  for (int64_t s = 0; s < nsec; ++s) {
    std::array<QT, 3> quple;
    for (size_t i = 0; i < nq; ++i) {
      quple[0][i] = qi[s * nq + i];
      quple[1][i] = qout[s * nq + i];
      quple[2][i] = qm[s * nq + i];
    }
    MA blocks(boost::extents[adim[s]][1]);
    for (int a = 0; a < adim[s]; ++a, ++slot) {
      auto& M = blocks[a][0];
      M.resize(rows[slot], cols[slot]);
      if (rows[slot] * cols[slot] > 0) std::memcpy(M.data(), data.data() + off, sizeof(double) * M.size());
      off += rows[slot] * cols[slot];
    }
    Enew.push_back(quple, blocks);
  }
}

}  // namespace vmps_adapter

// The seven-argument contract_L / contract_R of the reference
// (tensors/DmrgContractions.h:34-44, 179-189) for real double blocks.  The
// environment and the MPS blocks share one matrix type here, which makes
// these templates more specialised than the reference's (where the Tripod's
// MatrixType2 is free); the enable_if removes them for any other scalar, so
// the complex calls of TDVP (solvers/TDVPPropagator.h:876-897) and the
// reference's extra-argument forms still bind the generic templates.
template <typename Symmetry, typename MatrixType, typename MpoMatrixType,
          typename std::enable_if<std::is_same<typename MatrixType::Scalar, double>::value, int>::type = 0>
void contract_L(const Tripod<Symmetry, MatrixType>& Lold, const std::vector<Biped<Symmetry, MatrixType>>& Abra,
                const std::vector<std::vector<std::vector<Biped<Symmetry, MpoMatrixType>>>>& W,
                const std::vector<Biped<Symmetry, MatrixType>>& Aket,
                const std::vector<qarray<Symmetry::Nq>>& qloc, const std::vector<qarray<Symmetry::Nq>>& qOp,
                Tripod<Symmetry, MatrixType>& Lnew) {
  vmps_adapter::contract_env<Symmetry>(0, Lold, Abra, W, Aket, qloc, qOp, Lnew);
}

template <typename Symmetry, typename MatrixType, typename MpoMatrixType,
          typename std::enable_if<std::is_same<typename MatrixType::Scalar, double>::value, int>::type = 0>
void contract_R(const Tripod<Symmetry, MatrixType>& Rold, const std::vector<Biped<Symmetry, MatrixType>>& Abra,
                const std::vector<std::vector<std::vector<Biped<Symmetry, MpoMatrixType>>>>& W,
                const std::vector<Biped<Symmetry, MatrixType>>& Aket,
                const std::vector<qarray<Symmetry::Nq>>& qloc, const std::vector<qarray<Symmetry::Nq>>& qOp,
                Tripod<Symmetry, MatrixType>& Rnew) {
  vmps_adapter::contract_env<Symmetry>(1, Rold, Abra, W, Aket, qloc, qOp, Rnew);
}
