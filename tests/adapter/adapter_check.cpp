// Drives the reference-shaped HxV(PivotMatrix, PivotVector, PivotVector&)
// calls through include/vmps_hxv/pivot_adapter.hpp (test program).
//
//   adapter_check <pivot.txt> <vout.txt>
//
// reads a pivot (format written by tests/test_adapter.py), builds the
// reference-shaped objects, calls HxV and writes Vout; with a leading 10 / 11
// instead of the pivot's site count it reads Lold / Rold, Abra, Aket and W,
// calls contract_L / contract_R and writes the new environment.  Exit status: 0 on
// success, 2 when the generic (non-adapter) template was selected, 3 when the
// adapter reported a library error (message on stderr), 1 on bad input.
#include <cstdlib>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "mock_reference.hpp"
#include "vmps_hxv/pivot_adapter.hpp"

namespace {
struct Reader {
  std::ifstream in;
  explicit Reader(const char* path) : in(path) {
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  }
  long long i() {
    long long x;
    if (!(in >> x)) throw std::runtime_error("bad input (int)");
    return x;
  }
  double d() {
    double x;
    if (!(in >> x)) throw std::runtime_error("bad input (double)");
    return x;
  }
};

template <typename Sym>
typename Sym::qType rq(Reader& r) {
  typename Sym::qType q;
  for (size_t i = 0; i < Sym::Nq; ++i) q[i] = (int)r.i();
  return q;
}

template <typename Sym>
std::vector<qarray<Sym::Nq>> rqs(Reader& r) {
  std::vector<qarray<Sym::Nq>> v(r.i());
  for (auto& q : v) q = rq<Sym>(r);
  return v;
}

template <typename Sym>
Tripod<Sym, MatrixXd> read_env(Reader& r) {
  Tripod<Sym, MatrixXd> E;
  E.dim = r.i();
  for (size_t q = 0; q < E.dim; ++q) {
    std::array<typename Sym::qType, 3> idx = {rq<Sym>(r), rq<Sym>(r), rq<Sym>(r)};
    E.index.push_back(idx);
    MultiArray2<MatrixXd> b;
    size_t adim = r.i();
    b.s[0] = adim;
    b.d.resize(adim);
    for (size_t a = 0; a < adim; ++a) {
      long long rows = r.i(), cols = r.i();
      MatrixXd m(rows, cols);
      for (long long x = 0; x < rows * cols; ++x) m.v[x] = r.d();
      b.d[a].push_back(m);
    }
    E.block.push_back(b);
  }
  return E;
}

template <typename Sym>
WTensor<Sym> read_W(Reader& r, size_t nloc, size_t nop) {
  WTensor<Sym> W(nloc, std::vector<std::vector<Biped<Sym, SparseMatrix<double>>>>(
                           nloc, std::vector<Biped<Sym, SparseMatrix<double>>>(nop)));
  long long nb = r.i();
  for (long long b = 0; b < nb; ++b) {
    size_t s1 = r.i(), s2 = r.i(), k = r.i();
    auto qi = rq<Sym>(r), qo = rq<Sym>(r);
    SparseMatrix<double> m;
    m.r = (int)r.i();
    m.c = (int)r.i();
    m.colv.resize(m.c);
    long long nnz = r.i();
    for (long long x = 0; x < nnz; ++x) {
      int row = (int)r.i(), col = (int)r.i();
      double v = r.d();
      m.colv[col].push_back({row, v});  // written in column-major order
    }
    W[s1][s2][k].push_back(qi, qo, m);
  }
  return W;
}

template <typename Sym>
PivotVector<Sym, double> read_layout(Reader& r) {
  PivotVector<Sym, double> V;
  V.data.resize(r.i());
  for (auto& B : V.data) {
    long long nb = r.i();
    for (long long b = 0; b < nb; ++b) {
      auto qi = rq<Sym>(r), qo = rq<Sym>(r);
      long long rows = r.i(), cols = r.i();
      B.push_back(qi, qo, MatrixXd(rows, cols));
    }
  }
  return V;
}

template <typename Sym>
void fill(Reader& r, PivotVector<Sym, double>& V) {
  for (auto& B : V.data)
    for (auto& m : B.block)
      for (auto& x : m.v) x = r.d();
}

template <typename HT, size_t W>
void read_bs(Reader& r, HT& H) {
  long long nl = r.i();
  for (long long l = 0; l < nl; ++l) {
    H.qlhs.push_back({(size_t)r.i(), (size_t)r.i()});
    long long nr = r.i();
    H.qrhs.emplace_back();
    H.factor_cgcs.emplace_back();
    for (long long p = 0; p < nr; ++p) {
      std::array<size_t, W> e;
      for (size_t c = 0; c < 12; ++c) {
        size_t x = (size_t)r.i();
        if (c < W) e[c] = x;
      }
      H.qrhs.back().push_back(e);
      H.factor_cgcs.back().push_back(r.d());
    }
  }
}

template <typename Sym>
void write(const char* path, const PivotVector<Sym, double>& V) {
  FILE* f = std::fopen(path, "w");
  if (!f) throw std::runtime_error("cannot write output");
  for (auto& B : V.data)
    for (auto& m : B.block)
      for (double x : m.v) std::fprintf(f, "%.17g\n", x);
  std::fclose(f);
}

// site tensor A[s]: per s, nblocks of (qin qout rows cols data)
template <typename Sym>
std::vector<Biped<Sym, MatrixXd>> read_A(Reader& r) {
  std::vector<Biped<Sym, MatrixXd>> A(r.i());
  for (auto& B : A) {
    long long nb = r.i();
    for (long long b = 0; b < nb; ++b) {
      auto qi = rq<Sym>(r), qo = rq<Sym>(r);
      long long rows = r.i(), cols = r.i();
      MatrixXd m(rows, cols);
      for (auto& x : m.v) x = r.d();
      B.push_back(qi, qo, m);
    }
  }
  return A;
}

// contract_L (which = 0) / contract_R (which = 1) through the seven-argument call
template <typename Sym>
void run_env(Reader& r, int which, const char* outpath) {
  auto qloc = rqs<Sym>(r);
  auto qOp = rqs<Sym>(r);
  auto Eold = read_env<Sym>(r);
  auto Abra = read_A<Sym>(r);
  auto Aket = read_A<Sym>(r);
  auto W = read_W<Sym>(r, qloc.size(), qOp.size());
  Tripod<Sym, MatrixXd> Enew;
  if (which == 0)
    contract_L(Eold, Abra, W, Aket, qloc, qOp, Enew);
  else
    contract_R(Eold, Abra, W, Aket, qloc, qOp, Enew);
  FILE* f = std::fopen(outpath, "w");
  if (!f) throw std::runtime_error("cannot write output");
  std::fprintf(f, "%zu\n", Enew.dim);
  for (size_t q = 0; q < Enew.dim; ++q) {
    for (int leg = 0; leg < 3; ++leg)
      for (size_t i = 0; i < Sym::Nq; ++i) std::fprintf(f, "%d ", Enew.index[q][leg][i]);
    std::fprintf(f, "%zu\n", Enew.block[q].shape()[0]);
    for (size_t a = 0; a < Enew.block[q].shape()[0]; ++a) {
      const auto& M = Enew.block[q][a][0];
      std::fprintf(f, "%lld %lld", (long long)M.rows(), (long long)M.cols());
      for (double x : M.v) std::fprintf(f, " %.17g", x);
      std::fprintf(f, "\n");
    }
  }
  std::fclose(f);
}

template <typename Sym>
void run(Reader& r, int nsite, const char* outpath) {
  if (nsite >= 10) {
    run_env<Sym>(r, nsite - 10, outpath);
    return;
  }
  PivotVector<Sym, double> Vin, Vout;
  if (nsite == 2) {
    PivotMatrix2<Sym, double, double> H;
    auto qloc12 = rqs<Sym>(r);
    auto qloc34 = rqs<Sym>(r);
    long long nt = r.i();
    for (long long t = 0; t < nt; ++t) {
      PivotMatrix2Terms<Sym, double, double> T;
      T.qloc12 = qloc12;
      T.qloc34 = qloc34;
      T.qOp12 = rqs<Sym>(r);
      T.qOp34 = rqs<Sym>(r);
      T.L = read_env<Sym>(r);
      T.R = read_env<Sym>(r);
      T.W12 = read_W<Sym>(r, qloc12.size(), T.qOp12.size());
      T.W34 = read_W<Sym>(r, qloc34.size(), T.qOp34.size());
      H.Terms.push_back(T);
    }
    Vin = read_layout<Sym>(r);
    fill(r, Vin);
    read_bs<decltype(H), 12>(r, H);
    HxV(H, Vin, Vout);
    HxV(H, Vin, Vout);  // second call reuses the cached plan
    if (std::getenv("VMPS_CHECK_INPLACE")) {
      // change one interior value of an environment block in place: the
      // cached plan must not be served, the result must equal a fresh plan's
      bool changed = false;
      for (auto& T : H.Terms) {
        for (size_t q = 0; q < T.L.dim && !changed; ++q)
          for (size_t a = 0; a < T.L.block[q].shape()[0] && !changed; ++a) {
            auto& M = T.L.block[q][a][0];
            if (M.size() >= 3) {
              M.v[M.size() / 2] += 0.5;
              changed = true;
            }
          }
        if (changed) break;
      }
      if (!changed) throw std::runtime_error("no environment block with three entries");
      auto before = Vout;
      HxV(H, Vin, Vout);
      vmps_adapter::invalidate();
      PivotVector<Sym, double> fresh;
      HxV(H, Vin, fresh);
      bool same_as_fresh = true, same_as_before = true;
      for (size_t s = 0; s < Vout.data.size(); ++s)
        for (size_t q = 0; q < Vout.data[s].dim; ++q)
          for (int64_t i = 0; i < Vout.data[s].block[q].size(); ++i) {
            const double x = Vout.data[s].block[q].v[i];
            if (x != fresh.data[s].block[q].v[i]) same_as_fresh = false;
            if (x != before.data[s].block[q].v[i]) same_as_before = false;
          }
      if (!same_as_fresh || same_as_before) {
        std::cerr << "stale plan after an in-place environment change\n";
        std::exit(4);
      }
    }
  } else if (nsite == 1) {
    PivotMatrix1<Sym, double, double> H;
    auto qloc = rqs<Sym>(r);
    long long nt = r.i();
    for (long long t = 0; t < nt; ++t) {
      PivotMatrix1Terms<Sym, double, double> T;
      T.qloc = qloc;
      T.qOp = rqs<Sym>(r);
      T.L = read_env<Sym>(r);
      T.R = read_env<Sym>(r);
      T.W = read_W<Sym>(r, qloc.size(), T.qOp.size());
      H.Terms.push_back(T);
    }
    Vin = read_layout<Sym>(r);
    fill(r, Vin);
    read_bs<decltype(H), 6>(r, H);
    HxV(H, Vin, Vout);
  } else {
    PivotMatrix0<Sym, double, double> H;
    H.L = read_env<Sym>(r);
    H.R = read_env<Sym>(r);
    Vin = read_layout<Sym>(r);
    fill(r, Vin);
    HxV(H, Vin, Vout);
  }
  write(outpath, Vout);
}
}  // namespace

// The complex calls of TDVP (solvers/TDVPPropagator.h:876-897) must bind the
// reference's generic contract_L / contract_R, not the FP64 adapter.
int complex_dispatch() {
  using Sym = MockSym<1, S>;
  Tripod<Sym, MatrixXcd> Eold, Enew;
  std::vector<Biped<Sym, MatrixXcd>> A;
  WTensor<Sym> W;
  std::vector<qarray<1>> qloc, qOp;
  int generic = 0;
  try {
    contract_L(Eold, A, W, A, qloc, qOp, Enew);
  } catch (const GenericSelected&) {
    ++generic;
  }
  try {
    contract_R(Eold, A, W, A, qloc, qOp, Enew);
  } catch (const GenericSelected&) {
    ++generic;
  }
  return generic == 2 ? 0 : 5;
}

int main(int argc, char** argv) {
  if (argc == 2 && std::string(argv[1]) == "--complex-dispatch") return complex_dispatch();
  if (argc != 3) {
    std::cerr << "usage: adapter_check <pivot.txt> <vout.txt>\n";
    return 1;
  }
  try {
    Reader r(argv[1]);
    int nsite = (int)r.i();
    int nq = (int)r.i();
    std::vector<int> kinds(nq);
    for (auto& k : kinds) k = (int)r.i();
    // kinds: 1 = U(1) (as KIND::M / KIND::N), 2 = SU(2) (KIND::S)
    if (nq == 1 && kinds[0] == 1)
      run<MockSym<1, M>>(r, nsite, argv[2]);
    else if (nq == 1 && kinds[0] == 2)
      run<MockSym<1, S>>(r, nsite, argv[2]);
    else if (nq == 2 && kinds[0] == 1 && kinds[1] == 1)
      run<MockSym<2, N, M>>(r, nsite, argv[2]);
    else if (nq == 2 && kinds[0] == 2 && kinds[1] == 1)
      run<MockSym<2, S, N>>(r, nsite, argv[2]);
    else
      throw std::runtime_error("unsupported symmetry in test input");
  } catch (const GenericSelected& e) {
    std::cerr << e.what() << "\n";
    return 2;
  } catch (const std::runtime_error& e) {
    std::cerr << e.what() << "\n";
    return std::string(e.what()).rfind("vmps_hxv adapter:", 0) == 0 ? 3 : 1;
  }
  return 0;
}
