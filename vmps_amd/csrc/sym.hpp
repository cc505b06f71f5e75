// Quantum numbers, symmetry algebra and SU(2) recoupling coefficients (host, FP64).
//
// The symmetry is a runtime value: up to two components, each U(1) or SU(2).
// This covers U1 (HeisenbergU1), U1xU1 (HubbardU1xU1), SU2 (HeisenbergSU2)
// and SU2xU1 (HubbardSU2xU1).  Conventions follow the reference:
//   * U(1) component: additive charge; validate {a,b,c} <=> a+b == c
//     (symmetry/U1.h:171-177, :477-482), flip(q) = -q (U1.h:63).
//   * SU(2) component: q = 2S+1; reduce = |q1-q2|+1 .. q1+q2-1 step 2
//     (symmetry/SU2.h:176-183), triangle rule (SU2.h:547-554), flip(q) = q.
//   * products: Cartesian reduction, component-wise validate, product of
//     coefficients (symmetry/S1xS2.h:195-210, :310-315, :650-671).
//   * ordering of quantum-number arrays is lexicographic (symmetry/qarray.h:60,
//     U1.h:465-473, S1xS2.h:701-708).
#pragma once
#include <array>
#include <cstdint>
#include <cstddef>
#include <functional>
#include <vector>

namespace vmps {

constexpr int kMaxNq = 2;

enum SymKind : int { kU1 = 1, kSU2 = 2 };

struct Qn {
  std::array<int, kMaxNq> v{{0, 0}};
  int operator[](int i) const { return v[i]; }
  int& operator[](int i) { return v[i]; }
  bool operator==(const Qn& o) const { return v == o.v; }
  bool operator!=(const Qn& o) const { return v != o.v; }
  bool operator<(const Qn& o) const { return v < o.v; }
};

struct QnHash {
  size_t operator()(const Qn& q) const {
    return std::hash<int64_t>()((int64_t(q.v[0]) << 32) ^ (uint32_t)q.v[1]);
  }
};

// Irreps of one tensor product, held inline (reduce is called in the inner
// loops of the block-structure searches, where heap vectors dominated).
struct QnList {
  static constexpr int kCap = 128;
  Qn q[kCap];
  int n = 0;
  const Qn* begin() const { return q; }
  const Qn* end() const { return q + n; }
  size_t size() const { return (size_t)n; }
  bool empty() const { return n == 0; }
  const Qn& operator[](size_t i) const { return q[i]; }
};

struct Sym {
  int nq = 1;
  int kind[kMaxNq] = {kU1, kU1};

  bool non_abelian() const {
    for (int i = 0; i < nq; ++i)
      if (kind[i] == kSU2) return true;
    return false;
  }
  Qn vacuum() const {
    Qn q;
    for (int i = 0; i < nq; ++i) q[i] = (kind[i] == kSU2) ? 1 : 0;
    return q;
  }
  Qn flip(const Qn& q) const {
    Qn r = q;
    for (int i = 0; i < nq; ++i)
      if (kind[i] == kU1) r[i] = -q[i];
    return r;
  }
  // reduceSilent(ql, qr): all irreps of ql (x) qr, first component outermost.
  QnList reduce(const Qn& a, const Qn& b) const;
  // validate(qarray3{a,b,c}): c in a (x) b.
  bool validate3(const Qn& a, const Qn& b, const Qn& c) const;
  // degeneracy of the irrep (product of the SU(2) dimensions).
  int degeneracy(const Qn& q) const {
    int d = 1;
    for (int i = 0; i < nq; ++i)
      if (kind[i] == kSU2) d *= q[i];
    return d;
  }

  // Coefficients (all 1 for abelian components).
  double coeff_dot(const Qn& q) const;  // SU2.h:284-290
  double coeff_HPsi(const Qn&, const Qn&, const Qn&, const Qn&, const Qn&, const Qn&,
                    const Qn&, const Qn&, const Qn&) const;  // SU2.h:508-519
  double coeff_tensorProd(const Qn&, const Qn&, const Qn&, const Qn&, const Qn&, const Qn&,
                          const Qn&, const Qn&, const Qn&) const;  // SU2.h:455-466
  double coeff_buildL(const Qn&, const Qn&, const Qn&, const Qn&, const Qn&, const Qn&,
                      const Qn&, const Qn&, const Qn&) const;  // SU2.h:481-492
  double coeff_buildR(const Qn&, const Qn&, const Qn&, const Qn&, const Qn&, const Qn&,
                      const Qn&, const Qn&, const Qn&) const;  // SU2.h:494-506
  double coeff_Apair(const Qn&, const Qn&, const Qn&, const Qn&, const Qn&, const Qn&) const;  // SU2.h:388-397
};

// Wigner symbols in the q = 2j+1 convention used by the reference's wrappers
// (symmetry/SU2Wrappers.h:72-94 pass q-1 = 2j to GSL).  Racah's single-sum
// formula for 6j, and the 6j-sum representation of 9j, evaluated in long double.
double wigner6j_q(int q1, int q2, int q3, int q4, int q5, int q6);
double wigner9j_q(int q1, int q2, int q3, int q4, int q5, int q6, int q7, int q8, int q9);
// Same symbols with integer 2j arguments (GSL convention).
double wigner6j_2j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6);
double wigner9j_2j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6, int tj7, int tj8, int tj9);

// Threshold under which a recoupling factor counts as zero and the block
// triple is dropped (the reference uses mynumeric_limits<double>::epsilon(),
// DmrgIndexGymnastics.h:558,646,780,812; its value lives in an external header).
constexpr double kCgcEps = 1e-14;

}  // namespace vmps
