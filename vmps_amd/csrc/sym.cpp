// Symmetry algebra and SU(2) recoupling coefficients; see sym.hpp.
#include "sym.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <stdexcept>

namespace vmps {

QnList Sym::reduce(const Qn& a, const Qn& b) const {
  // Per-component irreps, then the Cartesian product with the first
  // component outermost (S1xS2.h:195-210).
  int lo[kMaxNq] = {0, 0}, cnt[kMaxNq] = {1, 1}, step[kMaxNq] = {0, 0};
  for (int i = 0; i < nq; ++i) {
    if (kind[i] == kU1) {
      lo[i] = a[i] + b[i];
      cnt[i] = 1;
    } else {
      const int qmin = std::abs(a[i] - b[i]) + 1, qmax = std::abs(a[i] + b[i]) - 1;
      lo[i] = qmin;
      cnt[i] = qmax >= qmin ? (qmax - qmin) / 2 + 1 : 0;
      step[i] = 2;
    }
  }
  QnList out;
  const int n1 = nq > 1 ? cnt[1] : 1;
  if ((int64_t)cnt[0] * n1 > QnList::kCap) throw std::length_error("too many irreps in one tensor product");
  for (int x = 0; x < cnt[0]; ++x)
    for (int y = 0; y < n1; ++y) {
      Qn q;
      q[0] = lo[0] + step[0] * x;
      if (nq > 1) q[1] = lo[1] + step[1] * y;
      out.q[out.n++] = q;
    }
  return out;
}

bool Sym::validate3(const Qn& a, const Qn& b, const Qn& c) const {
  for (int i = 0; i < nq; ++i) {
    if (kind[i] == kU1) {
      if (a[i] + b[i] != c[i]) return false;
    } else {
      // SU2::triangle, SU2.h:547-554 (q = 2S+1 convention)
      if (!(c[i] - 1 >= std::abs(a[i] - b[i]) && c[i] - 1 <= a[i] + b[i] - 2)) return false;
    }
  }
  return true;
}

namespace {

// phase<Scalar>(q) of symmetry/functions.h:50-55, with C++ remainder semantics.
inline double phase(int q) { return (q % 2) ? -1.0 : 1.0; }

const long double* fact_table() {
  static long double tab[1200];
  static std::once_flag once;
  std::call_once(once, [] {
    tab[0] = 1.0L;
    for (int i = 1; i < 1200; ++i) tab[i] = tab[i - 1] * (long double)i;
  });
  return tab;
}

inline bool tri2(int a, int b, int c) {
  // triangle condition for doubled angular momenta
  if (a < 0 || b < 0 || c < 0) return false;
  if ((a + b + c) & 1) return false;
  return c >= std::abs(a - b) && c <= a + b;
}

inline long double delta2(int a, int b, int c) {
  const long double* f = fact_table();
  return std::sqrt(f[(a + b - c) / 2] * f[(a - b + c) / 2] * f[(-a + b + c) / 2] / f[(a + b + c) / 2 + 1]);
}

}  // namespace

double wigner6j_2j(int j1, int j2, int j3, int j4, int j5, int j6) {
  if (!tri2(j1, j2, j3) || !tri2(j1, j5, j6) || !tri2(j4, j2, j6) || !tri2(j4, j5, j3)) return 0.0;
  const long double* f = fact_table();
  int a1 = (j1 + j2 + j3) / 2, a2 = (j1 + j5 + j6) / 2, a3 = (j4 + j2 + j6) / 2, a4 = (j4 + j5 + j3) / 2;
  int b1 = (j1 + j2 + j4 + j5) / 2, b2 = (j2 + j3 + j5 + j6) / 2, b3 = (j3 + j1 + j6 + j4) / 2;
  int tmin = std::max(std::max(a1, a2), std::max(a3, a4));
  int tmax = std::min(b1, std::min(b2, b3));
  long double sum = 0.0L;
  for (int t = tmin; t <= tmax; ++t) {
    long double den = f[t - a1] * f[t - a2] * f[t - a3] * f[t - a4] * f[b1 - t] * f[b2 - t] * f[b3 - t];
    long double term = f[t + 1] / den;
    sum += (t & 1) ? -term : term;
  }
  long double pre = delta2(j1, j2, j3) * delta2(j1, j5, j6) * delta2(j4, j2, j6) * delta2(j4, j5, j3);
  return (double)(pre * sum);
}

double wigner9j_2j(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8, int j9) {
  // rows and columns must each satisfy the triangle rule
  if (!tri2(j1, j2, j3) || !tri2(j4, j5, j6) || !tri2(j7, j8, j9) || !tri2(j1, j4, j7) ||
      !tri2(j2, j5, j8) || !tri2(j3, j6, j9))
    return 0.0;
  int xmin = std::max(std::abs(j1 - j9), std::max(std::abs(j4 - j8), std::abs(j2 - j6)));
  int xmax = std::min(j1 + j9, std::min(j4 + j8, j2 + j6));
  long double sum = 0.0L;
  for (int x = xmin; x <= xmax; x += 2) {
    long double s = (long double)wigner6j_2j(j1, j4, j7, j8, j9, x) *
                    (long double)wigner6j_2j(j2, j5, j8, j4, x, j6) *
                    (long double)wigner6j_2j(j3, j6, j9, x, j1, j2);
    long double w = (long double)(x + 1);
    sum += (x & 1) ? -w * s : w * s;
  }
  return (double)sum;
}

double wigner6j_q(int q1, int q2, int q3, int q4, int q5, int q6) {
  return wigner6j_2j(q1 - 1, q2 - 1, q3 - 1, q4 - 1, q5 - 1, q6 - 1);
}

double wigner9j_q(int q1, int q2, int q3, int q4, int q5, int q6, int q7, int q8, int q9) {
  return wigner9j_2j(q1 - 1, q2 - 1, q3 - 1, q4 - 1, q5 - 1, q6 - 1, q7 - 1, q8 - 1, q9 - 1);
}

double Sym::coeff_dot(const Qn& q) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i)
    if (kind[i] == kSU2) out *= (double)q[i];
  return out;
}

#define VMPS_NINE_ARGS const Qn &q1, const Qn &q2, const Qn &q3, const Qn &q4, const Qn &q5, \
                       const Qn &q6, const Qn &q7, const Qn &q8, const Qn &q9

double Sym::coeff_HPsi(VMPS_NINE_ARGS) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i) {
    if (kind[i] != kSU2) continue;
    out *= wigner9j_q(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i], q7[i], q8[i], q9[i]) *
           std::sqrt((double)(q7[i] * q8[i] * q3[i] * q6[i]));
  }
  return out;
}

double Sym::coeff_tensorProd(VMPS_NINE_ARGS) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i) {
    if (kind[i] != kSU2) continue;
    out *= wigner9j_q(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i], q7[i], q8[i], q9[i]) *
           std::sqrt((double)(q7[i] * q8[i] * q3[i] * q6[i]));
  }
  return out;
}

double Sym::coeff_buildL(VMPS_NINE_ARGS) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i) {
    if (kind[i] != kSU2) continue;
    out *= wigner9j_q(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i], q7[i], q8[i], q9[i]) *
           std::sqrt((double)(q7[i] * q8[i] * q3[i] * q6[i]));
  }
  return out;
}

double Sym::coeff_buildR(VMPS_NINE_ARGS) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i) {
    if (kind[i] != kSU2) continue;
    out *= wigner9j_q(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i], q7[i], q8[i], q9[i]) *
           std::sqrt((double)(q7[i] * q8[i] * q3[i] * q6[i])) * (double)q9[i] / (double)q7[i];
  }
  return out;
}

#undef VMPS_NINE_ARGS

double Sym::coeff_Apair(const Qn& q1, const Qn& q2, const Qn& q3, const Qn& q4, const Qn& q5,
                        const Qn& q6) const {
  double out = 1.0;
  for (int i = 0; i < nq; ++i) {
    if (kind[i] != kSU2) continue;
    out *= wigner6j_q(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i]) * std::sqrt((double)(q3[i] * q6[i])) *
           phase((q1[i] + q2[i] + q4[i] + q5[i] - 4) / 2);
  }
  return out;
}

}  // namespace vmps
