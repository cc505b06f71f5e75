// Test-only stand-ins for the reference types the OxV adapter touches, with
// the reference's member names and meaning (the real headers need Eigen and
// Boost, which are not available here):
//   qarray<Nq>                 symmetry/qarray.h:25-43     data, operator[]
//   Biped<Sym, Matrix>         tensors/Biped.h:62-104      dim, in, out, block
//   Multipede<3> / Tripod      tensors/Multipede.h:31-160  dim, index, block[q][a][0], block[q].shape()
//   SparseMatrix::InnerIterator (Eigen)                    outerSize, row, col, value
//   PivotVector                pivot/DmrgPivotVector.h:8-117  data
//   PivotMatrix2 / 1 / 0       pivot/DmrgPivotMatrix{2,1,0}.h  Terms, qlhs, qrhs, factor_cgcs, A0proj, Epenalty
//   contract_L / contract_R    tensors/DmrgContractions.h:34-44,179-189 (generic signature with defaults)
// plus generic HxV / OxV templates shaped like the reference's (HxV assigns
// Vout = Vin and calls OxV unqualified).  The generic OxV only reports that it
// was selected: the adapter's overloads must win overload resolution.
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <tuple>
#include <utility>
#include <vector>

// boost::extents[n0][n1] for the multi_array stand-in below
namespace boost {
struct extent_gen2 {
  size_t n0, n1;
};
struct extent_gen1 {
  size_t n0;
  extent_gen2 operator[](size_t n1) const { return {n0, n1}; }
};
struct extent_gen {
  extent_gen1 operator[](size_t n0) const { return {n0}; }
};
static const extent_gen extents{};
}  // namespace boost

enum KIND { S, Salt, T, N, M, Nup, Ndn, Nparity, K };

template <size_t Nq>
struct qarray {
  std::array<int, Nq> data{};
  int& operator[](size_t i) { return data[i]; }
  const int& operator[](size_t i) const { return data[i]; }
};
template <size_t Nq>
bool operator==(const qarray<Nq>& a, const qarray<Nq>& b) {
  return a.data == b.data;
}

// Symmetry with Nq components of the given kinds
template <size_t Nq_, KIND k0, KIND k1 = N>
struct MockSym {
  static constexpr size_t Nq = Nq_;
  typedef qarray<Nq> qType;
  static std::array<KIND, Nq> kind() {
    std::array<KIND, Nq> k;
    k[0] = k0;
    if (Nq > 1) k[Nq - 1] = k1;
    return k;
  }
};

struct MatrixXd {
  typedef double Scalar;
  int64_t r = 0, c = 0;
  std::vector<double> v;
  MatrixXd() = default;
  MatrixXd(int64_t rows, int64_t cols) : r(rows), c(cols), v(rows * cols, 0.0) {}
  int64_t rows() const { return r; }
  int64_t cols() const { return c; }
  int64_t size() const { return r * c; }
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  void resize(int64_t rows, int64_t cols) {
    r = rows;
    c = cols;
    v.assign(rows * cols, 0.0);
  }
};

// complex blocks (the TDVP instantiation, Scalar = complex<double>); only
// overload resolution looks at it
struct MatrixXcd {
  typedef std::complex<double> Scalar;
  int64_t r = 0, c = 0;
  std::vector<Scalar> v;
  int64_t rows() const { return r; }
  int64_t cols() const { return c; }
  int64_t size() const { return r * c; }
};

// column-major sparse matrix with Eigen's InnerIterator interface
template <typename Scalar>
struct SparseMatrix {
  int r = 0, c = 0;
  std::vector<std::vector<std::pair<int, Scalar>>> colv;  // per column: (row, value), rows ascending
  int rows() const { return r; }
  int cols() const { return c; }
  int outerSize() const { return c; }
  struct InnerIterator {
    const SparseMatrix* m;
    int col_;
    size_t i = 0;
    InnerIterator(const SparseMatrix& mat, int outer) : m(&mat), col_(outer) {}
    explicit operator bool() const { return i < m->colv[col_].size(); }
    InnerIterator& operator++() {
      ++i;
      return *this;
    }
    int row() const { return m->colv[col_][i].first; }
    int col() const { return col_; }
    Scalar value() const { return m->colv[col_][i].second; }
  };
};

template <typename Symmetry, typename MatrixType>
struct Biped {
  size_t dim = 0;
  std::vector<typename Symmetry::qType> in, out;
  std::vector<MatrixType> block;
  void push_back(const typename Symmetry::qType& qin, const typename Symmetry::qType& qout, const MatrixType& m) {
    in.push_back(qin);
    out.push_back(qout);
    block.push_back(m);
    ++dim;
  }
};

// boost::multi_array<MatrixType, 2> of shape [adim][1]
template <typename MatrixType>
struct MultiArray2 {
  std::vector<std::vector<MatrixType>> d;
  size_t s[2] = {0, 1};
  MultiArray2() = default;
  MultiArray2(const boost::extent_gen2& e) : d(e.n0, std::vector<MatrixType>(e.n1)), s{e.n0, e.n1} {}
  const size_t* shape() const { return s; }
  std::vector<MatrixType>& operator[](size_t a) { return d[a]; }
  const std::vector<MatrixType>& operator[](size_t a) const { return d[a]; }
};

template <size_t Nlegs, typename Symmetry, typename MatrixType>
struct Multipede {
  size_t dim = 0;
  std::vector<std::array<typename Symmetry::qType, Nlegs>> index;
  std::vector<MultiArray2<MatrixType>> block;
  void clear() {
    dim = 0;
    index.clear();
    block.clear();
  }
  void push_back(std::array<typename Symmetry::qType, Nlegs> quple, const MultiArray2<MatrixType>& M) {
    index.push_back(quple);
    block.push_back(M);
    ++dim;
  }
};
template <typename Symmetry, typename MatrixType>
using Tripod = Multipede<3, Symmetry, MatrixType>;

template <typename Symmetry, typename Scalar>
struct PivotVector {
  std::vector<Biped<Symmetry, MatrixXd>> data;
};

template <typename Symmetry>
using WTensor = std::vector<std::vector<std::vector<Biped<Symmetry, SparseMatrix<double>>>>>;

template <typename Symmetry, typename Scalar, typename MpoScalar = double>
struct PivotMatrix2Terms {
  Tripod<Symmetry, MatrixXd> L, R;
  WTensor<Symmetry> W12, W34;
  std::vector<qarray<Symmetry::Nq>> qloc12, qloc34, qOp12, qOp34;
};
template <typename Symmetry, typename Scalar, typename MpoScalar = double>
struct PivotMatrix2 {
  std::vector<PivotMatrix2Terms<Symmetry, Scalar, MpoScalar>> Terms;
  std::vector<std::array<size_t, 2>> qlhs;
  std::vector<std::vector<std::array<size_t, 12>>> qrhs;
  std::vector<std::vector<Scalar>> factor_cgcs;
  std::vector<std::vector<Biped<Symmetry, MatrixXd>>> A0proj;
  double Epenalty = 0;
};

template <typename Symmetry, typename Scalar, typename MpoScalar = double>
struct PivotMatrix1Terms {
  Tripod<Symmetry, MatrixXd> L, R;
  WTensor<Symmetry> W;
  std::vector<qarray<Symmetry::Nq>> qloc, qOp;
};
template <typename Symmetry, typename Scalar, typename MpoScalar = double>
struct PivotMatrix1 {
  std::vector<PivotMatrix1Terms<Symmetry, Scalar, MpoScalar>> Terms;
  std::vector<std::array<size_t, 2>> qlhs;
  std::vector<std::vector<std::array<size_t, 6>>> qrhs;
  std::vector<std::vector<Scalar>> factor_cgcs;
  std::vector<std::vector<Biped<Symmetry, MatrixXd>>> A0proj;
  double Epenalty = 0;
};

template <typename Symmetry, typename Scalar, typename MpoScalar = double>
struct PivotMatrix0 {
  Tripod<Symmetry, MatrixXd> L, R;
};

struct GenericSelected : std::logic_error {
  using std::logic_error::logic_error;
};

// generic templates in the reference's shape
template <typename Symmetry, typename Scalar, typename MpoScalar>
void OxV(const PivotMatrix2<Symmetry, Scalar, MpoScalar>&, const PivotVector<Symmetry, Scalar>&,
         PivotVector<Symmetry, Scalar>&) {
  throw GenericSelected("generic OxV(PivotMatrix2) selected");
}
template <typename Symmetry, typename Scalar, typename MpoScalar>
void HxV(const PivotMatrix2<Symmetry, Scalar, MpoScalar>& H, const PivotVector<Symmetry, Scalar>& Vin,
         PivotVector<Symmetry, Scalar>& Vout) {
  Vout = Vin;
  OxV(H, Vin, Vout);
}
template <typename Symmetry, typename Scalar, typename MpoScalar>
void OxV(const PivotMatrix1<Symmetry, Scalar, MpoScalar>&, const PivotVector<Symmetry, Scalar>&,
         PivotVector<Symmetry, Scalar>&) {
  throw GenericSelected("generic OxV(PivotMatrix1) selected");
}
template <typename Symmetry, typename Scalar, typename MpoScalar>
void HxV(const PivotMatrix1<Symmetry, Scalar, MpoScalar>& H, const PivotVector<Symmetry, Scalar>& Vin,
         PivotVector<Symmetry, Scalar>& Vout) {
  Vout = Vin;
  OxV(H, Vin, Vout);
}
template <typename Symmetry, typename Scalar, typename MpoScalar>
void HxV(const PivotMatrix0<Symmetry, Scalar, MpoScalar>&, const PivotVector<Symmetry, Scalar>&,
         PivotVector<Symmetry, Scalar>&) {
  throw GenericSelected("generic HxV(PivotMatrix0) selected");
}

enum CONTRACT_LR_MODE { FULL, FIXED_ROWS, FIXED_COLS };

template <typename Symmetry, typename MatrixType, typename MatrixType2, typename MpoMatrixType>
void contract_L(const Tripod<Symmetry, MatrixType2>&, const std::vector<Biped<Symmetry, MatrixType>>&,
                const std::vector<std::vector<std::vector<Biped<Symmetry, MpoMatrixType>>>>&,
                const std::vector<Biped<Symmetry, MatrixType>>&, const std::vector<qarray<Symmetry::Nq>>&,
                const std::vector<qarray<Symmetry::Nq>>&, Tripod<Symmetry, MatrixType2>&, bool RANDOMIZE = false,
                std::tuple<CONTRACT_LR_MODE, size_t> MODE_input = std::make_pair(FULL, 0),
                const void* basis_order_map = nullptr) {
  throw GenericSelected("generic contract_L selected");
}
template <typename Symmetry, typename MatrixType, typename MatrixType2, typename MpoMatrixType>
void contract_R(const Tripod<Symmetry, MatrixType2>&, const std::vector<Biped<Symmetry, MatrixType>>&,
                const std::vector<std::vector<std::vector<Biped<Symmetry, MpoMatrixType>>>>&,
                const std::vector<Biped<Symmetry, MatrixType>>&, const std::vector<qarray<Symmetry::Nq>>&,
                const std::vector<qarray<Symmetry::Nq>>&, Tripod<Symmetry, MatrixType2>&, bool RANDOMIZE = false,
                std::tuple<CONTRACT_LR_MODE, size_t> MODE_input = std::make_pair(FULL, 0),
                const void* basis_order_map = nullptr) {
  throw GenericSelected("generic contract_R selected");
}
