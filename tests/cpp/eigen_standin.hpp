// A small stand-in for the part of Eigen's dense API that SurfATTPP's
// forward_for_event uses around the solver calls (src/preproc.cpp:88-131):
// dynamic column- and row-major matrices and vectors (Eigen::MatrixX,
// Eigen::VectorX, Matrix<..., RowMajor>), element-wise expressions without
// storage (unary minus, .array() products and differences), cwiseInverse,
// scalar products and VectorX::Ones.  Expressions convert implicitly into
// matrices as Eigen's do (through a common base), and matrices expose the
// members the shim looks for: rows(), cols(), size(), data(), IsRowMajor,
// innerStride(), outerStride(), operator()(i, j), operator[](i).  It lets the
// shim's templates be compiled against Eigen-shaped types in an image without
// Eigen; it is test code, not a substitute for Eigen.
#pragma once
#include <cstddef>
#include <utility>
#include <vector>

namespace Eigen {

using Index = std::ptrdiff_t;
constexpr int Dynamic = -1;
enum StorageOptions { ColMajor = 0, RowMajor = 1 };

template <class D>
struct DenseBase {
  const D& derived() const { return static_cast<const D&>(*this); }
};

template <class Op, class A>
struct CwiseUnary : DenseBase<CwiseUnary<Op, A>> {
  static constexpr bool IsRowMajor = A::IsRowMajor;
  A a;
  Op op;
  CwiseUnary(const A& a_, Op op_) : a(a_), op(op_) {}
  Index rows() const { return a.rows(); }
  Index cols() const { return a.cols(); }
  double operator()(Index i, Index j) const { return op(a(i, j)); }
  const CwiseUnary& array() const { return *this; }
};

template <class Op, class A, class B>
struct CwiseBinary : DenseBase<CwiseBinary<Op, A, B>> {
  static constexpr bool IsRowMajor = A::IsRowMajor;
  A a;
  B b;
  Op op;
  CwiseBinary(const A& a_, const B& b_, Op op_) : a(a_), b(b_), op(op_) {}
  Index rows() const { return a.rows(); }
  Index cols() const { return a.cols(); }
  double operator()(Index i, Index j) const { return op(a(i, j), b(i, j)); }
  const CwiseBinary& array() const { return *this; }
};

// A non-owning element view of a matrix (what .array() and expression
// operands hold, so that expressions never copy their matrices).
template <class M>
struct Ref : DenseBase<Ref<M>> {
  static constexpr bool IsRowMajor = M::IsRowMajor;
  const M* m;
  explicit Ref(const M& m_) : m(&m_) {}
  Index rows() const { return m->rows(); }
  Index cols() const { return m->cols(); }
  double operator()(Index i, Index j) const { return (*m)(i, j); }
  const Ref& array() const { return *this; }
};

struct NegOp {
  double operator()(double x) const { return -x; }
};
struct SubOp {
  double operator()(double x, double y) const { return x - y; }
};
struct MulOp {
  double operator()(double x, double y) const { return x * y; }
};

template <class T, int R, int C, int Opt = ((R == 1 && C != 1) ? RowMajor : ColMajor)>
class Matrix : public DenseBase<Matrix<T, R, C, Opt>> {
 public:
  static constexpr bool IsRowMajor = Opt == RowMajor;
  Matrix() = default;
  Matrix(Index r, Index c) : r_(r), c_(c), d_((size_t)(r * c)) {}
  explicit Matrix(Index n)
    requires(C == 1)
      : r_(n), c_(1), d_((size_t)n) {}
  template <class D>
  Matrix(const DenseBase<D>& e) {  // NOLINT: implicit, as Eigen's EigenBase constructor
    assign(e.derived());
  }
  template <class D>
  Matrix& operator=(const DenseBase<D>& e) {
    Matrix t;
    t.assign(e.derived());
    *this = std::move(t);
    return *this;
  }
  Matrix(const Matrix&) = default;
  Matrix(Matrix&&) = default;
  Matrix& operator=(const Matrix&) = default;
  Matrix& operator=(Matrix&&) = default;

  Index rows() const { return r_; }
  Index cols() const { return c_; }
  Index size() const { return r_ * c_; }
  T* data() { return d_.data(); }
  const T* data() const { return d_.data(); }
  Index innerStride() const { return 1; }
  Index outerStride() const { return IsRowMajor ? c_ : r_; }
  T& operator()(Index i, Index j) { return d_[at(i, j)]; }
  T operator()(Index i, Index j) const { return d_[at(i, j)]; }
  T& operator[](Index i) { return d_[(size_t)i]; }
  T operator[](Index i) const { return d_[(size_t)i]; }
  T& operator()(Index i) { return d_[(size_t)i]; }
  T operator()(Index i) const { return d_[(size_t)i]; }
  void resize(Index n) {
    r_ = (C == 1) ? n : 1;
    c_ = (C == 1) ? 1 : n;
    d_.resize((size_t)n);
  }
  void resize(Index r, Index c) {
    r_ = r;
    c_ = c;
    d_.resize((size_t)(r * c));
  }
  static Matrix Ones(Index n)
    requires(C == 1)
  {
    Matrix m(n);
    for (auto& x : m.d_) x = T(1);
    return m;
  }

  Matrix cwiseInverse() const {
    Matrix m = *this;
    for (auto& x : m.d_) x = T(1) / x;
    return m;
  }
  friend Matrix operator*(const Matrix& m, T s) {
    Matrix o = m;
    for (auto& x : o.d_) x *= s;
    return o;
  }
  CwiseUnary<NegOp, Ref<Matrix>> operator-() const { return {Ref<Matrix>(*this), NegOp{}}; }
  Ref<Matrix> array() const { return Ref<Matrix>(*this); }

 private:
  size_t at(Index i, Index j) const { return (size_t)(IsRowMajor ? i * c_ + j : i + j * r_); }
  template <class E>
  void assign(const E& e) {
    r_ = e.rows();
    c_ = e.cols();
    d_.resize((size_t)(r_ * c_));
    for (Index j = 0; j < c_; j++)
      for (Index i = 0; i < r_; i++) d_[at(i, j)] = T(e(i, j));
  }
  Index r_ = 0, c_ = 0;
  std::vector<T> d_;
};

template <class T>
using MatrixX = Matrix<T, Dynamic, Dynamic>;
template <class T>
using VectorX = Matrix<T, Dynamic, 1>;

// Operands of the element-wise operators: matrices by reference, expressions by value.
template <class D>
auto hold(const DenseBase<D>& e) {
  if constexpr (requires { e.derived().data(); }) return Ref<D>(e.derived());
  else return e.derived();
}

template <class A, class B>
auto operator-(const DenseBase<A>& a, const DenseBase<B>& b) {
  return CwiseBinary<SubOp, decltype(hold(a)), decltype(hold(b))>(hold(a), hold(b), SubOp{});
}
template <class A, class B>
auto operator*(const DenseBase<A>& a, const DenseBase<B>& b) {
  return CwiseBinary<MulOp, decltype(hold(a)), decltype(hold(b))>(hold(a), hold(b), MulOp{});
}
template <class A>
  requires(!requires(const A& x) { x.data(); })
auto operator-(const DenseBase<A>& a) {
  return CwiseUnary<NegOp, A>(a.derived(), NegOp{});
}

}  // namespace Eigen
