// Minimal column-major dense matrices for builds without Eigen.
//
// include/nabo/nabo.h uses Eigen when <Eigen/Core> is available. Without it,
// this file supplies the handful of matrix operations libnabo's interface
// and its examples use (sizes, coefficient access, columns, Random/Constant,
// resize, Map), under namespace nabo_amd::dense, which nabo.h aliases as
// Eigen. It is not an Eigen implementation: no expression templates, no
// arithmetic beyond what those call sites need.
#ifndef NABO_AMD_DENSE_H
#define NABO_AMD_DENSE_H

#include <cassert>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace nabo_amd {
namespace dense {

constexpr int Dynamic = -1;
enum { Aligned = 1, Unaligned = 0 };

template <typename T>
class Column;

// Column-major matrix; Rows/Cols fix a size when not Dynamic.
template <typename T, int Rows = Dynamic, int Cols = Dynamic>
class Matrix {
 public:
  typedef T Scalar;
  typedef long Index;

  Matrix() : r_(Rows == Dynamic ? 0 : Rows), c_(Cols == Dynamic ? 0 : Cols) { storage(); }
  // Vector constructor (one dynamic dimension) or matrix constructor.
  explicit Matrix(Index n) : r_(Cols == 1 ? n : (Rows == Dynamic ? 1 : Rows)),
                             c_(Cols == 1 ? 1 : n) {
    storage();
  }
  Matrix(Index rows, Index cols) : r_(rows), c_(cols) { storage(); }

  Index rows() const { return r_; }
  Index cols() const { return c_; }
  Index size() const { return r_ * c_; }
  Index outerStride() const { return r_; }
  T* data() { return d_.data(); }
  const T* data() const { return d_.data(); }

  void resize(Index rows, Index cols) {
    r_ = rows;
    c_ = cols;
    storage();
  }
  void resize(Index n) {
    if (Cols == 1)
      resize(n, 1);
    else
      resize(1, n);
  }

  T& operator()(Index r, Index c) { return d_[size_t(c * r_ + r)]; }
  const T& operator()(Index r, Index c) const { return d_[size_t(c * r_ + r)]; }
  T& operator()(Index i) { return d_[size_t(i)]; }
  const T& operator()(Index i) const { return d_[size_t(i)]; }
  T& operator[](Index i) { return d_[size_t(i)]; }
  const T& operator[](Index i) const { return d_[size_t(i)]; }
  T& coeffRef(Index r, Index c) { return (*this)(r, c); }
  const T& coeff(Index r, Index c) const { return (*this)(r, c); }
  T& coeffRef(Index i) { return d_[size_t(i)]; }
  const T& coeff(Index i) const { return d_[size_t(i)]; }

  Column<T> col(Index c) { return Column<T>(d_.data() + c * r_, r_); }
  Column<const T> col(Index c) const { return Column<const T>(d_.data() + c * r_, r_); }

  template <typename Src>
  Matrix& operator=(const Src& src) {
    resize(src.rows(), src.cols());
    for (Index c = 0; c < c_; ++c)
      for (Index r = 0; r < r_; ++r) (*this)(r, c) = src.coeff(r, c);
    return *this;
  }
  Matrix& operator=(const Matrix&) = default;
  Matrix(const Matrix&) = default;

  void setZero() { std::fill(d_.begin(), d_.end(), T(0)); }
  void setConstant(T v) { std::fill(d_.begin(), d_.end(), v); }

  // Uniform in [-1, 1], as Eigen's Random.
  static Matrix Random(Index rows, Index cols) {
    Matrix m(rows, cols);
    for (auto& v : m.d_) v = T(2.0 * (double(std::rand()) / double(RAND_MAX)) - 1.0);
    return m;
  }
  static Matrix Random(Index n) {
    Matrix m(n);
    for (auto& v : m.d_) v = T(2.0 * (double(std::rand()) / double(RAND_MAX)) - 1.0);
    return m;
  }
  static Matrix Constant(Index rows, Index cols, T v) {
    Matrix m(rows, cols);
    m.setConstant(v);
    return m;
  }
  static Matrix Constant(Index n, T v) {
    Matrix m(n);
    m.setConstant(v);
    return m;
  }
  static Matrix Zero(Index rows, Index cols) { return Constant(rows, cols, T(0)); }
  static Matrix Zero(Index n) { return Constant(n, T(0)); }

 private:
  void storage() { d_.assign(size_t(r_ * c_), T()); }
  Index r_, c_;
  std::vector<T> d_;
};

// A contiguous column (or any strided vector view) of a matrix.
template <typename T>
class Column {
 public:
  typedef long Index;
  Column(T* p, Index n) : p_(p), n_(n) {}
  Index rows() const { return n_; }
  Index cols() const { return 1; }
  Index size() const { return n_; }
  T& operator[](Index i) const { return p_[i]; }
  T& operator()(Index i) const { return p_[i]; }
  T& coeffRef(Index i) const { return p_[i]; }
  T coeff(Index i) const { return p_[i]; }
  T coeff(Index r, Index) const { return p_[r]; }
  T* data() const { return p_; }

 private:
  T* p_;
  Index n_;
};

// Non-owning view of external column-major storage.
template <typename M, int Options = Unaligned>
class Map {
 public:
  typedef typename std::remove_const<typename M::Scalar>::type Scalar;
  typedef long Index;
  typedef typename std::conditional<std::is_const<M>::value, const Scalar*, Scalar*>::type Ptr;
  Map(Ptr p, Index rows, Index cols) : p_(p), r_(rows), c_(cols) {}
  Index rows() const { return r_; }
  Index cols() const { return c_; }
  Index outerStride() const { return r_; }
  Ptr data() const { return p_; }
  const Scalar& coeff(Index r, Index c) const { return p_[c * r_ + r]; }

 private:
  Ptr p_;
  Index r_, c_;
};

typedef Matrix<float, Dynamic, Dynamic> MatrixXf;
typedef Matrix<double, Dynamic, Dynamic> MatrixXd;
typedef Matrix<int, Dynamic, Dynamic> MatrixXi;
typedef Matrix<float, Dynamic, 1> VectorXf;
typedef Matrix<double, Dynamic, 1> VectorXd;
typedef Matrix<int, Dynamic, 1> VectorXi;
typedef Matrix<float, 3, Dynamic> Matrix3Xf;
typedef Matrix<double, 3, Dynamic> Matrix3Xd;

}  // namespace dense
}  // namespace nabo_amd

#endif  // NABO_AMD_DENSE_H
