/// @file tinymatrix.hpp
/// Fixed-size row-major matrix used for gate matrices, with the interface of the
/// reference's iqs::TinyMatrix (include/tinymatrix.hpp): m(i, j), m[i][j], numRows/numCols,
/// a `name` label, construction from nested initializer lists or C arrays.
#pragma once

#include <cstring>
#include <initializer_list>
#include <iostream>
#include <string>

namespace iqs {

template <class ValueType, unsigned M, unsigned N = M, unsigned align = alignof(ValueType)>
class TinyMatrix
{
 public:
  using value_type = ValueType;
  using size_type = unsigned;
  using RowType = ValueType[N];

  TinyMatrix()
  {
    static_assert(M * N != 0, "empty TinyMatrix");
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < N; ++j) d_[i][j] = ValueType(0);
  }
  template <class U>
  TinyMatrix(U init[M][N])
  {
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < N; ++j) d_[i][j] = init[i][j];
  }
  template <class U>
  TinyMatrix(std::initializer_list<std::initializer_list<U>> const& rows)
  {
    unsigned i = 0;
    for (auto const& r : rows) {
      unsigned j = 0;
      for (auto const& e : r) d_[i][j++] = e;
      ++i;
    }
  }
  template <class U, unsigned a2>
  TinyMatrix(TinyMatrix<U, M, N, a2> const& o)
  {
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < N; ++j) d_[i][j] = o(i, j);
    name = o.name;
  }
  TinyMatrix(TinyMatrix const&) = default;
  TinyMatrix& operator=(TinyMatrix const&) = default;

  ValueType& operator()(unsigned i, unsigned j) { return d_[i][j]; }
  ValueType const& operator()(unsigned i, unsigned j) const { return d_[i][j]; }
  RowType& operator[](unsigned i) { return d_[i]; }
  RowType const& operator[](unsigned i) const { return d_[i]; }

  static constexpr unsigned numRows() { return M; }
  static constexpr unsigned numCols() { return N; }
  static constexpr unsigned numElements() { return M * N; }
  ValueType* getPtrToData() { return &d_[0][0]; }
  ValueType const* getPtrToData() const { return &d_[0][0]; }

  bool operator==(TinyMatrix const& o) const
  {
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < N; ++j)
        if (!(d_[i][j] == o.d_[i][j])) return false;
    return true;
  }
  bool operator!=(TinyMatrix const& o) const { return !(*this == o); }

  std::string name;

 private:
  alignas(align) ValueType d_[M][N];
};

template <class V, unsigned M, unsigned N, unsigned a>
std::ostream& operator<<(std::ostream& os, TinyMatrix<V, M, N, a> const& m)
{
  for (unsigned i = 0; i < M; ++i) {
    for (unsigned j = 0; j < N; ++j) os << m(i, j) << " ";
    os << "\n";
  }
  return os;
}

}  // namespace iqs
