/// @file chi_matrix.hpp
/// Chi-matrix of a 1- or 2-qubit quantum channel with its eigensystem, the interface of the
/// reference's iqs::ChiMatrix (include/chi_matrix.hpp:49-238, src/chi_matrix.cpp:57-234):
///   rho' = sum_ij chi_ij sigma_i rho sigma_j^dagger,
/// Pauli basis {id, X, Y, Z}; for two qubits sigma_{4a+b} = sigma_a (first qubit) x sigma_b.
///
/// The eigensystem chi = sum_k E_k |E_k><E_k| is what ApplyChannel samples from: branch k
/// with probability |E_k| / sum_k |E_k|, its eigenvector (Pauli coefficients, scaled by
/// sqrt(sum_k |E_k|)) read as the matrix applied to the state.
///
/// Eigen-solver.  The reference calls Eigen's ComplexEigenSolver (Eigen is not a dependency
/// here).  chi is Hermitian for every Hermiticity-preserving channel, so this header uses the
/// cyclic complex Jacobi method (Jacobi 1846; Golub & Van Loan, "Matrix Computations",
/// sec. 8.5), in double precision, converging to |off-diagonal| <= 1e-15 |chi|.
/// Conventions chosen to agree with the reference's solver where it is determined:
///   * eigenpairs ordered by ascending |E_k|, by the same selection (first minimum of the
///     remaining tail, swapped into place) that ComplexEigenSolver::sortEigenvalues uses;
///   * unit eigenvectors with their largest component real and positive.
/// For a diagonal chi (Pauli channels) the Jacobi method performs no rotation, so the
/// eigenvectors are the unit vectors and the whole eigensystem -- hence every sampled
/// branch -- equals the reference's.  For a general chi the eigenvectors agree up to a
/// phase (and a basis choice inside degenerate eigenspaces), which leaves the channel, the
/// branch probabilities and every expectation value unchanged.
/// SolveEigenSystem throws std::runtime_error for a non-Hermitian matrix (the reference
/// asserts on complex eigenvalues).
#pragma once

#include <cmath>
#include <complex>
#include <initializer_list>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "tinymatrix.hpp"
#include "utils.hpp"

namespace iqs {

namespace chi_detail {

/// Eigen-decomposition of the Hermitian M x M matrix `a` (row-major): eigenvalues `w` and
/// unit eigenvectors `v[k]`, in the order of the Jacobi diagonal (not yet sorted).
inline void hermitian_jacobi(unsigned M, std::vector<std::complex<double>> a, std::vector<double>& w,
                             std::vector<std::vector<std::complex<double>>>& v)
{
  using C = std::complex<double>;
  std::vector<C> q(M * M, C(0));
  for (unsigned i = 0; i < M; ++i) q[i * M + i] = C(1);
  double fro = 0;
  for (const C& x : a) fro += std::norm(x);
  const double tol = 1e-30 * fro;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0;
    for (unsigned p = 0; p < M; ++p)
      for (unsigned r = p + 1; r < M; ++r) off += std::norm(a[p * M + r]);
    if (off <= tol) break;
    for (unsigned p = 0; p < M; ++p)
      for (unsigned r = p + 1; r < M; ++r) {
        const C apr = a[p * M + r];
        const double mag = std::abs(apr);
        if (mag == 0) continue;
        // Rotation in the (p, r) plane: V = diag(1, e) * [[c, s], [-s, c]] with
        // e = exp(-i arg a_pr) makes the pivot real, then the real rotation zeroes it:
        // tan(2 theta) = 2 |a_pr| / (a_rr - a_pp).
        const C e = std::conj(apr) / mag;
        const double th = 0.5 * std::atan2(2 * mag, a[r * M + r].real() - a[p * M + p].real());
        const double c = std::cos(th), s = std::sin(th);
        const C vpp = c, vpr = s, vrp = -s * e, vrr = c * e;
        // a <- a V (columns p, r), then a <- V^dagger a (rows p, r); q <- q V.
        for (unsigned i = 0; i < M; ++i) {
          const C x = a[i * M + p], y = a[i * M + r];
          a[i * M + p] = x * vpp + y * vrp;
          a[i * M + r] = x * vpr + y * vrr;
          const C u = q[i * M + p], t = q[i * M + r];
          q[i * M + p] = u * vpp + t * vrp;
          q[i * M + r] = u * vpr + t * vrr;
        }
        for (unsigned j = 0; j < M; ++j) {
          const C x = a[p * M + j], y = a[r * M + j];
          a[p * M + j] = std::conj(vpp) * x + std::conj(vrp) * y;
          a[r * M + j] = std::conj(vpr) * x + std::conj(vrr) * y;
        }
        a[p * M + r] = a[r * M + p] = C(0);
      }
  }
  w.assign(M, 0);
  v.assign(M, std::vector<C>(M));
  for (unsigned k = 0; k < M; ++k) {
    w[k] = a[k * M + k].real();
    unsigned big = 0;
    for (unsigned i = 0; i < M; ++i) {
      v[k][i] = q[i * M + k];
      if (std::abs(v[k][i]) > std::abs(v[k][big])) big = i;
    }
    const C ph = std::conj(v[k][big]) / std::abs(v[k][big]);
    for (unsigned i = 0; i < M; ++i) v[k][i] *= ph;
    v[k][big] = C(v[k][big].real(), 0);
  }
}

}  // namespace chi_detail

template <class ValueType, unsigned M, unsigned align = alignof(ValueType)>
class ChiMatrix : public TinyMatrix<ValueType, M, M, align>
{
 public:
  using value_type = ValueType;
  typedef typename extract_value_type<ValueType>::value_type BaseType;
  using base_type = BaseType;
  using size_type = unsigned;
  using Base = TinyMatrix<ValueType, M, M, align>;

  ChiMatrix() : Base() {}
  template <class U>
  ChiMatrix(U init[M][M]) : Base(init) {}
  template <class U>
  ChiMatrix(std::initializer_list<std::initializer_list<U>> const& init) : Base(init) {}
  template <class U, unsigned alignrhs>
  ChiMatrix(ChiMatrix<U, M, alignrhs> const& rhs) : Base(rhs)
  {
    for (auto const& x : rhs.GetEigenValues()) evalues_.push_back(value_type(x));
    for (auto const& vec : rhs.GetEigenVectors()) evectors_.emplace_back(vec.begin(), vec.end());
    for (auto x : rhs.GetEigenProbabilities()) eprobs_.push_back(base_type(x));
    for (auto x : rhs.GetEigenCumulativeProbabilities()) ecumprobs_.push_back(base_type(x));
  }
  ChiMatrix(ChiMatrix const&) = default;
  ChiMatrix& operator=(ChiMatrix const&) = default;

  ValueType* GetPtrToData() { return this->getPtrToData(); }

  /// Eigensystem of the ideal Hadamard channel chi = 1/2 (X + Z)(X + Z): eigenvalue 1 on
  /// (X + Z)/sqrt(2), 0 on (X - Z)/sqrt(2), id and Y -- the same vectors, in the same order,
  /// as the reference's specialization (src/chi_matrix.cpp:210-234).  Throws unless the
  /// matrix is exactly that chi.
  void EigensystemOfIdealHadamardChannel()
  {
    if (M != 4) throw std::runtime_error("EigensystemOfIdealHadamardChannel: 1-qubit chi-matrix only");
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < M; ++j) {
        const bool xz = (i == 1 || i == 3) && (j == 1 || j == 3);
        if (xz ? !((*this)(i, j) == value_type(0.5, 0)) : std::norm((*this)(i, j)) != 0)
          throw std::runtime_error("EigensystemOfIdealHadamardChannel: not the Hadamard chi-matrix");
      }
    const value_type one(1, 0), zero(0, 0), h(1 / std::sqrt(base_type(2)), 0);
    evalues_.assign(M, zero);
    evalues_[0] = one;
    evectors_ = {{zero, h, zero, h}, {zero, h, zero, -h}, {one, zero, zero, zero}, {zero, zero, one, zero}};
    NormalizeEigenProbAndRenormalizeEigenVect();
  }

  /// Decompose chi into its eigensystem, standardize the eigenvectors (so that
  /// chi = sum_k E_k |E_k><E_k|) and build the branch probabilities
  /// (src/chi_matrix.cpp:66-127).
  void SolveEigenSystem()
  {
    std::vector<std::complex<double>> a(M * M);
    double scale = 0;
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j < M; ++j) {
        a[i * M + j] = std::complex<double>((*this)(i, j));
        scale = std::max(scale, std::abs(a[i * M + j]));
      }
    for (unsigned i = 0; i < M; ++i)
      for (unsigned j = 0; j <= i; ++j)
        if (std::abs(a[i * M + j] - std::conj(a[j * M + i])) > 1e-12 * std::max(scale, 1.0))
          throw std::runtime_error("ChiMatrix::SolveEigenSystem: the chi-matrix is not Hermitian");
    std::vector<double> w;
    std::vector<std::vector<std::complex<double>>> v;
    chi_detail::hermitian_jacobi(M, a, w, v);
    // Ascending |E_k|: the first minimum of the remaining tail is swapped into place.
    for (unsigned i = 0; i < M; ++i) {
      unsigned k = i;
      for (unsigned j = i + 1; j < M; ++j)
        if (std::abs(w[j]) < std::abs(w[k])) k = j;
      if (k != i) {
        std::swap(w[i], w[k]);
        std::swap(v[i], v[k]);
      }
    }
    evalues_.assign(M, value_type(0));
    evectors_.assign(M, std::vector<value_type>(M));
    for (unsigned k = 0; k < M; ++k) {
      evalues_[k] = value_type(base_type(w[k]), 0);
      for (unsigned i = 0; i < M; ++i) evectors_[k][i] = value_type(v[k][i]);
    }
    // Standardization: |E'_k> = sqrt(E_k / G_k) |E_k> with G_k = <E_k| chi |E_k>.
    for (unsigned k = 0; k < M; ++k) {
      value_type gk(0);
      for (unsigned i = 0; i < M; ++i) {
        value_type gi(0);
        for (unsigned j = 0; j < M; ++j) gi += (*this)(i, j) * evectors_[k][j];
        gk += std::conj(evectors_[k][i]) * gi;
      }
      if (std::abs(gk) > 0)
        for (unsigned i = 0; i < M; ++i) evectors_[k][i] *= std::sqrt(evalues_[k] / gk);
    }
    NormalizeEigenProbAndRenormalizeEigenVect();
  }

  /// Print the matrix and, optionally, its eigensystem (the reference's layout,
  /// src/chi_matrix.cpp:131-158).
  void Print(bool with_eigensystem = true)
  {
    std::cout << "chi_matrix :\n";
    for (size_type i = 0; i < M; ++i) {
      for (size_type j = 0; j < M; ++j) std::cout << (*this)(i, j) << "\t";
      std::cout << "\n";
    }
    if (!with_eigensystem) return;
    std::cout << "eigenvalues :\n";
    for (auto const& x : evalues_) std::cout << x << "\t";
    std::cout << "\neigenprobs :\n";
    for (auto const& x : eprobs_) std::cout << x << "\t";
    for (size_type i = 0; i < evectors_.size(); ++i) {
      std::cout << "\neigenvector " << i << " :\n";
      for (auto const& x : evectors_[i]) std::cout << x << "\t";
    }
    std::cout << "\n";
  }

  value_type GetEigenValue(size_type k) const { return evalues_[k]; }
  std::vector<value_type> GetEigenValues() const { return evalues_; }
  base_type GetEigenProbability(size_type k) const { return eprobs_[k]; }
  std::vector<base_type> GetEigenProbabilities() const { return eprobs_; }
  base_type GetEigenCumulativeProbability(size_type k) const { return ecumprobs_[k]; }
  std::vector<base_type> GetEigenCumulativeProbabilities() const { return ecumprobs_; }
  std::vector<value_type> GetEigenVector(size_type k) const { return evectors_[k]; }
  std::vector<std::vector<value_type>> GetEigenVectors() const { return evectors_; }
  /// True once SolveEigenSystem (or the Hadamard eigensystem) has run.
  bool HasEigenSystem() const { return evalues_.size() == M && ecumprobs_.size() == M; }

 protected:
  std::vector<value_type> evalues_;
  std::vector<std::vector<value_type>> evectors_;
  std::vector<base_type> eprobs_;
  std::vector<base_type> ecumprobs_;

  /// Branch probabilities p_k = |E_k| / sum |E_k| and eigenvectors scaled by
  /// sqrt(sum |E_k|), which moves the normalization of the sampling into the branch
  /// matrices (src/chi_matrix.cpp:166-205).
  void NormalizeEigenProbAndRenormalizeEigenVect()
  {
    base_type normalization = 0;
    eprobs_.clear();
    ecumprobs_.clear();
    for (auto const& x : evalues_) {
      if (std::abs(std::imag(x)) >= 1.0e-14) throw std::runtime_error("ChiMatrix: eigenvalues must be real");
      const base_type p = std::abs(std::real(x));
      eprobs_.push_back(p);
      normalization += p;
      ecumprobs_.push_back(normalization);
    }
    if (normalization != 0 && normalization != 1) {
      for (size_type i = 0; i < eprobs_.size(); ++i) {
        eprobs_[i] /= normalization;
        ecumprobs_[i] /= normalization;
      }
      for (auto& vec : evectors_)
        for (auto& x : vec) x *= std::sqrt(normalization);
    }
  }
};

}  // namespace iqs
