/// @file QubitRegisterMetric.hpp
/// QubitRegister that counts the logical gates it applies and tracks the circuit's parallel
/// depth, the interface of the reference's iqs::QubitRegisterMetric
/// (include/QubitRegisterMetric.hpp:13-113): a 1-qubit gate adds one layer on its qubit;
/// a 2-qubit gate puts both qubits one layer above the deeper of the two.  Only the gates
/// listed here are counted, as in the reference.
#pragma once

#include <algorithm>
#include <vector>

namespace iqs {

template <class Type = ComplexDP>
class QubitRegisterMetric : public QubitRegister<Type>
{
 public:
  explicit QubitRegisterMetric(int num_qubits) : QubitRegister<Type>(num_qubits), depth_(num_qubits, 0) {}

  int GetTotalQubitGateCount() { return one_ + two_; }
  int GetOneQubitGateCount() { return one_; }
  int GetTwoQubitGateCount() { return two_; }
  int GetParallelDepth() { return depth_.empty() ? 0 : *std::max_element(depth_.begin(), depth_.end()); }

  void ApplyHadamard(int q)
  {
    QubitRegister<Type>::ApplyHadamard(q);
    one_qubit_(q);
  }
  void ApplyRotationX(int q, double theta)
  {
    QubitRegister<Type>::ApplyRotationX(q, theta);
    one_qubit_(q);
  }
  void ApplyRotationY(int q, double theta)
  {
    QubitRegister<Type>::ApplyRotationY(q, theta);
    one_qubit_(q);
  }
  void ApplyRotationZ(int q, double theta)
  {
    QubitRegister<Type>::ApplyRotationZ(q, theta);
    one_qubit_(q);
  }
  void ApplyCPauliX(int q1, int q2)
  {
    QubitRegister<Type>::ApplyCPauliX(q1, q2);
    two_qubit_(q1, q2);
  }
  void ApplyControlled1QubitGate(int q1, int q2, TM2x2<Type> V)
  {
    QubitRegister<Type>::ApplyControlled1QubitGate(q1, q2, V);
    two_qubit_(q1, q2);
  }

 private:
  void one_qubit_(int q)
  {
    ++one_;
    ++depth_.at(q);
  }
  void two_qubit_(int q1, int q2)
  {
    ++two_;
    const int d = std::max(depth_.at(q1), depth_.at(q2)) + 1;
    depth_[q1] = d;
    depth_[q2] = d;
  }

  int one_ = 0, two_ = 0;
  std::vector<int> depth_;
};

}  // namespace iqs
