/// @file gate_counter.hpp
/// Gate counts and greedy depth per program qubit (reference include/gate_counter.hpp:18-95).
#pragma once

#include <algorithm>
#include <cstdio>
#include <vector>

namespace iqs {

class GateCounter
{
 public:
  explicit GateCounter(int num_qubits) : n_(num_qubits), depth_(num_qubits, 0) {}
  void Reset()
  {
    depth_.assign(n_, 0);
    total_ = one_ = two_ = 0;
  }
  int GetTotalGateCount() const { return total_; }
  int GetOneQubitGateCount() const { return one_; }
  int GetTwoQubitGateCount() const { return two_; }
  int GetParallelDepth() const { return depth_.empty() ? 0 : *std::max_element(depth_.begin(), depth_.end()); }
  void OneQubitIncrement(int q)
  {
    ++total_;
    ++one_;
    ++depth_[q];
  }
  void TwoQubitIncrement(int a, int b)
  {
    ++total_;
    ++two_;
    const int d = std::max(depth_[a], depth_[b]) + 1;
    depth_[a] = depth_[b] = d;
  }
  void Breakdown() const
  {
    printf("The quantum circuit is composed of %d one-qubit gates and %d two-qubitgates, for a total of %d gates.\n"
           "The greedy depth (all gates lasting one clock cycle) is %d.\n",
           one_, two_, total_, GetParallelDepth());
  }

 private:
  int n_;
  int total_ = 0, one_ = 0, two_ = 0;
  std::vector<int> depth_;
};

}  // namespace iqs
