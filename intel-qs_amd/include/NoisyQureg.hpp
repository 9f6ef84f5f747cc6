/// @file NoisyQureg.hpp
/// QubitRegister that inserts a stochastic noise gate before every logical ("experimental")
/// gate, the interface and noise model of the reference's iqs::NoisyQureg
/// (include/NoisyQureg.hpp:26-509).
///
/// Model.  Each qubit accumulates idle time: every 1-qubit gate advances all qubits by the
/// 1-qubit gate duration, every 2-qubit gate by the 2-qubit duration, and the qubits a gate
/// acts on restart from zero.  Before a gate, each qubit it acts on receives the Pauli-twirl
/// noise rotation for its idle time t:
///   p_X = p_Y = (1 - e^{-t/T1}) / 4,  p_Z = (1 - e^{-t/T2}) / 2 + (1 - e^{-t/T1}) / 4,
///   v_P = g_P sqrt(-log(1 - p_P)) / 2 with g_P standard normal (drawn X, Y, Z),
///   U = exp(-i v_X X) exp(-i v_Y Y) exp(-i v_Z Z),
/// applied through the register's 1-qubit gate kernel.  The normal draws come from
/// std::default_random_engine seeded with the constructor's seed and one
/// std::normal_distribution<BaseType>, in the reference's order, so a trajectory is the
/// reference's trajectory for the same seed on the same standard library.
/// Differences: a non-positive noise probability (T1 or T2 not positive) throws
/// std::invalid_argument instead of asserting; the reference's NoiseGate_OLD, documented
/// there as not to be used, is not provided.
#pragma once

#include <cmath>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace iqs {

template <class Type = ComplexDP>
class NoisyQureg : public QubitRegister<Type>
{
 public:
  using BaseType = typename QubitRegister<Type>::BaseType;

  NoisyQureg(unsigned num_qubits, unsigned RNG_seed = 12345, BaseType T1 = 2000, BaseType T2 = 1000)
      : QubitRegister<Type>(num_qubits), T_1(T1), T_2(T2), idle_(num_qubits, BaseType(0)),
        counts_(std::size_t(num_qubits) * num_qubits, 0u)
  {
    generator.seed(RNG_seed);
  }

  /// Initialize the state (without noise) and restart the gate counts and idle times.
  void Initialize(std::string style, std::size_t base_index)
  {
    one_qubit_count_ = 0;
    two_qubit_count_ = 0;
    ResetTimeForAllQubits();
    QubitRegister<Type>::Initialize(style, base_index);
  }

  void ResetTimeForAllQubits() { std::fill(idle_.begin(), idle_.end(), BaseType(0)); }

  /// Noise for the idle time of every qubit (e.g. before a final measurement), then reset.
  void ApplyNoiseGatesOnAllQubits()
  {
    for (unsigned q = 0; q < nq_(); ++q) NoiseGate(q);
    ResetTimeForAllQubits();
  }

  void SetDecoherenceTime(BaseType T1, BaseType T2)
  {
    T_1 = T1;
    T_2 = T2;
  }
  void SetGateDurations(BaseType Ts, BaseType Td)
  {
    t_one_ = Ts;
    t_two_ = Td;
  }

  unsigned GetTotalExperimentalGateCount() { return one_qubit_count_ + two_qubit_count_; }
  unsigned GetOneQubitExperimentalGateCount() { return one_qubit_count_; }
  unsigned GetTwoQubitExperimentalGateCount() { return two_qubit_count_; }
  /// Gates that involved qubit q and each other qubit (the diagonal counts 1-qubit gates).
  std::vector<unsigned> GetExperimentalGateCount(unsigned q)
  {
    return std::vector<unsigned>(counts_.begin() + std::size_t(q) * nq_(), counts_.begin() + std::size_t(q + 1) * nq_());
  }
  unsigned GetExperimentalGateCount(unsigned q1, unsigned q2) { return counts_[std::size_t(q1) * nq_() + q2]; }

  /// Noise before a 1-qubit gate on `qubit`, then advance the clock.
  void AddNoiseOneQubitGate(unsigned const qubit)
  {
    NoiseGate(qubit);
    advance_(t_one_);
    idle_[qubit] = 0;
    ++one_qubit_count_;
    ++counts_[std::size_t(qubit) * nq_() + qubit];
  }

  /// Noise before a 2-qubit gate on (q1, q2), then advance the clock.
  void AddNoiseTwoQubitGate(unsigned const q1, unsigned const q2)
  {
    NoiseGate(q1);
    NoiseGate(q2);
    advance_(t_two_);
    idle_[q1] = 0;
    idle_[q2] = 0;
    ++two_qubit_count_;
    ++counts_[std::size_t(q1) * nq_() + q2];
    ++counts_[std::size_t(q2) * nq_() + q1];
  }

  /// The noise rotation for the idle time of `qubit` (nothing for zero idle time).
  void NoiseGate(unsigned const qubit)
  {
    const BaseType t = idle_[qubit];
    if (t == 0) return;
    // Same expressions (and double-precision intermediates) as the reference.
    const BaseType p_X = (1. - std::exp(-t / T_1)) / 4.;
    const BaseType p_Y = (1. - std::exp(-t / T_1)) / 4.;
    const BaseType p_Z = (1. - std::exp(-t / T_2)) / 2. + (1. - std::exp(-t / T_1)) / 4.;
    if (!(p_X > 0 && p_Y > 0 && p_Z > 0)) throw std::invalid_argument("NoisyQureg: non-positive noise probability");
    const BaseType s_X = std::sqrt(-std::log(1. - p_X));
    const BaseType s_Y = std::sqrt(-std::log(1. - p_Y));
    const BaseType s_Z = std::sqrt(-std::log(1. - p_Z));
    const BaseType v_X = gaussian_RNG(generator) * s_X / 2.;
    const BaseType v_Y = gaussian_RNG(generator) * s_Y / 2.;
    const BaseType v_Z = gaussian_RNG(generator) * s_Z / 2.;
    // exp(-i v_X X) exp(-i v_Y Y) exp(-i v_Z Z) = [[A B, -A* C*], [A C, A* B*]] with
    // A = e^{-i v_Z}, B = cos v_X cos v_Y - i sin v_X sin v_Y, C = cos v_X sin v_Y - i sin v_X cos v_Y.
    const Type A(std::cos(v_Z), -std::sin(v_Z));
    const Type B(std::cos(v_X) * std::cos(v_Y), -std::sin(v_X) * std::sin(v_Y));
    const Type C(std::cos(v_X) * std::sin(v_Y), -std::sin(v_X) * std::cos(v_Y));
    TM2x2<Type> u;
    u(0, 0) = A * B;
    u(0, 1) = -std::conj(A) * std::conj(C);
    u(1, 0) = A * C;
    u(1, 1) = std::conj(A) * std::conj(B);
    QubitRegister<Type>::Apply1QubitGate(qubit, u);
  }

  // Logical gates: noise first, then the gate of the underlying register.
  void Apply1QubitGate(unsigned const q, TM2x2<Type> V)
  {
    AddNoiseOneQubitGate(q);
    QubitRegister<Type>::Apply1QubitGate(q, V);
  }
  void ApplyHadamard(unsigned const q)
  {
    AddNoiseOneQubitGate(q);
    QubitRegister<Type>::ApplyHadamard(q);
  }
  void ApplyRotationX(unsigned const q, BaseType theta)
  {
    AddNoiseOneQubitGate(q);
    QubitRegister<Type>::ApplyRotationX(q, theta);
  }
  void ApplyRotationY(unsigned const q, BaseType theta)
  {
    AddNoiseOneQubitGate(q);
    QubitRegister<Type>::ApplyRotationY(q, theta);
  }
  void ApplyRotationZ(unsigned const q, BaseType theta)
  {
    AddNoiseOneQubitGate(q);
    QubitRegister<Type>::ApplyRotationZ(q, theta);
  }
  void ApplyCPauliX(unsigned const q1, unsigned const q2)
  {
    AddNoiseTwoQubitGate(q1, q2);
    QubitRegister<Type>::ApplyCPauliX(q1, q2);
  }
  void ApplyControlled1QubitGate(unsigned const q1, unsigned const q2, TM2x2<Type> V)
  {
    AddNoiseTwoQubitGate(q1, q2);
    QubitRegister<Type>::ApplyControlled1QubitGate(q1, q2, V);
  }

 private:
  unsigned nq_() const { return (unsigned)this->num_qubits; }
  void advance_(BaseType dt)
  {
    for (auto& t : idle_) t += dt;
  }

  BaseType T_1, T_2;                  // decoherence times (chosen time unit)
  BaseType t_one_ = 1., t_two_ = 1.;  // gate durations (same unit)
  std::vector<BaseType> idle_;        // time since each qubit's last gate
  std::vector<unsigned> counts_;      // gates per (qubit, qubit) pair
  unsigned one_qubit_count_ = 0, two_qubit_count_ = 0;
  std::default_random_engine generator;
  std::normal_distribution<BaseType> gaussian_RNG{0., 1.};
};

}  // namespace iqs
