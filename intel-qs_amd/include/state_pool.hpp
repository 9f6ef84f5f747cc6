/// @file state_pool.hpp
/// A pool of independent states of one size on one GPU: the single-process form of the
/// reference's pool of states, where each MPI rank group holds one state and runs the same
/// program (src/mpi_env.cpp:239-404, tutorials/get_started_with_noisy_IQS.cpp), and
/// Environment::IncoherentSumOverAllStatesOfPool combines one value per state
/// (src/mpi_env.cpp:468-486).
///
/// Layout.  All states live in one QubitRegister of num_qubits + k qubits (2^k >=
/// num_states): qubits 0 .. num_qubits-1 are the state's qubits, the top k select the
/// state.  A gate called on Register() with qubits < num_qubits is applied to every state
/// by one kernel launch; the top k qubits must not be touched by gates.
///
/// Per-state operations (one launch for the whole pool each): noise gates and channel
/// branches, where every state draws from its own generator -- state s gets the streams
/// that pool state s gets from SetSeedStreamPtrs(seed) in the reference (state stream
/// seeded with seed + 1 + s), so state s follows the trajectory of the reference's pool
/// state s -- and reductions that return one value per state.
#pragma once

#include <string>
#include <vector>

#include "qureg.hpp"

namespace iqs {

template <class Type = ComplexDP>
class StatePool
{
 public:
  using BaseType = typename QubitRegister<Type>::BaseType;

  StatePool(unsigned num_qubits, unsigned num_states, std::string style = "base", std::size_t base_index = 0);

  unsigned NumQubits() const { return n_; }
  unsigned NumStates() const { return ns_; }
  /// The register holding every state (num_qubits + k qubits).
  QubitRegister<Type>& Register() { return reg_; }

  /// Every state to |base_index> ("base") or to the uniform superposition ("++++").
  void Initialize(std::string style, std::size_t base_index);

  // ---- per-state randomness and noise ----
  /// Seed the generator of every state: state s as pool state s of the reference.
  void SetSeedStreamPtrs(std::size_t seed);
  RandomNumberGenerator<BaseType>& Generator(unsigned s) { return gens_.at(s); }
  void SetNoiseTimescales(BaseType T1, BaseType T2);
  /// ApplyNoiseGate (src/qureg_noisysimul.cpp:33-85) on every state, each with its own draws.
  void ApplyNoiseGate(unsigned qubit, BaseType duration);
  /// ApplyChannel (src/qureg_apply_channel.cpp:17-58) on every state, each drawing its branch.
  void ApplyChannel(unsigned qubit, CM4x4<Type>& chi);
  /// 2-qubit ApplyChannel (src/qureg_apply_channel.cpp:66-106) on every state.
  void ApplyChannel(unsigned qubit1, unsigned qubit2, CM16x16<Type>& chi);
  std::vector<BaseType> GetOverallSignOfChannels() const { return sign_; }

  // ---- one value per state ----
  std::vector<BaseType> GetProbability(unsigned qubit);
  std::vector<BaseType> ComputeNorm();
  /// <Z_q1 Z_q2 ...> of every state (unnormalized, as ExpectationValue of Z strings).
  std::vector<BaseType> ExpectationValueZString(const std::vector<unsigned>& qubits);
  /// <psi|state s> for every s: ComputeOverlap(psi) called on each state.
  std::vector<Type> ComputeOverlap(QubitRegister<Type>& psi);
  /// Amplitudes of state s (host copy).
  std::vector<Type> GetState(unsigned s);

  /// Sum over states of one value per state (IncoherentSumOverAllStatesOfPool).
  static BaseType IncoherentSumOverAllStatesOfPool(const std::vector<BaseType>& values);

 private:
  iqsh_state* dev_();
  unsigned pos_(unsigned qubit) const;

  unsigned n_, ns_;
  QubitRegister<Type> reg_;
  std::vector<RandomNumberGenerator<BaseType>> gens_;
  std::vector<BaseType> sign_;
  BaseType T_1_ = 0, T_2_ = 0;
};

}  // namespace iqs
