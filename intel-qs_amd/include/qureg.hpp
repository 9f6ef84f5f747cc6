/// @file qureg.hpp
/// iqs::QubitRegister<ComplexSP|ComplexDP>: drop-in for the reference class of the same name
/// (include/qureg.hpp:127-417 of intel-qs), with the amplitudes resident in one GPU's HBM.
///
/// Public names, signatures and conventions follow the reference: qubit 0 is the least
/// significant bit; gates act on program qubits, mapped to data positions through
/// `qubit_permutation`; 4x4 matrices are indexed 2*b(first qubit) + b(second qubit).
/// Every gate and reduction runs as a gfx950 kernel through the C ABI of libiqs_hip.so
/// (include/iqs_hip.h).  Host access to amplitudes goes through a host mirror:
///   * operator[] / RawState() download the state into the mirror (once) and mark it as
///     possibly modified; the next device operation uploads it back;
///   * GetGlobalAmplitude / SetGlobalAmplitude move a single amplitude.
/// Differences from the reference that a caller can observe are documented at each method.
#pragma once

// The reference's qureg.hpp includes these standard headers (include/qureg.hpp:6-17), and
// its callers rely on them without including them.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "chi_matrix.hpp"
#include "gate_counter.hpp"
#include "gate_spec.hpp"
#include "mpi_env.hpp"
#include "permutation.hpp"
#include "rng_utils.hpp"
#include "timer.hpp"
#include "tinymatrix.hpp"
#include "utils.hpp"

template <class Type>
using TM2x2 = iqs::TinyMatrix<Type, 2, 2, 32>;
template <class Type>
using TM4x4 = iqs::TinyMatrix<Type, 4, 4, 32>;
template <class Type>
using CM4x4 = iqs::ChiMatrix<Type, 4, 32>;
template <class Type>
using CM16x16 = iqs::ChiMatrix<Type, 16, 32>;

struct iqsh_state;

namespace iqs {

template <class Type = ComplexDP>
class QubitRegister
{
 public:
  using value_type = Type;
  typedef typename extract_value_type<Type>::value_type BaseType;

  // ---- construction (src/qureg_init.cpp) ----
  QubitRegister();
  QubitRegister(std::size_t num_qubits, std::string style = "", std::size_t base_index = 0,
                std::size_t tmp_spacesize = 0);
  QubitRegister(const QubitRegister& in);
  /// Imports the amplitudes of `state` (2^n entries).  The buffer becomes the host mirror:
  /// it is refreshed on host access and when the register is destroyed.
  QubitRegister(std::size_t num_qubits, Type* state, std::size_t tmp_spacesize = 0);
  ~QubitRegister();

  void AllocateAdditionalQubit();
  void Allocate(std::size_t new_num_qubits, std::size_t tmp_spacesize);
  void Initialize(std::size_t new_num_qubits, std::size_t tmp_spacesize);
  void Initialize(std::string style, std::size_t base_index);

  // ---- amplitude access ----
  Type& operator[](std::size_t index);
  Type& operator[](std::size_t index) const;
  Type GetGlobalAmplitude(std::size_t index) const;
  void SetGlobalAmplitude(std::size_t index, Type value);
  std::size_t LocalSize() const { return global_size_; }
  std::size_t GlobalSize() const { return global_size_; }
  std::size_t size() const { return global_size_; }
  std::size_t NumQubits() const { return num_qubits; }
  size_t TmpSize() const { return 0; }
  /// Host pointer to the amplitudes (host mirror, refreshed from HBM).
  Type* RawState();

  bool check_bit(std::size_t v, std::size_t p) const { return (v >> p) & 1; }
  std::size_t set_bit(std::size_t v, std::size_t p) const { return v | (UL(1) << p); }
  std::size_t clear_bit(std::size_t v, std::size_t p) const { return v & ~(UL(1) << p); }

  // ---- statistics ----
  void EnableStatistics();
  void GetStatistics();
  void DisableStatistics();
  void ResetStatistics();

  // ---- permutation of the qubit order (src/qureg_permute.cpp) ----
  void PermuteQubits(std::vector<std::size_t> new_map, std::string style_of_map = "direct");
  void PermuteLocalQubits(std::vector<std::size_t> new_map, std::string style_of_map = "direct");
  // With the whole state on one GPU every qubit is local: the reference's global-qubit
  // permutations (src/qureg_permute.cpp:110-229) are PermuteQubits here.
  void PermuteGlobalQubits(std::vector<std::size_t> new_map, std::string style_of_map = "direct")
  {
    PermuteQubits(new_map, style_of_map);
  }
  void PermuteByLocalGlobalExchangeOfQubitPairs(std::vector<std::size_t> new_map, std::string style_of_map = "direct")
  {
    PermuteQubits(new_map, style_of_map);
  }
  // One process holds the state: data rank = communicator rank = 0 (include/qureg.hpp:204-209).
  int GetDataRank(int) const { return 0; }
  int GetMyDataRank() const { return 0; }
  int GetCommRank(int) const { return 0; }
  void EmulateSwap(unsigned qubit1, unsigned qubit2);

  // ---- one-qubit gates (src/qureg_apply1qubitgate.cpp) ----
  bool Apply1QubitGate_helper(unsigned qubit, TM2x2<Type> const& m, std::size_t sind, std::size_t eind,
                              GateSpec1Q spec = GateSpec1Q::None, BaseType angle = 0);
  void Apply1QubitGate(unsigned qubit, TM2x2<Type> const& m, GateSpec1Q spec = GateSpec1Q::None, BaseType angle = 0);
  // ---- controlled gates (src/qureg_applyctrl1qubitgate.cpp) ----
  bool ApplyControlled1QubitGate_helper(unsigned control_qubit, unsigned target_qubit, TM2x2<Type> const& m,
                                        std::size_t sind, std::size_t eind, GateSpec2Q spec = GateSpec2Q::None,
                                        BaseType angle = 0);
  void ApplyControlled1QubitGate(unsigned control_qubit, unsigned target_qubit, TM2x2<Type> const& m,
                                 GateSpec2Q spec = GateSpec2Q::None, BaseType angle = 0);
  // ---- swap family (src/qureg_applyswap.cpp) ----
  bool ApplySwap_helper(unsigned qubit1, unsigned qubit2, TM2x2<Type> const& m);
  void ApplySwap(unsigned qubit1, unsigned qubit2);
  void ApplyISwap(unsigned qubit1, unsigned qubit2);
  void Apply4thRootISwap(unsigned qubit1, unsigned qubit2);
  void ApplySqrtISwap(unsigned qubit1, unsigned qubit2);
  void ApplyISwapRotation(unsigned qubit1, unsigned qubit2, TM2x2<Type> const& m);
  // ---- diagonal and dense two-qubit gates ----
  void ApplyDiagSimp(unsigned qubit1, unsigned qubit2, TM4x4<Type> const& m);
  void ApplyDiag(unsigned qubit1, unsigned qubit2, TM4x4<Type> const& m);
  void Apply2QubitGate(unsigned const qubit_high, unsigned const qubit_low, TM4x4<Type> const& m);
  // ---- named gates ----
  void ApplyRotationX(unsigned const qubit, BaseType theta);
  void ApplyRotationY(unsigned const qubit, BaseType theta);
  void ApplyRotationZ(unsigned const qubit, BaseType theta);
  void ApplyPauliX(unsigned const qubit);
  void ApplyPauliY(unsigned const qubit);
  void ApplyPauliZ(unsigned const qubit);
  void ApplyPauliSqrtX(unsigned const qubit);
  void ApplyPauliSqrtY(unsigned const qubit);
  void ApplyPauliSqrtZ(unsigned const qubit);
  void ApplyT(unsigned const qubit);
  void ApplyToffoli(unsigned const qubit1, unsigned const qubit2, unsigned const qubit3);
  void ApplyHadamard(unsigned const qubit);
  void ApplyRotationXY(unsigned const qubit, BaseType phi, BaseType theta);
  void ApplyCRotationX(unsigned const control_qubit, unsigned const target_qubit, BaseType theta);
  void ApplyCRotationY(unsigned const control_qubit, unsigned const target_qubit, BaseType theta);
  void ApplyCRotationZ(unsigned const control_qubit, unsigned const target_qubit, BaseType theta);
  void ApplyCPauliX(unsigned const control_qubit, unsigned const target_qubit);
  void ApplyCPauliY(unsigned const control_qubit, unsigned const target_qubit);
  void ApplyCPauliZ(unsigned const control_qubit, unsigned const target_qubit);
  void ApplyCPauliSqrtZ(unsigned const control_qubit, unsigned const target_qubit);
  void ApplyCHadamard(unsigned const control_qubit, unsigned const target_qubit);
  void ApplyCPhaseRotation(unsigned const qubit, unsigned const qubit2, BaseType theta);

  // ---- fusion ----
  /// Gates are queued and applied as dense blocks of up to FusionBlockQubits() qubits,
  /// one HBM pass per block (the reference's fusion replays queued gates cache block by
  /// cache block, src/qureg_fusion.cpp).  log2llc is kept for API compatibility.
  void TurnOnFusion(unsigned log2llc = 20);
  void TurnOffFusion();
  bool IsFusionEnabled();
  void ApplyFusedGates();
  void SetFusionBlockQubits(unsigned k);
  unsigned FusionBlockQubits() const { return fuse_k_; }
  /// (queued gates applied so far, blocks launched for them) since construction.
  std::pair<std::size_t, std::size_t> FusionStats() const { return {fused_gates_, fused_blocks_}; }

  // ---- specialization ----
  /// Matrix structure (identity, diagonal, X-type) is always exploited exactly by the
  /// kernels; TurnOnSpecialize/TurnOffSpecialize only record the flag.
  void TurnOnSpecialize();
  void TurnOffSpecialize();
  /// Gate-typed bodies for named gates (constants in BaseType precision).
  void TurnOnSpecializeV2();
  void TurnOffSpecializeV2();

  // ---- measurement (src/qureg_measure.cpp) ----
  bool GetClassicalValue(unsigned qubit, BaseType tolerance = 1.e-13) const;
  bool IsClassicalBit(unsigned qubit, BaseType tolerance = 1.e-13) const;
  void CollapseQubit(unsigned qubit, bool value);
  /// Sum of |psi_i|^2 with the qubit's bit = 1, accumulated in FP64.
  BaseType GetProbability(unsigned qubit);

  // ---- expectation values (src/qureg_expectval.cpp) ----
  /// One read-only pass per Pauli string; the state is left bit-for-bit unchanged (the
  /// reference applies and undoes basis changes).  Single-qubit forms return
  /// coeff * (1 - 2 P(1)) in the rotated basis, as the reference does.
  BaseType ExpectationValueX(unsigned const qubit, BaseType coeff = 1.);
  BaseType ExpectationValueY(unsigned const qubit, BaseType coeff = 1.);
  BaseType ExpectationValueZ(unsigned const qubit, BaseType coeff = 1.);
  BaseType ExpectationValueXX(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueXY(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueXZ(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueYX(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueYY(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueYZ(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueZX(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueZY(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValueZZ(unsigned const qubit, unsigned const qubit2, BaseType coeff = 1.);
  BaseType ExpectationValue(std::vector<unsigned>& qubits, std::vector<unsigned>& observables, BaseType coeff = 1.);

  // ---- utilities (src/qureg_utils.cpp) ----
  bool operator==(const QubitRegister& rhs);
  BaseType MaxAbsDiff(QubitRegister& x, Type sfactor = Type(1, 0));
  BaseType MaxL2NormDiff(QubitRegister& x);
  void Normalize();
  BaseType ComputeNorm();
  void InitializationWithSameAmplitudeEverywhere(Type amplitude);
  void AmplitudeWiseScalarMultiplication(Type factor);
  void AmplitudeWiseSum(QubitRegister<Type>& psi, Type factor = Type(1, 0));
  Type ComputeOverlap(QubitRegister<Type>& psi);
  double Entropy();
  std::vector<double> GoogleStats();
  void Print(std::string x, std::vector<std::size_t> qbits = {});
  void ExportAmplitudes(std::string ofname);
  void dumpbin(std::string fn);
  void loadbin(std::string fn);

  // ---- random numbers ----
  RandomNumberGenerator<BaseType>* GetRngPtr() { return rng_ptr_; }
  void ResetRngPtr() { rng_ptr_ = nullptr; }

  // ---- noisy simulation (src/qureg_noisysimul.cpp) ----
  BaseType GetT1() { return T_1_; }
  BaseType GetT2() { return T_2_; }
  BaseType GetTphi() { return T_phi_; }
  void SetNoiseTimescales(BaseType T1, BaseType T2);
  void ApplyNoiseGate(const unsigned qubit, const BaseType duration);
  // ---- quantum channels (src/qureg_apply_channel.cpp) ----
  void ApplyChannel(const unsigned qubit, CM4x4<Type>& chi);
  void ApplyChannel(const unsigned qubit1, const unsigned qubit2, CM16x16<Type>& chi);
  BaseType GetOverallSignOfChannels() { return overall_sign_of_channels; }
  void SetRngPtr(RandomNumberGenerator<BaseType>* p) { rng_ptr_ = p; }
  void SetSeedRngPtr(std::size_t seed)
  {
    assert(rng_ptr_);
    rng_ptr_->SetSeedStreamPtrs(seed);
  }

  static void SetDoPrintExtraInfo(bool value) { do_print_extra_info = value; }

  // ---- device interface (extension) ----
  /// Handle of the device state for direct C-ABI calls (flushes pending host/fusion work).
  iqsh_state* DeviceHandle();
  /// Make pending work visible: apply queued fused gates and upload a modified host mirror.
  void Sync() const;

  // ---- public members, as in the reference (include/qureg.hpp:381-398) ----
  std::size_t num_qubits = 0;
  Permutation* qubit_permutation = nullptr;
  Timer* timer = nullptr;
  GateCounter* gate_counter = nullptr;
  bool imported_state = false;
  bool specialize = false;
  bool specialize2 = false;
  BaseType overall_sign_of_channels = 1;
  bool fusion = false;
  unsigned log2llc = 20;
  std::vector<std::tuple<std::string, TM2x2<Type>, unsigned, unsigned>> fwindow;

 private:
  struct Pending;  // queued gate for fusion
  void create_(std::size_t n);
  void release_();
  iqsh_state* dev_() const;  // flush host mirror + fusion queue, return device state
  Type* host_() const;       // host mirror, downloaded if stale; marks it modified
  unsigned pos_(unsigned qubit) const;
  void enqueue_(unsigned npos, unsigned p0, unsigned p1, const Type* m);
  void nest_(const char* who, std::size_t sind, std::size_t eind, const uint64_t c[9], const double md[8], double frac);
  void gate1_(unsigned position, const Type m[4], GateSpec1Q spec, BaseType angle);
  void cgate1_(unsigned cpos, unsigned tpos, const Type m[4], GateSpec2Q spec, BaseType angle);
  BaseType expval_(const std::vector<unsigned>& qubits, const std::vector<unsigned>& obs, BaseType coeff);

  std::size_t global_size_ = 0;
  iqsh_state* h_ = nullptr;
  mutable std::vector<Type> mirror_;
  Type* external_ = nullptr;      // imported host buffer (mirror)
  mutable bool host_valid_ = false;
  mutable bool host_dirty_ = false;
  unsigned fuse_k_ = 5;
  std::size_t fused_gates_ = 0, fused_blocks_ = 0;
  std::shared_ptr<std::vector<Pending>> queue_;
  RandomNumberGenerator<BaseType>* rng_ptr_ = nullptr;
  BaseType T_1_ = 0, T_2_ = 0, T_phi_ = 0;  // noise timescales (chosen time unit)
  static bool do_print_extra_info;

  QubitRegister& operator=(const QubitRegister&) = delete;
};

template <typename Type>
bool QubitRegister<Type>::do_print_extra_info = false;

template <typename Type>
using BaseType = typename QubitRegister<Type>::BaseType;

/// The Pauli-twirl noise rotation of ApplyNoiseGate for an idle time `duration` under
/// timescales T1, T2 (src/qureg_noisysimul.cpp:33-85): three standard normals drawn from
/// the state stream of `rng` (X, Y, Z in that order).  Throws std::invalid_argument for
/// non-positive noise probabilities.
template <class Type>
TM2x2<Type> NoiseGateMatrix(typename extract_value_type<Type>::value_type T1,
                            typename extract_value_type<Type>::value_type T2,
                            typename extract_value_type<Type>::value_type duration,
                            RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng);

/// One sampled branch of a 1-qubit channel, as ApplyChannel applies it
/// (src/qureg_apply_channel.cpp:17-58): one uniform number from the state stream of `rng`;
/// *negative tells whether the branch's eigenvalue is negative.
template <class Type>
TM2x2<Type> ChannelBranchMatrix(CM4x4<Type>& chi, RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng,
                                bool* negative);
/// The same for a 2-qubit channel (src/qureg_apply_channel.cpp:66-106): a 4x4 in the index
/// 2 b(qubit1) + b(qubit2).
template <class Type>
TM4x4<Type> ChannelBranchMatrix(CM16x16<Type>& chi, RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng,
                                bool* negative);

}  // namespace iqs

// Derived registers that count gates and depth, and that insert noise gates automatically
// (the reference includes both here too, include/qureg.hpp:429-433).
#include "QubitRegisterMetric.hpp"
#include "NoisyQureg.hpp"

// The reference's QubitRegisterMetric.hpp and NoisyQureg.hpp, included at the end of its
// qureg.hpp, declare `using namespace std;` at file scope, and callers depend on it
// (benchmarks/basic_code_for_scaling.cpp:156-157 uses ofstream and ios unqualified).
using namespace std;
