// iqs::StatePool: a pool of states in one GPU register (include/state_pool.hpp), over the
// pool entry points of the C ABI (include/iqs_hip.h).
#include "state_pool.hpp"

#include <stdexcept>
#include <string>

#include "../../include/iqs_hip.h"

namespace iqs {

namespace {

void pool_ck(int rc)
{
  if (rc != IQSH_OK) throw std::runtime_error(std::string("iqs_hip: ") + iqsh_last_error());
}

unsigned state_index_bits(unsigned num_states)
{
  if (num_states == 0) throw std::invalid_argument("StatePool: at least one state");
  unsigned k = 0;
  while ((1u << k) < num_states) ++k;
  return k;
}

}  // namespace

template <class Type>
StatePool<Type>::StatePool(unsigned num_qubits, unsigned num_states, std::string style, std::size_t base_index)
    : n_(num_qubits), ns_(num_states), reg_(num_qubits + state_index_bits(num_states), "base", 0),
      gens_(num_states), sign_(num_states, BaseType(1))
{
  if (num_qubits == 0) throw std::invalid_argument("StatePool: at least one qubit per state");
  Initialize(style, base_index);
}

template <class Type>
iqsh_state* StatePool<Type>::dev_()
{
  // The top qubits index the states: a permutation must leave them where they are.
  const auto& map = reg_.qubit_permutation->map;
  for (std::size_t q = n_; q < reg_.num_qubits; ++q)
    if (map[q] != q) throw std::logic_error("StatePool: the state-index qubits were permuted");
  return reg_.DeviceHandle();
}

template <class Type>
unsigned StatePool<Type>::pos_(unsigned qubit) const
{
  if (qubit >= n_) throw std::out_of_range("StatePool: qubit index out of range");
  return (unsigned)reg_.qubit_permutation->map[qubit];
}

template <class Type>
void StatePool<Type>::Initialize(std::string style, std::size_t base_index)
{
  iqsh_state* h = dev_();
  if (style == "base") {
    if (base_index >> n_) throw std::out_of_range("StatePool::Initialize: index out of range");
    pool_ck(iqsh_pool_init(h, n_, 0, reg_.qubit_permutation->program2data_(base_index), ns_));
  } else if (style == "++++") {
    pool_ck(iqsh_pool_init(h, n_, 1, 0, ns_));
  } else {
    throw std::invalid_argument("StatePool::Initialize: style must be \"base\" or \"++++\"");
  }
  std::fill(sign_.begin(), sign_.end(), BaseType(1));
}

template <class Type>
void StatePool<Type>::SetSeedStreamPtrs(std::size_t seed)
{
  for (unsigned s = 0; s < ns_; ++s) gens_[s].SetSeedStreamPtrs(seed, (int)s, (int)ns_, (int)s);
}

template <class Type>
void StatePool<Type>::SetNoiseTimescales(BaseType T1, BaseType T2)
{
  if (!(T2 >= T1 / 2.)) throw std::invalid_argument("SetNoiseTimescales: T2 must be >= T1/2");
  T_1_ = T1;
  T_2_ = T2;
}

// Pack one 2x2 matrix per state as 8 doubles each.
template <class Type>
static void put(std::vector<double>& buf, unsigned s, const TM2x2<Type>& m)
{
  for (int e = 0; e < 4; ++e) {
    buf[8 * s + 2 * e] = (double)std::real(m(e / 2, e % 2));
    buf[8 * s + 2 * e + 1] = (double)std::imag(m(e / 2, e % 2));
  }
}

template <class Type>
void StatePool<Type>::ApplyNoiseGate(unsigned qubit, BaseType duration)
{
  const unsigned p = pos_(qubit);
  if (duration == 0) return;
  std::vector<double> mats(8 * std::size_t(ns_));
  for (unsigned s = 0; s < ns_; ++s) put(mats, s, NoiseGateMatrix<Type>(T_1_, T_2_, duration, gens_[s]));
  pool_ck(iqsh_pool_apply_1q_each(dev_(), n_, p, mats.data(), ns_));
}

template <class Type>
void StatePool<Type>::ApplyChannel(unsigned qubit, CM4x4<Type>& chi)
{
  const unsigned p = pos_(qubit);
  std::vector<double> mats(8 * std::size_t(ns_));
  for (unsigned s = 0; s < ns_; ++s) {
    bool negative = false;
    put(mats, s, ChannelBranchMatrix(chi, gens_[s], &negative));
    if (negative) sign_[s] *= -1;
  }
  pool_ck(iqsh_pool_apply_1q_each(dev_(), n_, p, mats.data(), ns_));
}

template <class Type>
void StatePool<Type>::ApplyChannel(unsigned qubit1, unsigned qubit2, CM16x16<Type>& chi)
{
  const unsigned p1 = pos_(qubit1), p2 = pos_(qubit2);
  if (qubit1 == qubit2) throw std::invalid_argument("StatePool::ApplyChannel: qubits coincide");
  std::vector<double> mats(32 * std::size_t(ns_));
  for (unsigned s = 0; s < ns_; ++s) {
    bool negative = false;
    const TM4x4<Type> m = ChannelBranchMatrix(chi, gens_[s], &negative);
    if (negative) sign_[s] *= -1;
    for (int e = 0; e < 16; ++e) {
      mats[32 * s + 2 * e] = (double)std::real(m(e / 4, e % 4));
      mats[32 * s + 2 * e + 1] = (double)std::imag(m(e / 4, e % 4));
    }
  }
  pool_ck(iqsh_pool_apply_2q_each(dev_(), n_, p1, p2, mats.data(), ns_));
}

template <class Type>
std::vector<typename StatePool<Type>::BaseType> StatePool<Type>::GetProbability(unsigned qubit)
{
  const unsigned p = pos_(qubit);
  std::vector<double> out(ns_);
  pool_ck(iqsh_pool_reduce(dev_(), n_, IQSH_POOL_PROB1, p, out.data(), ns_));
  return std::vector<BaseType>(out.begin(), out.end());
}

template <class Type>
std::vector<typename StatePool<Type>::BaseType> StatePool<Type>::ComputeNorm()
{
  std::vector<double> out(ns_);
  pool_ck(iqsh_pool_reduce(dev_(), n_, IQSH_POOL_NORM2, 0, out.data(), ns_));
  std::vector<BaseType> r(ns_);
  for (unsigned s = 0; s < ns_; ++s) r[s] = (BaseType)std::sqrt(out[s]);
  return r;
}

template <class Type>
std::vector<typename StatePool<Type>::BaseType> StatePool<Type>::ExpectationValueZString(const std::vector<unsigned>& qubits)
{
  uint64_t mask = 0;
  for (unsigned q : qubits) mask ^= uint64_t(1) << pos_(q);
  std::vector<double> out(ns_);
  pool_ck(iqsh_pool_reduce(dev_(), n_, IQSH_POOL_ZSTRING, mask, out.data(), ns_));
  return std::vector<BaseType>(out.begin(), out.end());
}

template <class Type>
std::vector<Type> StatePool<Type>::ComputeOverlap(QubitRegister<Type>& psi)
{
  if (psi.num_qubits != n_) throw std::invalid_argument("StatePool::ComputeOverlap: register size differs");
  for (unsigned q = 0; q < n_; ++q)
    if (psi.qubit_permutation->map[q] != reg_.qubit_permutation->map[q])
      throw std::invalid_argument("StatePool::ComputeOverlap: qubit permutations differ");
  std::vector<double> out(2 * std::size_t(ns_));
  pool_ck(iqsh_pool_overlap(dev_(), n_, psi.DeviceHandle(), out.data(), ns_));
  std::vector<Type> r(ns_);
  for (unsigned s = 0; s < ns_; ++s) r[s] = Type((BaseType)out[2 * s], (BaseType)out[2 * s + 1]);
  return r;
}

template <class Type>
std::vector<Type> StatePool<Type>::GetState(unsigned s)
{
  if (s >= ns_) throw std::out_of_range("StatePool::GetState: state index out of range");
  const uint64_t len = uint64_t(1) << n_;
  std::vector<Type> out(len);
  pool_ck(iqsh_download(dev_(), out.data(), uint64_t(s) * len, len));
  return out;
}

template <class Type>
typename StatePool<Type>::BaseType StatePool<Type>::IncoherentSumOverAllStatesOfPool(const std::vector<BaseType>& values)
{
  BaseType sum = 0;
  for (BaseType v : values) sum += v;
  return sum;
}

template class StatePool<ComplexSP>;
template class StatePool<ComplexDP>;

}  // namespace iqs
