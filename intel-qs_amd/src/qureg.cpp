// iqs::QubitRegister over the C ABI of libiqs_hip.so.  Host bookkeeping only: program
// qubit -> data position, named-gate matrices (built exactly as the reference builds
// them), gate counting, the fusion queue and the host mirror.  Every amplitude update
// and reduction is a kernel launch.
#include "qureg.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "iqs_hip.h"

namespace iqs {

namespace {

void ck(int rc)
{
  if (rc != IQSH_OK) throw std::runtime_error(std::string("iqs_hip: ") + iqsh_last_error());
}

template <class Type>
constexpr int dtype_of();
template <>
constexpr int dtype_of<ComplexSP>() { return IQSH_C64; }
template <>
constexpr int dtype_of<ComplexDP>() { return IQSH_C128; }

template <class Type>
void to_doubles(const Type* m, int count, double* out)
{
  for (int i = 0; i < count; ++i) {
    out[2 * i] = (double)m[i].real();
    out[2 * i + 1] = (double)m[i].imag();
  }
}

template <class Type>
void m2(const TM2x2<Type>& m, Type out[4])
{
  out[0] = m(0, 0);
  out[1] = m(0, 1);
  out[2] = m(1, 0);
  out[3] = m(1, 1);
}

using cd = std::complex<double>;

}  // namespace

// A queued gate: a 2x2 (np = 1) or 4x4 (np = 2, row index 2*b(p[0]) + b(p[1])) matrix on
// data positions, resolved when the gate was queued.
template <class Type>
struct QubitRegister<Type>::Pending {
  unsigned np;
  unsigned p[2];
  cd m[16];
};

// ---------------------------------------------------------------------------------------
// lifetime

template <class Type>
void QubitRegister<Type>::create_(std::size_t n)
{
  ck(iqsh_create((unsigned)n, dtype_of<Type>(), &h_));
  num_qubits = n;
  global_size_ = std::size_t(1) << n;
  delete qubit_permutation;
  qubit_permutation = new Permutation(n);
  mirror_.clear();
  host_valid_ = host_dirty_ = false;
  if (!queue_) queue_ = std::make_shared<std::vector<Pending>>();
  queue_->clear();
  fwindow.clear();
}

template <class Type>
void QubitRegister<Type>::release_()
{
  if (h_) {
    iqsh_destroy(h_);
    h_ = nullptr;
  }
}

template <class Type>
QubitRegister<Type>::QubitRegister()
{
  // Reference default constructor: one amplitude equal to 1 (src/qureg_init.cpp:22-47).
  create_(0);
  Initialize("base", 0);
}

template <class Type>
QubitRegister<Type>::QubitRegister(std::size_t n, std::string style, std::size_t base_index, std::size_t)
{
  create_(n);
  Initialize(style, base_index);
}

template <class Type>
QubitRegister<Type>::QubitRegister(const QubitRegister& in)
{
  create_(in.num_qubits);
  ck(iqsh_copy(h_, in.dev_()));
  *qubit_permutation = *in.qubit_permutation;
}

template <class Type>
QubitRegister<Type>::QubitRegister(std::size_t n, Type* state, std::size_t)
{
  create_(n);
  imported_state = true;
  external_ = state;
  ck(iqsh_upload(h_, state, 0, global_size_));
  host_valid_ = true;
}

template <class Type>
QubitRegister<Type>::~QubitRegister()
{
  try {
    if (external_ && h_) {
      dev_();
      ck(iqsh_download(h_, external_, 0, global_size_));
    }
  } catch (...) {
  }
  release_();
  delete timer;
  delete gate_counter;
  delete qubit_permutation;
}

template <class Type>
void QubitRegister<Type>::Allocate(std::size_t n, std::size_t)
{
  release_();
  external_ = nullptr;
  imported_state = specialize = fusion = false;
  create_(n);
  ck(iqsh_init_uniform(h_, 0, 0));
}

template <class Type>
void QubitRegister<Type>::Initialize(std::size_t n, std::size_t tmp)
{
  Allocate(n, tmp);
}

template <class Type>
void QubitRegister<Type>::AllocateAdditionalQubit()
{
  Allocate(num_qubits + 1, 0);
}

template <class Type>
void QubitRegister<Type>::Initialize(std::string style, std::size_t base_index)
{
  // src/qureg_init.cpp:225-359
  if (queue_) queue_->clear();
  fwindow.clear();
  host_dirty_ = false;
  host_valid_ = false;
  if (style == "base") {
    if (base_index >= global_size_) throw std::out_of_range("Initialize(\"base\"): index out of range");
    ck(iqsh_init_basis(h_, qubit_permutation->program2data_(base_index)));
  } else if (style == "++++") {
    const BaseType a = BaseType(1. / std::sqrt((double)global_size_));
    ck(iqsh_init_uniform(h_, (double)a, 0.0));
  } else if (style == "rand") {
    // Same streams as the reference: base_index == 0 uses the pool stream (common state),
    // base_index == number of states uses the local stream; uniform in [-1, 1), normalized.
    if (!rng_ptr_) throw std::logic_error("Initialize(\"rand\") needs an RNG (SetRngPtr)");
    if (rng_ptr_->GetSeed() == 0) throw std::logic_error("Initialize(\"rand\") needs a non-zero seed");
    if (base_index != 0 && base_index != (std::size_t)mpi::Environment::GetNumStates())
      throw std::invalid_argument("Initialize(\"rand\"): base_index must be 0 or the number of states");
    const std::string stream = base_index > 0 ? "local" : "pool";
    RandomNumberGenerator<BaseType> thread_rng(rng_ptr_);
    std::vector<Type> buf(global_size_);
    thread_rng.UniformRandomNumbers(reinterpret_cast<BaseType*>(buf.data()), 2 * global_size_, BaseType(-1),
                                    BaseType(1), stream);
    rng_ptr_->SkipAhead(2 * global_size_, stream);
    ck(iqsh_upload(h_, buf.data(), 0, global_size_));
    Normalize();
  } else {
    ck(iqsh_init_uniform(h_, 0.0, 0.0));
  }
}

// ---------------------------------------------------------------------------------------
// residency

template <class Type>
iqsh_state* QubitRegister<Type>::dev_() const
{
  auto* self = const_cast<QubitRegister<Type>*>(this);
  if (host_dirty_) {
    ck(iqsh_upload(h_, external_ ? external_ : mirror_.data(), 0, global_size_));
    host_dirty_ = false;
  }
  if (queue_ && !queue_->empty()) self->ApplyFusedGates();
  host_valid_ = false;
  return h_;
}

template <class Type>
Type* QubitRegister<Type>::host_() const
{
  if (queue_ && !queue_->empty()) dev_();
  Type* p = external_;
  if (!p) {
    if (mirror_.size() != global_size_) mirror_.assign(global_size_, Type(0));
    p = mirror_.data();
  }
  if (!host_valid_) {
    ck(iqsh_download(h_, p, 0, global_size_));
    host_valid_ = true;
  }
  host_dirty_ = true;
  return p;
}

template <class Type>
void QubitRegister<Type>::Sync() const
{
  dev_();
  ck(iqsh_sync());
}

template <class Type>
iqsh_state* QubitRegister<Type>::DeviceHandle()
{
  return dev_();
}

template <class Type>
Type& QubitRegister<Type>::operator[](std::size_t i)
{
  return host_()[i];
}

template <class Type>
Type& QubitRegister<Type>::operator[](std::size_t i) const
{
  return host_()[i];
}

template <class Type>
Type* QubitRegister<Type>::RawState()
{
  return host_();
}

template <class Type>
Type QubitRegister<Type>::GetGlobalAmplitude(std::size_t index) const
{
  if (index >= global_size_) throw std::out_of_range("GetGlobalAmplitude: index out of range");
  const std::size_t d = qubit_permutation->program2data_(index);
  if (host_dirty_ || host_valid_) return (external_ ? external_ : mirror_.data())[d];
  Type v;
  ck(iqsh_download(dev_(), &v, d, 1));
  return v;
}

template <class Type>
void QubitRegister<Type>::SetGlobalAmplitude(std::size_t index, Type value)
{
  if (index >= global_size_) throw std::out_of_range("SetGlobalAmplitude: index out of range");
  const std::size_t d = qubit_permutation->program2data_(index);
  if (host_dirty_) {
    (external_ ? external_ : mirror_.data())[d] = value;
    return;
  }
  ck(iqsh_upload(dev_(), &value, d, 1));
}

template <class Type>
unsigned QubitRegister<Type>::pos_(unsigned qubit) const
{
  if (qubit >= num_qubits) throw std::out_of_range("qubit index " + toString(qubit) + " out of range");
  return (unsigned)qubit_permutation->map[qubit];
}

// ---------------------------------------------------------------------------------------
// statistics

template <class Type>
void QubitRegister<Type>::EnableStatistics()
{
  if (!gate_counter) gate_counter = new GateCounter((int)num_qubits);
  if (!timer) timer = new Timer(num_qubits);
}

template <class Type>
void QubitRegister<Type>::GetStatistics()
{
  if (timer) timer->Breakdown();
  if (gate_counter) gate_counter->Breakdown();
}

template <class Type>
void QubitRegister<Type>::DisableStatistics()
{
  delete timer;
  delete gate_counter;
  timer = nullptr;
  gate_counter = nullptr;
}

template <class Type>
void QubitRegister<Type>::ResetStatistics()
{
  if (timer) timer->Reset();
  if (gate_counter) gate_counter->Reset();
}

// ---------------------------------------------------------------------------------------
// permutation

template <class Type>
void QubitRegister<Type>::PermuteQubits(std::vector<std::size_t> new_map, std::string style)
{
  if (new_map.size() != num_qubits) throw std::invalid_argument("PermuteQubits: map of the wrong size");
  Permutation np(new_map, style);
  if (np.imap == qubit_permutation->imap) return;
  // New position b holds program qubit np.imap[b], which sits at old position map[np.imap[b]].
  std::vector<unsigned> src(num_qubits);
  for (std::size_t b = 0; b < num_qubits; ++b) src[b] = (unsigned)qubit_permutation->map[np.imap[b]];
  ck(iqsh_permute(dev_(), src.data()));
  *qubit_permutation = np;
}

template <class Type>
void QubitRegister<Type>::PermuteLocalQubits(std::vector<std::size_t> new_map, std::string style)
{
  PermuteQubits(new_map, style);
}

template <class Type>
void QubitRegister<Type>::EmulateSwap(unsigned q1, unsigned q2)
{
  pos_(q1);
  pos_(q2);
  qubit_permutation->ExchangeTwoElements(q1, q2);
}

// ---------------------------------------------------------------------------------------
// gates

template <class Type>
void QubitRegister<Type>::gate1_(unsigned p, const Type m[4], GateSpec1Q spec, BaseType angle)
{
  iqsh_state* h = dev_();
  if (timer) timer->Start("SQG(" + toString(p) + ")", p);
  if (specialize2 && spec != GateSpec1Q::None) {
    ck(iqsh_apply_1q_spec(h, p, (int)spec, (double)angle));
  } else {
    double md[8];
    to_doubles(m, 4, md);
    ck(iqsh_apply_1q(h, p, md));
  }
  if (timer) {
    ck(iqsh_sync());
    timer->Stop(2.0 * global_size_ * sizeof(Type));
  }
}

template <class Type>
void QubitRegister<Type>::cgate1_(unsigned c, unsigned t, const Type m[4], GateSpec2Q spec, BaseType angle)
{
  iqsh_state* h = dev_();
  if (timer) timer->Start("CSQG(" + toString(c) + "," + toString(t) + ")", c, t);
  if (specialize2 && spec != GateSpec2Q::None) {
    ck(iqsh_apply_c1q_spec(h, c, t, (int)spec, (double)angle));
  } else {
    double md[8];
    to_doubles(m, 4, md);
    ck(iqsh_apply_c1q(h, c, t, md));
  }
  if (timer) {
    ck(iqsh_sync());
    timer->Stop(1.0 * global_size_ * sizeof(Type));
  }
}

template <class Type>
void QubitRegister<Type>::enqueue_(unsigned np, unsigned p0, unsigned p1, const Type* m)
{
  Pending g;
  g.np = np;
  g.p[0] = p0;
  g.p[1] = p1;
  const int cnt = np == 1 ? 4 : 16;
  for (int i = 0; i < cnt; ++i) g.m[i] = cd((double)m[i].real(), (double)m[i].imag());
  queue_->push_back(g);
}

// The 2x2 (as 8 doubles) a sub-range helper applies: the SpecializeV2 constants when that
// path is on and the kind has a body, else the caller's matrix.
template <class Type>
static void range_matrix(const Type mm[4], bool spec_on, int spec, bool controlled, double angle, double md[8])
{
  if (spec_on) ck(iqsh_spec_matrix(dtype_of<Type>(), spec, controlled ? 1 : 0, angle, md));
  else to_doubles(mm, 4, md);
}

template <class Type>
void QubitRegister<Type>::nest_(const char* who, std::size_t sind, std::size_t eind, const uint64_t c[9],
                                const double md[8], double frac)
{
  if (sind > eind || eind > global_size_)
    throw std::invalid_argument(std::string(who) + ": index range outside the register");
  iqsh_state* h = dev_();
  if (timer) timer->Start(std::string(who), 0);
  ck(iqsh_apply_nest(h, c, md));
  if (timer) {
    ck(iqsh_sync());
    timer->Stop(2.0 * frac * double(eind - sind) * sizeof(Type));
  }
}

template <class Type>
bool QubitRegister<Type>::Apply1QubitGate_helper(unsigned qubit, TM2x2<Type> const& m, std::size_t sind,
                                                 std::size_t eind, GateSpec1Q spec, BaseType angle)
{
  Type mm[4];
  m2(m, mm);
  if (sind == 0 && eind == global_size_) {
    gate1_(pos_(qubit), mm, spec, angle);
    return true;
  }
  // Loop_DN over [sind, eind) (src/qureg_apply1qubitgate.cpp:196-202): groups of 2^(P+1)
  // from sind, pairs (i, i + 2^P) in the first half of each.
  const uint64_t P = pos_(qubit);
  double md[8];
  range_matrix(mm, specialize2 && spec != GateSpec1Q::None, (int)spec, false, (double)angle, md);
  const uint64_t c[9] = {sind, eind, uint64_t(2) << P, 0, 1, 1, 0, uint64_t(1) << P, uint64_t(1) << P};
  nest_("Apply1QubitGate_helper", sind, eind, c, md, 1.0);
  return true;
}

template <class Type>
void QubitRegister<Type>::Apply1QubitGate(unsigned qubit, TM2x2<Type> const& m, GateSpec1Q spec, BaseType angle)
{
  const unsigned p = pos_(qubit);
  if (gate_counter) gate_counter->OneQubitIncrement(qubit);
  Type mm[4];
  m2(m, mm);
  if (fusion) {
    enqueue_(1, p, 0, mm);
    fwindow.push_back(std::make_tuple(std::string("sqg"), m, qubit, 0U));
    return;
  }
  gate1_(p, mm, spec, angle);
}

template <class Type>
bool QubitRegister<Type>::ApplyControlled1QubitGate_helper(unsigned c, unsigned t, TM2x2<Type> const& m,
                                                           std::size_t sind, std::size_t eind, GateSpec2Q spec,
                                                           BaseType angle)
{
  if (c == t) throw std::invalid_argument("control and target qubit coincide");
  Type mm[4];
  m2(m, mm);
  if (sind == 0 && eind == global_size_) {
    cgate1_(pos_(c), pos_(t), mm, spec, angle);
    return true;
  }
  // The reference's sub-range nests (src/qureg_applyctrl1qubitgate.cpp:262-346; every
  // position is local on one GPU).
  const uint64_t C = pos_(c), T = pos_(t), one = 1;
  const bool spec_on = specialize2 && spec != GateSpec2Q::None;
  double md[8];
  if (C > T && C >= log2llc && global_size_ > eind - sind) {
    // control stride beyond the cache block: the whole block has bit C = 0 or 1
    if (!((sind >> C) & 1)) return false;
    const GateSpec1Q s1 = ConvertSpec2to1(spec);
    range_matrix(mm, specialize2 && s1 != GateSpec1Q::None, (int)s1, false, (double)angle, md);
    const uint64_t nest[9] = {sind, eind, one << (T + 1), 0, 1, 1, 0, one << T, one << T};
    nest_("ApplyControlled1QubitGate_helper", sind, eind, nest, md, 1.0);
    return true;
  }
  range_matrix(mm, spec_on, (int)spec, true, (double)angle, md);
  if (C > T) {
    const uint64_t nest[9] = {sind, eind, one << (C + 1), one << C, one << (C + 1), one << (T + 1), 0, one << T, one << T};
    nest_("ApplyControlled1QubitGate_helper", sind, eind, nest, md, 0.5);
  } else {
    const uint64_t nest[9] = {sind, eind, one << (T + 1), 0, one << T, one << (C + 1), one << C, one << (C + 1), one << T};
    nest_("ApplyControlled1QubitGate_helper", sind, eind, nest, md, 0.5);
  }
  return true;
}

template <class Type>
void QubitRegister<Type>::ApplyControlled1QubitGate(unsigned c, unsigned t, TM2x2<Type> const& m, GateSpec2Q spec,
                                                    BaseType angle)
{
  const unsigned cp = pos_(c), tp = pos_(t);
  if (c == t) throw std::invalid_argument("control and target qubit coincide");
  if (gate_counter) gate_counter->TwoQubitIncrement(c, t);
  Type mm[4];
  m2(m, mm);
  if (fusion) {
    Type m4[16] = {};
    m4[0] = m4[5] = Type(1, 0);
    m4[10] = mm[0];
    m4[11] = mm[1];
    m4[14] = mm[2];
    m4[15] = mm[3];
    enqueue_(2, cp, tp, m4);
    fwindow.push_back(std::make_tuple(std::string("cqg"), m, c, t));
    return;
  }
  cgate1_(cp, tp, mm, spec, angle);
}

template <class Type>
bool QubitRegister<Type>::ApplySwap_helper(unsigned q1, unsigned q2, TM2x2<Type> const& m)
{
  if (gate_counter) gate_counter->TwoQubitIncrement(q1, q2);
  unsigned p1 = pos_(q1), p2 = pos_(q2);
  if (p1 == p2) throw std::invalid_argument("swap on a single qubit");
  if (p1 > p2) {
    std::swap(p1, p2);
    if (!(m(0, 1) == m(1, 0))) throw std::invalid_argument("swap-type gate must have m01 == m10");
  }
  iqsh_state* h = dev_();
  double md[8];
  Type mm[4];
  m2(m, mm);
  to_doubles(mm, 4, md);
  if (timer) timer->Start("TQG(" + toString(p1) + "," + toString(p2) + ")", p1, p2);
  ck(iqsh_apply_swaplike(h, p1, p2, md));
  if (timer) {
    ck(iqsh_sync());
    timer->Stop(1.0 * global_size_ * sizeof(Type));
  }
  return true;
}

template <class Type>
void QubitRegister<Type>::ApplySwap(unsigned q1, unsigned q2)
{
  TM2x2<Type> g;
  g(0, 0) = g(1, 1) = Type(0, 0);
  g(0, 1) = g(1, 0) = Type(1, 0);
  ApplySwap_helper(q1, q2, g);
}

template <class Type>
void QubitRegister<Type>::ApplyISwap(unsigned q1, unsigned q2)
{
  TM2x2<Type> g;
  g(0, 0) = g(1, 1) = Type(0, 0);
  g(0, 1) = g(1, 0) = Type(0, 1);
  ApplySwap_helper(q1, q2, g);
}

template <class Type>
void QubitRegister<Type>::ApplySqrtISwap(unsigned q1, unsigned q2)
{
  const BaseType f = 1. / std::sqrt(2.);
  TM2x2<Type> g;
  g(0, 0) = g(1, 1) = Type(f, 0);
  g(0, 1) = g(1, 0) = Type(0, f);
  ApplySwap_helper(q1, q2, g);
}

template <class Type>
void QubitRegister<Type>::ApplyISwapRotation(unsigned q1, unsigned q2, TM2x2<Type> const& m)
{
  if (!(m(0, 1) == m(1, 0))) throw std::invalid_argument("ApplyISwapRotation: m01 must equal m10");
  ApplySwap_helper(q1, q2, m);
}

template <class Type>
void QubitRegister<Type>::Apply4thRootISwap(unsigned q1, unsigned q2)
{
  // qureg_applyswap.cpp:93-108: a = polar(1/2, pi/8), b = polar(1/2, 7 pi/8).
  const std::complex<double> a(.5 * std::cos(M_PI / 8.), .5 * std::sin(M_PI / 8.));
  const std::complex<double> b(.5 * std::cos(7. * M_PI / 8.), .5 * std::sin(7. * M_PI / 8.));
  TM2x2<Type> g;
  g(0, 0) = g(1, 1) = Type(a - b);
  g(0, 1) = g(1, 0) = Type(a + b);
  ApplySwap_helper(q1, q2, g);
}

template <class Type>
void QubitRegister<Type>::ApplyDiag(unsigned q1, unsigned q2, TM4x4<Type> const& m)
{
  const unsigned p1 = pos_(q1), p2 = pos_(q2);
  if (gate_counter) gate_counter->TwoQubitIncrement(q1, q2);
  Type d[4] = {m(0, 0), m(1, 1), m(2, 2), m(3, 3)};
  double dd[8];
  to_doubles(d, 4, dd);
  iqsh_state* h = dev_();
  if (timer) timer->Start("DIAG(" + toString(p1) + "," + toString(p2) + ")", p1, p2);
  ck(iqsh_apply_diag2(h, p1, p2, dd));
  if (timer) {
    ck(iqsh_sync());
    timer->Stop(2.0 * global_size_ * sizeof(Type));
  }
}

template <class Type>
void QubitRegister<Type>::ApplyDiagSimp(unsigned q1, unsigned q2, TM4x4<Type> const& m)
{
  // Same diagonal action as ApplyDiag (qureg_applydiag.cpp:17-51), without counting.
  const unsigned p1 = pos_(q1), p2 = pos_(q2);
  Type d[4] = {m(0, 0), m(1, 1), m(2, 2), m(3, 3)};
  double dd[8];
  to_doubles(d, 4, dd);
  ck(iqsh_apply_diag2(dev_(), p1, p2, dd));
}

template <class Type>
void QubitRegister<Type>::Apply2QubitGate(unsigned const qh, unsigned const ql, TM4x4<Type> const& m)
{
  const unsigned ph = pos_(qh), pl = pos_(ql);
  if (qh == ql) throw std::invalid_argument("Apply2QubitGate: qubits coincide");
  Type mm[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) mm[4 * i + j] = m(i, j);
  if (fusion) {
    enqueue_(2, ph, pl, mm);
  } else {
    double md[32];
    to_doubles(mm, 16, md);
    iqsh_state* h = dev_();
    if (timer) timer->Start("2QG(" + toString(ph) + "," + toString(pl) + ")", ph, pl);
    ck(iqsh_apply_2q(h, ph, pl, md));
    if (timer) {
      ck(iqsh_sync());
      timer->Stop(2.0 * global_size_ * sizeof(Type));
    }
  }
  if (gate_counter) gate_counter->TwoQubitIncrement(qh, ql);
}

// ---- named 1q gates: matrices as in src/qureg_apply1qubitgate.cpp:277-501 ----
template <class Type>
void QubitRegister<Type>::ApplyRotationX(unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 1) = m(1, 0) = Type(0, -std::sin(theta / 2.));
  m(0, 0) = m(1, 1) = Type(std::cos(theta / 2.), 0);
  Apply1QubitGate(q, m, GateSpec1Q::RotationX, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyRotationY(unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 1) = Type(-std::sin(theta / 2.), 0.);
  m(1, 0) = Type(std::sin(theta / 2.), 0.);
  m(0, 0) = m(1, 1) = Type(std::cos(theta / 2.), 0);
  Apply1QubitGate(q, m, GateSpec1Q::RotationY, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyRotationZ(unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 0) = Type(std::cos(theta / 2.), -std::sin(theta / 2.));
  m(1, 1) = Type(std::cos(theta / 2.), std::sin(theta / 2.));
  m(0, 1) = m(1, 0) = Type(0., 0.);
  Apply1QubitGate(q, m, GateSpec1Q::RotationZ, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliX(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 1) = m(1, 0) = Type(1., 0.);
  Apply1QubitGate(q, m, GateSpec1Q::PauliX);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliSqrtX(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = m(1, 1) = Type(0.5, 0.5);
  m(0, 1) = m(1, 0) = Type(0.5, -0.5);
  Apply1QubitGate(q, m);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliY(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 1) = Type(0., -1.);
  m(1, 0) = Type(0., 1.);
  Apply1QubitGate(q, m, GateSpec1Q::PauliY);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliSqrtY(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = m(1, 0) = m(1, 1) = Type(0.5, 0.5);
  m(0, 1) = Type(-0.5, -0.5);
  Apply1QubitGate(q, m);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliZ(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1., 0.);
  m(1, 1) = Type(-1., 0.);
  Apply1QubitGate(q, m, GateSpec1Q::PauliZ);
}

template <class Type>
void QubitRegister<Type>::ApplyPauliSqrtZ(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1., 0.);
  m(1, 1) = Type(0., 1.);
  Apply1QubitGate(q, m);
}

template <class Type>
void QubitRegister<Type>::ApplyHadamard(unsigned const q)
{
  const BaseType f = 1. / std::sqrt(2.);
  TM2x2<Type> m;
  m(0, 0) = m(0, 1) = m(1, 0) = Type(f, 0.);
  m(1, 1) = Type(-f, 0.);
  Apply1QubitGate(q, m, GateSpec1Q::Hadamard);
}

template <class Type>
void QubitRegister<Type>::ApplyRotationXY(unsigned const q, BaseType phi, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 0) = m(1, 1) = Type(std::cos(theta / 2.), 0);
  m(0, 1) = Type(-std::sin(theta / 2.) * std::sin(phi), -std::sin(theta / 2.) * std::cos(phi));
  m(1, 0) = Type(std::sin(theta / 2.) * std::sin(phi), -std::sin(theta / 2.) * std::cos(phi));
  Apply1QubitGate(q, m);
}

template <class Type>
void QubitRegister<Type>::ApplyT(unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1.0, 0.0);
  m(1, 1) = Type(std::cos(M_PI / 4.0), std::sin(M_PI / 4.0));
  Apply1QubitGate(q, m, GateSpec1Q::T);
}

// ---- named controlled gates: src/qureg_applyctrl1qubitgate.cpp:458-618 ----
template <class Type>
void QubitRegister<Type>::ApplyCRotationX(unsigned const c, unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 1) = m(1, 0) = Type(0, -std::sin(theta / 2.));
  m(0, 0) = m(1, 1) = Type(std::cos(theta / 2.), 0);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CRotationX, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyCRotationY(unsigned const c, unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 1) = Type(-std::sin(theta / 2.), 0.);
  m(1, 0) = Type(std::sin(theta / 2.), 0.);
  m(0, 0) = m(1, 1) = Type(std::cos(theta / 2.), 0);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CRotationY, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyCRotationZ(unsigned const c, unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 0) = Type(std::cos(theta / 2.), -std::sin(theta / 2.));
  m(1, 1) = Type(std::cos(theta / 2.), std::sin(theta / 2.));
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CRotationZ, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyCPauliX(unsigned const c, unsigned const q)
{
  TM2x2<Type> m;
  m(0, 1) = m(1, 0) = Type(1., 0.);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CPauliX);
}

template <class Type>
void QubitRegister<Type>::ApplyCPauliY(unsigned const c, unsigned const q)
{
  TM2x2<Type> m;
  m(0, 1) = Type(0., -1.);
  m(1, 0) = Type(0., 1.);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CPauliY);
}

template <class Type>
void QubitRegister<Type>::ApplyCPauliZ(unsigned const c, unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1., 0.);
  m(1, 1) = Type(-1., 0.);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CPauliZ);
}

template <class Type>
void QubitRegister<Type>::ApplyCPauliSqrtZ(unsigned const c, unsigned const q)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1., 0.);
  m(1, 1) = Type(0., 1.);
  ApplyControlled1QubitGate(c, q, m);
}

template <class Type>
void QubitRegister<Type>::ApplyCHadamard(unsigned const c, unsigned const q)
{
  const BaseType f = 1. / std::sqrt(2.);
  TM2x2<Type> m;
  m(0, 0) = m(0, 1) = m(1, 0) = Type(f, 0.);
  m(1, 1) = Type(-f, 0.);
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CHadamard);
}

template <class Type>
void QubitRegister<Type>::ApplyCPhaseRotation(unsigned const c, unsigned const q, BaseType theta)
{
  TM2x2<Type> m;
  m(0, 0) = Type(1, 0);
  m(1, 1) = Type(std::cos(theta), std::sin(theta));
  ApplyControlled1QubitGate(c, q, m, GateSpec2Q::CPhase, theta);
}

template <class Type>
void QubitRegister<Type>::ApplyToffoli(unsigned const c1, unsigned const c2, unsigned const t)
{
  // The reference applies five controlled gates (qureg_applytoffoli.cpp:24-47); the result
  // is the doubly-controlled X, applied here as one exact move on the quarter of the state
  // with both controls set.  The gate counter records the same five two-qubit gates.
  const unsigned p1 = pos_(c1), p2 = pos_(c2), pt = pos_(t);
  if (c1 == c2 || c1 == t || c2 == t) throw std::invalid_argument("ApplyToffoli: qubits must differ");
  if (gate_counter) {
    gate_counter->TwoQubitIncrement(c1, t);
    gate_counter->TwoQubitIncrement(c2, c1);
    gate_counter->TwoQubitIncrement(c1, t);
    gate_counter->TwoQubitIncrement(c2, c1);
    gate_counter->TwoQubitIncrement(c2, t);
  }
  const double x[8] = {0, 0, 1, 0, 1, 0, 0, 0};
  ck(iqsh_apply_mc1q(dev_(), (uint64_t(1) << p1) | (uint64_t(1) << p2), pt, x));
}

// ---------------------------------------------------------------------------------------
// fusion: queued gates are multiplied into dense blocks of <= fuse_k_ positions

template <class Type>
void QubitRegister<Type>::TurnOnFusion(unsigned l2)
{
  log2llc = l2;
  fusion = true;
  if (do_print_extra_info) printf("Fusion is enabled: blocks of up to %u qubits\n", fuse_k_);
}

template <class Type>
void QubitRegister<Type>::TurnOffFusion()
{
  ApplyFusedGates();
  fusion = false;
}

template <class Type>
bool QubitRegister<Type>::IsFusionEnabled()
{
  return fusion;
}

template <class Type>
void QubitRegister<Type>::SetFusionBlockQubits(unsigned k)
{
  if (k < 1 || k > 5) throw std::invalid_argument("fusion block size must be in [1, 5]");
  ApplyFusedGates();
  fuse_k_ = k;
}

template <class Type>
void QubitRegister<Type>::ApplyFusedGates()
{
  if (!queue_ || queue_->empty()) {
    fwindow.clear();
    return;
  }
  std::vector<Pending> q;
  q.swap(*queue_);
  fused_gates_ += q.size();
  fwindow.clear();
  if (host_dirty_) {
    ck(iqsh_upload(h_, external_ ? external_ : mirror_.data(), 0, global_size_));
    host_dirty_ = false;
  }
  host_valid_ = false;

  // Plan the blocks over the gates' dependency graph (gate j depends on the previous gate on
  // each of its positions; gates on disjoint positions commute).  A block starts empty and
  // repeatedly absorbs the ready gate (every predecessor already placed) that adds the fewest
  // new positions, oldest first, while it stays within fuse_k_ positions; the gates of a
  // block multiply in absorption order, which respects every dependency.  200 Haar-random
  // 1q/2q gates on 32 positions (seed 20240603) take 26 blocks of <= 5 positions this way,
  // against 59 when a block is closed at the first gate that does not fit.
  const std::size_t ng = q.size();
  std::vector<std::vector<std::size_t>> succ(ng);
  std::vector<unsigned> npred(ng, 0);
  {
    std::vector<long> last(64, -1);
    for (std::size_t i = 0; i < ng; ++i) {
      long prev = -1;
      for (unsigned j = 0; j < q[i].np; ++j) {
        const long l = last[q[i].p[j]];
        if (l >= 0 && l != prev) {
          succ[(std::size_t)l].push_back(i);
          ++npred[i];
          prev = l;
        }
        last[q[i].p[j]] = (long)i;
      }
    }
  }
  std::vector<std::size_t> ready;  // ascending gate index
  for (std::size_t i = 0; i < ng; ++i)
    if (!npred[i]) ready.push_back(i);
  std::vector<std::vector<std::size_t>> plan;
  while (!ready.empty()) {
    std::vector<std::size_t> blk;
    std::vector<unsigned> Q;
    for (;;) {
      std::size_t best = ready.size();
      unsigned best_add = ~0u;
      for (std::size_t r = 0; r < ready.size() && best_add != 0; ++r) {
        const Pending& g = q[ready[r]];
        unsigned add = 0;
        for (unsigned j = 0; j < g.np; ++j) add += std::find(Q.begin(), Q.end(), g.p[j]) == Q.end();
        // (a gate wider than fuse_k_ is a block of its own)
        if ((Q.size() + add <= fuse_k_ || Q.empty()) && add < best_add) {
          best = r;
          best_add = add;
        }
      }
      if (best == ready.size()) break;
      const std::size_t gi = ready[best];
      ready.erase(ready.begin() + (long)best);
      for (unsigned j = 0; j < q[gi].np; ++j)
        if (std::find(Q.begin(), Q.end(), q[gi].p[j]) == Q.end()) Q.push_back(q[gi].p[j]);
      blk.push_back(gi);
      for (std::size_t sj : succ[gi])
        if (--npred[sj] == 0) ready.insert(std::lower_bound(ready.begin(), ready.end(), sj), sj);
    }
    plan.push_back(std::move(blk));
  }

  std::vector<unsigned> S;  // block positions; row-index bit j <-> S[j]
  std::vector<cd> B;        // 2^|S| x 2^|S| row-major
  int ngates = 0;
  auto emit = [&]() {
    if (S.empty()) return;
    const unsigned k = (unsigned)S.size();
    const std::size_t D = std::size_t(1) << k;
    std::vector<double> md(2 * D * D);
    for (std::size_t i = 0; i < D * D; ++i) {
      md[2 * i] = B[i].real();
      md[2 * i + 1] = B[i].imag();
    }
    if (k == 1) ck(iqsh_apply_1q(h_, S[0], md.data()));
    else if (k == 2) ck(iqsh_apply_2q(h_, S[1], S[0], md.data()));  // row = 2*b(S[1]) + b(S[0])
    else ck(iqsh_apply_kq(h_, k, S.data(), md.data()));
    ++fused_blocks_;
    S.clear();
    B.clear();
    ngates = 0;
  };
  for (const std::vector<std::size_t>& blk : plan) {
    for (std::size_t gi : blk) {
      const Pending& g = q[gi];
      // Positions of the gate, in its own index order: for np = 2, gate row = 2*b(p[0]) + b(p[1]).
      std::vector<unsigned> U(S);
      for (unsigned j = 0; j < g.np; ++j)
        if (std::find(U.begin(), U.end(), g.p[j]) == U.end()) U.push_back(g.p[j]);
      // Grow B to the positions U (new positions appended above the old ones: B (x) I).
      const std::size_t Dold = std::size_t(1) << S.size(), D = std::size_t(1) << U.size();
      std::vector<cd> Bn(D * D, cd(0, 0));
      if (S.empty()) {
        for (std::size_t i = 0; i < D; ++i) Bn[i * D + i] = 1.0;
      } else {
        for (std::size_t r = 0; r < D; ++r)
          for (std::size_t c = 0; c < D; ++c)
            if ((r / Dold) == (c / Dold)) Bn[r * D + c] = B[(r % Dold) * Dold + (c % Dold)];
      }
      S = U;
      // Left-multiply by the gate expanded to S.
      unsigned bitof[2];
      for (unsigned j = 0; j < g.np; ++j)
        bitof[j] = (unsigned)(std::find(S.begin(), S.end(), g.p[j]) - S.begin());
      const std::size_t gmask = (g.np == 1) ? (std::size_t(1) << bitof[0])
                                            : ((std::size_t(1) << bitof[0]) | (std::size_t(1) << bitof[1]));
      auto gidx = [&](std::size_t r) -> unsigned {
        if (g.np == 1) return (unsigned)((r >> bitof[0]) & 1);
        return (unsigned)(2 * ((r >> bitof[0]) & 1) + ((r >> bitof[1]) & 1));
      };
      const unsigned gd = g.np == 1 ? 2 : 4;
      std::vector<cd> R(D * D, cd(0, 0));
      for (std::size_t r = 0; r < D; ++r) {
        const unsigned gr = gidx(r);
        for (unsigned gc = 0; gc < gd; ++gc) {
          const cd gv = g.m[gr * gd + gc];
          if (gv == cd(0, 0)) continue;
          // the index that differs from r only in the gate bits, with gate value gc
          std::size_t rr = r & ~gmask;
          if (g.np == 1) rr |= (std::size_t)(gc & 1) << bitof[0];
          else rr |= ((std::size_t)((gc >> 1) & 1) << bitof[0]) | ((std::size_t)(gc & 1) << bitof[1]);
          for (std::size_t c = 0; c < D; ++c) R[r * D + c] += gv * Bn[rr * D + c];
        }
      }
      B.swap(R);
      ++ngates;
    }
    emit();
  }
}

// ---------------------------------------------------------------------------------------
// specialization flags

template <class Type>
void QubitRegister<Type>::TurnOnSpecialize()
{
  specialize = true;
}
template <class Type>
void QubitRegister<Type>::TurnOffSpecialize()
{
  specialize = false;
}
template <class Type>
void QubitRegister<Type>::TurnOnSpecializeV2()
{
  specialize2 = true;
}
template <class Type>
void QubitRegister<Type>::TurnOffSpecializeV2()
{
  specialize2 = false;
}

// ---------------------------------------------------------------------------------------
// measurement

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::GetProbability(unsigned qubit)
{
  const unsigned p = pos_(qubit);
  double r;
  ck(iqsh_prob1(dev_(), p, &r));
  return (BaseType)r;
}

template <class Type>
void QubitRegister<Type>::CollapseQubit(unsigned qubit, bool value)
{
  const unsigned p = pos_(qubit);
  ck(iqsh_collapse(dev_(), p, value ? 1 : 0));
}

template <class Type>
bool QubitRegister<Type>::IsClassicalBit(unsigned qubit, BaseType tol) const
{
  const unsigned p = pos_(qubit);
  double m0, m1;
  ck(iqsh_maxnorm_half(dev_(), p, 0, &m0));
  ck(iqsh_maxnorm_half(dev_(), p, 1, &m1));
  return !((BaseType)m0 > tol && (BaseType)m1 > tol);
}

template <class Type>
bool QubitRegister<Type>::GetClassicalValue(unsigned qubit, BaseType tol) const
{
  const unsigned p = pos_(qubit);
  double m0, m1;
  ck(iqsh_maxnorm_half(dev_(), p, 0, &m0));
  ck(iqsh_maxnorm_half(dev_(), p, 1, &m1));
  const bool zero = (BaseType)m0 > tol, one = (BaseType)m1 > tol;
  if (zero && !one) return false;
  if (one && !zero) return true;
  throw std::logic_error("GetClassicalValue: the qubit is not in a classical state");
}

// ---------------------------------------------------------------------------------------
// expectation values

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::expval_(const std::vector<unsigned>& qs,
                                                                    const std::vector<unsigned>& os,
                                                                    BaseType coeff)
{
  if (qs.size() != os.size()) throw std::invalid_argument("ExpectationValue: qubits and observables differ in length");
  if (qs.empty()) return coeff;
  uint64_t x = 0, z = 0, seen = 0;
  unsigned ny = 0;
  for (std::size_t i = 0; i < qs.size(); ++i) {
    const unsigned p = pos_(qs[i]);
    if ((seen >> p) & 1) throw std::invalid_argument("ExpectationValue: repeated qubit");
    seen |= uint64_t(1) << p;
    if (os[i] == 1) x |= uint64_t(1) << p;
    else if (os[i] == 2) {
      x |= uint64_t(1) << p;
      z |= uint64_t(1) << p;
      ++ny;
    } else if (os[i] == 3) z |= uint64_t(1) << p;
    else throw std::invalid_argument("ExpectationValue: observables are 1 (X), 2 (Y), 3 (Z)");
  }
  double v, n2;
  ck(iqsh_expval_pauli(dev_(), x, z, ny, &v, &n2));
  // The reference rotates each X or Y factor into the Z basis and back with two one-qubit
  // gates (ApplyHadamard / Apply1QubitGate(G), qureg_expectval.cpp:146-156, 194-206); its
  // gate counter sees them, so it sees them here too.
  if (gate_counter)
    for (std::size_t i = 0; i < qs.size(); ++i)
      if (os[i] == 1 || os[i] == 2) {
        gate_counter->OneQubitIncrement(qs[i]);
        gate_counter->OneQubitIncrement(qs[i]);
      }
  if (qs.size() == 1) {
    // 1 - 2 P(1) in the rotated basis = 1 - |psi|^2 + <P>  (qureg_expectval.cpp:18-74)
    const BaseType e = (BaseType)(1.0 - n2 + v);
    return coeff * e;
  }
  return coeff * (BaseType)v;
}

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::ExpectationValue(std::vector<unsigned>& qubits,
                                                                             std::vector<unsigned>& observables,
                                                                             BaseType coeff)
{
  return expval_(qubits, observables, coeff);
}

#define IQS_EXPVAL1(NAME, OBS)                                                                      \
  template <class Type>                                                                             \
  typename QubitRegister<Type>::BaseType QubitRegister<Type>::NAME(unsigned const q, BaseType coeff) \
  {                                                                                                 \
    return expval_({q}, {OBS}, coeff);                                                              \
  }
IQS_EXPVAL1(ExpectationValueX, 1)
IQS_EXPVAL1(ExpectationValueY, 2)
IQS_EXPVAL1(ExpectationValueZ, 3)
#undef IQS_EXPVAL1

#define IQS_EXPVAL2(NAME, O1, O2)                                                                              \
  template <class Type>                                                                                        \
  typename QubitRegister<Type>::BaseType QubitRegister<Type>::NAME(unsigned const q, unsigned const q2,         \
                                                                   BaseType coeff)                            \
  {                                                                                                            \
    return expval_({q, q2}, {O1, O2}, coeff);                                                                  \
  }
IQS_EXPVAL2(ExpectationValueXX, 1, 1)
IQS_EXPVAL2(ExpectationValueXY, 1, 2)
IQS_EXPVAL2(ExpectationValueXZ, 1, 3)
IQS_EXPVAL2(ExpectationValueYX, 2, 1)
IQS_EXPVAL2(ExpectationValueYY, 2, 2)
IQS_EXPVAL2(ExpectationValueYZ, 2, 3)
IQS_EXPVAL2(ExpectationValueZX, 3, 1)
IQS_EXPVAL2(ExpectationValueZY, 3, 2)
IQS_EXPVAL2(ExpectationValueZZ, 3, 3)
#undef IQS_EXPVAL2

// ---------------------------------------------------------------------------------------
// utilities

template <class Type>
bool QubitRegister<Type>::operator==(const QubitRegister& rhs)
{
  if (rhs.GlobalSize() != GlobalSize() || rhs.qubit_permutation->map != qubit_permutation->map)
    throw std::invalid_argument("operator==: registers differ in size or qubit order");
  const Type* a = host_();
  const Type* b = rhs.host_();
  host_dirty_ = false;
  rhs.host_dirty_ = false;
  for (std::size_t i = 0; i < global_size_; ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::MaxAbsDiff(QubitRegister& x, Type s)
{
  if (x.qubit_permutation->map != qubit_permutation->map) throw std::invalid_argument("MaxAbsDiff: qubit orders differ");
  double r;
  ck(iqsh_maxabsdiff(dev_(), x.dev_(), (double)s.real(), (double)s.imag(), &r));
  return (BaseType)r;
}

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::MaxL2NormDiff(QubitRegister& x)
{
  if (x.qubit_permutation->map != qubit_permutation->map) throw std::invalid_argument("MaxL2NormDiff: qubit orders differ");
  double r;
  ck(iqsh_l2diff2(dev_(), x.dev_(), &r));
  return (BaseType)r;
}

template <class Type>
typename QubitRegister<Type>::BaseType QubitRegister<Type>::ComputeNorm()
{
  double r;
  ck(iqsh_norm2(dev_(), &r));
  return (BaseType)std::sqrt(r);
}

template <class Type>
void QubitRegister<Type>::Normalize()
{
  const BaseType nrm = ComputeNorm();
  AmplitudeWiseScalarMultiplication(Type(1 / nrm, 0));
}

template <class Type>
void QubitRegister<Type>::InitializationWithSameAmplitudeEverywhere(Type a)
{
  ck(iqsh_init_uniform(dev_(), (double)a.real(), (double)a.imag()));
}

template <class Type>
void QubitRegister<Type>::AmplitudeWiseScalarMultiplication(Type f)
{
  ck(iqsh_scale(dev_(), (double)f.real(), (double)f.imag()));
}

template <class Type>
void QubitRegister<Type>::AmplitudeWiseSum(QubitRegister<Type>& psi, Type f)
{
  ck(iqsh_axpy(dev_(), psi.dev_(), (double)f.real(), (double)f.imag()));
}

template <class Type>
Type QubitRegister<Type>::ComputeOverlap(QubitRegister<Type>& psi)
{
  if (psi.qubit_permutation->map != qubit_permutation->map) throw std::invalid_argument("ComputeOverlap: qubit orders differ");
  double re, im;
  ck(iqsh_overlap(dev_(), psi.dev_(), &re, &im));
  return Type((BaseType)re, (BaseType)im);
}

// State I/O.  The host mirror is synchronized once; amplitudes are listed in data order
// with program-order labels (Permutation::data2program), as src/qureg_utils.cpp:461-563
// (Print) and :566-657 (ExportAmplitudes) format them for a single rank.
template <class Type>
void QubitRegister<Type>::Print(std::string x, std::vector<std::size_t>)
{
  const Type* s = host_();
  host_dirty_ = false;
  BaseType cumulative_probability = 0;
  printf("qubit permutation: %s\n", qubit_permutation->GetMapStr().c_str());
  printf("%s=[\n", x.c_str());
  for (std::size_t i = 0; i < global_size_; ++i) {
    const std::string bin = qubit_permutation->data2program(i);
    printf("\t%-13.8lf + i * %-13.8lf   %% |%s> p=%lf\n", (double)std::real(s[i]), (double)std::imag(s[i]),
           bin.c_str(), (double)std::norm(s[i]));
    cumulative_probability += std::norm(s[i]);
  }
  printf("]; %% cumulative probability = %lf\n", (double)cumulative_probability);
}

template <class Type>
void QubitRegister<Type>::ExportAmplitudes(std::string ofname)
{
  const Type* s = host_();
  host_dirty_ = false;
  std::string str;
  char line[4096];
  for (std::size_t i = 0; i < global_size_; ++i) {
    const std::string bin = qubit_permutation->data2program(i);
    snprintf(line, sizeof line, "\t\t\"%s\" : [%-13.8lf, %-13.8lf, %lf],\n", bin.c_str(), (double)std::real(s[i]),
             (double)std::imag(s[i]), (double)std::norm(s[i]));
    str += line;
  }
  str = str.substr(0, str.size() - 2) + "\n";  // drop the last trailing comma
  std::FILE* f = std::fopen(ofname.c_str(), "a");
  if (!f) throw std::runtime_error("ExportAmplitudes: cannot open " + ofname);
  std::fputs("\t\"amplitudes\" :\n\t{\n", f);
  std::fputs(str.c_str(), f);
  std::fputs("\t}\n", f);
  std::fclose(f);
}

// Raw amplitudes in data order, sizeof(Type) bytes each.  The reference writes the same
// layout through MPI-IO only (src/qureg_utils.cpp:661-689); here one process writes it.
template <class Type>
void QubitRegister<Type>::dumpbin(std::string fn)
{
  const auto t0 = std::chrono::steady_clock::now();
  const Type* s = host_();
  host_dirty_ = false;
  std::FILE* f = std::fopen(fn.c_str(), "wb");
  if (!f) throw std::runtime_error("dumpbin: cannot open " + fn);
  const std::size_t w = std::fwrite(s, sizeof(Type), global_size_, f);
  std::fclose(f);
  if (w != global_size_) throw std::runtime_error("dumpbin: short write to " + fn);
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("Dumping state to %s took %lf sec (%lf MB/s)\n", fn.c_str(), dt,
         (double)(sizeof(Type) * global_size_) / dt / 1e6);
}

// Inverse of dumpbin (no counterpart in the reference): the file must hold exactly 2^n
// amplitudes of this precision, in data order for the register's current permutation.
template <class Type>
void QubitRegister<Type>::loadbin(std::string fn)
{
  std::FILE* f = std::fopen(fn.c_str(), "rb");
  if (!f) throw std::runtime_error("loadbin: cannot open " + fn);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (bytes < 0 || (std::size_t)bytes != sizeof(Type) * global_size_) {
    std::fclose(f);
    throw std::runtime_error("loadbin: " + fn + " does not hold " + std::to_string(global_size_) +
                             " amplitudes of this precision");
  }
  std::vector<Type> buf(global_size_);
  const std::size_t r = std::fread(buf.data(), sizeof(Type), global_size_, f);
  std::fclose(f);
  if (r != global_size_) throw std::runtime_error("loadbin: short read from " + fn);
  ck(iqsh_upload(dev_(), buf.data(), 0, global_size_));
  host_dirty_ = false;
  host_valid_ = false;
}

// Noise (src/qureg_noisysimul.cpp:18-80): a stochastic Pauli-twirl rotation per call,
// U = exp(-i vX X) exp(-i vY Y) exp(-i vZ Z) with Gaussian angles drawn from the state
// stream of the register's generator, built and applied exactly as the reference does.
template <class Type>
void QubitRegister<Type>::SetNoiseTimescales(BaseType T1, BaseType T2)
{
  if (!(T2 >= T1 / 2.)) throw std::invalid_argument("SetNoiseTimescales: T2 must be >= T1/2");
  T_1_ = T1;
  T_2_ = T2;
  T_phi_ = 1. / (1. / T2 - 1. / (2. * T1));
}

template <class Type>
TM2x2<Type> NoiseGateMatrix(typename extract_value_type<Type>::value_type T_1_,
                            typename extract_value_type<Type>::value_type T_2_,
                            typename extract_value_type<Type>::value_type duration,
                            RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng)
{
  using BaseType = typename extract_value_type<Type>::value_type;
  BaseType p_X, p_Y, p_Z;
  p_X = (1. - std::exp(-duration / T_1_)) / 4.;
  p_Y = (1. - std::exp(-duration / T_1_)) / 4.;
  p_Z = (1. - std::exp(-duration / T_2_)) / 2. + (1. - std::exp(-duration / T_1_)) / 4.;
  if (!(p_X > 0 && p_Y > 0 && p_Z > 0)) throw std::invalid_argument("ApplyNoiseGate: non-positive probabilities");
  BaseType s_X, s_Y, s_Z;
  s_X = std::sqrt(-std::log(1. - p_X));
  s_Y = std::sqrt(-std::log(1. - p_Y));
  s_Z = std::sqrt(-std::log(1. - p_Z));
  BaseType v_X, v_Y, v_Z;
  rng.GaussianRandomNumbers(&v_X, 1, "state");
  v_X *= s_X / 2.;
  rng.GaussianRandomNumbers(&v_Y, 1, "state");
  v_Y *= s_Y / 2.;
  rng.GaussianRandomNumbers(&v_Z, 1, "state");
  v_Z *= s_Z / 2.;
  Type A, B, C;
  A = {std::cos(v_Z), -std::sin(v_Z)};
  B = {std::cos(v_X) * std::cos(v_Y), -std::sin(v_X) * std::sin(v_Y)};
  C = {std::cos(v_X) * std::sin(v_Y), -std::sin(v_X) * std::cos(v_Y)};
  TM2x2<Type> U;
  U(0, 0) = A * B;
  U(0, 1) = -std::conj(A) * std::conj(C);
  U(1, 0) = A * C;
  U(1, 1) = std::conj(A) * std::conj(B);
  return U;
}

template <class Type>
void QubitRegister<Type>::ApplyNoiseGate(const unsigned qubit, const BaseType duration)
{
  if (!rng_ptr_) throw std::logic_error("ApplyNoiseGate: no random number generator (SetRngPtr)");
  if (duration == 0) return;
  Apply1QubitGate(qubit, NoiseGateMatrix<Type>(T_1_, T_2_, duration, *rng_ptr_));
}

// Branch index for a uniform draw r in [0, 1): the first k with r <= cumulative p_k
// (qureg_apply_channel.cpp:23-31).  The cumulative distribution ends at exactly 1 unless
// chi is zero, which is reported.
template <class Chi>
static unsigned channel_branch(const Chi& chi, double r, unsigned M)
{
  if (!chi.HasEigenSystem()) throw std::logic_error("ApplyChannel: call SolveEigenSystem() on the chi-matrix first");
  unsigned k = 0;
  while (r > chi.GetEigenCumulativeProbability(k))
    if (++k >= M) throw std::runtime_error("ApplyChannel: the chi-matrix has no non-zero eigenvalue");
  return k;
}

/// One stochastic step of a 1-qubit channel (qureg_apply_channel.cpp:17-58): draw a branch
/// k with probability |E_k| / sum |E_k| from the register's state stream, apply
/// E_k = e0 id + e1 X + e2 Y + e3 Z as a 1-qubit gate, and flip the overall sign of
/// channels for a negative eigenvalue.
template <class Type>
TM2x2<Type> ChannelBranchMatrix(CM4x4<Type>& chi, RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng,
                                bool* negative)
{
  typename extract_value_type<Type>::value_type r;
  rng.UniformRandomNumbers(&r, 1, 0, 1, "state");
  const unsigned k = channel_branch(chi, r, 4);
  const std::vector<Type> e = chi.GetEigenVector(k);
  const Type I(0., 1.);
  TM2x2<Type> m;
  m(0, 0) = e[0] + e[3];
  m(0, 1) = e[1] - I * e[2];
  m(1, 0) = e[1] + I * e[2];
  m(1, 1) = e[0] - e[3];
  *negative = std::real(chi.GetEigenValue(k)) < 0;
  return m;
}

template <class Type>
void QubitRegister<Type>::ApplyChannel(const unsigned qubit, CM4x4<Type>& chi)
{
  if (!rng_ptr_) throw std::logic_error("ApplyChannel: no random number generator (SetRngPtr)");
  bool negative = false;
  const TM2x2<Type> m = ChannelBranchMatrix(chi, *rng_ptr_, &negative);
  if (negative) overall_sign_of_channels *= -1;
  Apply1QubitGate(qubit, m);
}

/// Branch of a 2-qubit channel (qureg_apply_channel.cpp:66-106): draw k from the state
/// stream, then E_k = sum_{a,b} e_{4a+b} sigma_a (qubit1) x sigma_b (qubit2) as a 4x4 in the
/// index 2 b(qubit1) + b(qubit2).  The diagonal and upper triangle use the reference's
/// expressions; the lower triangle is evaluated from the same sum (the reference writes it
/// as the conjugate of the upper one, which equals the sum only for real coefficients) and,
/// as in the 1-qubit case, a negative eigenvalue is reported so the caller can flip the
/// overall sign of channels.
template <class Type>
TM4x4<Type> ChannelBranchMatrix(CM16x16<Type>& chi, RandomNumberGenerator<typename extract_value_type<Type>::value_type>& rng,
                                bool* negative)
{
  typename extract_value_type<Type>::value_type r;
  rng.UniformRandomNumbers(&r, 1, 0, 1, "state");
  const unsigned k = channel_branch(chi, r, 16);
  const std::vector<Type> e = chi.GetEigenVector(k);
  const Type I(0., 1.);
  TM4x4<Type> m;
  m(0, 0) = e[0] + e[3] + e[12] + e[15];
  m(1, 1) = e[0] - e[3] + e[12] - e[15];
  m(2, 2) = e[0] + e[3] - e[12] - e[15];
  m(3, 3) = e[0] - e[3] - e[12] + e[15];

  m(0, 1) = e[1] + e[13] - I * (e[2] + e[14]);
  m(1, 0) = e[1] + e[13] + I * (e[2] + e[14]);
  m(2, 3) = e[1] - e[13] - I * (e[2] - e[14]);
  m(3, 2) = e[1] - e[13] + I * (e[2] - e[14]);

  m(0, 2) = e[4] + e[7] - I * (e[8] + e[11]);
  m(2, 0) = e[4] + e[7] + I * (e[8] + e[11]);
  m(1, 3) = e[4] - e[7] - I * (e[8] - e[11]);
  m(3, 1) = e[4] - e[7] + I * (e[8] - e[11]);

  m(0, 3) = e[5] - e[10] - I * (e[9] + e[6]);
  m(3, 0) = e[5] - e[10] + I * (e[9] + e[6]);
  m(1, 2) = e[5] + e[10] - I * (e[9] - e[6]);
  m(2, 1) = e[5] + e[10] + I * (e[9] - e[6]);
  *negative = std::real(chi.GetEigenValue(k)) < 0;
  return m;
}

/// One stochastic step of a 2-qubit channel, applied with Apply2QubitGate(qubit1, qubit2, .).
template <class Type>
void QubitRegister<Type>::ApplyChannel(const unsigned qubit1, const unsigned qubit2, CM16x16<Type>& chi)
{
  if (!rng_ptr_) throw std::logic_error("ApplyChannel: no random number generator (SetRngPtr)");
  bool negative = false;
  const TM4x4<Type> m = ChannelBranchMatrix(chi, *rng_ptr_, &negative);
  if (negative) overall_sign_of_channels *= -1;
  Apply2QubitGate(qubit1, qubit2, m);
}

/// -sum p log2 p over the basis probabilities (src/qureg_utils.cpp:306-345).
template <class Type>
double QubitRegister<Type>::Entropy()
{
  double m[11];
  ck(iqsh_prob_moments(dev_(), m));
  return m[0] / (double)std::log(2.0);
}

/// [entropy (bits), average self-information (bits) / 2^n, then for k = 2..10 the moments
/// sum p^k scaled by 2^(n (k - 1)) / k!] (src/qureg_utils.cpp:352-450).
template <class Type>
std::vector<double> QubitRegister<Type>::GoogleStats()
{
  double m[11];
  ck(iqsh_prob_moments(dev_(), m));
  const double two2n = double(GlobalSize());
  std::vector<double> stats;
  stats.push_back(m[0] / (double)std::log(2.0));
  stats.push_back(m[1] / (double)std::log(2.0) / two2n);
  double factorial = 1.0;
  for (int k = 2; k <= 10; ++k) {
    factorial *= double(k);
    stats.push_back(m[k] * (std::pow(two2n, double(k - 1)) / factorial));
  }
  return stats;
}

template class QubitRegister<ComplexSP>;
template class QubitRegister<ComplexDP>;
template TM2x2<ComplexSP> NoiseGateMatrix<ComplexSP>(float, float, float, RandomNumberGenerator<float>&);
template TM2x2<ComplexDP> NoiseGateMatrix<ComplexDP>(double, double, double, RandomNumberGenerator<double>&);
template TM2x2<ComplexSP> ChannelBranchMatrix<ComplexSP>(CM4x4<ComplexSP>&, RandomNumberGenerator<float>&, bool*);
template TM2x2<ComplexDP> ChannelBranchMatrix<ComplexDP>(CM4x4<ComplexDP>&, RandomNumberGenerator<double>&, bool*);
template TM4x4<ComplexSP> ChannelBranchMatrix<ComplexSP>(CM16x16<ComplexSP>&, RandomNumberGenerator<float>&, bool*);
template TM4x4<ComplexDP> ChannelBranchMatrix<ComplexDP>(CM16x16<ComplexDP>&, RandomNumberGenerator<double>&, bool*);

}  // namespace iqs
