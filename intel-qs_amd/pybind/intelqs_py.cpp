// intelqs_py: drop-in for the reference's pybind11 module of the same name
// (pybind11/intelqs_py.cpp), over the GPU-resident iqs::QubitRegister.
//
// Same names and argument conventions as the reference module: QubitRegister (ComplexDP),
// RandomNumberGenerator, the QAOA functions, EnvInit/EnvFinalize and MPIEnvironment
// statics.  Additions: QubitRegisterSP (ComplexSP), Sync(), DeviceHandle(), fusion block size.
// The numpy buffer of a register is its host mirror (refreshed from HBM when the buffer is
// requested); take the view after the gates whose result it should show.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/iostream.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "qaoa_features.hpp"
#include "qureg.hpp"
#include "state_pool.hpp"

namespace py = pybind11;
using iqs::QubitRegister;
using Environment = iqs::mpi::Environment;

namespace {

template <class T>
TM2x2<T> tm2(py::array_t<T, py::array::c_style | py::array::forcecast> a, const char* what)
{
  auto b = a.request();
  if (b.ndim != 2) throw std::runtime_error("Number of dimensions must be two.");
  if (b.shape[0] != 2 || b.shape[1] != 2) throw std::runtime_error(what);
  const T* p = static_cast<const T*>(b.ptr);
  TM2x2<T> m;
  m(0, 0) = p[0];
  m(0, 1) = p[1];
  m(1, 0) = p[2];
  m(1, 1) = p[3];
  return m;
}

template <class T>
TM4x4<T> tm4(py::array_t<T, py::array::c_style | py::array::forcecast> a)
{
  auto b = a.request();
  if (b.ndim != 2 || b.shape[0] != 4 || b.shape[1] != 4) throw std::runtime_error("Input shape is not 4x4.");
  const T* p = static_cast<const T*>(b.ptr);
  TM4x4<T> m;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m(i, j) = p[4 * i + j];
  return m;
}

// ChiMatrix<ComplexDP, M> as CM4x4 / CM16x16 (pybind11/intelqs_py.cpp:94-163 of the
// reference): element access by (row, column) pairs, SolveEigenSystem, Print.  Additions: a
// constructor from an M x M array and read access to the eigensystem.
template <unsigned M>
void bind_chi(py::module_& m, const char* name, const char* repr)
{
  using X = iqs::ChiMatrix<ComplexDP, M, 32>;
  py::class_<X>(m, name)
      .def(py::init<>())
      .def(py::init([](py::array_t<ComplexDP, py::array::c_style | py::array::forcecast> a) {
        auto b = a.request();
        if (b.ndim != 2 || b.shape[0] != M || b.shape[1] != M)
          throw std::runtime_error("The shape of the chi-matrix is not " + std::to_string(M) + "x" + std::to_string(M) + ".");
        const ComplexDP* p = static_cast<const ComplexDP*>(b.ptr);
        X x;
        for (unsigned i = 0; i < M; ++i)
          for (unsigned j = 0; j < M; ++j) x(i, j) = p[M * i + j];
        return x;
      }))
      .def("__getitem__",
           [](const X& a, std::pair<py::ssize_t, py::ssize_t> i) {
             if (i.first < 0 || i.first >= M || i.second < 0 || i.second >= M) throw py::index_error();
             return a(i.first, i.second);
           },
           py::is_operator())
      .def("__setitem__",
           [](X& a, std::pair<py::ssize_t, py::ssize_t> i, ComplexDP value) {
             if (i.first < 0 || i.first >= M || i.second < 0 || i.second >= M) throw py::index_error();
             a(i.first, i.second) = value;
           },
           py::is_operator())
      .def("SolveEigenSystem", &X::SolveEigenSystem)
      .def("EigensystemOfIdealHadamardChannel", &X::EigensystemOfIdealHadamardChannel)
      .def("Print",
           [](X& a, bool with_eigensystem) {
             py::scoped_ostream_redirect stream(std::cout, py::module_::import("sys").attr("stdout"));
             a.Print(with_eigensystem);
           },
           py::arg("with_eigensystem") = true)
      .def("GetEigenValues", &X::GetEigenValues)
      .def("GetEigenProbabilities", &X::GetEigenProbabilities)
      .def("GetEigenCumulativeProbabilities", &X::GetEigenCumulativeProbabilities)
      .def("GetEigenVectors", &X::GetEigenVectors)
      .def("__repr__", [repr](const X&) { return std::string(repr); });
}

template <class B>
void bind_rng(py::module_& m, const char* name)
{
  using R = iqs::RandomNumberGenerator<B>;
  py::class_<R>(m, name)
      .def(py::init<>())
      .def("GetSeed", &R::GetSeed)
      .def("SetSeedStreamPtrs", py::overload_cast<std::size_t>(&R::SetSeedStreamPtrs))
      .def("SkipeAhead", &R::SkipAhead)
      // extension: the stream counters of the C++ class (include/rng_utils.hpp:82-89)
      .def("GetNumGeneratedOrSkippedLocalNumbers", &R::GetNumGeneratedOrSkippedLocalNumbers)
      .def("GetNumGeneratedOrSkippedStateNumbers", &R::GetNumGeneratedOrSkippedStateNumbers)
      .def("GetNumGeneratedOrSkippedPoolNumbers", &R::GetNumGeneratedOrSkippedPoolNumbers)
      .def("GetUniformRandomNumbers",
           [](R& r, std::size_t size, B a, B b, std::string shared) {
             std::vector<B> v(size);
             r.UniformRandomNumbers(v.data(), size, a, b, shared);
             return v;
           })
      // The reference binds these with raw pointers; here they fill a writable numpy array.
      .def("UniformRandomNumbers",
           [](R& r, py::array_t<B, py::array::c_style> out, std::size_t size, B a, B b, std::string shared) {
             if ((std::size_t)out.size() < size) throw std::runtime_error("output array too small");
             r.UniformRandomNumbers(out.mutable_data(), size, a, b, shared);
           },
           py::arg("out"), py::arg("size") = 1, py::arg("a") = B(0), py::arg("b") = B(1), py::arg("shared") = "local")
      .def("GaussianRandomNumbers",
           [](R& r, py::array_t<B, py::array::c_style> out, std::size_t size, std::string shared) {
             if ((std::size_t)out.size() < size) throw std::runtime_error("output array too small");
             r.GaussianRandomNumbers(out.mutable_data(), size, shared);
           },
           py::arg("out"), py::arg("size") = 1, py::arg("shared") = "local")
      .def("RandomIntegersInRange",
           [](R& r, py::array_t<int, py::array::c_style> out, std::size_t size, int a, int b, std::string shared) {
             if ((std::size_t)out.size() < size) throw std::runtime_error("output array too small");
             r.RandomIntegersInRange(out.mutable_data(), size, a, b, shared);
           },
           py::arg("out"), py::arg("size") = 1, py::arg("a") = 0, py::arg("b") = 2, py::arg("shared") = "local")
      .def("__repr__", [name](const R&) { return std::string("<") + name + ">"; });
}

template <class T>
void bind_register(py::module_& m, const char* name)
{
  using Q = QubitRegister<T>;
  py::class_<Q>(m, name, py::buffer_protocol(), py::dynamic_attr())
      .def(py::init<>())
      .def(py::init<const Q&>())
      .def(py::init<std::size_t, std::string, std::size_t, std::size_t>(), py::arg("num_qubits"),
           py::arg("style") = "", py::arg("base_index") = 0, py::arg("tmp_spacesize") = 0)
      .def("NumQubits", &Q::NumQubits)
      .def("GlobalSize", &Q::GlobalSize)
      .def("LocalSize", &Q::LocalSize)
      .def("__getitem__",
           [](const Q& a, std::size_t i) {
             if (i >= a.LocalSize()) throw py::index_error();
             return a.GetGlobalAmplitude(a.qubit_permutation->data2program_(i));
           },
           py::is_operator())
      .def("__setitem__",
           [](Q& a, std::size_t i, T v) {
             if (i >= a.LocalSize()) throw py::index_error();
             a.SetGlobalAmplitude(a.qubit_permutation->data2program_(i), v);
           },
           py::is_operator())
      .def_buffer([](Q& r) -> py::buffer_info {
        return py::buffer_info(r.RawState(), sizeof(T), py::format_descriptor<T>::format(), 1, {r.LocalSize()},
                               {sizeof(T)});
      })
      // one-qubit gates
      .def("ApplyRotationX", &Q::ApplyRotationX)
      .def("ApplyRotationY", &Q::ApplyRotationY)
      .def("ApplyRotationZ", &Q::ApplyRotationZ)
      .def("ApplyPauliX", &Q::ApplyPauliX)
      .def("ApplyPauliY", &Q::ApplyPauliY)
      .def("ApplyPauliZ", &Q::ApplyPauliZ)
      .def("ApplyPauliSqrtX", &Q::ApplyPauliSqrtX)
      .def("ApplyPauliSqrtY", &Q::ApplyPauliSqrtY)
      .def("ApplyPauliSqrtZ", &Q::ApplyPauliSqrtZ)
      .def("ApplyT", &Q::ApplyT)
      .def("ApplyRotationXY", &Q::ApplyRotationXY)
      .def("ApplyHadamard", &Q::ApplyHadamard)
      // two-qubit gates
      .def("ApplySwap", &Q::ApplySwap)
      .def("ApplyISwap", &Q::ApplyISwap)
      .def("ApplySqrtISwap", &Q::ApplySqrtISwap)
      .def("Apply4thRootISwap", &Q::Apply4thRootISwap)
      .def("ApplyISwapRotation",
           [](Q& a, unsigned q1, unsigned q2, py::array_t<T, py::array::c_style | py::array::forcecast> mm) {
             a.ApplyISwapRotation(q1, q2, tm2<T>(mm, "Input shape is not 2x2."));
           })
      .def("ApplyCRotationX", &Q::ApplyCRotationX)
      .def("ApplyCRotationY", &Q::ApplyCRotationY)
      .def("ApplyCRotationZ", &Q::ApplyCRotationZ)
      .def("ApplyCPauliX", &Q::ApplyCPauliX)
      .def("ApplyCPauliY", &Q::ApplyCPauliY)
      .def("ApplyCPauliZ", &Q::ApplyCPauliZ)
      .def("ApplyCPauliSqrtZ", &Q::ApplyCPauliSqrtZ)
      .def("ApplyCHadamard", &Q::ApplyCHadamard)
      .def("ApplyCPhaseRotation", &Q::ApplyCPhaseRotation)
      .def("Apply1QubitGate",
           [](Q& a, unsigned q, py::array_t<T, py::array::c_style | py::array::forcecast> mm) {
             a.Apply1QubitGate(q, tm2<T>(mm, "Input shape is not 2x2."));
           },
           "Apply custom 1-qubit gate.")
      .def("ApplyControlled1QubitGate",
           [](Q& a, unsigned c, unsigned q, py::array_t<T, py::array::c_style | py::array::forcecast> mm) {
             a.ApplyControlled1QubitGate(c, q, tm2<T>(mm, "The shape of the unitary-matrix is not 2x2."));
           },
           "Apply custom controlled-1-qubit gate.")
      // Extension: the C++ sub-range helpers (public in the reference's qureg.hpp, not bound
      // by its module); spec is the GateSpec1Q / GateSpec2Q number (8 = None).
      .def("Apply1QubitGate_helper",
           [](Q& a, unsigned q, py::array_t<T, py::array::c_style | py::array::forcecast> mm, std::size_t sind,
              std::size_t eind, int spec, double angle) {
             return a.Apply1QubitGate_helper(q, tm2<T>(mm, "Input shape is not 2x2."), sind, eind,
                                             (iqs::GateSpec1Q)spec, (typename Q::BaseType)angle);
           },
           py::arg("qubit"), py::arg("m"), py::arg("sind"), py::arg("eind"), py::arg("spec") = 8,
           py::arg("angle") = 0.0)
      .def("ApplyControlled1QubitGate_helper",
           [](Q& a, unsigned c, unsigned q, py::array_t<T, py::array::c_style | py::array::forcecast> mm,
              std::size_t sind, std::size_t eind, int spec, double angle) {
             return a.ApplyControlled1QubitGate_helper(c, q, tm2<T>(mm, "Input shape is not 2x2."), sind, eind,
                                                       (iqs::GateSpec2Q)spec, (typename Q::BaseType)angle);
           },
           py::arg("control"), py::arg("qubit"), py::arg("m"), py::arg("sind"), py::arg("eind"),
           py::arg("spec") = 8, py::arg("angle") = 0.0)
      .def("ApplyDiag",
           [](Q& a, unsigned q1, unsigned q2, py::array_t<T, py::array::c_style | py::array::forcecast> mm) {
             a.ApplyDiag(q1, q2, tm4<T>(mm));
           })
      .def("Apply2QubitGate",
           [](Q& a, unsigned qh, unsigned ql, py::array_t<T, py::array::c_style | py::array::forcecast> mm) {
             a.Apply2QubitGate(qh, ql, tm4<T>(mm));
           })
      .def("ApplyToffoli", &Q::ApplyToffoli)
      // permutation
      .def("PermuteQubits", &Q::PermuteQubits, py::arg("new_map"), py::arg("style_of_map") = "direct")
      .def("EmulateSwap", &Q::EmulateSwap)
      .def("GetQubitPermutation", [](const Q& a) { return a.qubit_permutation->map; })
      // state initialization
      .def("Initialize", (void (Q::*)(std::string, std::size_t)) & Q::Initialize)
      // specialization and fusion
      .def("TurnOnSpecialize", &Q::TurnOnSpecialize)
      .def("TurnOffSpecialize", &Q::TurnOffSpecialize)
      .def("TurnOnSpecializeV2", &Q::TurnOnSpecializeV2)
      .def("TurnOffSpecializeV2", &Q::TurnOffSpecializeV2)
      .def("TurnOnFusion", &Q::TurnOnFusion, py::arg("log2llc") = 20)
      .def("TurnOffFusion", &Q::TurnOffFusion)
      .def("IsFusionEnabled", &Q::IsFusionEnabled)
      .def("ApplyFusedGates", &Q::ApplyFusedGates)
      .def("SetFusionBlockQubits", &Q::SetFusionBlockQubits)
      .def("FusionStats", &Q::FusionStats)
      // random numbers
      .def("ResetRngPtr", &Q::ResetRngPtr)
      .def("SetRngPtr", &Q::SetRngPtr)
      .def("GetT1", &Q::GetT1)
      .def("GetT2", &Q::GetT2)
      .def("GetTphi", &Q::GetTphi)
      .def("SetNoiseTimescales", &Q::SetNoiseTimescales)
      .def("ApplyNoiseGate", &Q::ApplyNoiseGate)
      .def("GetOverallSignOfChannels", &Q::GetOverallSignOfChannels)
      .def("ApplyChannel",
           [](Q& a, unsigned qubit, const iqs::ChiMatrix<ComplexDP, 4, 32>& chi) {
             CM4x4<T> c(chi);
             a.ApplyChannel(qubit, c);
           },
           "Apply 1-qubit channel provided via its chi-matrix.")
      .def("ApplyChannel",
           [](Q& a, unsigned qubit1, unsigned qubit2, const iqs::ChiMatrix<ComplexDP, 16, 32>& chi) {
             CM16x16<T> c(chi);
             a.ApplyChannel(qubit1, qubit2, c);
           },
           "Apply 2-qubit channel provided via its chi-matrix.")
      .def("SetSeedRngPtr", &Q::SetSeedRngPtr)
      // measurement, norms, expectation values
      .def("GetProbability", &Q::GetProbability)
      .def("CollapseQubit", &Q::CollapseQubit)
      .def("IsClassicalBit", &Q::IsClassicalBit, py::arg("qubit"), py::arg("tolerance") = 1.e-13)
      .def("GetClassicalValue", &Q::GetClassicalValue, py::arg("qubit"), py::arg("tolerance") = 1.e-13)
      .def("Normalize", &Q::Normalize)
      .def("AmplitudeWiseScalarMultiplication", &Q::AmplitudeWiseScalarMultiplication)
      .def("AmplitudeWiseSum", &Q::AmplitudeWiseSum, py::arg("psi"), py::arg("factor") = T(1, 0))
      .def("ExpectationValue", &Q::ExpectationValue, py::arg("qubits"), py::arg("observables"), py::arg("coeff") = 1.)
      .def("ExpectationValueX", &Q::ExpectationValueX, py::arg("qubit"), py::arg("coeff") = 1.)
      .def("ExpectationValueY", &Q::ExpectationValueY, py::arg("qubit"), py::arg("coeff") = 1.)
      .def("ExpectationValueZ", &Q::ExpectationValueZ, py::arg("qubit"), py::arg("coeff") = 1.)
      .def("ComputeNorm", &Q::ComputeNorm)
      .def("ComputeOverlap", &Q::ComputeOverlap)
      .def("MaxAbsDiff", &Q::MaxAbsDiff, py::arg("x"), py::arg("sfactor") = T(1, 0))
      .def("MaxL2NormDiff", &Q::MaxL2NormDiff)
      .def("GetGlobalAmplitude", &Q::GetGlobalAmplitude)
      .def("SetGlobalAmplitude", &Q::SetGlobalAmplitude)
      // statistics
      .def("EnableStatistics", &Q::EnableStatistics)
      .def("GetStatistics", &Q::GetStatistics)
      .def("DisableStatistics", &Q::DisableStatistics)
      .def("ResetStatistics", &Q::ResetStatistics)
      .def("GetGateCounts",
           [](const Q& a) {
             if (!a.gate_counter) throw std::runtime_error("statistics are disabled");
             return py::make_tuple(a.gate_counter->GetOneQubitGateCount(), a.gate_counter->GetTwoQubitGateCount(),
                                   a.gate_counter->GetTotalGateCount(), a.gate_counter->GetParallelDepth());
           })
      // device interface
      .def("Sync", &Q::Sync)
      .def("DeviceHandle", [](Q& a) { return reinterpret_cast<std::uintptr_t>(a.DeviceHandle()); })
      .def("Print",
           [](Q& a, std::string description) {
             py::scoped_ostream_redirect stream(std::cout, py::module_::import("sys").attr("stdout"));
             std::cout << "<<the output has been redirected to the terminal>>\n";
             a.Print(description, {});
           },
           "Print the quantum state with an initial description.")
      .def("Entropy", &Q::Entropy)
      .def("GoogleStats", &Q::GoogleStats)
      .def("PermuteGlobalQubits", &Q::PermuteGlobalQubits, py::arg("new_map"), py::arg("style_of_map") = "direct")
      .def("ExportAmplitudes", &Q::ExportAmplitudes)
      .def("dumpbin", &Q::dumpbin)
      .def("loadbin", &Q::loadbin)
      .def("__repr__", [name](const Q& a) {
        return std::string("<") + name + " of " + std::to_string(a.NumQubits()) + " qubits in GPU memory>";
      });
}

}  // namespace

PYBIND11_MODULE(intelqs_py, m)
{
  m.doc() = "Intel-QS drop-in: state-vector simulator with the register in MI355X HBM";

  m.def("EnvInit", []() { Environment::Init(); });
  m.def("EnvFinalize", []() { Environment::Finalize(); });
  m.def("EnvFinalizeDummyRanks", []() {});

  bind_rng<double>(m, "RandomNumberGenerator");
  // extension: the C++ utility ShuffleFisherYates (src/rng_utils.cpp:403-423) on a list of ints
  m.def(
      "ShuffleFisherYates",
      [](std::vector<int> a, iqs::RandomNumberGenerator<double>& rng, std::string shared) {
        iqs::ShuffleFisherYates(a, &rng, shared);
        return a;
      },
      py::arg("array"), py::arg("rng"), py::arg("shared") = "local");
  bind_rng<float>(m, "RandomNumberGeneratorSP");  // extension: the generator QubitRegisterSP takes

  bind_chi<4>(m, "CM4x4", "<ChiMatrix for 1-qubit channel>");
  bind_chi<16>(m, "CM16x16", "<ChiMatrix for 2-qubit channel>");

  bind_register<ComplexDP>(m, "QubitRegister");
  bind_register<ComplexSP>(m, "QubitRegisterSP");

  // Extension: a pool of states in one GPU register (include/state_pool.hpp).
  using SP = iqs::StatePool<ComplexDP>;
  py::class_<SP>(m, "StatePool")
      .def(py::init<unsigned, unsigned, std::string, std::size_t>(), py::arg("num_qubits"), py::arg("num_states"),
           py::arg("style") = "base", py::arg("base_index") = 0)
      .def("NumQubits", &SP::NumQubits)
      .def("NumStates", &SP::NumStates)
      .def("Register", &SP::Register, py::return_value_policy::reference_internal)
      .def("Initialize", &SP::Initialize)
      .def("SetSeedStreamPtrs", &SP::SetSeedStreamPtrs)
      .def("SetNoiseTimescales", &SP::SetNoiseTimescales)
      .def("ApplyNoiseGate", &SP::ApplyNoiseGate)
      .def("ApplyChannel",
           [](SP& a, unsigned qubit, const iqs::ChiMatrix<ComplexDP, 4, 32>& chi) {
             CM4x4<ComplexDP> c(chi);
             a.ApplyChannel(qubit, c);
           })
      .def("ApplyChannel",
           [](SP& a, unsigned qubit1, unsigned qubit2, const iqs::ChiMatrix<ComplexDP, 16, 32>& chi) {
             CM16x16<ComplexDP> c(chi);
             a.ApplyChannel(qubit1, qubit2, c);
           })
      .def("GetOverallSignOfChannels", &SP::GetOverallSignOfChannels)
      .def("GetProbability", &SP::GetProbability)
      .def("ComputeNorm", &SP::ComputeNorm)
      .def("ExpectationValueZString", &SP::ExpectationValueZString)
      .def("ComputeOverlap", &SP::ComputeOverlap)
      .def("GetState",
           [](SP& a, unsigned s) {
             std::vector<ComplexDP> v = a.GetState(s);
             return py::array_t<ComplexDP>(v.size(), v.data());
           })
      .def_static("IncoherentSumOverAllStatesOfPool", &SP::IncoherentSumOverAllStatesOfPool)
      .def("__repr__", [](const SP& a) {
        return "<StatePool of " + std::to_string(a.NumStates()) + " states of " + std::to_string(a.NumQubits()) +
               " qubits in GPU memory>";
      });

  // Extension: the C++ QubitRegisterMetric (include/QubitRegisterMetric.hpp).
  using QM = iqs::QubitRegisterMetric<ComplexDP>;
  py::class_<QM, QubitRegister<ComplexDP>>(m, "QubitRegisterMetric", py::buffer_protocol(), py::dynamic_attr())
      .def(py::init<int>())
      .def("GetTotalQubitGateCount", &QM::GetTotalQubitGateCount)
      .def("GetOneQubitGateCount", &QM::GetOneQubitGateCount)
      .def("GetTwoQubitGateCount", &QM::GetTwoQubitGateCount)
      .def("GetParallelDepth", &QM::GetParallelDepth)
      .def("ApplyHadamard", &QM::ApplyHadamard)
      .def("ApplyRotationX", &QM::ApplyRotationX)
      .def("ApplyRotationY", &QM::ApplyRotationY)
      .def("ApplyRotationZ", &QM::ApplyRotationZ)
      .def("ApplyCPauliX", &QM::ApplyCPauliX)
      .def("ApplyControlled1QubitGate",
           [](QM& a, int q1, int q2, py::array_t<ComplexDP, py::array::c_style | py::array::forcecast> m) {
             a.ApplyControlled1QubitGate(q1, q2, tm2<ComplexDP>(m, "Input shape is not 2x2."));
           });

  // Extension: the C++ NoisyQureg (include/NoisyQureg.hpp) for Python callers.
  using NQ = iqs::NoisyQureg<ComplexDP>;
  py::class_<NQ, QubitRegister<ComplexDP>>(m, "NoisyQureg", py::buffer_protocol(), py::dynamic_attr())
      .def(py::init<unsigned, unsigned, double, double>(), py::arg("num_qubits"), py::arg("RNG_seed") = 12345,
           py::arg("T1") = 2000., py::arg("T2") = 1000.)
      .def("Initialize", &NQ::Initialize)
      .def("ResetTimeForAllQubits", &NQ::ResetTimeForAllQubits)
      .def("ApplyNoiseGatesOnAllQubits", &NQ::ApplyNoiseGatesOnAllQubits)
      .def("SetDecoherenceTime", &NQ::SetDecoherenceTime)
      .def("SetGateDurations", &NQ::SetGateDurations)
      .def("GetTotalExperimentalGateCount", &NQ::GetTotalExperimentalGateCount)
      .def("GetOneQubitExperimentalGateCount", &NQ::GetOneQubitExperimentalGateCount)
      .def("GetTwoQubitExperimentalGateCount", &NQ::GetTwoQubitExperimentalGateCount)
      .def("GetExperimentalGateCount", py::overload_cast<unsigned>(&NQ::GetExperimentalGateCount))
      .def("GetExperimentalGateCount", py::overload_cast<unsigned, unsigned>(&NQ::GetExperimentalGateCount))
      .def("AddNoiseOneQubitGate", &NQ::AddNoiseOneQubitGate)
      .def("AddNoiseTwoQubitGate", &NQ::AddNoiseTwoQubitGate)
      .def("NoiseGate", &NQ::NoiseGate)
      .def("Apply1QubitGate",
           [](NQ& a, unsigned q, py::array_t<ComplexDP, py::array::c_style | py::array::forcecast> m) {
             a.Apply1QubitGate(q, tm2<ComplexDP>(m, "Input shape is not 2x2."));
           })
      .def("ApplyHadamard", &NQ::ApplyHadamard)
      .def("ApplyRotationX", &NQ::ApplyRotationX)
      .def("ApplyRotationY", &NQ::ApplyRotationY)
      .def("ApplyRotationZ", &NQ::ApplyRotationZ)
      .def("ApplyCPauliX", &NQ::ApplyCPauliX)
      .def("ApplyControlled1QubitGate",
           [](NQ& a, unsigned q1, unsigned q2, py::array_t<ComplexDP, py::array::c_style | py::array::forcecast> m) {
             a.ApplyControlled1QubitGate(q1, q2, tm2<ComplexDP>(m, "Input shape is not 2x2."));
           })
      .def("__repr__", [](const NQ& a) {
        return "<NoisyQureg of " + std::to_string(a.NumQubits()) + " qubits in GPU memory>";
      });

  m.def("InitializeVectorAsMaxCutCostFunction", &iqs::qaoa::InitializeVectorAsMaxCutCostFunction<ComplexDP>,
        "Use IQS vector to store a large real vector and not as a quantum state.");
  m.def("InitializeVectorAsWeightedMaxCutCostFunction",
        &iqs::qaoa::InitializeVectorAsWeightedMaxCutCostFunction<ComplexDP>,
        "Use IQS vector to store a large real vector and not as a quantum state.");
  m.def("ImplementQaoaLayerBasedOnCostFunction", &iqs::qaoa::ImplementQaoaLayerBasedOnCostFunction<ComplexDP>,
        "Implement exp(-i gamma C)|psi>.");
  m.def("GetExpectationValueFromCostFunction", &iqs::qaoa::GetExpectationValueFromCostFunction<ComplexDP>,
        "Get expectation value from the cost function.");
  m.def("GetHistogramFromCostFunction", &iqs::qaoa::GetHistogramFromCostFunction<ComplexDP>,
        "Get histogram instead of just the expectation value.");
  m.def("GetHistogramFromCostFunctionWithWeightsRounded",
        &iqs::qaoa::GetHistogramFromCostFunctionWithWeightsRounded<ComplexDP>,
        "Get histogram instead of just the expectation value for the weighted graphs, rounded down.");
  m.def("GetHistogramFromCostFunctionWithWeightsBinned",
        &iqs::qaoa::GetHistogramFromCostFunctionWithWeightsBinned<ComplexDP>,
        "Get histogram instead of just the expectation value for the weighted graphs, with bins of given width.");
  m.def("GetExpectationValueSquaredFromCostFunction",
        &iqs::qaoa::GetExpectationValueSquaredFromCostFunction<ComplexDP>,
        "Get expectation value squared from the cost function.");
  // The reference module binds these for ComplexDP registers only (pybind11/intelqs_py.cpp:
  // 359-391); its C++ templates accept both precisions (src/qaoa_features.cpp), so the same
  // names also take QubitRegisterSP registers (pybind11 picks the overload by type).
  m.def("InitializeVectorAsMaxCutCostFunction", &iqs::qaoa::InitializeVectorAsMaxCutCostFunction<ComplexSP>);
  m.def("InitializeVectorAsWeightedMaxCutCostFunction",
        &iqs::qaoa::InitializeVectorAsWeightedMaxCutCostFunction<ComplexSP>);
  m.def("ImplementQaoaLayerBasedOnCostFunction", &iqs::qaoa::ImplementQaoaLayerBasedOnCostFunction<ComplexSP>);
  m.def("GetExpectationValueFromCostFunction", &iqs::qaoa::GetExpectationValueFromCostFunction<ComplexSP>);
  m.def("GetHistogramFromCostFunction", &iqs::qaoa::GetHistogramFromCostFunction<ComplexSP>);
  m.def("GetHistogramFromCostFunctionWithWeightsRounded",
        &iqs::qaoa::GetHistogramFromCostFunctionWithWeightsRounded<ComplexSP>);
  m.def("GetHistogramFromCostFunctionWithWeightsBinned",
        &iqs::qaoa::GetHistogramFromCostFunctionWithWeightsBinned<ComplexSP>);
  m.def("GetExpectationValueSquaredFromCostFunction",
        &iqs::qaoa::GetExpectationValueSquaredFromCostFunction<ComplexSP>);

  py::class_<Environment>(m, "MPIEnvironment")
      .def(py::init<>())
      .def_static("GetRank", &Environment::GetRank)
      .def_static("IsUsefulRank", &Environment::IsUsefulRank)
      .def_static("GetSizeWorldComm", []() { return 1; })
      .def_static("GetPoolRank", &Environment::GetPoolRank)
      .def_static("GetStateRank", &Environment::GetStateRank)
      .def_static("GetPoolSize", &Environment::GetPoolSize)
      .def_static("GetStateSize", &Environment::GetStateSize)
      .def_static("GetNumRanksPerNode", &Environment::GetNumRanksPerNode)
      .def_static("GetNumNodes", &Environment::GetNumNodes)
      .def_static("GetStateId", &Environment::GetStateId)
      .def_static("GetNumStates", &Environment::GetNumStates)
      .def_static("Barrier", &iqs::mpi::Barrier)
      .def_static("PoolBarrier", &iqs::mpi::PoolBarrier)
      .def_static("StateBarrier", &iqs::mpi::StateBarrier)
      .def_static("IncoherentSumOverAllStatesOfPool", &Environment::IncoherentSumOverAllStatesOfPool<double>)
      .def_static("UpdateStateComm", &Environment::UpdateStateComm);
}
