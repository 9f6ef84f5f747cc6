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

  py::class_<iqs::RandomNumberGenerator<double>>(m, "RandomNumberGenerator")
      .def(py::init<>())
      .def("GetSeed", &iqs::RandomNumberGenerator<double>::GetSeed)
      .def("SetSeedStreamPtrs", &iqs::RandomNumberGenerator<double>::SetSeedStreamPtrs)
      .def("SkipeAhead", &iqs::RandomNumberGenerator<double>::SkipAhead)
      .def("GetUniformRandomNumbers",
           [](iqs::RandomNumberGenerator<double>& r, std::size_t size, double a, double b, std::string shared) {
             std::vector<double> v(size);
             r.UniformRandomNumbers(v.data(), size, a, b, shared);
             return v;
           })
      .def("__repr__", []() { return "<RandomNumberGenerator>"; });

  bind_register<ComplexDP>(m, "QubitRegister");
  bind_register<ComplexSP>(m, "QubitRegisterSP");

  m.def("InitializeVectorAsMaxCutCostFunction", &iqs::qaoa::InitializeVectorAsMaxCutCostFunction<ComplexDP>,
        "Use IQS vector to store a large real vector and not as a quantum state.");
  m.def("InitializeVectorAsWeightedMaxCutCostFunction",
        &iqs::qaoa::InitializeVectorAsWeightedMaxCutCostFunction<ComplexDP>,
        "Use IQS vector to store a large real vector and not as a quantum state.");
  m.def("ImplementQaoaLayerBasedOnCostFunction", &iqs::qaoa::ImplementQaoaLayerBasedOnCostFunction<ComplexDP>,
        "Implement exp(-i gamma C)|psi>.");
  m.def("GetExpectationValueFromCostFunction", &iqs::qaoa::GetExpectationValueFromCostFunction<ComplexDP>,
        "Get expectation value from the cost function.");
  m.def("GetExpectationValueSquaredFromCostFunction",
        &iqs::qaoa::GetExpectationValueSquaredFromCostFunction<ComplexDP>,
        "Get expectation value squared from the cost function.");

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
