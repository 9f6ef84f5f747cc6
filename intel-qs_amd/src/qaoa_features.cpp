// QAOA MaxCut helpers of the drop-in, over the C ABI (see qaoa_features.hpp).
#include "qaoa_features.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

#include "iqs_hip.h"

namespace iqs {
namespace qaoa {

namespace {
void ck(int rc)
{
  if (rc != IQSH_OK) throw std::runtime_error(std::string("iqs_hip: ") + iqsh_last_error());
}

// Edges (v < u, adjacency != 0) of a symmetric adjacency matrix, with the vertices mapped
// to data positions of `reg` (vertex k is program qubit k).
template <typename Type, typename A>
void edges_of(const QubitRegister<Type>& reg, const std::vector<A>& adj, std::vector<unsigned>& eu,
              std::vector<unsigned>& ev, std::vector<double>& ew)
{
  const std::size_t n = reg.NumQubits();
  if (adj.size() != n * n) throw std::invalid_argument("adjacency matrix must be num_qubits x num_qubits");
  for (std::size_t v = 0; v < n; ++v) {
    if (adj[v * n + v] != 0) throw std::invalid_argument("adjacency matrix must have a zero diagonal");
    for (std::size_t u = v + 1; u < n; ++u) {
      if (adj[v * n + u] != adj[u * n + v]) throw std::invalid_argument("adjacency matrix must be symmetric");
      if (adj[v * n + u] != 0) {
        eu.push_back((unsigned)reg.qubit_permutation->map[v]);
        ev.push_back((unsigned)reg.qubit_permutation->map[u]);
        ew.push_back((double)adj[v * n + u]);
      }
    }
  }
}

template <typename Type>
void same_layout(const QubitRegister<Type>& a, const QubitRegister<Type>& b)
{
  if (a.NumQubits() != b.NumQubits() || a.qubit_permutation->map != b.qubit_permutation->map)
    throw std::invalid_argument("psi and diag must have the same size and qubit order");
}
}  // namespace

template <typename Type>
int InitializeVectorAsMaxCutCostFunction(QubitRegister<Type>& diag, std::vector<int>& adjacency)
{
  std::vector<unsigned> eu, ev;
  std::vector<double> ew;
  edges_of(diag, adjacency, eu, ev, ew);
  for (double w : ew)
    if (w != 1.0) throw std::invalid_argument("unweighted MaxCut expects a 0/1 adjacency matrix");
  double mx = 0;
  ck(iqsh_maxcut_diag(diag.DeviceHandle(), (unsigned)eu.size(), eu.data(), ev.data(), nullptr, &mx));
  return (int)mx;
}

template <typename Type>
typename QubitRegister<Type>::BaseType InitializeVectorAsWeightedMaxCutCostFunction(
    QubitRegister<Type>& diag, std::vector<typename QubitRegister<Type>::BaseType>& adjacency)
{
  std::vector<unsigned> eu, ev;
  std::vector<double> ew;
  edges_of(diag, adjacency, eu, ev, ew);
  double mx = 0;
  ck(iqsh_maxcut_diag(diag.DeviceHandle(), (unsigned)eu.size(), eu.data(), ev.data(), ew.data(), &mx));
  return (typename QubitRegister<Type>::BaseType)mx;
}

template <typename Type>
void ImplementQaoaLayerBasedOnCostFunction(QubitRegister<Type>& psi, QubitRegister<Type>& diag,
                                           typename QubitRegister<Type>::BaseType gamma)
{
  same_layout(psi, diag);
  ck(iqsh_qaoa_phase_diag(psi.DeviceHandle(), diag.DeviceHandle(), (double)gamma));
}

template <typename Type>
typename QubitRegister<Type>::BaseType GetExpectationValueFromCostFunction(const QubitRegister<Type>& psi,
                                                                           const QubitRegister<Type>& diag)
{
  same_layout(psi, diag);
  double r;
  ck(iqsh_qaoa_expect_diag(const_cast<QubitRegister<Type>&>(psi).DeviceHandle(),
                           const_cast<QubitRegister<Type>&>(diag).DeviceHandle(), 0, &r));
  return (typename QubitRegister<Type>::BaseType)r;
}

template <typename Type>
typename QubitRegister<Type>::BaseType GetExpectationValueSquaredFromCostFunction(const QubitRegister<Type>& psi,
                                                                                  const QubitRegister<Type>& diag)
{
  same_layout(psi, diag);
  double r;
  ck(iqsh_qaoa_expect_diag(const_cast<QubitRegister<Type>&>(psi).DeviceHandle(),
                           const_cast<QubitRegister<Type>&>(diag).DeviceHandle(), 1, &r));
  return (typename QubitRegister<Type>::BaseType)r;
}

template <typename Type>
static std::vector<typename QubitRegister<Type>::BaseType> histogram(const QubitRegister<Type>& psi,
                                                                     const QubitRegister<Type>& diag, int mode,
                                                                     double max_value, double width, int nbins)
{
  same_layout(psi, diag);
  if (nbins < 1) throw std::invalid_argument("cost histogram: max_value must be positive");
  std::vector<double> h(nbins);
  ck(iqsh_cost_histogram(const_cast<QubitRegister<Type>&>(psi).DeviceHandle(),
                         const_cast<QubitRegister<Type>&>(diag).DeviceHandle(), mode, max_value, width, h.data(), nbins));
  return std::vector<typename QubitRegister<Type>::BaseType>(h.begin(), h.end());
}

template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunction(const QubitRegister<Type>& psi,
                                                                                 const QubitRegister<Type>& diag,
                                                                                 int max_value)
{
  return histogram(psi, diag, 0, (double)max_value, 1.0, max_value + 1);
}

template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunctionWithWeightsRounded(
    const QubitRegister<Type>& psi, const QubitRegister<Type>& diag, double max_value)
{
  return histogram(psi, diag, 1, max_value, 1.0, (int)std::floor(max_value) + 1);
}

template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunctionWithWeightsBinned(
    const QubitRegister<Type>& psi, const QubitRegister<Type>& diag, double max_value, double bin_width)
{
  if (!(bin_width > 0)) throw std::invalid_argument("cost histogram: bin_width must be positive");
  return histogram(psi, diag, 2, max_value, bin_width, (int)std::ceil(max_value / bin_width) + 1);
}

template int InitializeVectorAsMaxCutCostFunction<ComplexDP>(QubitRegister<ComplexDP>&, std::vector<int>&);
template int InitializeVectorAsMaxCutCostFunction<ComplexSP>(QubitRegister<ComplexSP>&, std::vector<int>&);
template double InitializeVectorAsWeightedMaxCutCostFunction<ComplexDP>(QubitRegister<ComplexDP>&, std::vector<double>&);
template float InitializeVectorAsWeightedMaxCutCostFunction<ComplexSP>(QubitRegister<ComplexSP>&, std::vector<float>&);
template void ImplementQaoaLayerBasedOnCostFunction<ComplexDP>(QubitRegister<ComplexDP>&, QubitRegister<ComplexDP>&, double);
template void ImplementQaoaLayerBasedOnCostFunction<ComplexSP>(QubitRegister<ComplexSP>&, QubitRegister<ComplexSP>&, float);
template double GetExpectationValueFromCostFunction<ComplexDP>(const QubitRegister<ComplexDP>&, const QubitRegister<ComplexDP>&);
template float GetExpectationValueFromCostFunction<ComplexSP>(const QubitRegister<ComplexSP>&, const QubitRegister<ComplexSP>&);
template double GetExpectationValueSquaredFromCostFunction<ComplexDP>(const QubitRegister<ComplexDP>&, const QubitRegister<ComplexDP>&);
template float GetExpectationValueSquaredFromCostFunction<ComplexSP>(const QubitRegister<ComplexSP>&, const QubitRegister<ComplexSP>&);
template std::vector<double> GetHistogramFromCostFunction<ComplexDP>(const QubitRegister<ComplexDP>&,
                                                                    const QubitRegister<ComplexDP>&, int);
template std::vector<float> GetHistogramFromCostFunction<ComplexSP>(const QubitRegister<ComplexSP>&,
                                                                   const QubitRegister<ComplexSP>&, int);
template std::vector<double> GetHistogramFromCostFunctionWithWeightsRounded<ComplexDP>(const QubitRegister<ComplexDP>&,
                                                                                      const QubitRegister<ComplexDP>&, double);
template std::vector<float> GetHistogramFromCostFunctionWithWeightsRounded<ComplexSP>(const QubitRegister<ComplexSP>&,
                                                                                     const QubitRegister<ComplexSP>&, double);
template std::vector<double> GetHistogramFromCostFunctionWithWeightsBinned<ComplexDP>(const QubitRegister<ComplexDP>&,
                                                                                     const QubitRegister<ComplexDP>&,
                                                                                     double, double);
template std::vector<float> GetHistogramFromCostFunctionWithWeightsBinned<ComplexSP>(const QubitRegister<ComplexSP>&,
                                                                                    const QubitRegister<ComplexSP>&,
                                                                                    double, double);

}  // namespace qaoa
}  // namespace iqs
