/// @file qaoa_features.hpp
/// QAOA MaxCut helpers with the interface of the reference (include/qaoa_features.hpp,
/// src/qaoa_features.cpp:73-326).  The cost register ("diag") holds cut values in the real
/// part of each amplitude, in data order; it is filled on the GPU from the edge list
/// (at most 64 edges), and the phase layer / expectation kernels read it.
#pragma once

#include <vector>

#include "qureg.hpp"

namespace iqs {
namespace qaoa {

template <typename Type>
int InitializeVectorAsMaxCutCostFunction(QubitRegister<Type>& diag, std::vector<int>& adjacency);

template <typename Type>
typename QubitRegister<Type>::BaseType InitializeVectorAsWeightedMaxCutCostFunction(
    QubitRegister<Type>& diag, std::vector<typename QubitRegister<Type>::BaseType>& adjacency);

/// psi_i *= exp(-i gamma diag_i)
template <typename Type>
void ImplementQaoaLayerBasedOnCostFunction(QubitRegister<Type>& psi, QubitRegister<Type>& diag,
                                           typename QubitRegister<Type>::BaseType gamma);

/// sum_i diag_i |psi_i|^2 (FP64 accumulation)
template <typename Type>
typename QubitRegister<Type>::BaseType GetExpectationValueFromCostFunction(const QubitRegister<Type>& psi,
                                                                           const QubitRegister<Type>& diag);

/// sum_i diag_i^2 |psi_i|^2 (FP64 accumulation)
template <typename Type>
typename QubitRegister<Type>::BaseType GetExpectationValueSquaredFromCostFunction(const QubitRegister<Type>& psi,
                                                                                  const QubitRegister<Type>& diag);

/// Probability of each integer cost value 0..max_value: bin (int)diag_i
/// (src/qaoa_features.cpp:335-393).
template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunction(const QubitRegister<Type>& psi,
                                                                                 const QubitRegister<Type>& diag,
                                                                                 int max_value);

/// Weighted costs rounded down: bin floor(diag_i + 1e-7), floor(max_value) + 1 bins
/// (src/qaoa_features.cpp:402-462).
template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunctionWithWeightsRounded(
    const QubitRegister<Type>& psi, const QubitRegister<Type>& diag, double max_value);

/// Weighted costs in bins of bin_width: bin floor(diag_i / bin_width + 1e-7),
/// ceil(max_value / bin_width) + 1 bins (src/qaoa_features.cpp:471-529).
template <typename Type>
std::vector<typename QubitRegister<Type>::BaseType> GetHistogramFromCostFunctionWithWeightsBinned(
    const QubitRegister<Type>& psi, const QubitRegister<Type>& diag, double max_value, double bin_width);

}  // namespace qaoa
}  // namespace iqs
