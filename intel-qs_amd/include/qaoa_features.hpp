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

}  // namespace qaoa
}  // namespace iqs
