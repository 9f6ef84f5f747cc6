/// @file utils.hpp
/// Basic types and helpers of the drop-in API (include/utils.hpp of the reference:
/// ComplexSP/ComplexDP, UL, the TODO/INFO compile-time messages, sec(), xstr,
/// extract_value_type, time_in_seconds, WhatCompileDefinitions).  Callers of the reference
/// use the macros (e.g. examples/quantum_fourier_transform.cpp:131 calls TODO), so they keep
/// their names and meaning.
#pragma once

#include <chrono>
#include <complex>
#include <cstddef>
#include <iostream>
#include <sstream>
#include <string>

// Compile-time notes (include/utils.hpp:7-15 of the reference).
#define DO_PRAGMA(x) _Pragma(#x)
#ifndef TODO
#define TODO(x) DO_PRAGMA(message("\033[30;43mTODO\033[0m - " #x))
#endif
#ifndef INFO
#define INFO(x) DO_PRAGMA(message("\033[30;46mINFO\033[0m - " #x))
#endif

#define UL(x) ((std::size_t)(x))
#define sec() time_in_seconds()
#define xstr(s) __str__(s)
#define __str__(s) #s

using ComplexSP = std::complex<float>;
using ComplexDP = std::complex<double>;

namespace iqs {

template <typename T>
struct extract_value_type { typedef T value_type; };
template <template <typename> class X, typename T>
struct extract_value_type<X<T>> { typedef T value_type; };

inline double time_in_seconds()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <typename T>
std::string toString(const T& v)
{
  std::ostringstream os;
  os << v;
  return os.str();
}

inline unsigned ilog2(std::size_t x)
{
  unsigned r = 0;
  while (x >>= 1) ++r;
  return r;
}

inline bool isPowerOf2(std::size_t x) { return x && !(x & (x - 1)); }

/// The build's configuration, in the reference's format (src/utils.cpp): this engine has
/// no MPI and no MKL; the amplitudes live in GPU memory (no _mm_malloc).
inline void WhatCompileDefinitions()
{
  std::cout << "Compiler flags:\n"
            << "          INTELQS_HAS_MPI --> [ NO]\n"
            << "                  USE_MKL --> [ NO]\n"
#ifdef _OPENMP
            << "                  _OPENMP --> [YES]\n"
#else
            << "                  _OPENMP --> [ NO]\n"
#endif
            << "         __INTEL_COMPILER --> [ NO]\n"
            << "                    __ICC --> [ NO]\n"
            << "            USE_MM_MALLOC --> [ NO]\n"
            << "               STANDALONE --> [ NO]\n"
#ifdef NDEBUG
            << "                   NDEBUG --> [YES]\n";
#else
            << "                   NDEBUG --> [ NO]\n";
#endif
}

}  // namespace iqs
