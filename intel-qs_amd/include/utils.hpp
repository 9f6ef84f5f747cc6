/// @file utils.hpp
/// Basic types of the drop-in API (include/utils.hpp of the reference: ComplexSP/ComplexDP,
/// UL, extract_value_type, time_in_seconds).
#pragma once

#include <chrono>
#include <complex>
#include <cstddef>
#include <sstream>
#include <string>

#define UL(x) ((std::size_t)(x))

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

}  // namespace iqs
