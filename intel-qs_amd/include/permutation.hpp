/// @file permutation.hpp
/// Program-qubit <-> data-position map with the interface of the reference's
/// iqs::Permutation (include/permutation.hpp:51-264):
///   map[q]  = data position of program qubit q,   imap[p] = program qubit at position p.
/// data2program_/program2data_ move the bits of a basis index between the two orders.
#pragma once

#include <cstddef>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "utils.hpp"

namespace iqs {

class Permutation
{
 public:
  std::vector<std::size_t> map, imap;
  std::size_t num_elements = 0;

  explicit Permutation(std::size_t n) : num_elements(n)
  {
    std::vector<std::size_t> id(n);
    std::iota(id.begin(), id.end(), 0);
    SetNewPermutationFromMap(id, "direct");
  }
  Permutation(std::vector<std::size_t> m, std::string style = "direct") : num_elements(m.size())
  {
    SetNewPermutationFromMap(m, style);
  }

  unsigned operator[](std::size_t i) const { return (unsigned)at(i); }
  unsigned operator[](unsigned i) const { return (unsigned)at(i); }
  int operator[](int i) const { return (int)at((std::size_t)i); }
  std::size_t size() const { return map.size(); }

  /// Program qubit stored at a data position (= imap[position]).
  std::size_t Find(std::size_t position) const
  {
    for (std::size_t q = 0; q < map.size(); ++q)
      if (map[q] == position) return q;
    throw std::out_of_range("Permutation::Find: position not present");
  }

  void SetNewPermutationFromMap(std::vector<std::size_t> m, std::string style = "direct")
  {
    if (m.size() != num_elements) throw std::invalid_argument("permutation of the wrong size");
    std::vector<char> seen(m.size(), 0);
    for (std::size_t v : m) {
      if (v >= m.size() || seen[v]) throw std::invalid_argument("not a permutation");
      seen[v] = 1;
    }
    if (style == "direct") {
      map = m;
      imap.assign(m.size(), 0);
      for (std::size_t q = 0; q < m.size(); ++q) imap[m[q]] = q;
    } else if (style == "inverse") {
      imap = m;
      map.assign(m.size(), 0);
      for (std::size_t p = 0; p < m.size(); ++p) map[m[p]] = p;
    } else {
      throw std::invalid_argument("style of map must be \"direct\" or \"inverse\"");
    }
  }

  /// Swap the positions of two program qubits (EmulateSwap).
  void ExchangeTwoElements(std::size_t e1, std::size_t e2)
  {
    const std::size_t p1 = map[e1], p2 = map[e2];
    map[e1] = p2;
    map[e2] = p1;
    imap[p1] = e2;
    imap[p2] = e1;
  }

  /// Basis index in data order -> program order (bit map[q] of v becomes bit q).
  std::size_t data2program_(std::size_t v) const
  {
    std::size_t r = 0;
    for (std::size_t q = 0; q < num_elements; ++q) r |= ((v >> map[q]) & 1) << q;
    return r;
  }
  /// Basis index in program order -> data order (bit imap[p] of v becomes bit p).
  std::size_t program2data_(std::size_t v) const
  {
    std::size_t r = 0;
    for (std::size_t p = 0; p < num_elements; ++p) r |= ((v >> imap[p]) & 1) << p;
    return r;
  }

  /// "i0i1i2..." with the least significant bit first.
  static std::string dec2bin(std::size_t v, std::size_t nbits)
  {
    std::string s(nbits, '0');
    for (std::size_t i = 0; i < nbits; ++i) s[i] = ((v >> i) & 1) ? '1' : '0';
    return s;
  }
  static std::size_t bin2dec(const std::string& s)
  {
    std::size_t v = 0;
    for (std::size_t i = 0; i < s.size(); ++i) v |= (std::size_t)(s[i] - '0') << i;
    return v;
  }
  std::string data2program(std::size_t v) const { return data2program(dec2bin(v, num_elements)); }
  std::string data2program(const std::string& s) const
  {
    std::string r(s);
    for (std::size_t q = 0; q < s.size(); ++q) r[q] = s[map[q]];
    return r;
  }
  std::string program2data(std::size_t v) const { return program2data(dec2bin(v, num_elements)); }
  std::string program2data(const std::string& s) const
  {
    std::string r(s);
    for (std::size_t p = 0; p < s.size(); ++p) r[p] = s[imap[p]];
    return r;
  }

  std::string GetMapStr() const
  {
    std::string s;
    for (auto v : map) s += " " + toString(v);
    return s;
  }
  std::string GetImapStr() const
  {
    std::string s;
    for (auto v : imap) s += " " + toString(v);
    return s;
  }
  void Print() const { printf("qubit permutation: %s\n", GetMapStr().c_str()); }

 private:
  std::size_t at(std::size_t i) const
  {
    if (i >= num_elements) throw std::out_of_range("Permutation index out of range");
    return map[i];
  }
};

}  // namespace iqs
