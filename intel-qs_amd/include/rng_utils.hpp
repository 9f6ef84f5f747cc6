/// @file rng_utils.hpp
/// Host random numbers with the stream scheme of the reference (include/rng_utils.hpp,
/// src/rng_utils.cpp:84-164): three std::mt19937 streams (pool, state, local) seeded with
/// seed, seed + 1 + state_id and seed + 1 + num_states + pool_rank, uniform numbers from
/// std::uniform_real_distribution<Type>(0, 1).  Used for Initialize("rand").
#pragma once

#include <cassert>
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mpi_env.hpp"

namespace iqs {

template <typename Type>
class RandomNumberGenerator
{
 public:
  RandomNumberGenerator() = default;
  explicit RandomNumberGenerator(RandomNumberGenerator* src) { *this = *src; }

  std::size_t GetSeed() const { return seed_; }
  /// Numbers drawn or skipped per stream since the last seeding (a Gaussian counts twice),
  /// as include/rng_utils.hpp:82-89 of the reference.
  std::size_t GetNumGeneratedOrSkippedLocalNumbers() const { return n_local_; }
  std::size_t GetNumGeneratedOrSkippedStateNumbers() const { return n_state_; }
  std::size_t GetNumGeneratedOrSkippedPoolNumbers() const { return n_pool_; }
  void SetSeedStreamPtrs(std::size_t seed)
  {
    SetSeedStreamPtrs(seed, mpi::Environment::GetStateId(), mpi::Environment::GetNumStates(),
                      mpi::Environment::GetPoolRank());
  }
  /// The seeding a rank with these pool coordinates gets from SetSeedStreamPtrs(seed)
  /// (src/rng_utils.cpp:84-99); used by iqs::StatePool to give each of its states the
  /// streams of the corresponding pool state.
  void SetSeedStreamPtrs(std::size_t seed, int state_id, int num_states, int pool_rank)
  {
    seed_ = seed;
    n_local_ = n_state_ = n_pool_ = 0;
    pool_.seed(seed + 0);
    state_.seed(seed + 1 + state_id);
    local_.seed(seed + 1 + num_states + pool_rank);
  }
  void SkipAhead(std::size_t num_skip, std::string shared = "local")
  {
    std::vector<Type> sink(num_skip);
    UniformRandomNumbers(sink.data(), num_skip, Type(0), Type(1), shared);
  }
  void UniformRandomNumbers(Type* out, std::size_t size, Type a = 0, Type b = 1, std::string shared = "local")
  {
    std::mt19937* g = select(size, shared);
    for (std::size_t i = 0; i < size; ++i) out[i] = a + (b - a) * u_(*g);
  }
  void GaussianRandomNumbers(Type* out, std::size_t size, std::string shared = "local")
  {
    // Counted as two uniforms per Gaussian (src/rng_utils.cpp:196-210).
    std::mt19937* g = select(2 * size, shared);
    for (std::size_t i = 0; i < size; ++i) out[i] = n_(*g);
  }
  /// Integers in [a, b): floor(a + u (b - a)) with u from the uniform stream, as the
  /// reference computes them (src/rng_utils.cpp:170-188).
  void RandomIntegersInRange(int* out, std::size_t size = 1, int a = 0, int b = 2, std::string shared = "local")
  {
    std::mt19937* g = select(size, shared);
    for (std::size_t i = 0; i < size; ++i) {
      Type r = u_(*g);
      r = (Type)a + r * Type(b - a);
      out[i] = (int)std::floor(r);
    }
  }

 private:
  std::mt19937* select(std::size_t size, const std::string& s)
  {
    if (s == "local") { n_local_ += size; return &local_; }
    if (s == "state") { n_state_ += size; return &state_; }
    if (s == "pool") { n_pool_ += size; return &pool_; }
    throw std::invalid_argument("RNG stream must be local, state or pool");
  }
  std::size_t seed_ = 0;
  std::size_t n_local_ = 0, n_state_ = 0, n_pool_ = 0;
  std::mt19937 local_, state_, pool_;
  std::uniform_real_distribution<Type> u_{Type(0), Type(1)};
  std::normal_distribution<Type> n_{Type(0), Type(1)};
};

/// Fisher-Yates shuffle in the reference's draw order (src/rng_utils.cpp:403-423): for
/// ra = size-1 down to 1, one integer rb in [0, ra + 1) from `shared`, then swap a[ra], a[rb].
template <typename Type, typename TypeFloat>
void ShuffleFisherYates(std::vector<Type>& array, RandomNumberGenerator<TypeFloat>* rng, std::string shared = "local")
{
  if (!rng) throw std::invalid_argument("ShuffleFisherYates: null generator");
  for (int ra = (int)array.size() - 1; ra > 0; --ra) {
    int rb;
    rng->RandomIntegersInRange(&rb, 1, 0, ra + 1, shared);
    if (ra != rb) std::swap(array[ra], array[rb]);
  }
}

}  // namespace iqs
