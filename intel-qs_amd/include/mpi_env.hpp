/// @file mpi_env.hpp
/// Single-process environment with the interface of the reference's non-MPI build
/// (include/mpi_env.hpp:46-145, src/mpi_env.cpp:33-81): one state rank, one state, no
/// communicators.  The whole register lives in one GPU's HBM; nothing here communicates.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>

namespace iqs {
namespace mpi {

class Environment
{
 public:
  Environment() = default;
  Environment(int&, char**&, bool = false) {}
  static void Init() { initialized_() = true; }
  static void Init(int&, char**&) { initialized_() = true; }
  static void Finalize() { initialized_() = false; }
  static Environment* GetSharedInstance() { static Environment e; return &e; }
  static bool IsUsefulRank() { return true; }
  static int GetRank() { return 0; }
  static int GetPoolRank() { return 0; }
  static int GetStateRank() { return 0; }
  static int GetPoolSize() { return 1; }
  static int GetStateSize() { return 1; }
  static int GetSize() { return 1; }
  static int GetNumRanksPerNode() { return 1; }
  static int GetNumNodes() { return 1; }
  static int GetNodeId() { return 0; }
  static int GetStateId() { return 0; }
  static int GetNumStates() { return 1; }
  static void RemapStateRank(int) {}
  static void UpdateStateComm(int num_states)
  {
    if (num_states != 1)
      throw std::invalid_argument("one process holds one state here; a pool of states is iqs::StatePool (state_pool.hpp)");
    std::printf("MPI not enabled.\n");  // as the reference's single-process build (src/mpi_env.cpp:43-44)
  }
  template <class T>
  static T IncoherentSumOverAllStatesOfPool(T v) { return v; }

 private:
  static bool& initialized_() { static bool b = false; return b; }
};

inline void Barrier() {}
inline void PoolBarrier() {}
inline void StateBarrier() {}

/// Print from state rank 0 (src/mpi_env.cpp:66-87, non-MPI build): the string alone, or
/// with all = true prefixed by "[|state id>:rank]".
inline void StatePrint(std::string s, bool all = false)
{
  if (all) std::printf("[|%d>:%3d] %s\n", Environment::GetStateId(), Environment::GetStateRank(), s.c_str());
  else std::printf("%s\n", s.c_str());
  std::fflush(stdout);
}
inline void PoolPrint(std::string s, bool all = false) { StatePrint(s, all); }
inline void Print(std::string s, bool all = false) { StatePrint(s, all); }

}  // namespace mpi
}  // namespace iqs
