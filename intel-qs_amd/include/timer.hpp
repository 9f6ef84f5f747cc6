/// @file timer.hpp
/// Per-gate statistics of the drop-in (the reference's Timer, include/timer.hpp:129-341,
/// records host time and a bandwidth model per gate).  Here each Start/Stop pair
/// synchronizes the device and records wall time and the algorithmic bytes of the gate.
/// Breakdown() prints what the reference's single-process build prints (its per-gate table
/// needs MPI, include/timer.hpp:320-323), so programs print the same text through the
/// drop-in; BreakdownRates() prints the achieved GB/s per gate name.  Enabled by
/// EnableStatistics().
#pragma once

#include <cstdio>
#include <map>
#include <string>

#include "utils.hpp"

namespace iqs {

class Timer
{
 public:
  explicit Timer(std::size_t num_qubits) : n_(num_qubits) {}
  void Start(const std::string& name, std::size_t /*pos*/, std::size_t /*pos2*/ = 0)
  {
    cur_ = name;
    t0_ = time_in_seconds();
  }
  void Stop(double bytes = 0)
  {
    Rec& r = recs_[cur_];
    r.n += 1;
    r.t += time_in_seconds() - t0_;
    r.bytes += bytes;
  }
  void Reset() { recs_.clear(); }
  void Breakdown() const
  {
    printf(" *** The statistics (i.e. time used in computation and bandwidth) are available only when MPI is used inside Intel QS.\n");
    printf(" *** This is not the case for this simulation. If needed, rebuild the IQS adding the CMake option '-DIqsMPI=ON'.\n");
  }
  void BreakdownRates() const
  {
    printf("%-40s %8s %12s %10s\n", "gate", "calls", "time[s]", "GB/s");
    for (auto const& kv : recs_)
      printf("%-40s %8ld %12.6f %10.1f\n", kv.first.c_str(), kv.second.n, kv.second.t,
             kv.second.t > 0 ? kv.second.bytes / kv.second.t / 1e9 : 0.0);
  }

 private:
  struct Rec { long n = 0; double t = 0, bytes = 0; };
  std::size_t n_;
  std::string cur_;
  double t0_ = 0;
  std::map<std::string, Rec> recs_;
};

}  // namespace iqs
