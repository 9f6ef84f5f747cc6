// Runs a NoisyQureg program on the COMPILED REFERENCE (iqs::NoisyQureg of
// include/NoisyQureg.hpp over oracle/_ref/libiqs_ref.so).  Test infrastructure only: it
// writes golden fixtures in this container and is never part of the product.
//
// Program lines (whitespace separated, '#' comments):
//   NoisyQureg n seed T1 T2          construct (must come first)
//   Initialize base|++++ index        SetGateDurations ts td      SetDecoherenceTime T1 T2
//   ResetTimeForAllQubits             ApplyNoiseGatesOnAllQubits
//   ApplyHadamard q                   ApplyRotationX|Y|Z q theta
//   ApplyCPauliX q1 q2                Apply1QubitGate q m00r m00i m01r m01i m10r m10i m11r m11i
//   ApplyControlled1QubitGate q1 q2 <8 numbers as above>
//   Counts                            prints "R Counts total one two" then "R Row q c0 c1 ..."
// The final state is written as raw complex128 amplitudes with --state FILE.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "qureg.hpp"

int main(int argc, char** argv)
{
  const char* prog = nullptr;
  const char* state_out = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--state") && i + 1 < argc) state_out = argv[++i];
    else prog = argv[i];
  }
  if (!prog) { std::fprintf(stderr, "usage: noisy_driver PROGRAM [--state F]\n"); return 1; }
  std::ifstream in(prog);
  std::string line;
  iqs::NoisyQureg<ComplexDP>* psi = nullptr;
  auto mat = [](std::istringstream& s) {
    TM2x2<ComplexDP> m;
    double v[8];
    for (double& x : v) s >> x;
    m(0, 0) = {v[0], v[1]}; m(0, 1) = {v[2], v[3]}; m(1, 0) = {v[4], v[5]}; m(1, 1) = {v[6], v[7]};
    return m;
  };
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream s(line);
    std::string op;
    s >> op;
    if (op == "NoisyQureg") {
      unsigned n, seed; double t1, t2;
      s >> n >> seed >> t1 >> t2;
      psi = new iqs::NoisyQureg<ComplexDP>(n, seed, t1, t2);
      continue;
    }
    if (!psi) { std::fprintf(stderr, "first line must construct the register\n"); return 1; }
    unsigned q = 0, q2 = 0; double a = 0, b = 0; std::string st;
    if (op == "Initialize") { std::size_t idx; s >> st >> idx; psi->Initialize(st, idx); }
    else if (op == "SetGateDurations") { s >> a >> b; psi->SetGateDurations(a, b); }
    else if (op == "SetDecoherenceTime") { s >> a >> b; psi->SetDecoherenceTime(a, b); }
    else if (op == "ResetTimeForAllQubits") psi->ResetTimeForAllQubits();
    else if (op == "ApplyNoiseGatesOnAllQubits") psi->ApplyNoiseGatesOnAllQubits();
    else if (op == "ApplyHadamard") { s >> q; psi->ApplyHadamard(q); }
    else if (op == "ApplyRotationX") { s >> q >> a; psi->ApplyRotationX(q, a); }
    else if (op == "ApplyRotationY") { s >> q >> a; psi->ApplyRotationY(q, a); }
    else if (op == "ApplyRotationZ") { s >> q >> a; psi->ApplyRotationZ(q, a); }
    else if (op == "ApplyCPauliX") { s >> q >> q2; psi->ApplyCPauliX(q, q2); }
    else if (op == "Apply1QubitGate") { s >> q; psi->Apply1QubitGate(q, mat(s)); }
    else if (op == "ApplyControlled1QubitGate") { s >> q >> q2; psi->ApplyControlled1QubitGate(q, q2, mat(s)); }
    else if (op == "Counts") {
      std::printf("R Counts %u %u %u\n", psi->GetTotalExperimentalGateCount(), psi->GetOneQubitExperimentalGateCount(),
                  psi->GetTwoQubitExperimentalGateCount());
      for (unsigned r = 0; r < psi->num_qubits; ++r) {
        std::printf("R Row %u", r);
        for (unsigned c : psi->GetExperimentalGateCount(r)) std::printf(" %u", c);
        std::printf("\n");
      }
    } else { std::fprintf(stderr, "unknown op %s\n", op.c_str()); return 1; }
  }
  if (psi && state_out) {
    FILE* f = std::fopen(state_out, "wb");
    std::fwrite(psi->RawState(), sizeof(ComplexDP), psi->LocalSize(), f);
    std::fclose(f);
  }
  delete psi;
  return 0;
}
