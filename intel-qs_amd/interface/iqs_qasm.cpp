// QASM front end of the engine: reads "<op> <args>" lines from standard input and drives
// one iqs::QubitRegister<ComplexDP> on the GPU.  Operations, qubit-id assignment and
// messages are those of the reference's QASM interface (interface/src/intelqs_interface_main
// .cpp:17-55, interface_api_qasm.cpp:21-140, interface_api_memory.cpp:14-73,
// interface_api_qubitid.cpp:41-60, interface_api_version.cpp:7-20), so a client that drives
// the reference through this wire format can drive this engine unchanged.
#include <complex>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "qureg.hpp"

namespace {

using Type = ComplexDP;

// Version strings reported by ".version" and ".iversion": the interface level this
// program is compatible with (interface_api_version.h:16, qureg_version.hpp:21).
constexpr const char* kInterfaceVersion = "RC-1.1.0";
constexpr const char* kLibraryVersion = "SDK-RC-2.1.0";
// Angle of the T-dagger gate, -pi/4 as the reference writes it (interface_api_qasm.cpp:18).
constexpr double kTdagTheta = -0.785398163397448;

struct Session {
  iqs::QubitRegister<Type>* psi = nullptr;
  // Qubit names get register indices 0, 1, ... in order of first use.
  std::unordered_map<std::string, int> ids;
  int next = 0;

  int qubit(const std::string& name)
  {
    auto it = ids.find(name);
    if (it == ids.end()) it = ids.emplace(name, next++).first;
    return it->second;
  }

  unsigned long malloc_(const std::string& args)
  {
    if (psi) {
      std::cerr << ".malloc called before .free - memory leak detected." << std::endl;
      return 1;
    }
    int n = 0;
    try {
      n = std::stoi(args, nullptr, 10);
    } catch (const std::invalid_argument&) {
      std::cerr << ".malloc (size) - size parameter was not an integer." << std::endl;
    } catch (const std::out_of_range&) {
      std::cerr << ".malloc (size) - size parameter was too large to fit in an integer." << std::endl;
    }
    if (n > 0 && n <= 43) {
      try {
        psi = new iqs::QubitRegister<Type>(n);
      } catch (const std::exception& e) {
        std::cerr << ".malloc (size) - allocation of " << n << " qubits failed: " << e.what() << std::endl;
        psi = nullptr;
        return 1;
      }
      (*psi)[0] = 1;
      std::cout << "Allocated [" << n << "] qubits." << std::endl;
      return 0;
    }
    std::cerr << ".malloc (size) - allocations in the range from 1..43 only." << std::endl;
    return 1;
  }

  unsigned long free_(const std::string&)
  {
    if (psi) {
      delete psi;
      psi = nullptr;
      std::cout << "Qubits successfully free." << std::endl;
      return 0;
    }
    std::cerr << "Harmless attempt to .free before .malloc. Nothing done." << std::endl;
    return 1;
  }

  // Gate on one named qubit: echo "<op> [<args>]", then apply.
  unsigned long gate1(const char* op, const std::string& args, const std::function<void(unsigned)>& f)
  {
    std::cout << op << " [" << args << "]" << std::endl;
    f(qubit(args));
    return 0;
  }

  unsigned long execute(const std::string& op, const std::string& args)
  {
    if (op == ".malloc") return malloc_(args);
    if (op == ".free") return free_(args);
    if (op == ".iversion") {
      std::cout << kInterfaceVersion << std::endl;
      return 0;
    }
    if (op == ".version") {
      std::cout << kLibraryVersion << std::endl;
      return 0;
    }
    const bool needs_register = op == "H" || op == "X" || op == "T" || op == "S" || op == "Tdag" ||
                                op == "CNOT" || op == "MeasZ";
    if (needs_register && !psi) {
      // (the reference dereferences its unallocated register here)
      std::cerr << op << ": no qubits allocated (.malloc first)" << std::endl;
      return 1;
    }
    if (op == "H") return gate1("H", args, [&](unsigned q) { psi->ApplyHadamard(q); });
    if (op == "X") return gate1("X", args, [&](unsigned q) { psi->ApplyPauliX(q); });
    if (op == "T") return gate1("T", args, [&](unsigned q) { psi->ApplyT(q); });
    if (op == "S") return gate1("S", args, [&](unsigned q) { psi->ApplyPauliSqrtZ(q); });
    if (op == "Tdag") return gate1("Tdag", args, [&](unsigned q) { psi->ApplyRotationZ(q, kTdagTheta); });
    if (op == "CNOT") {
      const std::size_t comma = args.find_first_of(',');
      const int a = qubit(args.substr(0, comma));
      const int b = qubit(args.substr(comma + 1, args.length()));
      std::cout << "CNOT" << " [" << args << "]" << std::endl;
      psi->ApplyCPauliX(a, b);
      return 0;
    }
    if (op == "MeasZ") {
      std::cout << "MeasZ" << " [" << args << "]" << std::endl;
      const Type p = psi->GetProbability(qubit(args));
      std::cout << p << std::endl;
      return 0;
    }
    if (op == "PrepZ") {
      std::cout << "PrepZ" << " [" << args << "]" << std::endl;
      return 0;
    }
    if (op == "qubit") {
      std::cout << "Qubit (scaffold syntax): " << args << std::endl;
      return 0;
    }
    if (op == "cbit") {
      std::cout << "Cbit (scaffold syntax): " << args << std::endl;
      return 0;
    }
    return 1;  // "*" and unknown operations
  }
};

}  // namespace

int main(int argc, char* argv[])
{
  iqs::mpi::Environment env(argc, argv);
  Session s;
  std::string line;
  while (true) {
    std::getline(std::cin, line);
    if (line.length() < 1) break;
    const std::size_t end = line.find_first_of(' ');
    const std::string token = line.substr(0, end);
    if (token.empty()) continue;
    // With no space, end + 1 wraps to 0 and the arguments are the whole line, as in the
    // reference.
    unsigned long result = 1;
    try {
      result = s.execute(token, line.substr(end + 1, line.length()));
    } catch (const std::exception& e) {
      std::cerr << token << ": " << e.what() << std::endl;
      result = 1;
    }
    if (result > 0) std::cerr << "Qasm Op failed - [" << token << "]" << std::endl;
  }
  delete s.psi;
  return 0;
}
