/// @file gate_spec.hpp
/// Gate kinds used by SpecializeV2 (reference include/gate_spec.hpp:10-25).  The numeric
/// values cross the C ABI (iqsh_spec).
#pragma once

namespace iqs {

enum class GateSpec1Q { Hadamard = 0, RotationX, RotationY, RotationZ, PauliX, PauliY, PauliZ, T, None };
enum class GateSpec2Q { CHadamard = 0, CRotationX, CRotationY, CRotationZ, CPauliX, CPauliY, CPauliZ, CPhase, None };

inline GateSpec1Q ConvertSpec2to1(GateSpec2Q s)
{
  switch (s) {
    case GateSpec2Q::CHadamard: return GateSpec1Q::Hadamard;
    case GateSpec2Q::CRotationX: return GateSpec1Q::RotationX;
    case GateSpec2Q::CRotationY: return GateSpec1Q::RotationY;
    case GateSpec2Q::CRotationZ: return GateSpec1Q::RotationZ;
    case GateSpec2Q::CPauliX: return GateSpec1Q::PauliX;
    case GateSpec2Q::CPauliY: return GateSpec1Q::PauliY;
    case GateSpec2Q::CPauliZ: return GateSpec1Q::PauliZ;
    default: return GateSpec1Q::None;
  }
}

}  // namespace iqs
