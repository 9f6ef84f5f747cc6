"""MI355X-native state-vector engine for the gate-application path of intel-qs.

Modules:
  engine    device-resident StateVector over the C ABI of libiqs_hip.so
  circuits  serialized gate lists and deterministic workload generators
  gates     the reference's named-gate matrices (program-level helpers)
"""
from .engine import StateVector, sync  # noqa: F401
