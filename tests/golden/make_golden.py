"""Write the golden fixtures in this folder by running circuits through the COMPILED
REFERENCE (oracle/_ref/ref_driver, built by `make -C oracle ref` from the reference's own
sources).  Run in the development container, where the reference sources exist:

    python tests/golden/make_golden.py

Each case writes <name>.circ (the gate list), <name>.res (the reference's read-out lines)
and <name>.npz with either the full final state ("state") or, for large registers, sampled
amplitudes ("idx", "amp") plus FP64 checksums ("norm2", "sum").  Reductions are made with
OMP_NUM_THREADS=1 so that their order is the reference's serial order.
"""
from __future__ import annotations

import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "intel-qs_amd"))

from iqs_amd.circuits import (Circuit, SplitMix64, brick_circuit, format_expectation,  # noqa: E402
                              haar_unitary, qaoa_circuit, random_1q2q_circuit,
                              random_regular_graph, random_su2)

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def prep(c: Circuit, rng: SplitMix64) -> None:
    """Generic entangled start state: H and random rotations everywhere, CNOT chain."""
    n = c.num_qubits
    for q in range(n):
        c.add("ApplyHadamard", q)
        c.add("Apply1QubitGate", q, random_su2(rng))
    for q in range(n - 1):
        c.add("ApplyCPauliX", q, q + 1)
    for q in range(n):
        c.add("ApplyRotationY", q, rng.uniform(0, 2 * math.pi))


def readout(c: Circuit) -> None:
    c.add("ZAll")
    c.add("ComputeNorm")
    c.add("GetGlobalAmplitude", 5)


def case_gates1q(prec: str, spec2: bool = False) -> Circuit:
    rng = SplitMix64(101)
    c = Circuit(6, precision=prec)
    if spec2:
        c.add("TurnOnSpecializeV2")
    prep(c, rng)
    for q in range(6):
        for name in ("ApplyHadamard", "ApplyPauliX", "ApplyPauliY", "ApplyPauliZ", "ApplyPauliSqrtX",
                     "ApplyPauliSqrtY", "ApplyPauliSqrtZ", "ApplyT"):
            c.add(name, q)
        for name in ("ApplyRotationX", "ApplyRotationY", "ApplyRotationZ"):
            c.add(name, q, rng.uniform(-7, 7))
        c.add("ApplyRotationXY", q, rng.uniform(0, 6), rng.uniform(0, 6))
        c.add("Apply1QubitGate", q, haar_unitary(2, rng))
    readout(c)
    return c


def case_gates2q(prec: str, spec2: bool = False) -> Circuit:
    rng = SplitMix64(202)
    n = 6
    c = Circuit(n, precision=prec)
    if spec2:
        c.add("TurnOnSpecializeV2")
    prep(c, rng)
    pairs = [(0, 1), (1, 0), (0, 5), (5, 0), (2, 4), (4, 2), (3, 1)]
    for (a, b) in pairs:
        for name in ("ApplyCPauliX", "ApplyCPauliY", "ApplyCPauliZ", "ApplyCPauliSqrtZ", "ApplyCHadamard"):
            c.add(name, a, b)
        for name in ("ApplyCRotationX", "ApplyCRotationY", "ApplyCRotationZ", "ApplyCPhaseRotation"):
            c.add(name, a, b, rng.uniform(-7, 7))
        c.add("ApplyControlled1QubitGate", a, b, haar_unitary(2, rng))
        for name in ("ApplySwap", "ApplyISwap", "ApplySqrtISwap", "Apply4thRootISwap"):
            c.add(name, a, b)
        th = rng.uniform(0, 6)
        c.add("ApplyISwapRotation", a, b, np.array([[math.cos(th), 1j * math.sin(th)], [1j * math.sin(th), math.cos(th)]]))
        d = np.exp(1j * np.array([rng.uniform(0, 6) for _ in range(4)]))
        c.add("ApplyDiag", a, b, d)
        c.add("Apply2QubitGate", a, b, haar_unitary(4, rng))
    for (a, b, t) in [(0, 1, 2), (5, 3, 0), (2, 4, 5), (1, 5, 3)]:
        c.add("ApplyToffoli", a, b, t)
    readout(c)
    return c


def case_permute(prec: str) -> Circuit:
    rng = SplitMix64(303)
    n = 7
    c = Circuit(n, precision=prec)
    prep(c, rng)
    c.add("PermuteQubits", "direct", [1, 0, 2, 6, 4, 5, 3])
    c.add("ApplyHadamard", 3)
    c.add("ApplyCPauliX", 6, 0)
    c.add("Apply2QubitGate", 4, 1, haar_unitary(4, rng))  # position order == qubit order
    c.add("ApplyDiag", 2, 5, np.exp(1j * np.array([0.1, 0.7, -1.3, 2.2])))
    c.add("PermuteQubits", "inverse", [3, 1, 4, 0, 6, 2, 5])
    c.add("EmulateSwap", 0, 6)
    c.add("ApplyRotationX", 0, 0.4)
    c.add("ApplySwap", 2, 4)
    c.add("PermuteQubits", "direct", list(range(n)))
    for k in (0, 5, 77, 127):
        c.add("GetGlobalAmplitude", k)
    readout(c)
    return c


def case_expect(prec: str) -> Circuit:
    rng = SplitMix64(404)
    n = 6
    c = Circuit(n, precision=prec)
    prep(c, rng)
    for q in range(n):
        c.add("GetProbability", q)
    terms = [
        (1.0, [(0, 1)]), (1.0, [(2, 2)]), (0.5, [(3, 3)]), (2.0, [(0, 1), (1, 1)]),
        (1.0, [(0, 2), (5, 2)]), (-1.5, [(1, 3), (4, 3)]), (1.0, [(0, 1), (2, 2), (4, 3)]),
        (1.0, [(5, 1), (3, 2), (1, 3), (0, 1)]), (3.0, []),
    ]
    for coeff, pairs in terms:
        c.ops.append(format_expectation(coeff, pairs))
    c.add("CollapseQubit", 2, 1)
    c.add("ComputeNorm")
    c.add("Normalize")
    c.add("GetProbability", 2)
    c.add("CollapseQubit", 4, 0)
    c.add("Normalize")
    readout(c)
    return c


def case_qaoa(prec: str, weighted: bool = False) -> Circuit:
    rng = SplitMix64(505)
    if weighted:
        n = 6
        g = [(0, 1, 0.5), (1, 2, 1.25), (2, 3, 2.0), (3, 4, 0.75), (4, 5, 1.5), (5, 0, 0.3), (0, 3, 1.1)]
    else:
        n = 8
        g = random_regular_graph(n, 3, 20240604)
    gam = [rng.uniform(0, math.pi) for _ in range(2)]
    bet = [rng.uniform(0, math.pi) for _ in range(2)]
    c = qaoa_circuit(n, g, gam, bet, precision=prec)
    c.add("ZAll")
    return c


CASES = {
    "gates1q_dp": lambda: case_gates1q("dp"),
    "gates1q_sp": lambda: case_gates1q("sp"),
    "spec2_1q_dp": lambda: case_gates1q("dp", True),
    "spec2_1q_sp": lambda: case_gates1q("sp", True),
    "gates2q_dp": lambda: case_gates2q("dp"),
    "gates2q_sp": lambda: case_gates2q("sp"),
    "spec2_2q_dp": lambda: case_gates2q("dp", True),
    "spec2_2q_sp": lambda: case_gates2q("sp", True),
    "permute_dp": lambda: case_permute("dp"),
    "permute_sp": lambda: case_permute("sp"),
    "expect_dp": lambda: case_expect("dp"),
    "expect_sp": lambda: case_expect("sp"),
    "qaoa_dp": lambda: case_qaoa("dp"),
    "qaoa_sp": lambda: case_qaoa("sp"),
    "qaoaw_dp": lambda: case_qaoa("dp", True),
    "random1q2q_dp": lambda: random_1q2q_circuit(8, 60, 20240603),
    "random1q2q_sp": lambda: random_1q2q_circuit(8, 60, 20240603, precision="sp"),
    "brick10_dp": lambda: brick_circuit(10, 8, 77),
    "brick10_sp": lambda: brick_circuit(10, 8, 77, precision="sp"),
    # The reference's CPU-runnable headline case: 20 qubits, depth 20, seed 20240601.
    "brick20_dp": lambda: brick_circuit(20, 20, 20240601),
}

FULL_STATE_MAX_QUBITS = 10
SAMPLE_SEED = 99


def sample_indices(n: int, k: int = 4096) -> np.ndarray:
    rng = SplitMix64(SAMPLE_SEED)
    return np.array(sorted({rng.randint(1 << n) for _ in range(k)}), dtype=np.int64)


def run_reference(text: str):
    with tempfile.TemporaryDirectory() as d:
        cf, sf, rf = (os.path.join(d, x) for x in ("c.txt", "s.bin", "r.txt"))
        with open(cf, "w") as f:
            f.write(text)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cf, "--state", sf, "--results", rf], check=True, env=env)
        prec = "sp" if "precision sp" in text else "dp"
        st = np.fromfile(sf, dtype=np.complex64 if prec == "sp" else np.complex128)
        res = open(rf).read()
    return st, res


def case_io(prec: str) -> Circuit:
    """State I/O: Print (stdout text) and ExportAmplitudes (JSON fragment) after a qubit
    permutation, so labels are program-order strings of data-order amplitudes.  The
    ExportAmplitudes argument OUT is replaced by a file path when the circuit runs."""
    rng = SplitMix64(909)
    c = Circuit(4, precision=prec)
    prep(c, rng)
    c.add("PermuteQubits", "direct", [2, 0, 3, 1])
    c.add("ApplyHadamard", 1)
    c.add("ApplyCPauliX", 3, 0)
    c.add("Print", "state")
    c.add("ExportAmplitudes", "OUT")
    return c


def make_io() -> None:
    """io_<prec>.iocirc (kept apart from the state fixtures), io_print_<prec>.txt (the
    reference's stdout) and io_export_<prec>.json."""
    for prec in ("dp", "sp"):
        c = case_io(prec)
        text = c.to_text()
        with tempfile.TemporaryDirectory() as d:
            cf, out = os.path.join(d, "c.txt"), os.path.join(d, "export.json")
            with open(cf, "w") as f:
                f.write(text.replace("ExportAmplitudes OUT", f"ExportAmplitudes {out}"))
            r = subprocess.run([DRIVER, cf, "--results", os.path.join(d, "r.txt")], check=True,
                               env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True, text=True)
            export = open(out).read()
        for suffix, data in ((".iocirc", text), ("_print.txt", r.stdout), ("_export.json", export)):
            name = f"io_{prec}{suffix}" if suffix == ".iocirc" else f"io{suffix.replace('.', '_' + prec + '.', 1)}"
            with open(os.path.join(HERE, name), "w") as f:
                f.write(data)
        print(f"io_{prec}: print {len(r.stdout.splitlines())} lines, export {len(export.splitlines())} lines")


def case_noise() -> Circuit:
    """Noise gates (src/qureg_noisysimul.cpp): seeded Gaussian Pauli-twirl rotations on an
    entangled state.  The reference module binds the generator for ComplexDP only."""
    rng = SplitMix64(515)
    c = Circuit(6, precision="dp")
    prep(c, rng)
    c.add("SetRngSeed", 777)
    c.add("SetNoiseTimescales", 50.0, 30.0)
    for k in range(18):
        c.add("ApplyNoiseGate", k % 6, [0.5, 2.0, 7.5][k % 3])
        if k % 5 == 4:
            c.add("ApplyCPauliX", k % 6, (k + 1) % 6)
    readout(c)
    return c


def make_noise() -> None:
    """noise_dp.noisecirc, .noiseres (read-outs) and noise_dp_state.npz (kept apart from the
    oracle's fixture set: the oracle restates no random number generator)."""
    c = case_noise()
    text = c.to_text()
    st, res = run_reference(text)
    with open(os.path.join(HERE, "noise_dp.noisecirc"), "w") as f:
        f.write(text)
    with open(os.path.join(HERE, "noise_dp.noiseres"), "w") as f:
        f.write(res)
    np.savez_compressed(os.path.join(HERE, "noise_dp_state.npz"), state=st)
    print(f"noise_dp: ops={len(c.ops)} results={len(res.splitlines())}")


QASM_REF = os.path.join(ROOT, "oracle", "_ref", "iqs_qasm_ref")
QASM_SCRIPTS = {
    # every operation of the interface, qubit ids assigned by first use (the scaffold
    # "qubit"/"cbit" lines and PrepZ assign none), failing and unknown operations
    "qasm_mixed": [".iversion", ".version", ".malloc 5", "qubit q0", "cbit c0", "H q0", "CNOT q0,q1", "T q1",
                   "Tdag q2", "H q2", "S q3", "X q4", "H q3", "CNOT q3,q4", "PrepZ q0", "MeasZ q0", "MeasZ q1",
                   "MeasZ q2", "MeasZ q3", "MeasZ q4", "Foo bar", "*", ".malloc 3", ".free", ".free",
                   ".malloc 100", ".malloc abc"],
    # GHZ-like circuit on named qubits with T/Tdag phases
    "qasm_ghz": [".malloc 6"] + [f"H r{i}" for i in range(6)] + [f"CNOT r{i},r{i + 1}" for i in range(5)]
                + ["T r0", "Tdag r5", "H r0", "H r5", "S r2", "H r2"] + [f"MeasZ r{i}" for i in range(6)] + [".free"],
}


def make_qasm() -> None:
    """qasm_<name>.qasm and the reference program's stdout/stderr (.out/.err)."""
    for name, lines in QASM_SCRIPTS.items():
        text = "\n".join(lines) + "\n"
        r = subprocess.run([QASM_REF], input=text, capture_output=True, text=True, check=True,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        for suffix, data in ((".qasm", text), (".out", r.stdout), (".err", r.stderr)):
            with open(os.path.join(HERE, name + suffix), "w") as f:
                f.write(data)
        print(f"{name}: {len(lines)} lines, stdout {len(r.stdout.splitlines())} lines, stderr {len(r.stderr.splitlines())}")


def make_stats() -> None:
    """stats_<prec>.statcirc and the reference's Entropy / GoogleStats read-outs
    (.statres): a 10-qubit brick state and a sparse one (three basis states, so that zero
    probabilities are skipped as the reference skips them)."""
    for prec in ("dp", "sp"):
        c = brick_circuit(10, 6, 1234, precision=prec, readout=False)
        c.add("Entropy")
        c.add("GoogleStats")
        sparse = Circuit(10, precision=prec)
        sparse.add("ApplyHadamard", 3)
        sparse.add("ApplyCHadamard", 3, 7)
        sparse.add("Entropy")
        sparse.add("GoogleStats")
        for tag, circ in (("", c), ("_sparse", sparse)):
            text = circ.to_text()
            _, res = run_reference(text)
            with open(os.path.join(HERE, f"stats{tag}_{prec}.statcirc"), "w") as f:
                f.write(text)
            with open(os.path.join(HERE, f"stats{tag}_{prec}.statres"), "w") as f:
                f.write(res)
            print(f"stats{tag}_{prec}: {res.splitlines()[0][:60]}")


POOL_STATES, POOL_SEED = 5, 77777


def case_pool(seed: int) -> Circuit:
    """One pool state's program in the reference's noisy-pool tutorial
    (tutorials/get_started_with_noisy_IQS.cpp:160-215) at 6 qubits: noise after state
    preparation, then X, Y and Z rotations qubit by qubit, each followed by a noise gate on
    every qubit, a CNOT ladder in between, and the read-outs.  Pool state s of the
    reference draws from the state stream seeded with seed + 1 + s, which a single-state run
    gets from SetRngSeed(seed + s)."""
    n = 6
    ang = SplitMix64(4040)
    angles = [[ang.uniform() * math.pi for _ in range(n)] for _ in range(3)]
    c = Circuit(n, precision="dp")
    c.add("SetRngSeed", seed)
    c.add("SetNoiseTimescales", 40.0, 20.0)
    for q in range(n):
        c.add("ApplyNoiseGate", q, 1.5)
    for axis, gate in enumerate(("ApplyRotationX", "ApplyRotationY", "ApplyRotationZ")):
        for qubit in range(n):
            c.add(gate, qubit, angles[axis][qubit])
            for q in range(n):
                c.add("ApplyNoiseGate", q, 1.5)
        if axis == 0:
            for q in range(n - 1):
                c.add("ApplyCPauliX", q, q + 1)
    c.add("GetProbability", 0)
    c.add("GetProbability", 3)
    return c


def make_pool() -> None:
    """pool_dp.poolcirc (the program of pool state 0), and per state s the reference's final
    state and read-outs for seed POOL_SEED + s (pool_dp_states.npz, pool_dp.poolres)."""
    states, res = [], []
    for s in range(POOL_STATES):
        text = case_pool(POOL_SEED + s).to_text()
        if s == 0:
            with open(os.path.join(HERE, "pool_dp.poolcirc"), "w") as f:
                f.write(text)
        st, r = run_reference(text)
        states.append(st)
        res.append(f"# state {s}\n" + r)
    np.savez_compressed(os.path.join(HERE, "pool_dp_states.npz"), states=np.array(states), seed=POOL_SEED)
    with open(os.path.join(HERE, "pool_dp.poolres"), "w") as f:
        f.write("".join(res))
    print(f"pool_dp: {POOL_STATES} states, {len(case_pool(0).ops)} ops each")


NOISY_REF = os.path.join(ROOT, "oracle", "_ref", "noisy_driver")


def noisy_program() -> str:
    """Two noisy trajectories of one 6-qubit circuit on iqs::NoisyQureg (the pattern of the
    reference's examples/circuit_with_noise_gates.cpp): logical 1-qubit gates of every kind,
    CNOTs and controlled gates on qubit pairs, final idle noise, gate counts; the second
    trajectory re-initializes and continues the same generator."""
    rs = np.random.default_rng(20240610)
    q, _ = np.linalg.qr(rs.normal(size=(2, 2)) + 1j * rs.normal(size=(2, 2)))
    g = " ".join(f"{x:.17g}" for v in q.reshape(-1) for x in (v.real, v.imag))
    lines = ["NoisyQureg 6 4321 40 25", "SetGateDurations 1 3"]
    for traj in range(2):
        lines += ["Initialize base 0", "ResetTimeForAllQubits"]
        for i, (a, b) in enumerate([(0, 1), (2, 3), (4, 5), (1, 2), (3, 4), (5, 0)]):
            lines.append(f"ApplyCPauliX {a} {b}" if i % 2 else f"ApplyControlled1QubitGate {a} {b} {g}")
        for k in range(6):
            lines.append(f"Apply1QubitGate {k} {g}")
        lines += ["ApplyHadamard 2", "ApplyRotationX 0 0.37", "ApplyRotationY 4 -1.1", "ApplyRotationZ 5 2.2",
                  "ApplyNoiseGatesOnAllQubits", "Counts"]
        if traj == 0:
            lines.append("SetDecoherenceTime 60 35")
    return "\n".join(lines) + "\n"


def make_noisy() -> None:
    """noisy_dp.nqprog, the reference's read-outs (.nqres) and final state (noisy_dp_state.npz)."""
    text = noisy_program()
    prog = os.path.join(HERE, "noisy_dp.nqprog")
    with open(prog, "w") as f:
        f.write(text)
    with tempfile.TemporaryDirectory() as td:
        sf = os.path.join(td, "state.bin")
        r = subprocess.run([NOISY_REF, prog, "--state", sf], capture_output=True, text=True, check=True,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        st = np.fromfile(sf, dtype=np.complex128)
    with open(os.path.join(HERE, "noisy_dp.nqres"), "w") as f:
        f.write(r.stdout)
    np.savez_compressed(os.path.join(HERE, "noisy_dp_state.npz"), state=st)
    print(f"noisy_dp: {len(text.splitlines())} lines, norm {np.linalg.norm(st):.15f}")


def main() -> None:
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} missing: run `make -C oracle ref` first")
    make_io()
    make_noise()
    make_noisy()
    make_pool()
    make_stats()
    make_qasm()
    for name, fn in CASES.items():
        c = fn()
        text = c.to_text()
        st, res = run_reference(text)
        with open(os.path.join(HERE, name + ".circ"), "w") as f:
            f.write(text)
        with open(os.path.join(HERE, name + ".res"), "w") as f:
            f.write(res)
        if c.num_qubits <= FULL_STATE_MAX_QUBITS:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), state=st)
        else:
            idx = sample_indices(c.num_qubits)
            a = st.astype(np.complex128)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), idx=idx, amp=st[idx],
                                norm2=np.array([np.vdot(a, a).real]), sum=np.array([a.sum()]))
        print(f"{name}: n={c.num_qubits} ops={len(c.ops)} results={len(res.splitlines())}")


if __name__ == "__main__":
    main()
