"""Host side of the HEA batch: the seeded batch generator, the parameter layout and the
gate lists handed to the oracle; the oracle's cost of an HEA instance against a plain
numpy state-vector simulation of the same circuit (RY, RZ, CZ ladder, sum_q <Z_q>)."""
import math

import numpy as np

import oracle
from iqs_amd.batch import instance_circuit, param_offsets, random_batch


def test_random_batch_is_seeded_and_in_range():
    a = random_batch()
    b = random_batch()
    assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2])
    nq, dp, ang = a
    assert len(nq) == 512 and min(nq) >= 10 and max(nq) <= 22 and min(dp) >= 10 and max(dp) <= 50
    assert len(set(nq)) == 13 and ang.size == 2 * sum(n * d for n, d in zip(nq, dp))
    assert ang.min() >= 0 and ang.max() < 2 * math.pi
    assert param_offsets(nq, dp)[-1] == ang.size


def test_instance_circuit_ops():
    c = instance_circuit(4, 2, np.arange(16, dtype=float))
    names = [op[0] for op in c.ops]
    assert names.count("ApplyRotationY") == 8 and names.count("ApplyRotationZ") == 8
    assert names.count("ApplyCPauliZ") == 6 and names[-1] == "ZAll"
    assert c.ops[0] == ("ApplyRotationY", 0, 0.0) and c.ops[1] == ("ApplyRotationZ", 0, 1.0)


def numpy_hea(n, d, p):
    p = np.asarray(p).reshape(d, n, 2)
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    x = np.arange(1 << n)
    for l in range(d):
        for q in range(n):
            th, ph = p[l, q]
            c, s = math.cos(th / 2), math.sin(th / 2)
            u = np.array([[c, -s], [s, c]]) * 1.0
            u = np.diag([np.exp(-0.5j * ph), np.exp(0.5j * ph)]) @ u
            b = (x >> q) & 1
            i0 = x[b == 0]
            a0, a1 = psi[i0].copy(), psi[i0 | (1 << q)].copy()
            psi[i0], psi[i0 | (1 << q)] = u[0, 0] * a0 + u[0, 1] * a1, u[1, 0] * a0 + u[1, 1] * a1
        par = np.zeros(1 << n, int)
        for i in range(n - 1):
            par ^= ((x >> i) & 1) & ((x >> (i + 1)) & 1)
        psi = np.where(par, -psi, psi)
    o = n - 2 * np.array([bin(v).count("1") for v in x])
    return psi, float(np.sum(o * np.abs(psi) ** 2))


def test_oracle_hea_matches_numpy():
    rng = np.random.default_rng(2)
    for n, d in [(1, 2), (3, 2), (6, 3)]:
        p = rng.uniform(0, 2 * math.pi, 2 * n * d)
        st, res = oracle.run(instance_circuit(n, d, p).to_text())
        ref, cost = numpy_hea(n, d, p)
        assert np.max(np.abs(st - ref)) < 1e-13
        assert abs(sum(float(v) for v in res[-1].split()[2:]) - cost) < 1e-12
