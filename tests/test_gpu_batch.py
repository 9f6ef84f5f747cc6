"""Batched hardware-efficient-ansatz circuits on the GPU (one persistent launch):
final states and costs against the CPU oracle running the same gate lists
(ApplyRotationY / ApplyRotationZ / ApplyCPauliZ / <Z_q>), adjoint gradients against
central finite differences of the oracle cost and against the GPU parameter-shift rule,
complex64 against complex128, graph replays against direct launches, tuning invariance,
the full 512-instance batch, and the torch.autograd.Function."""
import math

import numpy as np
import pytest

import oracle
from iqs_amd._lib import IqsError
from iqs_amd.batch import HeaBatch, instance_circuit, param_offsets, random_batch, torch_function

pytestmark = pytest.mark.gpu

TEST_VARIANT = 3  # 8 tile qubits: whole instances up to 8 qubits, tile passes for 9..16


def oracle_run(n, d, p, precision="dp"):
    st, res = oracle.run(instance_circuit(n, d, p, precision).to_text())
    z = [float(v) for v in res[-1].split()[2:]]
    return st, sum(z)


def small_batch(seed=7):
    # whole (n <= 8) and tiled (9..16, incl. n = 2*8 where B tiles have no low bits) instances
    nq = [1, 2, 3, 5, 8, 9, 10, 12, 13, 15, 16, 11]
    dp = [1, 2, 3, 2, 3, 1, 2, 3, 2, 4, 2, 5]
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * math.pi, 2 * sum(n * d for n, d in zip(nq, dp)))
    return nq, dp, ang


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 1e-5)])
@pytest.mark.parametrize("variant", [TEST_VARIANT, 0])
def test_final_states_and_costs_match_oracle(dtype, tol, variant):
    nq, dp, ang = small_batch()
    off = param_offsets(nq, dp)
    b = HeaBatch(nq, dp, dtype, tune=(variant, 0, 0))
    b.set_angles(ang)
    b.run(grad=False, keep_state=True)
    cost = b.results(grad=False)
    prec = "dp" if dtype == "c128" else "sp"
    for i, (n, d) in enumerate(zip(nq, dp)):
        st_ref, c_ref = oracle_run(n, d, ang[off[i]:off[i + 1]], prec)
        st = b.state(i)
        err = np.max(np.abs(st.astype(np.complex128) - st_ref.astype(np.complex128)))
        print(f"instance {i}: n={n} d={d} max-abs {err:.2e} cost {cost[i]:.15f} vs {c_ref:.15f}")
        assert err < tol, (i, n, d, err)
        assert abs(cost[i] - c_ref) < tol * n * 4, (i, cost[i], c_ref)
    b.close()


def test_gradient_matches_oracle_finite_differences():
    nq = [3, 4, 9, 10]
    dp = [2, 3, 2, 2]
    rng = np.random.default_rng(3)
    ang = rng.uniform(0, 2 * math.pi, 2 * sum(n * d for n, d in zip(nq, dp)))
    off = param_offsets(nq, dp)
    b = HeaBatch(nq, dp, "c128", tune=(TEST_VARIANT, 0, 0))
    cost, g = b.evaluate(ang)
    h = 1e-5
    for i, (n, d) in enumerate(zip(nq, dp)):
        p = ang[off[i]:off[i + 1]]
        assert abs(cost[i] - oracle_run(n, d, p)[1]) < 1e-12
        for k in range(0, p.size, max(1, p.size // 12)):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (oracle_run(n, d, pp)[1] - oracle_run(n, d, pm)[1]) / (2 * h)
            assert abs(g[off[i] + k] - fd) < 1e-7, (i, k, g[off[i] + k], fd)
    b.close()


@pytest.mark.parametrize("variant", [TEST_VARIANT, 0])
def test_gradient_matches_parameter_shift(variant):
    nq, dp, ang = small_batch(11)
    b = HeaBatch(nq, dp, "c128", tune=(variant, 0, 0))
    _, g = b.evaluate(ang)
    off = b.offsets
    rng = np.random.default_rng(1)
    for _ in range(4):
        slots = [int(rng.integers(0, off[i + 1] - off[i])) for i in range(b.count)]
        ps = b.parameter_shift(ang, slots)
        adj = g[off[:-1] + np.array(slots)]
        assert np.max(np.abs(adj - ps)) < 1e-11, np.max(np.abs(adj - ps))
    b.close()


def test_complex64_against_complex128():
    nq, dp, ang = small_batch(5)
    b64 = HeaBatch(nq, dp, "c64", tune=(TEST_VARIANT, 0, 0))
    b128 = HeaBatch(nq, dp, "c128", tune=(TEST_VARIANT, 0, 0))
    c64, g64 = b64.evaluate(ang)
    c128, g128 = b128.evaluate(ang)
    print("cost err", np.max(np.abs(c64 - c128)), "grad err", np.max(np.abs(g64 - g128)))
    assert np.max(np.abs(c64 - c128)) < 1e-4
    assert np.max(np.abs(g64 - g128)) < 1e-4
    b64.close()
    b128.close()


def test_graph_replays_equal_direct_launches():
    nq, dp, ang = small_batch(9)
    b = HeaBatch(nq, dp, "c64", tune=(TEST_VARIANT, 0, 2))
    b.set_angles(ang)
    b.run(grad=True, graph=False)
    c0, g0 = b.results()
    for rep in range(4):
        b.run(grad=True, graph=True)  # replay
        c, g = b.results()
        assert np.array_equal(c, c0) and np.array_equal(g, g0), rep
        b.run(grad=False, graph=False)  # other eager work between replays
        b.results(grad=False)
    b.close()


def test_tunings_agree():
    """Workers and task order change only the schedule: bitwise equal results.  Other
    variants change the summation tree: equal to rounding."""
    nq, dp, ang = small_batch(13)
    b = HeaBatch(nq, dp, "c128", tune=(TEST_VARIANT, 0, 0))
    c0, g0 = b.evaluate(ang)
    for w in (1, 2):
        for o in range(4):
            b.tune = (TEST_VARIANT, w, o)
            c, g = b.evaluate(ang)
            assert np.array_equal(c, c0) and np.array_equal(g, g0), (w, o)
        for gm in (1, 4):  # groups of a few instances (states of 2^16 x 16 B = 1 MiB each)
            b.tune = (TEST_VARIANT, w, 0, gm)
            c, g = b.evaluate(ang)
            assert np.array_equal(c, c0) and np.array_equal(g, g0), (w, gm)
    for v in range(3):
        b.tune = (v, 0, 0)
        c, g = b.evaluate(ang)
        assert np.max(np.abs(c - c0)) < 1e-12 and np.max(np.abs(g - g0)) < 1e-12, v
    b.close()


def test_autotune_small_batch():
    nq, dp, ang = random_batch(24, seed=17, qubits=(6, 16), depth=(2, 6))
    b = HeaBatch(nq, dp, "c64")
    b.set_angles(ang)
    c0, g0 = b.evaluate(ang)
    tune, ms, trials = b.autotune(reps=1)
    assert len(trials) == 3 * 2 * 6 and ms > 0 and tune in [t for t, _ in trials]
    c, g = b.evaluate(ang)
    assert np.max(np.abs(c - c0)) < 1e-5 and np.max(np.abs(g - g0)) < 1e-5
    b.close()


def test_full_batch_512_instances():
    """512 instances, n ~ U{10..22}, d ~ U{10..50}: complex64 against complex128, and
    parameter-shift spot checks of the complex128 gradient."""
    nq, dp, ang = random_batch()
    assert len(nq) == 512 and min(nq) == 10 and max(nq) == 22
    b128 = HeaBatch(nq, dp, "c128")
    c128, g128 = b128.evaluate(ang)
    slots = [0] * 512
    ps = b128.parameter_shift(ang, slots)
    assert np.max(np.abs(g128[b128.offsets[:-1]] - ps)) < 1e-10
    b64 = HeaBatch(nq, dp, "c64")
    c64, g64 = b64.evaluate(ang)
    ec, eg = np.max(np.abs(c64 - c128)), np.max(np.abs(g64 - g128))
    print(f"512 instances: cost err {ec:.2e}  grad err {eg:.2e}")
    # measured: cost 9.8e-7, gradient 1.9e-6 (FP32 amplitudes, FP64 sums)
    assert ec < 1e-5 and eg < 1e-5
    b64.close()
    b128.close()


def test_torch_function_backward():
    import torch

    nq, dp, ang = small_batch(21)
    b = HeaBatch(nq, dp, "c128", tune=(TEST_VARIANT, 0, 0))
    f = torch_function()
    a = torch.tensor(ang, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.arange(1, b.count + 1, dtype=torch.float64, device="cuda")
    cost = f.apply(a, b)
    (cost * w).sum().backward()
    c_ref, g_ref = b.evaluate(ang)
    assert np.allclose(cost.detach().cpu().numpy(), c_ref, atol=0, rtol=0)
    scale = np.repeat(np.arange(1, b.count + 1, dtype=np.float64), np.diff(b.offsets))
    assert np.array_equal(a.grad.cpu().numpy(), g_ref * scale)
    b.close()


def test_dependency_wait_gives_up_cleanly():
    """Fault injection from the host: every tile dependency wait is made unsatisfiable with
    a 64-poll limit.  The run must end (every workgroup exits), both result forms must
    report the abort and copy nothing, the torch Function must raise; restoring the
    default limit gives the normal results again."""
    import torch

    nq, dp, ang = small_batch(5)
    b = HeaBatch(nq, dp, "c128", tune=(TEST_VARIANT, 0, 0))
    c0, g0 = b.evaluate(ang)
    b.set_wait_limit(64, unmet=True)
    b.set_angles(ang)
    b.run(grad=True)
    with pytest.raises(IqsError, match="aborted"):
        b.results()
    cost = torch.full((b.count,), float("nan"), dtype=torch.float64, device="cuda")
    grad = torch.full((b.num_params,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(IqsError, match="aborted"):
        b.results_device(cost.data_ptr(), grad.data_ptr())
    assert torch.isnan(cost).all() and torch.isnan(grad).all()
    f = torch_function()
    a = torch.tensor(ang, dtype=torch.float64, device="cuda", requires_grad=True)
    with pytest.raises(IqsError, match="aborted"):
        f.apply(a, b)
    b.set_wait_limit()
    c, g = b.evaluate(ang)
    assert np.array_equal(c, c0) and np.array_equal(g, g0)
    b.close()
