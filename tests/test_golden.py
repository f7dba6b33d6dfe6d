"""Committed FP64 fixtures (tests/golden/make_golden.py) against the oracle (CPU) and the
GPU engine (gpu).  Engine tolerance: the order-2 beta = 2 truncation bound (max |dw| 4e-3);
at its defaults the engine builds the restated reference's tree, so it also equals the CPU
traversal of that tree within 2e-5 in w (and a single triangle is exact)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "winding_golden.npz")
NAMES = ["single_triangle", "cube", "icosphere_nu4", "open_displaced_nu8", "torus_24x8"]


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_golden(oracle, golden, name):
    V, F, Q, om = (golden[f"{name}/{k}"] for k in ("V", "F", "Q", "omega"))
    np.testing.assert_allclose(oracle.bruteforce(V, F, Q), om, rtol=0, atol=1e-11)
    np.testing.assert_allclose(oracle.bruteforce_numpy(V, F, Q), om, rtol=0, atol=1e-11)


@pytest.mark.parametrize("name", NAMES)
def test_restated_engine_on_golden(oracle, golden, name):
    V, F, Q, om = (golden[f"{name}/{k}"] for k in ("V", "F", "Q", "omega"))
    got = oracle.Restated(V, F).solid_angle(Q)
    assert np.abs(got - om).max() / (4 * np.pi) <= 4e-3


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_engine_on_golden(oracle, golden, name):
    from lagrange_amd.winding import FastWindingNumber
    V, F, Q, om = (golden[f"{name}/{k}"] for k in ("V", "F", "Q", "omega"))
    e = FastWindingNumber((V, F))
    got, ins = e.solid_angle_batch(Q, with_inside=True)
    err = np.abs(got.astype(np.float64) - om) / (4 * np.pi)
    assert err.max() <= 4e-3
    if F.shape[0] == 1:  # a single triangle has no far-field term near the points
        assert err.max() <= 1e-6
    R = oracle.Restated(V, F)
    tau = float(e.export_tree()[3][0])
    ref = R.solid_angle(Q, 2.0, tau=tau)
    assert np.abs(got.astype(np.float64) - ref).max() / (4 * np.pi) <= 2e-5
    np.testing.assert_allclose(e.bruteforce_f64(Q), om, rtol=0, atol=1e-9)
    w = om / (4 * np.pi)
    ok = np.abs(w - 0.5) > 5e-3
    assert np.array_equal(ins[ok], (w > 0.5)[ok])
