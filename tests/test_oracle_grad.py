"""Pins the oracle's FP64 gradient (oracle_sf_grad_f64) with central finite
differences of the FP64 forward oracle."""
import numpy as np
import pytest

import oracle
from tests.test_gpu_adjoint import small_problem


def test_oracle_gradient_central_differences():
    """Pins oracle.sf_grad_f64 with FP64 central differences of oracle.sf_f64
    (positions perturbed in FP32-representable steps)."""
    Z, pos, B, q = small_problem(natoms=30, n=12, seed=3)
    g = np.random.default_rng(2).normal(size=q.shape[0])
    gp, gb = oracle.sf_grad_f64(q, Z, pos, g, atom_b=B)
    L = lambda p, b: float(np.dot(g, oracle.sf_f64(q, Z, p, atom_b=b)[1]))
    for j in (0, 7, 29):
        for c in range(3):
            h = np.float32(2e-12)
            pp, pm = pos.copy(), pos.copy()
            pp[j, c] += h
            pm[j, c] -= h
            step = float(pp[j, c]) - float(pm[j, c])
            fd = (L(pp, B) - L(pm, B)) / step
            assert fd == pytest.approx(gp[j, c], rel=2e-4, abs=1e-6 * np.abs(gp).max())
        bp, bm = B.copy(), B.copy()
        bp[j] += np.float32(0.01)
        bm[j] -= np.float32(0.01)
        fd = (L(pos, bp) - L(pos, bm)) / (float(bp[j]) - float(bm[j]))
        assert fd == pytest.approx(gb[j], rel=1e-5, abs=1e-8 * np.abs(gb).max())


def test_oracle_subset_gradient_equals_full_gradient():
    """The subset entry point given F (used at the full differentiation
    workload) equals the full gradient on the same atoms, bit for bit."""
    Z, pos, B, q = small_problem(natoms=40, n=10, seed=5)
    Z[3] = 0
    g = np.random.default_rng(6).normal(size=q.shape[0])
    gp, gb = oracle.sf_grad_f64(q, Z, pos, g, atom_b=B, B=1.5)
    F, _ = oracle.sf_f64(q, Z, pos, atom_b=B, B=1.5)
    sub = np.array([39, 3, 0, 17], np.int32)
    sp, sb = oracle.sf_grad_given_f_f64(q, F, Z, pos, g, sub, atom_b=B, B=1.5)
    assert np.array_equal(sp, gp[sub]) and np.array_equal(sb, gb[sub])
    assert np.all(sp[1] == 0) and sb[1] == 0
