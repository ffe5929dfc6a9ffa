"""The oracle (CPU) and the HIP path (GPU) against the committed golden
fixtures of tests/golden (numpy/scipy formulation, see gen_golden.py)."""
import os

import numpy as np
import pytest

from tests import oracle_lib as orc

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_2x4x4x4.npz"))
DIMS = tuple(int(d) for d in G["dims"])


def check_obs(o, tol):
    plaq, energy, topc = G["obs"]
    assert abs(o[0] - plaq) < tol
    assert abs(o[1] / energy - 1) < tol
    assert abs(o[2] - topc) < tol * o[3]


def test_oracle_exp_against_golden():
    for q, e in zip(G["exp_q"], G["exp"]):
        got, _ = orc.su3_exp(q)
        assert np.abs(got - e).max() < 2e-14


def test_oracle_against_golden():
    L = G["links"]
    check_obs(orc.observables(L, DIMS, 0), 1e-13)
    for mu in range(4):
        assert np.abs(orc.derivative(L, DIMS, mu) - G["force"][mu]).max() < 1e-14
    assert np.abs(orc.flow_step(L, DIMS, float(G["step_eps"])) - G["step"]).max() < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_gpu_against_golden(precision, tol):
    from latticeyangmills_amd import LatticeContext
    with LatticeContext(DIMS, precision) as ctx:
        ctx.upload(G["links"])
        check_obs(ctx.observables(), tol)
        for mu in range(4):
            assert np.abs(ctx.derivative(mu) - G["force"][mu]).max() < tol * 10
        ctx.flow_steps(float(G["step_eps"]), 1)
        assert np.abs(ctx.download() - G["step"]).max() < tol * 10
