"""Pins of the oracle's SU(3) algebra and exponential map.

  - bit-exact against the reference's own src/Math/su3.cpp, compiled from the
    reference tree into oracle/_ref (skipped where the tree is absent);
  - against numpy complex matmul and scipy.linalg.expm / eigendecomposition;
  - closed forms: diagonal Q, Q = 0, unitarity and det = 1;
  - branch coverage of the exponential: c0 < 0, |w| < 0.05, near-degenerate theta.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

from tests import oracle_lib as orc

rng = np.random.default_rng(2018)


def rand_herm(scale):
    H = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    H = H + H.conj().T
    H -= np.trace(H) / 3 * np.eye(3)
    return H * scale / np.linalg.norm(H)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.skipif(orc.ref() is None, reason="reference su3.cpp not built (reference tree absent)")
def test_bit_exact_against_reference_su3():
    R, O = orc.ref(), orc.lib()
    for i in range(3000):
        a, b = rng.normal(size=18), rng.normal(size=18)
        r1, r2 = np.empty(18), np.empty(18)
        O.or_su3_mul(_p(a), _p(b), _p(r1))
        R.ref_su3_mul(_p(a), _p(b), _p(r2))
        assert np.array_equal(r1, r2)
        O.or_su3_adj(_p(a), _p(r1))
        R.ref_su3_adj(_p(a), _p(r2))
        assert np.array_equal(r1, r2)
        assert O.or_su3_retr_mul(_p(a), _p(b)) == R.ref_su3_retr_mul(_p(a), _p(b))
        q = orc.to18(rand_herm(rng.uniform(1e-3, 3.0)))
        O.or_su3_exp(_p(q), _p(r1))
        R.ref_su3_exp(_p(q), _p(r2))
        assert np.array_equal(r1, r2), i


def test_mul_matches_numpy():
    for _ in range(200):
        A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        B = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        got = orc.from18(orc.su3_mul(orc.to18(A), orc.to18(B)))
        assert np.abs(got - A @ B).max() < 1e-14


@pytest.mark.parametrize("scale", [1e-6, 1e-3, 0.03, 0.3, 1.0, 2.5])
def test_exp_matches_expm_and_eigh(scale):
    for _ in range(100):
        H = rand_herm(scale)
        got, flag = orc.su3_exp(orc.to18(H))
        E = orc.from18(got)
        assert flag == 0
        assert np.abs(E - sl.expm(1j * H)).max() < 2e-14
        w, V = np.linalg.eigh(H)
        assert np.abs(E - V @ np.diag(np.exp(1j * w)) @ V.conj().T).max() < 2e-14
        assert np.abs(E @ E.conj().T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(E) - 1) < 1e-14


def test_exp_diagonal_closed_form():
    for _ in range(50):
        lam = rng.normal(size=3)
        lam -= lam.mean()
        got, _ = orc.su3_exp(orc.to18(np.diag(lam).astype(complex)))
        assert np.abs(orc.from18(got) - np.diag(np.exp(1j * lam))).max() < 1e-14


def test_exp_branches():
    # c0 < 0 (det Q < 0) and c0 > 0
    for sign in (1, -1):
        H = np.diag([0.3, 0.5, -0.8]).astype(complex) * sign
        H[0, 1], H[1, 0] = 0.1j, -0.1j
        got, _ = orc.su3_exp(orc.to18(H))
        assert np.abs(orc.from18(got) - sl.expm(1j * H)).max() < 1e-14
    # |w| < 0.05: two nearly equal eigenvalues -> theta near 0, the xi series
    H = np.diag([0.5, 0.5 + 1e-3, -1.0 - 1e-3]).astype(complex)
    got, _ = orc.su3_exp(orc.to18(H))
    assert np.abs(orc.from18(got) - sl.expm(1j * H)).max() < 1e-13
    # exactly degenerate eigenvalues: c0/c0max rounds to >= 1 (clamped)
    H = np.diag([0.5, 0.5, -1.0]).astype(complex)
    got, flag = orc.su3_exp(orc.to18(H))
    assert flag == 0 and np.all(np.isfinite(got))
    assert np.abs(orc.from18(got) - sl.expm(1j * H)).max() < 1e-12


def test_exp_of_zero_is_identity():
    got, flag = orc.su3_exp(np.zeros(18))
    assert flag == 1
    assert np.array_equal(orc.from18(got), np.eye(3))
