"""Pins of the oracle's random numbers, candidates and Metropolis update.

  - Philox-4x32-10 against the known-answer vectors of its authors' Random123
    distribution (kat_vectors);
  - candidates: unitarity, det = 1, the R*S*T structure of randSU3Transf
    (|x| = eps per SU(2) block), symmetry of exp(i eps H) under u -> 1-u;
  - Delta S of the update against the brute-force change of the Wilson action;
  - statistics: strong-coupling expansion of the plaquette, and the Haar
    moments of the hot start.
"""
import numpy as np
import pytest

from tests import npref as nr
from tests import oracle_lib as orc

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert orc.philox(ctr, key) == want


def _cand(kind, u, eps):
    import ctypes
    x = np.empty(18)
    f = orc.lib().or_rand_su3_transf if kind == 0 else orc.lib().or_rand_su3_exph
    f(u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), eps, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return orc.from18(x)


@pytest.mark.parametrize("kind", [0, 1])
def test_candidates_are_su3(kind):
    rng = np.random.default_rng(1)
    for _ in range(200):
        X = _cand(kind, rng.random(12), 0.24)
        assert np.abs(X @ X.conj().T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(X) - 1) < 1e-14


def test_su2_product_structure():
    """R*S*T with R, S, T SU(2) blocks of rows/cols (0,1), (0,2), (1,2), each
    [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 - i x3]], |x_vec| = eps."""
    rng = np.random.default_rng(2)
    eps = 0.24
    for _ in range(50):
        u = rng.random(12)
        blocks = []
        for b in range(3):
            x = u[4 * b:4 * b + 4] - 0.5
            v = x[1:] * eps / np.linalg.norm(x[1:])
            x0 = np.sign(x[0]) * np.sqrt(1 - eps * eps) if x[0] != 0 else np.sqrt(1 - eps * eps)
            m = np.array([[x0 + 1j * v[2], v[1] + 1j * v[0]], [-v[1] + 1j * v[0], x0 - 1j * v[2]]])
            blocks.append(m)
        R, S, T = np.eye(3, dtype=complex), np.eye(3, dtype=complex), np.eye(3, dtype=complex)
        R[np.ix_([0, 1], [0, 1])] = blocks[0]
        S[np.ix_([0, 2], [0, 2])] = blocks[1]
        T[np.ix_([1, 2], [1, 2])] = blocks[2]
        assert np.abs(_cand(0, u, eps) - R @ S @ T).max() < 1e-15


def test_exph_candidate_is_symmetric():
    """u -> 1-u flips H, so X(1-u) = X(u)^+ and the proposal is symmetric."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        u = rng.random(9)
        X, Y = _cand(1, u, 0.24), _cand(1, 1.0 - u, 0.24)
        assert np.abs(X @ Y - np.eye(3)).max() < 1e-13


def test_delta_s_equals_action_difference():
    """-(beta/3) Re Tr[(X U - U) Omega] = S(U') - S(U) for the Wilson action
    S = -(beta/3) sum Re Tr P (src/Actions/puregauge.cpp:28)."""
    dims = (4, 4, 4, 4)
    L = orc.hot_start(dims, 12)
    U = nr.to_field(L, dims)
    beta = 6.0
    rng = np.random.default_rng(4)
    S0 = -beta / 3 * nr.plaq_sum(U)
    for mu, site in [(0, 3), (1, 77), (2, 130), (3, 254)]:
        om = orc.from18(orc.staples(L, dims, mu))[site]
        X = _cand(1, rng.random(9), 0.24)
        W = U.copy().reshape(4, -1, 3, 3)
        Uold = W[mu, site].copy()
        W[mu, site] = X @ Uold
        dS = -beta / 3 * np.real(np.trace((W[mu, site] - Uold) @ om))
        S1 = -beta / 3 * nr.plaq_sum(W.reshape(U.shape))
        assert abs(dS - (S1 - S0)) < 1e-10


def test_metropolis_uses_predrawn_stream():
    """A pre-drawn stream equal to the Philox stream reproduces the Philox run."""
    dims = (2, 2, 2, 4)
    V = orc.volume(dims)
    nhit, nsw, seed, first = 3, 2, 77, 5
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for cand, nu in ((0, 13), (1, 9)):
        u = np.empty((nsw, 4 * V, nhit, nu))
        for w in range(nsw):
            for lk in range(4 * V):
                for h in range(nhit):
                    r = sum((orc.philox([lk, first + w, h, b], key) for b in range(4)), [])
                    u[w, lk, h] = np.array(r[:nu]) / 2.0 ** 32
        L0 = orc.hot_start(dims, 1)
        a, acc_a = orc.metropolis(L0, dims, 5.5, 0.24, nhit, nsw, first, cand, seed)
        b, acc_b = orc.metropolis(L0, dims, 5.5, 0.24, nhit, nsw, 0, cand, 0, u.ravel())
        assert acc_a == acc_b and np.array_equal(a, b)


def test_metropolis_rejects_odd_extent():
    L = orc.cold_start((3, 4, 4, 4))
    import ctypes
    acc = ctypes.c_int64(0)
    rc = orc.lib().or_metropolis(L.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), (ctypes.c_int * 4)(3, 4, 4, 4),
                                 6.0, 0.2, 1, 1, 0, 1, 0, None, ctypes.byref(acc))
    assert rc == -1


def test_strong_coupling_plaquette():
    """<Re Tr P>/3 = beta/18 + beta^2/216 + O(beta^3) for SU(3) at small beta."""
    dims = (6, 6, 6, 6)
    beta = 0.6
    L = orc.hot_start(dims, 31)
    L, _ = orc.metropolis(L, dims, beta, 0.9, 8, 10, 0, 1, 31)
    vals = []
    for k in range(12):
        L, _ = orc.metropolis(L, dims, beta, 0.9, 8, 2, 10 + 2 * k, 1, 31)
        vals.append(orc.observables(L, dims)[0])
    want = beta / 18 + beta ** 2 / 216
    err = np.std(vals) / np.sqrt(len(vals)) + 1.0 / np.sqrt(6 * orc.volume(dims) * len(vals)) / 3
    assert abs(np.mean(vals) - want) < 5 * err + 2e-3, (np.mean(vals), want, err)


def test_hot_start_is_haar():
    """Haar SU(3): E[Tr U] = 0, E[|Tr U|^2] = 1, E[Tr U^2] = 0."""
    dims = (8, 8, 8, 8)
    M = orc.from18(orc.hot_start(dims, 2024)).reshape(-1, 3, 3)
    n = M.shape[0]
    assert np.abs(M @ nr.dag(M) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    tr = np.trace(M, axis1=1, axis2=2)
    tol = 5 / np.sqrt(n)
    assert abs(tr.mean()) < tol
    assert abs(np.mean(np.abs(tr) ** 2) - 1) < 3 * tol
    assert abs(np.trace(M @ M, axis1=1, axis2=2).mean()) < 2 * tol
