"""Pins of the oracle's staples, force, observables and flow against the
independent numpy formulation (tests/npref.py), invariances and the
mathematics of the RK3 integrator."""
import numpy as np
import pytest

from tests import npref as nr
from tests import oracle_lib as orc

DIMS = [(4, 4, 4, 4), (4, 6, 2, 4), (3, 4, 5, 2)]


@pytest.fixture(scope="module", params=DIMS, ids=lambda d: "x".join(map(str, d)))
def hot(request):
    dims = request.param
    return dims, orc.hot_start(dims, 77)


def test_cold_start(hot):
    dims, _ = hot
    L = orc.cold_start(dims)
    plaq, energy, topc, scale = orc.observables(L, dims)
    assert plaq == 1.0 and energy == 0.0 and topc == 0.0 and scale == 0.0
    for mu in range(4):
        om = orc.from18(orc.staples(L, dims, mu))
        assert np.array_equal(om, np.broadcast_to(6 * np.eye(3), om.shape))
        assert not np.any(orc.derivative(L, dims, mu))


def test_staples_and_force_match_numpy(hot):
    dims, L = hot
    U = nr.to_field(L, dims)
    for mu in range(4):
        om = orc.from18(orc.staples(L, dims, mu))
        assert np.abs(om - nr.staple_sum(U, mu).reshape(-1, 3, 3)).max() < 1e-14
        z = orc.from18(orc.derivative(L, dims, mu))
        assert np.abs(z - nr.force(U, mu).reshape(-1, 3, 3)).max() < 1e-14


def test_force_is_lie_derivative_of_plaquette_sum():
    """Tr(H Z_mu(x)) = d/ds sum Re Tr P for U_mu(x) -> exp(isH) U_mu(x): pins the
    factor 1/2, the sign and the staple orientation."""
    dims = (4, 4, 4, 4)
    L = orc.hot_start(dims, 5)
    U = nr.to_field(L, dims)
    rng = np.random.default_rng(0)
    h = 1e-5
    for mu, site in [(0, 0), (1, 37), (2, 100), (3, 255)]:
        z = orc.from18(orc.derivative(L, dims, mu))[site]
        H = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        H = H + H.conj().T
        H -= np.trace(H) / 3 * np.eye(3)
        vals = []
        for s in (h, -h):
            W = U.copy().reshape(4, -1, 3, 3)
            W[mu, site] = nr.expi(s * H) @ W[mu, site]
            vals.append(nr.plaq_sum(W.reshape(U.shape)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - np.real(np.trace(H @ z))) < 1e-7 * max(1.0, abs(fd))


def test_observables_match_numpy_clover(hot):
    dims, L = hot
    plaq, energy, topc = nr.observables(nr.to_field(L, dims))
    for mode in (0, 1):
        o = orc.observables(L, dims, mode)
        assert abs(o[0] - plaq) < 1e-14
        assert abs(o[1] / energy - 1) < 1e-13
        assert abs(o[2] - topc) < 1e-13 * o[3]


def test_gauge_invariance_and_covariance(hot):
    dims, L = hot
    U = nr.to_field(L, dims)
    rng = np.random.default_rng(3)
    G = nr.haar_su3(rng, orc.volume(dims)).reshape(U[0].shape)
    Lg = nr.to_links(nr.gauge_transform(U, G))
    a, b = orc.observables(L, dims), orc.observables(Lg, dims)
    assert abs(a[0] - b[0]) < 1e-13 and abs(a[1] / b[1] - 1) < 1e-12 and abs(a[2] - b[2]) < 1e-12 * a[3]
    for mu in range(4):
        z = orc.from18(orc.derivative(L, dims, mu))
        zg = orc.from18(orc.derivative(Lg, dims, mu))
        g = G.reshape(-1, 3, 3)
        assert np.abs(zg - g @ z @ nr.dag(g)).max() < 1e-12


def test_translation_invariance(hot):
    dims, L = hot
    U = nr.to_field(L, dims)
    Ut = np.roll(np.roll(U, 1, axis=1 + nr.ax(0)), -1, axis=1 + nr.ax(3))
    a, b = orc.observables(L, dims), orc.observables(nr.to_links(Ut), dims)
    assert abs(a[0] - b[0]) < 1e-13 and abs(a[1] / b[1] - 1) < 1e-12 and abs(a[2] - b[2]) < 1e-12 * a[3]


def test_topological_charge_odd_under_reflection():
    """t -> -t: U'_t(x') = U_t(x - t)^+, U'_i(x') = U_i(x) with x' = (x, y, z, -t)."""
    dims = (4, 4, 4, 6)
    L = orc.hot_start(dims, 9)
    L, _ = orc.metropolis(L, dims, 6.0, 0.24, 10, 4, 0, 1, 9)
    U = nr.to_field(L, dims)
    refl = lambda A: np.roll(np.flip(A, axis=0), 1, axis=0)  # t -> -t (mod Lt) on a (Lt, ...) field
    Ur = np.stack([refl(U[0]), refl(U[1]), refl(U[2]), refl(nr.dag(nr.bwd(U[3], 3)))])
    a, b = orc.observables(L, dims), orc.observables(nr.to_links(Ur), dims)
    assert abs(a[0] - b[0]) < 1e-13 and abs(a[1] / b[1] - 1) < 1e-12
    assert abs(a[2] + b[2]) < 1e-12 * a[3] and abs(a[2]) > 1e-6


def test_flow_step_matches_numpy_rk3(hot):
    dims, L = hot
    got = orc.flow_step(L, dims, 0.02)
    want = nr.to_links(nr.rk3_step(nr.to_field(L, dims), 0.02))
    assert np.abs(got - want).max() < 1e-13


def test_rk3_is_third_order():
    """Global error at fixed flow time falls by 2^3 when eps halves."""
    dims = (4, 4, 4, 4)
    L, _ = orc.metropolis(orc.hot_start(dims, 2), dims, 6.0, 0.24, 10, 5, 0, 1, 2)
    T = 0.4

    def run(eps):
        W = L.copy()
        for _ in range(int(round(T / eps))):
            W = orc.flow_step(W, dims, eps)
        return W

    ref = run(0.005)
    errs = [np.abs(run(e) - ref).max() for e in (0.04, 0.02)]
    ratio = errs[0] / errs[1]
    assert 6.5 < ratio < 10.0, (errs, ratio)  # first / second order would give 2 / 4


def test_flow_decreases_action_and_keeps_su3():
    dims = (4, 4, 4, 4)
    L, _ = orc.metropolis(orc.hot_start(dims, 4), dims, 6.0, 0.24, 10, 5, 0, 1, 4)
    series, W = orc.flow(L, dims, 0.02, 0.5, 1)
    assert np.all(np.diff(series[:, 1]) > 0)  # Wilson action falls, plaquette rises
    M = orc.from18(W)
    assert np.abs(M @ nr.dag(M) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13


def test_flow_row_count_follows_reference():
    """rows = int(tau/eps), rows-1 steps, t accumulated by += eps (wilsonflow.cpp:117-126)."""
    assert orc.lib().or_flow_rows.__call__ is not None
    import ctypes
    f = orc.lib().or_flow_rows
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_double, ctypes.c_double]
    assert [f(t, e) for t, e in ((1.0, 0.01), (4.0, 0.01), (6.0, 0.02), (2.0, 0.01), (8.0, 0.02))] == \
        [100, 400, 300, 200, 400]


def weak_abelian_mode(dims, alpha, plane_fields):
    """Diagonal links U_mu = exp(i A_mu h) for smooth abelian fields A_mu(x):
    plane_fields maps mu -> (h, function of the site coordinates)."""
    V = orc.volume(dims)
    s = np.arange(V)
    coords = [s % dims[0], (s // dims[0]) % dims[1], (s // (dims[0] * dims[1])) % dims[2],
              s // (dims[0] * dims[1] * dims[2])]
    U = np.zeros((4, V, 3, 3), complex)
    U[:] = np.eye(3)
    for mu, (h, f) in plane_fields.items():
        a = alpha * f(coords)
        for k in range(3):
            U[mu, :, k, k] = np.exp(1j * a * h[k])
    return orc.to18(U)


def test_energy_density_normalisation_on_smooth_field():
    """a^4 E = -sum Re Tr C^2 / 64 is the continuum E = sum_{mu<nu} tr F^2:
    for A_0 = a sin(k y) h1 and A_2 = a cos(k t) h2 (tr h^2 = 2) the volume
    average is 2 a^2 k^2, and the plaquette form 36 (1 - plaq) agrees; both up
    to O(k^2) lattice corrections, so they converge as L grows."""
    h1 = np.array([1.0, -1.0, 0.0])
    h2 = np.array([1.0, 1.0, -2.0]) / np.sqrt(3)
    prev = None
    for L in (16, 32):
        dims = (2, L, 2, L)
        k = 2 * np.pi / L
        links = weak_abelian_mode(dims, 0.05, {0: (h1, lambda c: np.sin(k * c[1])),
                                               2: (h2, lambda c: np.cos(k * c[3]))})
        plaq, energy, _, _ = orc.observables(links, dims)
        V = orc.volume(dims)
        e_clover = -energy / (64 * V)
        e_cont = 2 * 0.05 ** 2 * k ** 2
        assert abs(e_clover / (36 * (1 - plaq)) - 1) < 0.04
        err = abs(e_clover / e_cont - 1)
        assert err < 0.06
        if prev is not None:
            assert err < prev / 3  # O(k^2): halving k quarters the deviation
        prev = err


def test_flow_time_is_heat_equation_time():
    """A weak abelian mode A_0 = a sin(k y) h1 obeys the heat equation under
    the flow, dA/dt = Laplacian A, so E(t) = E(0) exp(-2 khat^2 t) with
    khat^2 = 4 sin^2(k/2).  Pins the 1/2 in Z, the RK3 coefficients and the
    flow-time scale t (in units of a^2) together."""
    L = 16
    dims = (2, L, 2, 2)
    k = 2 * np.pi / L
    links = weak_abelian_mode(dims, 0.01, {0: (np.array([1.0, -1.0, 0.0]), lambda c: np.sin(k * c[1]))})
    series, _ = orc.flow(links, dims, 0.01, 1.01, 25, 0)
    kh2 = 4 * np.sin(k / 2) ** 2
    for t, _, _, e in series[::25]:
        assert abs(e / series[0, 3] - np.exp(-2 * kh2 * t)) < 1e-5, t


@pytest.mark.parametrize("dims,eps", [((2, 4, 4, 4), 0.02), ((4, 4, 4, 6), 0.3), ((3, 4, 5, 2), 0.05)])
def test_whole_lattice_structure_equals_per_site(dims, eps):
    """The oracle's restatement in the reference's whole-lattice structure
    (shift copies and lattice temporaries, puregauge.cpp:36-79,
    wilsonflow.cpp:141-172), the CPU baseline's code, gives the per-site
    oracle's RK3 step bit for bit, odd extents and periodic wraps included."""
    links = orc.thermalised(dims, beta=6.0, sweeps=2, seed=3) if all(d % 2 == 0 for d in dims) else orc.hot_start(dims, 3)
    np.testing.assert_array_equal(orc.flow_step_lattice(links, dims, eps), orc.flow_step(links, dims, eps))
