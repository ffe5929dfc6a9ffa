"""Independent numpy formulation of the lattice quantities, written from the
textbook definitions rather than from the reference's code path, used to pin
the oracle:

  - links as complex (4, Lt, Lz, Ly, Lx, 3, 3) arrays, shifts by np.roll;
  - plaquettes U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+;
  - the clover as the sum of the four counter-clockwise leaves,
    F_{mu nu} = (Q - Q^+)/8 made traceless;
  - the flow force from the Lie derivative of the Wilson action;
  - the Luescher RK3 integrator with scipy.linalg.expm.
"""
import numpy as np
import scipy.linalg as sl

from tests.oracle_lib import from18, to18


def to_field(links, dims):
    Lx, Ly, Lz, Lt = dims
    return from18(links).reshape(4, Lt, Lz, Ly, Lx, 3, 3)


def to_links(U):
    return to18(U.reshape(4, -1, 3, 3))


def ax(d):
    """array axis of direction d in a (Lt, Lz, Ly, Lx, 3, 3) field"""
    return 3 - d


def fwd(A, d):
    """A(x + e_d)"""
    return np.roll(A, -1, axis=ax(d))


def bwd(A, d):
    """A(x - e_d)"""
    return np.roll(A, 1, axis=ax(d))


def dag(A):
    return np.conj(np.swapaxes(A, -1, -2))


def mm(*ms):
    r = ms[0]
    for m in ms[1:]:
        r = r @ m
    return r


def plaquette(U, mu, nu):
    return mm(U[mu], fwd(U[nu], mu), dag(fwd(U[mu], nu)), dag(U[nu]))


def retr(A):
    return np.real(np.trace(A, axis1=-2, axis2=-1))


def plaq_sum(U):
    """sum over sites and planes mu < nu of Re Tr P"""
    return sum(retr(plaquette(U, m, n)).sum() for m in range(4) for n in range(m + 1, 4))


def staple_sum(U, mu):
    """Omega with U_mu(x) Omega(x) = sum of the six plaquettes through the link"""
    om = np.zeros_like(U[0])
    for nu in range(4):
        if nu == mu:
            continue
        om += mm(fwd(U[nu], mu), dag(fwd(U[mu], nu)), dag(U[nu]))
        lower = mm(dag(fwd(U[nu], mu)), dag(U[mu]), U[nu])  # evaluated at x - nu
        om += bwd(lower, nu)
    return om


def traceless(A):
    return A - np.trace(A, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)


def force(U, mu):
    """Z = (i/2)[X - X^+]_traceless, X = U Omega: d/ds sum Re Tr P(e^{isH} U) = Tr(H Z)."""
    X = U[mu] @ staple_sum(U, mu)
    return 0.5j * traceless(X - dag(X))


def clover_F(U, mu, nu):
    """F = (Q - Q^+)/8 traceless, Q = sum of the four counter-clockwise leaves at x"""
    Umu, Unu = U[mu], U[nu]
    q = mm(Umu, fwd(Unu, mu), dag(fwd(Umu, nu)), dag(Unu))
    q = q + mm(Unu, dag(bwd(fwd(Umu, nu), mu)), dag(bwd(Unu, mu)), bwd(Umu, mu))
    q = q + mm(dag(bwd(Umu, mu)), dag(bwd(bwd(Unu, mu), nu)), bwd(bwd(Umu, mu), nu), bwd(Unu, nu))
    q = q + mm(dag(bwd(Unu, nu)), bwd(Umu, nu), fwd(bwd(Unu, nu), mu), dag(Umu))
    return traceless(q - dag(q)) / 8.0


def observables(U):
    """(plaq, energy_raw, topc) in the reference's normalisation:
    plaq = sum Re Tr P / (18 V); energy_raw = sum_{mu<nu} Re Tr C^2 with C = 8F;
    topc = -sum [ReTr(C01 C23) + ReTr(C02 C13^+) + ReTr(C03 C12)] / (256 pi^2)."""
    V = U[0].size // 9
    F = {(m, n): clover_F(U, m, n) for m in range(4) for n in range(m + 1, 4)}
    plaq = plaq_sum(U) / (18.0 * V)
    energy = sum(retr(64.0 * F[k] @ F[k]).sum() for k in F)
    q = retr(64.0 * F[0, 1] @ F[2, 3]) + retr(64.0 * F[0, 2] @ dag(F[1, 3])) + retr(64.0 * F[0, 3] @ F[1, 2])
    return plaq, energy, -q.sum() / (256.0 * np.pi ** 2)


def expi(Q):
    """exp(iQ) per site with scipy"""
    flat = Q.reshape(-1, 3, 3)
    return np.stack([sl.expm(1j * q) for q in flat]).reshape(Q.shape)


def rk3_step(U, eps):
    """Luescher's RK3 (arXiv:1006.4518, App. C) with exp(i eps Z) left-multiplied."""
    W = U.copy()
    Z0 = np.stack([force(W, m) for m in range(4)])
    W = np.stack([expi(eps / 4 * Z0[m]) @ W[m] for m in range(4)])
    Z1 = np.stack([force(W, m) for m in range(4)])
    A = 8.0 / 9.0 * Z1 - 17.0 / 36.0 * Z0
    W = np.stack([expi(eps * A[m]) @ W[m] for m in range(4)])
    Z2 = np.stack([force(W, m) for m in range(4)])
    B = 3.0 / 4.0 * Z2 - A
    return np.stack([expi(eps * B[m]) @ W[m] for m in range(4)])


def haar_su3(rng, n):
    """n Haar-random SU(3) matrices (complex Gaussian QR with phase fix)"""
    Z = (rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))) / np.sqrt(2)
    Q, R = np.linalg.qr(Z)
    d = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(Q)
    return Q / det[:, None, None] ** (1.0 / 3.0)


def gauge_transform(U, G):
    """U_mu(x) -> G(x) U_mu(x) G(x+mu)^+"""
    return np.stack([G @ U[m] @ dag(fwd(G, m)) for m in range(4)])
