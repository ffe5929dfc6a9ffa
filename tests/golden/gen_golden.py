"""Writes tests/golden/golden_2x4x4x4.npz from the numpy/scipy formulation in
tests/npref.py (independent of the oracle and of the HIP path):

  links      Haar SU(3) links from numpy's PCG64 (seed 20180101), (4, V, 18)
  obs        plaq, energy_raw, topc in the reference's normalisation
  force      Z_mu for mu = 0..3, (4, V, 18)
  step       links after one RK3 step, eps = 0.05, (4, V, 18)
  exp_q/exp  20 Hermitian traceless Q (scale 1e-4 .. 3) and expm(iQ), (20, 18)

Run from the repository root: python -m tests.golden.gen_golden
"""
import os

import numpy as np
import scipy.linalg as sl

from tests import npref as nr
from tests.oracle_lib import to18

DIMS = (2, 4, 4, 4)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_2x4x4x4.npz")


def main():
    rng = np.random.default_rng(20180101)
    V = int(np.prod(DIMS))
    U = nr.haar_su3(rng, 4 * V).reshape((4,) + tuple(reversed(DIMS)) + (3, 3))
    links = nr.to_links(U)
    plaq, energy, topc = nr.observables(U)
    force = np.stack([to18(nr.force(U, mu).reshape(-1, 3, 3)) for mu in range(4)])
    step = nr.to_links(nr.rk3_step(U, 0.05))
    qs, es = [], []
    for k in range(20):
        H = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        H = H + H.conj().T
        H -= np.trace(H) / 3 * np.eye(3)
        H *= 10 ** (-4 + 4.5 * k / 19) / np.linalg.norm(H)
        qs.append(to18(H))
        es.append(to18(sl.expm(1j * H)))
    np.savez(OUT, dims=np.array(DIMS), links=links, obs=np.array([plaq, energy, topc]), force=force, step=step,
             step_eps=np.array(0.05), exp_q=np.array(qs), exp=np.array(es))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
