"""The host math API of the drop-in headers, used the way the reference's own
code uses it (include/Math/su3.h, include/Math/lattice.h:231-324,
include/Apps/wilsonflow.h:58-60): a C++ caller compiled against the headers
and linked with liblymhost.so, no GPU call.

exp(iQ) (Morningstar-Peardon, src/Math/su3.cpp:96-183) is pinned against
scipy's matrix exponential and against the oracle's restatement of the
reference's exp, on Hermitian traceless matrices that cover both signs of
det Q, the small-w series of xi0, a degenerate spectrum and Q = 0; the
site-wise lattice maps against the same values."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg

from tests import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "latticeyangmills_amd")

CALLER = r'''
#include <cstdio>
#include "Apps/gpuwilsonflow.h"
#include "Math/field.h"
#include "Math/lattice.h"
#include "Math/su3.h"
#include "Observables/gpusuperobs.h"
int main() {
    std::array<int, 4> size{{2, 2, 2, 2}};
    Lattice<SU3> q(size);
    for (int i = 0; i < q.sites; i++)
        for (int k = 0; k < 18; k++)
            if (scanf("%lf", &q.at(i).mat[k]) != 1) return 3;
    Lattice<SU3> e = exp(q);                 // site-wise, lattice.h:231-245 of the reference
    Lattice<double> rt = realTrace(e), it = imagTrace(e);
    SU3 one = exp(q.at(1));                  // SU3 exp(const SU3&), su3.h:57
    for (int i = 0; i < e.sites; i++) {
        for (int k = 0; k < 18; k++) printf("%.17g ", e.at(i).mat[k]);
        printf("%.17g %.17g\n", rt.at(i), it.at(i));
    }
    for (int k = 0; k < 18; k++) printf("%.17g ", one.mat[k]);
    printf("\n");
    GpuWilsonFlow app(1.0, 0.01);            // no lattice yet: no device context
    GpuSuperObs obs;
    app.addObservable(&obs);
    std::array<int, 4> &s = app.getSize();   // wilsonflow.h:58-60 of the reference
    std::vector<double> &v = app.getObsValues();
    std::vector<Observable *> &o = app.getObs();
    printf("%d %zu %zu %s\n", s[0], v.size(), o.size(), o[0]->getName());
    return 0;
}
'''


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    d = tmp_path_factory.mktemp("caller")
    (d / "caller.cpp").write_text(CALLER)
    exe = d / "caller"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(PKG, "host", "include"), str(d / "caller.cpp"), "-o", str(exe),
                    "-L", PKG, "-llymhost", "-llym", f"-Wl,-rpath,{PKG}"], check=True)
    return exe


def hermitian_traceless(rng, scale):
    a = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    h = (a + a.conj().T) / 2
    return scale * (h - np.trace(h) / 3 * np.eye(3))


def cases():
    rng = np.random.default_rng(2024)
    qs = [hermitian_traceless(rng, s) for s in (0.01, 0.3, 1.0, 2.5, 0.7, 4.0)]
    qs.append(-qs[2])                                          # det Q of the other sign
    qs.append(np.diag([0.5, 0.5, -1.0]).astype(complex))       # degenerate: w = 0, xi0 series
    qs.append(np.diag([0.8, -0.3, -0.5]).astype(complex))      # diagonal: exp = diag(e^{i lambda})
    v = np.linalg.qr(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))[0]
    qs.append(v @ np.diag([0.2, 0.2 + 1e-3, -0.4 - 1e-3]) @ v.conj().T)  # small w, rotated
    qs.append(np.zeros((3, 3), complex))                        # Q = 0: the identity
    while len(qs) < 16:
        qs.append(hermitian_traceless(rng, rng.uniform(0.05, 3.0)))
    return qs


def test_host_exp_and_lattice_maps(caller):
    qs = cases()
    q18 = orc.to18(np.array(qs))
    out = subprocess.run([str(caller)], input="\n".join(" ".join("%.17g" % x for x in row) for row in q18),
                         capture_output=True, text=True, check=True).stdout.splitlines()
    vals = np.array([[float(x) for x in line.split()] for line in out[:16]])
    got = orc.from18(vals[:, :18])
    for i, q in enumerate(qs):
        want = scipy.linalg.expm(1j * q)
        assert np.abs(got[i] - want).max() < 1e-14, (i, np.abs(got[i] - want).max())
        det = np.linalg.det(got[i])
        assert abs(det - 1) < 1e-14 and np.abs(got[i] @ got[i].conj().T - np.eye(3)).max() < 1e-14
        if np.abs(q).max() > 0:  # the reference's exp divides 0/0 at Q = 0
            ref, _ = orc.su3_exp(q18[i])
            assert np.abs(vals[i, :18] - ref).max() < 1e-14, i
        assert vals[i, 18] == sum(vals[i, [0, 8, 16]]) and vals[i, 19] == sum(vals[i, [1, 9, 17]])
    np.testing.assert_array_equal(np.array([float(x) for x in out[16].split()]), vals[1, :18])
    np.testing.assert_allclose(got[8], np.diag(np.exp(1j * np.array([0.8, -0.3, -0.5]))), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(got[10], np.eye(3))
    assert out[17].split() == ["0", "0", "1", "SuperObs"]
