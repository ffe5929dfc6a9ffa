"""Scale setting and lattice units (latticeyangmills_amd/scale.py and the
host library's Observables/flowscale.h, Utils/latticeunits.h)."""
import math

import numpy as np
import pytest

from latticeyangmills_amd import scale


@pytest.mark.parametrize("beta,r0_over_a", [(5.8, 3.673), (5.95, 4.898), (6.0, 5.368), (6.2, 7.380), (6.4, 9.74),
                                             (6.57, 12.38)])
def test_lattice_spacing_against_published_r0(beta, r0_over_a):
    """a(beta) against independently published r0/a (Necco and Sommer, Nucl.
    Phys. B622 (2002) 328, Table 2; r0 = 0.5 fm), within the 0.5 % the
    parametrisation claims."""
    assert abs(0.5 / scale.lattice_spacing_fm(beta) / r0_over_a - 1) < 0.005


def series_for(t2e_of_t, ts, volume):
    """A flow series whose t^2 <E> is the given function (energy column raw, as lym_flow)."""
    rows = []
    for t in ts:
        e = t2e_of_t(t) / t ** 2 if t > 0 else 0.0
        rows.append([t, 0.9, 0.0, -64.0 * volume * e])
    return np.array(rows)


def test_t0_w0_linear_t2e():
    """t^2 E = c t: t0 = 0.3 / c; t d/dt(t^2 E) = c t, so w0^2 = 0.3 / c."""
    V, c = 4096, 0.06
    s = series_for(lambda t: c * t, np.arange(0, 8.0, 0.02), V)
    assert abs(scale.t0(s, V) - 0.3 / c) < 1e-9
    assert abs(scale.w0(s, V) - math.sqrt(0.3 / c)) < 1e-9


def test_t0_w0_quadratic_t2e_and_skipped_rows():
    """t^2 E = b t^2: t0 = sqrt(0.3 / b); t d/dt = 2 b t^2, so w0 = (0.15 / b)^(1/4).
    Unmeasured (NaN) rows are skipped; unreached scales give None."""
    V, b = 1000, 0.02
    s = series_for(lambda t: b * t * t, np.arange(0, 6.0, 0.02), V)
    s[1::10, 1:] = np.nan
    assert abs(scale.t0(s, V) - math.sqrt(0.3 / b)) < 1e-3
    assert abs(scale.w0(s, V) - (0.15 / b) ** 0.25) < 1e-3
    assert scale.t0(s[:50], V) is None and scale.w0(s[:10], V) is None


def test_energy_normalisation_matches_clover_convention():
    assert scale.e_mean(-64.0 * 10, 10) == 1.0
