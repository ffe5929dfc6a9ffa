"""Gradient-flow scale setting and lattice units from a flow series.

Rows of a series are [t, plaq, topc, energy] as lym_flow returns them (rows
that were not measured hold NaN).  The reference writes the raw series only;
t0 and w0 are the standard gradient-flow scales, and lattice_spacing_fm is the
reference's LatticeUnits::calculateLatticeSpacing (src/Utils/latticeunits.cpp:29-32).
The host library has the same functions in C++ (Observables/flowscale.h,
Utils/latticeunits.h).
"""
import math

import numpy as np


def e_mean(energy_raw, volume):
    """<E> a^4 = -energy / (64 V) from the raw clover sum (superobs.cpp:109)."""
    return -energy_raw / (64.0 * volume)


def _crossing(t, y, target):
    for i in range(1, len(t)):
        if y[i - 1] < target <= y[i]:
            return t[i - 1] + (target - y[i - 1]) * (t[i] - t[i - 1]) / (y[i] - y[i - 1])
    return None


def t2e(series, volume):
    """(t, t^2 <E>) over the measured rows."""
    s = np.asarray(series, dtype=float)
    s = s[np.isfinite(s[:, 1]) & np.isfinite(s[:, 3])]
    return s[:, 0], s[:, 0] ** 2 * e_mean(s[:, 3], volume)


def t0(series, volume, target=0.3):
    """t0 / a^2 where t^2 <E> = target (Luscher, JHEP 08 (2010) 071), linear
    between the measured rows around the first crossing; None if not reached."""
    t, y = t2e(series, volume)
    return _crossing(t, y, target)


def w0(series, volume, target=0.3):
    """w0 / a where t d/dt (t^2 <E>) = target (BMW, JHEP 09 (2012) 010): the
    derivative by central differences at the midpoints of consecutive measured
    rows, then linear interpolation in t; None if not reached."""
    t, y = t2e(series, volume)
    tm = 0.5 * (t[1:] + t[:-1])
    w = tm * np.diff(y) / np.diff(t)
    x = _crossing(tm, w, target)
    return math.sqrt(x) if x is not None else None


def lattice_spacing_fm(beta):
    """a(beta) in fm: ln(a / r0) = -1.6805 - 1.7139 (b - 6) + 0.8155 (b - 6)^2
    - 0.6667 (b - 6)^3, r0 = 0.5 fm (src/Utils/latticeunits.cpp:29-32; the
    parametrisation of Guagnelli, Sommer and Wittig, Nucl. Phys. B535 (1998) 389)."""
    d = beta - 6.0
    return 0.5 * math.exp(-1.6805 - 1.7139 * d + 0.8155 * d * d - 0.6667 * d ** 3)
