// Scale setting from a Wilson-flow series: rows [t, plaq, topc, energy] as
// GpuWilsonFlow::applyWilsonFlow returns them (rows that were not measured
// hold NaN and are skipped).  The reference computes neither (it writes the
// raw series only); these are the standard gradient-flow scales.
#pragma once
#include <array>
#include <vector>

struct FlowScales {
    double t0;  // t0 / a^2 where t^2 <E> = target (Luscher, JHEP 08 (2010) 071); NaN if not reached
    double w0;  // w0 / a where t d/dt (t^2 <E>) = target (BMW, JHEP 09 (2012) 010); NaN if not reached
};

// <E> a^4 = -energy / (64 V) from the raw clover sum (superobs.cpp:109).
double meanEnergyDensity(double energyRaw, long volume);

// t0: linear interpolation of t^2 <E> between the measured rows around the
// first crossing.  w0: t d/dt (t^2 <E>) by central differences at the
// midpoints of consecutive measured rows, then the same interpolation.
FlowScales flowScales(const std::vector<std::array<double, 4>> &series, long volume, double target = 0.3);
