// LatticeUnits and the gradient-flow scales t0, w0 (see Utils/latticeunits.h,
// Observables/flowscale.h).
#include <cmath>
#include <limits>
#include <numeric>

#include "Observables/flowscale.h"
#include "Utils/latticeunits.h"

int LatticeUnits::Nc = 3;
int LatticeUnits::latticeSites = 0;
double LatticeUnits::beta = 0;
double LatticeUnits::latticeVolume = 0;
double LatticeUnits::latticeSpacing = 0;
std::array<int, 4> LatticeUnits::size{{0, 0, 0, 0}};

void LatticeUnits::initialize(double b, const std::array<int, 4> &sz)
{
    size = sz;
    initialize(b);
}

void LatticeUnits::initialize(double b)
{
    Nc = 3;
    beta = b;
    latticeSpacing = b > 0 ? calculateLatticeSpacing(b) : 0.0;  // no action (beta) yet: unknown
    latticeSites = std::accumulate(size.begin(), size.end(), 1, [](int x, int y) { return x * y; });
    latticeVolume = std::pow(latticeSpacing, 4) * latticeSites;
}

double LatticeUnits::calculateLatticeSpacing(double b)
{
    const double d = b - 6.0;
    return 0.5 * std::exp(-1.6805 - 1.7139 * d + 0.8155 * d * d - 0.6667 * d * d * d);
}

double LatticeUnits::plaquette(double value) { return value; }
double LatticeUnits::energyDensity(double value) { return meanEnergyDensity(value, latticeSites); }
double LatticeUnits::topologicalCharge(double value) { return value; }

double meanEnergyDensity(double energyRaw, long volume) { return -energyRaw / (64.0 * double(volume)); }

namespace {
// first t where y(t) crosses target upwards, linear between the samples
double crossing(const std::vector<double> &t, const std::vector<double> &y, double target)
{
    for (size_t i = 1; i < t.size(); i++)
        if (y[i - 1] < target && target <= y[i])
            return t[i - 1] + (target - y[i - 1]) * (t[i] - t[i - 1]) / (y[i] - y[i - 1]);
    return std::numeric_limits<double>::quiet_NaN();
}
}  // namespace

FlowScales flowScales(const std::vector<std::array<double, 4>> &series, long volume, double target)
{
    std::vector<double> t, y;
    for (const auto &r : series)
        if (std::isfinite(r[1]) && std::isfinite(r[3])) {
            t.push_back(r[0]);
            y.push_back(r[0] * r[0] * meanEnergyDensity(r[3], volume));
        }
    FlowScales s{crossing(t, y, target), std::numeric_limits<double>::quiet_NaN()};
    std::vector<double> tm, w;
    for (size_t i = 1; i < t.size(); i++) {
        const double mid = 0.5 * (t[i] + t[i - 1]);
        tm.push_back(mid);
        w.push_back(mid * (y[i] - y[i - 1]) / (t[i] - t[i - 1]));
    }
    const double w0sq = crossing(tm, w, target);
    if (std::isfinite(w0sq)) s.w0 = std::sqrt(w0sq);
    return s;
}
