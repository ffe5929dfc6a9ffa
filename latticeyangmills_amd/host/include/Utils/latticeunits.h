// LatticeUnits (include/Utils/latticeunits.h): beta, lattice spacing and
// volume of the run, as static members like the reference's.  initialize()
// is called by GpuPureGauge::initAction, as the reference's PureGauge calls it
// (src/Actions/puregauge.cpp:24).
#pragma once
#include <array>

struct LatticeUnits {
    // Sets beta, the lattice spacing a(beta) in fm and the volume a^4 * sites.
    static void initialize(double beta, const std::array<int, 4> &size);
    static void initialize(double beta);  // keeps the current size

    static double beta;
    static double latticeVolume;   // fm^4
    static double latticeSpacing;  // fm
    static std::array<int, 4> size;
    static int latticeSites;
    static int Nc;

    // The reference declares these three without defining them; here:
    static double plaquette(double value);          // Re Tr P / 3, already normalised: unchanged
    static double energyDensity(double value);      // raw clover sum (superobs.cpp:109) -> <E> a^4
    static double topologicalCharge(double value);  // already normalised: unchanged

    // a(beta) in fm from ln(a / r0) = -1.6805 - 1.7139 (b - 6) + 0.8155 (b - 6)^2
    // - 0.6667 (b - 6)^3 with r0 = 0.5 fm (src/Utils/latticeunits.cpp:29-32;
    // the parametrisation of Guagnelli, Sommer and Wittig, Nucl. Phys. B535
    // (1998) 389, for 5.7 <= beta <= 6.57).
    static double calculateLatticeSpacing(double beta);
};
