// Wilson plaquette action on the GPU: the drop-in for PureGauge
// (include/Actions/puregauge.h:50-69, src/Actions/puregauge.cpp).
#pragma once
#include "Utils/latticeunits.h"
#include "Actions/action.h"
#include "Math/devicefield.h"

class GpuPureGauge : public Action {
public:
    /// Sets LatticeUnits for beta, as the reference's constructor does (src/Actions/puregauge.cpp:24)
    explicit GpuPureGauge(double beta, int precision = LYM_PREC_FP64) : m_beta(beta), m_precision(precision)
    {
        LatticeUnits::initialize(beta);
    }

    /// -(beta/3) Re Tr[(newLink - U) Omega] with Omega the staple of the link
    /// from the last computeStaples(mu) (src/Actions/puregauge.cpp:28)
    double compute(int x, int y, int z, int t, int mu, SU3 &newLink) override;
    /// PureGauge::computeStaples (src/Actions/puregauge.cpp:36-54), via lym_staples
    void computeStaples(int mu) override;
    /// PureGauge::computeDerivative (src/Actions/puregauge.cpp:62-79), via lym_derivative
    Lattice<SU3> computeDerivative(int mu) override;

    double beta() const { return m_beta; }
    const Lattice<SU3> &staples() const { return m_omega; }
    int stapleDirection() const { return m_omegaMu; }
    DeviceField &device() { return DeviceField::of(m_lat, m_precision); }

private:
    double m_beta;
    int m_precision;
    Lattice<SU3> m_omega;
    int m_omegaMu = -1;
};
