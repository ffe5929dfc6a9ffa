// Action interface of the reference (include/Actions/action.h:54-68).
#pragma once
#include "Math/field.h"

class Action {
public:
    virtual ~Action() {}
    /// Action difference for replacing link (x,y,z,t,mu) by newLink
    virtual double compute(int x, int y, int z, int t, int mu, SU3 &newLink) = 0;
    /// Staple sums of all links in direction mu
    virtual void computeStaples(int mu) = 0;
    /// Flow force Z of all links in direction mu
    virtual Lattice<SU3> computeDerivative(int mu) = 0;
    /// Links the action to a gauge field
    void initAction(GluonField *field) { m_lat = field; }

protected:
    GluonField *m_lat = nullptr;
};
