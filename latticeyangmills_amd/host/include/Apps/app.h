// App interface of the reference (include/Apps/app.h:59-77).
#pragma once
#include "Utils/latticeunits.h"
#include <array>
#include <vector>

#include "Actions/action.h"
#include "Math/field.h"
#include "Observables/observable.h"

class App {
public:
    virtual ~App() {}
    void setAction(Action *action) { m_act = action; }
    virtual void createLattice(std::array<int, 4> latticeSize)
    {
        m_size = latticeSize;
        m_lat = new GluonField(m_size);
        LatticeUnits::initialize(LatticeUnits::beta, m_size);  // the volume is known from here on
    }
    void addObservable(Observable *observable) { m_obs.push_back(observable); }
    virtual void execute() = 0;
    GluonField *lattice() { return m_lat; }

protected:
    Action *m_act = nullptr;
    GluonField *m_lat = nullptr;
    std::array<int, 4> m_size{{0, 0, 0, 0}};
    std::vector<Observable *> m_obs;
};
