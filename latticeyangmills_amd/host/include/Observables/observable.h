// Observable interface of the reference (include/Observables/observable.h:52-66).
#pragma once
#include <string>

#include "Math/field.h"

class Observable {
public:
    virtual ~Observable() {}
    virtual void compute() = 0;
    virtual void initObservable(GluonField *field) = 0;
    const char *getName() { return m_name.c_str(); }
    double value() { return m_value; }

    double plaq = 0, energy = 0, topc = 0;

    /// sets the results from a measurement made elsewhere (a flow series row);
    /// value() is the plaquette
    void setResults(double p, double t, double e)
    {
        plaq = p;
        topc = t;
        energy = e;
        m_value = p;
    }

protected:
    GluonField *m_field = nullptr;
    double m_value = 0;
    std::string m_name;
};
