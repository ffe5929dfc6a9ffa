// Plaquette, clover energy and topological charge on the GPU: the drop-in
// for SuperObs (include/Observables/superobs.h, src/Observables/superobs.cpp:31-128).
#pragma once
#include "Math/devicefield.h"
#include "Observables/observable.h"

class GpuSuperObs : public Observable {
public:
    explicit GpuSuperObs(int precision = LYM_PREC_FP64) : m_precision(precision) { m_name = "SuperObs"; }  // the reference's name, superobs.cpp:15
    void initObservable(GluonField *field) override { m_field = field; }
    void compute() override;
    /// sum over sites of |q-terms|: the scale of the cancellations in topc
    double topcScale = 0;

private:
    int m_precision;
};
