// extern "C" entry points onto the reference's own SU3 code (include/Math/su3.h,
// src/Math/su3.cpp), compiled from the reference tree by oracle/Makefile into
// oracle/_ref/.  Test infrastructure: used only to pin the oracle's SU(3)
// algebra and exp map bit for bit.
#include <cstring>
#include "Math/su3.h"

static SU3 load(const double *p)
{
    SU3 m;
    std::memcpy(m.mat.data(), p, 18 * sizeof(double));
    return m;
}

static void store(const SU3 &m, double *p) { std::memcpy(p, m.mat.data(), 18 * sizeof(double)); }

extern "C" {
void ref_su3_mul(const double *a, const double *b, double *r)
{
    SU3 A = load(a), B = load(b);
    store(A * B, r);
}
void ref_su3_adj(const double *a, double *r) { store(~load(a), r); }
double ref_su3_retr_mul(const double *a, const double *b)
{
    SU3 A = load(a), B = load(b);
    return getMultSumTrace(A, B);
}
void ref_su3_exp(const double *q, double *r) { store(exp(load(q)), r); }
}
