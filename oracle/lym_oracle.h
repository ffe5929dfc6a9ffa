/*
 * lym_oracle.h -- CPU restatement of LatticeYangMills' hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (liblym.so, the C++ host
 * classes, the Python package) may include, link or call this code.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only as the checker / the timed CPU baseline.
 *
 * Conventions (follow the reference):
 *   SU(3) element = 18 doubles, m[2*(3i+j)] = Re U_ij, m[2*(3i+j)+1] = Im U_ij
 *                                           (include/Math/su3.h:63, src/Math/su3.cpp:30-34)
 *   site index    s = x + Lx*(y + Ly*(z + Lz*t))   (include/Math/lattice.h:87-89)
 *   gauge field   links[mu*V*18 + s*18 + k]       (GluonField = 4 Lattice<SU3>,
 *                                                  include/Math/field.h:52-76)
 */
#ifndef LYM_ORACLE_H
#define LYM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SU(3) algebra (src/Math/su3.cpp, include/Math/su3.h) ---- */
void   or_su3_mul(const double *a, const double *b, double *r);
void   or_su3_adj(const double *a, double *r);
double or_su3_retr_mul(const double *a, const double *b);
int    or_su3_exp(const double *q, double *r);

/* ---- action (src/Actions/puregauge.cpp) ---- */
void or_staples(const double *links, const int dims[4], int mu, double *omega);
void or_derivative(const double *links, const int dims[4], int mu, double *z);

/* ---- flow (src/Apps/wilsonflow.cpp) ---- */
int  or_flow_step(double *links, const int dims[4], double eps);
int  or_flow_rows(double tau, double eps);
/* or_flow_step in the reference's whole-lattice-temporary structure (equal
 * bit for bit; the CPU baseline) */
int  or_flow_step_lattice(double *links, const int dims[4], double eps);
int  or_flow(double *links, const int dims[4], double eps, double tau,
             int measure_every, int sum_mode, double *series);

/* ---- observables (src/Observables/superobs.cpp) ---- */
void or_observables(const double *links, const int dims[4], int sum_mode,
                    double out[4]);

/* ---- random numbers and Metropolis (src/Math/random.cpp,
 *      src/Apps/gaugefieldfactory.cpp) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_rand_su3_transf(const double *u12, double eps, double *x);
void or_rand_su3_exph(const double *u8, double eps, double *x);
void or_hot_start(double *links, const int dims[4], uint64_t seed);
int  or_metropolis(double *links, const int dims[4], double beta, double eps,
                   int nhit, int nsweeps, int sweep0, int candidate,
                   uint64_t seed, const double *predrawn, int64_t *accepted);
int  or_uniforms_per_hit(int candidate);

#ifdef __cplusplus
}
#endif
#endif
