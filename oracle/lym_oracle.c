/*
 * lym_oracle.c -- CPU restatement of LatticeYangMills' hot path, FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels and the timed CPU
 * baseline.  The product never links or calls it.
 *
 * Every function states the reference file:line it restates.  The arithmetic
 * is kept in the reference's operation order (built with -ffp-contract=off)
 * so that, where the reference's own source compiles here (src/Math/su3.cpp),
 * results agree bit for bit.  The MPI halo shift (include/Math/lattice.h:360-594)
 * is a pure copy; on one rank it is the periodic neighbour, which is what the
 * index helpers below compute.
 *
 * Deviations, each limited to inputs where the reference returns NaN:
 *   - or_su3_exp: Q with Tr Q^2 == 0 returns the identity (the reference
 *     divides 0/0 at src/Math/su3.cpp:111-112);
 *   - or_su3_exp: c0/c0max is clamped to 1 (acos of a rounded ratio > 1).
 */
#include "lym_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RE(i, j) (2 * (3 * (i) + (j)))
#define IM(i, j) (2 * (3 * (i) + (j)) + 1)

/* ------------------------------------------------------------------------ */
/* SU(3) algebra                                                            */
/* ------------------------------------------------------------------------ */

/* r = a*b, in the summation order of SU3::operator*= (include/Math/su3.h:143-183):
 * real part  ((((ar0 br0 - ai0 bi0) + ar1 br1) - ai1 bi1) + ar2 br2) - ai2 bi2
 * imag part  ar0 bi0 + ai0 br0 + ar1 bi1 + ai1 br1 + ar2 bi2 + ai2 br2  (left to right) */
void or_su3_mul(const double *a, const double *b, double *r)
{
    double t[18];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            double re = a[RE(i, 0)] * b[RE(0, j)];
            re = re - a[IM(i, 0)] * b[IM(0, j)];
            re = re + a[RE(i, 1)] * b[RE(1, j)];
            re = re - a[IM(i, 1)] * b[IM(1, j)];
            re = re + a[RE(i, 2)] * b[RE(2, j)];
            re = re - a[IM(i, 2)] * b[IM(2, j)];
            double im = a[RE(i, 0)] * b[IM(0, j)];
            im = im + a[IM(i, 0)] * b[RE(0, j)];
            im = im + a[RE(i, 1)] * b[IM(1, j)];
            im = im + a[IM(i, 1)] * b[RE(1, j)];
            im = im + a[RE(i, 2)] * b[IM(2, j)];
            im = im + a[IM(i, 2)] * b[RE(2, j)];
            t[RE(i, j)] = re;
            t[IM(i, j)] = im;
        }
    }
    memcpy(r, t, sizeof t);
}

/* r = a^dagger (src/Math/su3.cpp:28-37) */
void or_su3_adj(const double *a, double *r)
{
    double t[18];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            t[RE(i, j)] = a[RE(j, i)];
            t[IM(i, j)] = -a[IM(j, i)];
        }
    memcpy(r, t, sizeof t);
}

/* Re Tr(a*b) in the order of getMultSumTrace (src/Math/su3.cpp:82-87). */
double or_su3_retr_mul(const double *a, const double *b)
{
    double s = a[RE(0, 0)] * b[RE(0, 0)];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            if (i == 0 && k == 0) {
                s = s - a[IM(0, 0)] * b[IM(0, 0)];
                continue;
            }
            s = s + a[RE(i, k)] * b[RE(k, i)];
            s = s - a[IM(i, k)] * b[IM(k, i)];
        }
    return s;
}

static inline double retr3(const double *m) { return m[0] + m[8] + m[16]; }
static inline double imtr3(const double *m) { return m[1] + m[9] + m[17]; }

static void su3_identity(double *r)
{
    memset(r, 0, 18 * sizeof(double));
    r[0] = r[8] = r[16] = 1.0;
}

/* exp(iQ) for Hermitian traceless Q by the Cayley-Hamilton form of
 * Morningstar & Peardon (hep-lat/0311018); restates exp(const SU3&)
 * (src/Math/su3.cpp:96-183) in the same operation order.
 * Returns 0 normally, 1 when the Q == 0 policy applied. */
int or_su3_exp(const double *q, double *r)
{
    double q2[18], q3[18];
    or_su3_mul(q, q, q2);
    or_su3_mul(q2, q, q3);
    double c0 = retr3(q3) / 3.0;
    double c1 = retr3(q2) / 2.0;
    if (!(c1 > 0.0)) { /* reference: 0/0 -> NaN */
        su3_identity(r);
        return 1;
    }
    int c0neg = 0;
    if (c0 < 0) {
        c0neg = 1;
        c0 = -c0;
    }
    double c1sqrt = sqrt(c1);
    double inv_sqrt3 = 1.0 / sqrt(3);
    double c0max = c1 * c1sqrt * inv_sqrt3 * 2.0 / 3.0;
    double ratio = c0 / c0max;
    if (ratio > 1.0) ratio = 1.0; /* reference: acos(>1) -> NaN */
    double theta = acos(ratio);
    double w = c1sqrt * sin(theta / 3.0);
    double u = c1sqrt * inv_sqrt3 * cos(theta / 3.0);
    double u2 = u * u;
    double w2 = w * w;
    double xi;
    if (fabs(w) < 0.05)
        xi = 1.0 - w2 / 6.0 * (1 - w2 / 20.0 * (1 - w2 / 42.0));
    else
        xi = sin(w) / w;
    double cmu = cos(-u);
    double smu = sin(-u);
    double c2u = cmu * cmu - smu * smu;
    double s2u = -2.0 * cmu * smu;
    double cw = cos(w);
    double cwcmu = cw * cmu;
    double cwsmu = cw * smu;
    double xicmu = xi * cmu;
    double xismu = xi * smu;

    double h0r = (u2 - w2) * c2u + 8.0 * u2 * cwcmu - 2.0 * u * (3.0 * u2 + w2) * xismu;
    double h0i = (u2 - w2) * s2u + 8.0 * u2 * cwsmu + 2.0 * u * (3.0 * u2 + w2) * xicmu;
    double h1r = 2 * u * c2u - 2 * u * cwcmu - (3.0 * u2 - w2) * xismu;
    double h1i = 2 * u * s2u - 2 * u * cwsmu + (3.0 * u2 - w2) * xicmu;
    double h2r = c2u - cwcmu + 3 * u * xismu;
    double h2i = s2u - cwsmu - 3 * u * xicmu;
    if (c0neg) { /* f_j(-c0) = (-1)^j conj f_j(c0) */
        h0i *= -1;
        h1r *= -1;
        h2i *= -1;
    }
    double norm = 1.0 / (9.0 * u2 - w2);
    h0r *= norm;
    h0i *= norm;
    h1r *= norm;
    h1i *= norm;
    h2r *= norm;
    h2i *= norm;

    double t[18];
    for (int i = 0; i < 18; i += 2) {
        t[i] = h1r * q[i] - h1i * q[i + 1];
        t[i] += h2r * q2[i] - h2i * q2[i + 1];
        if (i % 8 == 0) t[i] += h0r;
        t[i + 1] = h1r * q[i + 1] + h1i * q[i];
        t[i + 1] += h2r * q2[i + 1] + h2i * q2[i];
        if (i % 8 == 0) t[i + 1] += h0i;
    }
    memcpy(r, t, sizeof t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* geometry: periodic neighbour = one-rank semantics of shift()             */
/* (include/Math/lattice.h:360-594)                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int L[4];
    long V;
    long stride[4];
} geom;

static geom make_geom(const int dims[4])
{
    geom g;
    g.V = 1;
    for (int d = 0; d < 4; d++) {
        g.L[d] = dims[d];
        g.stride[d] = g.V;
        g.V *= dims[d];
    }
    return g;
}

/* site index of s + step * e_dir, periodic */
static inline long nbr(const geom *g, long s, int dir, int step)
{
    long c = (s / g->stride[dir]) % g->L[dir];
    long cn = c + step;
    if (cn < 0) cn += g->L[dir];
    if (cn >= g->L[dir]) cn -= g->L[dir];
    return s + (cn - c) * g->stride[dir];
}

static inline const double *link(const double *links, const geom *g, int mu, long s)
{
    return links + ((long)mu * g->V + s) * 18;
}

/* ------------------------------------------------------------------------ */
/* action                                                                   */
/* ------------------------------------------------------------------------ */

/* Staple sum at one link (src/Actions/puregauge.cpp:36-54):
 *   omega = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
 *                          + (U_nu(y+mu)^+ U_mu(y)^+) U_nu(y) ]_{y = x-nu}
 * accumulated in the reference's order (upper then lower, nu ascending). */
static void staple_at(const double *links, const geom *g, int mu, long s, double *omega)
{
    double a[18], b[18], t[18], t2[18];
    memset(omega, 0, 18 * sizeof(double));
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        /* upper: (U_nu(x+mu) * U_mu(x+nu)^+) * U_nu(x)^+ */
        or_su3_adj(link(links, g, mu, nbr(g, s, nu, 1)), a);
        or_su3_mul(link(links, g, nu, nbr(g, s, mu, 1)), a, t);
        or_su3_adj(link(links, g, nu, s), b);
        or_su3_mul(t, b, t2);
        for (int k = 0; k < 18; k++) omega[k] += t2[k];
        /* lower, evaluated at y = x - nu */
        long y = nbr(g, s, nu, -1);
        or_su3_adj(link(links, g, nu, nbr(g, y, mu, 1)), a);
        or_su3_adj(link(links, g, mu, y), b);
        or_su3_mul(a, b, t);
        or_su3_mul(t, link(links, g, nu, y), t2);
        for (int k = 0; k < 18; k++) omega[k] += t2[k];
    }
}

void or_staples(const double *links, const int dims[4], int mu, double *omega)
{
    geom g = make_geom(dims);
    for (long s = 0; s < g.V; s++) staple_at(links, &g, mu, s, omega + s * 18);
}

/* Force Z for one link (src/Actions/puregauge.cpp:62-79):
 *   M = omega^+ U^+ ; D = M^+ - M ; subtract Im Tr D / 3 from the diagonal
 *   imaginary parts ; Z = (i/2) D.  Z is Hermitian and traceless. */
static void derivative_from_staple(const double *u, const double *omega, double *z)
{
    double oa[18], ua[18], m[18], ma[18], d[18];
    or_su3_adj(omega, oa);
    or_su3_adj(u, ua);
    or_su3_mul(oa, ua, m);
    or_su3_adj(m, ma);
    for (int k = 0; k < 18; k++) d[k] = ma[k] - m[k];
    double tr = imtr3(d) / 3.0;
    for (int k = 1; k < 18; k += 8) d[k] -= tr;
    for (int k = 0; k < 18; k += 2) {
        double tmp = d[k];
        z[k] = -d[k + 1] * 0.5;
        z[k + 1] = tmp * 0.5;
    }
}

void or_derivative(const double *links, const int dims[4], int mu, double *z)
{
    geom g = make_geom(dims);
    double omega[18];
    for (long s = 0; s < g.V; s++) {
        staple_at(links, &g, mu, s, omega);
        derivative_from_staple(link(links, &g, mu, s), omega, z + s * 18);
    }
}

/* ------------------------------------------------------------------------ */
/* Wilson flow, Luescher RK3 (src/Apps/wilsonflow.cpp:141-172)              */
/* ------------------------------------------------------------------------ */

/* U <- exp(i * Q) * U for every link of direction mu, Q = q_of(Z) */
static void exp_update(double *links, const geom *g, int mu, const double *z,
                       double scale1, double scale2)
{
    double q[18], e[18];
    for (long s = 0; s < g->V; s++) {
        const double *zs = z + s * 18;
        for (int k = 0; k < 18; k++) q[k] = zs[k] * scale1 * scale2;
        or_su3_exp(q, e);
        double *u = links + ((long)mu * g->V + s) * 18;
        or_su3_mul(e, u, u);
    }
}

int or_flow_step(double *links, const int dims[4], double eps)
{
    geom g = make_geom(dims);
    long n = g.V * 18;
    double *Z = (double *)malloc(sizeof(double) * 4 * n);
    double *D = (double *)malloc(sizeof(double) * n);
    if (!Z || !D) {
        free(Z);
        free(D);
        return -1;
    }
    /* Z0 = Z(W0); W1 = exp(i eps/4 Z0) W0 */
    for (int mu = 0; mu < 4; mu++) or_derivative(links, dims, mu, Z + mu * n);
    for (int mu = 0; mu < 4; mu++) exp_update(links, &g, mu, Z + mu * n, eps, 1.0 / 4.0);
    /* Z <- 8/9 Z(W1) - 17/36 Z ; W2 = exp(i eps Z) W1 */
    for (int mu = 0; mu < 4; mu++) {
        or_derivative(links, dims, mu, D);
        double *z = Z + mu * n;
        for (long k = 0; k < n; k++) z[k] = D[k] * (8.0 / 9.0) - z[k] * (17.0 / 36.0);
    }
    for (int mu = 0; mu < 4; mu++) exp_update(links, &g, mu, Z + mu * n, eps, 1.0);
    /* Z <- 3/4 Z(W2) - Z ; W3 = exp(i eps Z) W2 */
    for (int mu = 0; mu < 4; mu++) {
        or_derivative(links, dims, mu, D);
        double *z = Z + mu * n;
        for (long k = 0; k < n; k++) z[k] = D[k] * (3.0 / 4.0) - z[k];
    }
    for (int mu = 0; mu < 4; mu++) exp_update(links, &g, mu, Z + mu * n, eps, 1.0);
    free(Z);
    free(D);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The same flow step in the reference's whole-lattice structure: every      */
/* operation of computeStaples / computeDerivative / flowStep               */
/* (src/Actions/puregauge.cpp:36-79, src/Apps/wilsonflow.cpp:141-172) is a   */
/* pass over a whole-lattice temporary, the MPI halo shift                   */
/* (include/Math/lattice.h:360-594) a periodic copy.  Per site it does the   */
/* arithmetic of or_flow_step in the same order, so the results are equal    */
/* bit for bit; it exists to time the reference's memory structure as the    */
/* CPU baseline.                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    geom g;
    long n; /* V * 18 */
    double *shiftmu, *tmp, *tmp2, *omega, *m, *ma;
} lat_work;

/* out(x) = in(x + step e_dir) */
static void lat_shift(const double *in, const geom *g, int dir, int step, double *out)
{
    for (long s = 0; s < g->V; s++) memcpy(out + s * 18, in + nbr(g, s, dir, step) * 18, 18 * sizeof(double));
}
static void lat_adj(const double *in, long V, double *out)
{
    for (long s = 0; s < V; s++) or_su3_adj(in + s * 18, out + s * 18);
}
/* a(x) = a(x) b(x) */
static void lat_mul_into(double *a, const double *b, long V)
{
    double r[18];
    for (long s = 0; s < V; s++) {
        or_su3_mul(a + s * 18, b + s * 18, r);
        memcpy(a + s * 18, r, sizeof r);
    }
}
static void lat_add_into(double *a, const double *b, long n)
{
    for (long k = 0; k < n; k++) a[k] += b[k];
}
static void lat_scale(const double *in, double c, long n, double *out)
{
    for (long k = 0; k < n; k++) out[k] = in[k] * c;
}

/* computeStaples(mu) into w->omega (puregauge.cpp:36-54) */
static void lat_staples(const double *links, lat_work *w, int mu)
{
    const geom *g = &w->g;
    const double *U_mu = links + (long)mu * w->n;
    memset(w->omega, 0, sizeof(double) * w->n);
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        const double *U_nu = links + (long)nu * w->n;
        lat_shift(U_nu, g, mu, 1, w->shiftmu);          /* shiftmu = shift(U_nu, mu, +1) */
        memcpy(w->tmp, w->shiftmu, sizeof(double) * w->n);
        lat_shift(U_mu, g, nu, 1, w->tmp2);              /* tempStaple *= adj(shift(U_mu, nu, +1)) */
        lat_adj(w->tmp2, g->V, w->m);
        lat_mul_into(w->tmp, w->m, g->V);
        lat_adj(U_nu, g->V, w->m);                       /* tempStaple *= adj(U_nu) */
        lat_mul_into(w->tmp, w->m, g->V);
        lat_add_into(w->omega, w->tmp, w->n);
        lat_adj(w->shiftmu, g->V, w->tmp);               /* tempStaple = adj(shiftmu) */
        lat_adj(U_mu, g->V, w->m);                       /* tempStaple *= adj(U_mu) */
        lat_mul_into(w->tmp, w->m, g->V);
        memcpy(w->m, U_nu, sizeof(double) * w->n);       /* tempStaple *= U_nu */
        lat_mul_into(w->tmp, w->m, g->V);
        lat_shift(w->tmp, g, nu, -1, w->tmp2);           /* omega += shift(tempStaple, nu, -1) */
        lat_add_into(w->omega, w->tmp2, w->n);
    }
}

/* computeDerivative(mu) into z (puregauge.cpp:62-79) */
static void lat_derivative(const double *links, lat_work *w, int mu, double *z)
{
    const long V = w->g.V;
    lat_staples(links, w, mu);
    lat_adj(w->omega, V, w->m);                                /* m = adj(omega) * adj(U_mu) */
    lat_adj(links + (long)mu * w->n, V, w->tmp);
    lat_mul_into(w->m, w->tmp, V);
    lat_adj(w->m, V, w->ma);                                   /* m = adj(m) - m */
    for (long k = 0; k < w->n; k++) w->ma[k] -= w->m[k];
    for (long s = 0; s < V; s++) {
        double *d = w->ma + s * 18, *zs = z + s * 18;
        double tr = imtr3(d) / 3.0;
        for (int k = 1; k < 18; k += 8) d[k] -= tr;
        for (int k = 0; k < 18; k += 2) {
            double tmp = d[k];
            zs[k] = -d[k + 1] * 0.5;
            zs[k + 1] = tmp * 0.5;
        }
    }
}

/* U_mu = exp(Z_mu * c1 * c2) * U_mu over the whole lattice */
static void lat_exp_update(double *links, lat_work *w, int mu, const double *z, double c1, double c2)
{
    const long V = w->g.V;
    lat_scale(z, c1, w->n, w->tmp);
    lat_scale(w->tmp, c2, w->n, w->tmp2);
    for (long s = 0; s < V; s++) or_su3_exp(w->tmp2 + s * 18, w->m + s * 18);
    double *u = links + (long)mu * w->n;
    lat_mul_into(w->m, u, V);
    memcpy(u, w->m, sizeof(double) * w->n);
}

int or_flow_step_lattice(double *links, const int dims[4], double eps)
{
    lat_work w;
    w.g = make_geom(dims);
    w.n = w.g.V * 18;
    double *Z = (double *)malloc(sizeof(double) * 4 * w.n), *D = (double *)malloc(sizeof(double) * w.n);
    double **bufs[6] = {&w.shiftmu, &w.tmp, &w.tmp2, &w.omega, &w.m, &w.ma};
    int ok = Z && D;
    for (int i = 0; i < 6; i++) {
        *bufs[i] = ok ? (double *)malloc(sizeof(double) * w.n) : NULL;
        ok = ok && *bufs[i];
    }
    if (ok) {
        for (int mu = 0; mu < 4; mu++) lat_derivative(links, &w, mu, Z + mu * w.n);
        for (int mu = 0; mu < 4; mu++) lat_exp_update(links, &w, mu, Z + mu * w.n, eps, 1.0 / 4.0);
        for (int mu = 0; mu < 4; mu++) {
            double *z = Z + mu * w.n;
            lat_derivative(links, &w, mu, D);
            for (long k = 0; k < w.n; k++) z[k] = D[k] * (8.0 / 9.0) - z[k] * (17.0 / 36.0);
        }
        for (int mu = 0; mu < 4; mu++) lat_exp_update(links, &w, mu, Z + mu * w.n, eps, 1.0);
        for (int mu = 0; mu < 4; mu++) {
            double *z = Z + mu * w.n;
            lat_derivative(links, &w, mu, D);
            for (long k = 0; k < w.n; k++) z[k] = D[k] * (3.0 / 4.0) - z[k];
        }
        for (int mu = 0; mu < 4; mu++) lat_exp_update(links, &w, mu, Z + mu * w.n, eps, 1.0);
    }
    for (int i = 0; i < 6; i++) free(*bufs[i]);
    free(Z);
    free(D);
    return ok ? 0 : -1;
}

/* Number of series rows, int(tau/eps) (src/Apps/wilsonflow.cpp:117);
 * the flow runs rows-1 steps (wilsonflow.cpp:124). */
int or_flow_rows(double tau, double eps) { return (int)(tau / eps); }

/* Flow loop (src/Apps/wilsonflow.cpp:115-136).  series has rows*4 entries,
 * columns [t, plaq, topc, energy]; t accumulates by += eps.  With
 * measure_every = k only rows with t % k == 0 are measured (others NaN);
 * the reference measures every row (k = 1). */
int or_flow(double *links, const int dims[4], double eps, double tau,
            int measure_every, int sum_mode, double *series)
{
    int rows = or_flow_rows(tau, eps);
    if (rows < 1 || measure_every < 1) return -1;
    double obs[4];
    or_observables(links, dims, sum_mode, obs);
    series[0] = 0.0;
    series[1] = obs[0];
    series[2] = obs[2];
    series[3] = obs[1];
    for (int t = 1; t < rows; t++) {
        if (or_flow_step(links, dims, eps)) return -1;
        series[4 * t] = series[4 * (t - 1)] + eps;
        if (t % measure_every == 0) {
            or_observables(links, dims, sum_mode, obs);
            series[4 * t + 1] = obs[0];
            series[4 * t + 2] = obs[2];
            series[4 * t + 3] = obs[1];
        } else {
            series[4 * t + 1] = series[4 * t + 2] = series[4 * t + 3] = NAN;
        }
    }
    return rows;
}

/* ------------------------------------------------------------------------ */
/* observables (src/Observables/superobs.cpp:31-128)                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    double s, c;
    int comp;
} acc_t;

static inline void acc_add(acc_t *a, double x)
{
    if (!a->comp) {
        a->s += x;
        return;
    }
    double t = a->s + x; /* Neumaier's improved Kahan-Babuska step */
    if (fabs(a->s) >= fabs(x))
        a->c += (a->s - t) + x;
    else
        a->c += (x - t) + a->s;
    a->s = t;
}
static inline double acc_val(const acc_t *a) { return a->s + a->c; }

/* sum_mode 0: serial site-order sums as Lattice sum() (include/Math/lattice.h:264-296);
 * sum_mode 1: the same terms summed with Neumaier compensation.
 * out = { plaq, energy (raw, <= 0), topc, sum over sites and the three terms of
 *         |Re Tr(C C')| / (256 pi^2) }  -- the last is the scale of topc's cancellations */
void or_observables(const double *links, const int dims[4], int sum_mode, double out[4])
{
    geom g = make_geom(dims);
    long V = g.V;
    double *clov = (double *)malloc(sizeof(double) * 6 * V * 18);
    int plane_of[4][4];
    double plaqSum[3] = {0, 0, 0};
    double energy = 0.0;
    int p = 0;
    for (int mu = 0; mu < 4; mu++) {
        for (int nu = mu + 1; nu < 4; nu++) {
            plane_of[mu][nu] = p;
            double *C = clov + (long)p * V * 18;
            acc_t pr[3] = {{0, 0, sum_mode}, {0, 0, sum_mode}, {0, 0, sum_mode}};
            double t1[18], t2[18], a1[18], a2[18];
            /* leaf P_{mu nu}(x) = ((U_mu U_nu(x+mu)) U_mu(x+nu)^+) U_nu^+ (superobs.cpp:47-50) */
            for (long s = 0; s < V; s++) {
                double *c = C + s * 18;
                or_su3_mul(link(links, &g, mu, s), link(links, &g, nu, nbr(&g, s, mu, 1)), t1);
                or_su3_adj(link(links, &g, mu, nbr(&g, s, nu, 1)), a1);
                or_su3_mul(t1, a1, t2);
                or_su3_adj(link(links, &g, nu, s), a2);
                or_su3_mul(t2, a2, c);
                acc_add(&pr[0], c[0]);
                acc_add(&pr[1], c[8]);
                acc_add(&pr[2], c[16]);
            }
            plaqSum[0] += acc_val(&pr[0]); /* superobs.cpp:71 */
            plaqSum[1] += acc_val(&pr[1]);
            plaqSum[2] += acc_val(&pr[2]);
            for (long s = 0; s < V; s++) {
                double *c = C + s * 18;
                double lm[18], lp[18], lq[18];
                /* tempmm at y = x-mu-nu: ((U_mu(y+nu)^+ U_nu(y)^+) U_mu(y)) U_nu(y+mu) */
                long y = nbr(&g, nbr(&g, s, mu, -1), nu, -1);
                or_su3_adj(link(links, &g, mu, nbr(&g, y, nu, 1)), a1);
                or_su3_adj(link(links, &g, nu, y), a2);
                or_su3_mul(a1, a2, t1);
                or_su3_mul(t1, link(links, &g, mu, y), t2);
                or_su3_mul(t2, link(links, &g, nu, nbr(&g, y, mu, 1)), lm);
                /* tempmp at y = x-nu: ((U_mu(y+nu) U_nu(y+mu)^+) U_mu(y)^+) U_nu(y) */
                y = nbr(&g, s, nu, -1);
                or_su3_adj(link(links, &g, nu, nbr(&g, y, mu, 1)), a1);
                or_su3_mul(link(links, &g, mu, nbr(&g, y, nu, 1)), a1, t1);
                or_su3_adj(link(links, &g, mu, y), a2);
                or_su3_mul(t1, a2, t2);
                or_su3_mul(t2, link(links, &g, nu, y), lp);
                /* temppm at y = x-mu: ((U_mu(y)^+ U_nu(y)) U_mu(y+nu)) U_nu(y+mu)^+ */
                y = nbr(&g, s, mu, -1);
                or_su3_adj(link(links, &g, mu, y), a1);
                or_su3_mul(a1, link(links, &g, nu, y), t1);
                or_su3_mul(t1, link(links, &g, mu, nbr(&g, y, nu, 1)), t2);
                or_su3_adj(link(links, &g, nu, nbr(&g, y, mu, 1)), a2);
                or_su3_mul(t2, a2, lq);
                for (int k = 0; k < 18; k++) c[k] += lm[k]; /* superobs.cpp:72-74 */
                for (int k = 0; k < 18; k++) c[k] -= lp[k];
                for (int k = 0; k < 18; k++) c[k] -= lq[k];
                /* C = clover - clover^+, traceless (superobs.cpp:101-106) */
                double ca[18];
                or_su3_adj(c, ca);
                for (int k = 0; k < 18; k++) c[k] -= ca[k];
                double tr = imtr3(c) / 3.0;
                for (int k = 1; k < 18; k += 8) c[k] -= tr;
            }
            /* energy += sum_x Re Tr(C C) (superobs.cpp:109) */
            acc_t e = {0, 0, sum_mode};
            double cc[18];
            for (long s = 0; s < V; s++) {
                or_su3_mul(C + s * 18, C + s * 18, cc);
                acc_add(&e, retr3(cc));
            }
            energy += acc_val(&e);
            p++;
        }
    }
    /* topological charge (superobs.cpp:113-126) */
    double topc = 0.0;
    acc_t qa = {0, 0, 1};
    int mu = 0;
    for (int nu = 1; nu < 4; nu++) {
        int rho = nu % 3;
        rho++;
        int sig = rho % 3;
        sig++;
        const double *A = clov + (long)plane_of[mu][nu] * V * 18;
        const double *B;
        int adjB = 0;
        if (rho < sig)
            B = clov + (long)plane_of[rho][sig] * V * 18;
        else {
            B = clov + (long)plane_of[sig][rho] * V * 18;
            adjB = 1;
        }
        acc_t q = {0, 0, sum_mode};
        double bb[18], ab[18];
        for (long s = 0; s < V; s++) {
            const double *b = B + s * 18;
            if (adjB) {
                or_su3_adj(b, bb);
                b = bb;
            }
            or_su3_mul(A + s * 18, b, ab);
            double tr = retr3(ab);
            acc_add(&q, tr);
            acc_add(&qa, fabs(tr));
        }
        topc -= acc_val(&q);
    }
    double plaq = (plaqSum[0] + plaqSum[1] + plaqSum[2]) / 18.0 / g.L[0] / g.L[1] / g.L[2] / g.L[3];
    double qnorm = (16 * 16 * M_PI * M_PI);
    topc /= qnorm;
    out[0] = plaq;
    out[1] = energy;
    out[2] = topc;
    out[3] = acc_val(&qa) / qnorm;
    free(clov);
}

/* ------------------------------------------------------------------------ */
/* random numbers                                                           */
/* ------------------------------------------------------------------------ */

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static inline double u32_to_unit(uint32_t r) { return (double)r * (1.0 / 4294967296.0); }

/* One SU(2) subgroup element for randSU3Transf (src/Math/random.cpp:81-102):
 * x0 = +-sqrt(1-eps^2) by the sign of the first draw, (x1,x2,x3) rescaled to norm eps. */
static void su2_near_identity(const double *u4, double eps, double x[4])
{
    double norm = 0;
    x[0] = u4[0] - 0.5;
    for (int i = 1; i < 4; i++) {
        x[i] = u4[i] - 0.5;
        norm += x[i] * x[i];
    }
    norm = sqrt(norm);
    if (x[0] < 0)
        x[0] = -sqrt(1 - eps * eps);
    else
        x[0] = sqrt(1 - eps * eps);
    for (int i = 1; i < 4; i++) x[i] = x[i] * eps / norm;
}

/* X = R*S*T from 12 uniforms (src/Math/random.cpp:71-156): SU(2) blocks
 * [[x0+i x3, x2+i x1], [-x2+i x1, x0-i x3]] embedded in rows/cols (0,1), (0,2), (1,2). */
void or_rand_su3_transf(const double *u12, double eps, double *X)
{
    double R[18], S[18], T[18], x[4], RS[18];
    su3_identity(R);
    su3_identity(S);
    su3_identity(T);
    su2_near_identity(u12, eps, x);
    R[0] = x[0]; R[8] = x[0]; R[1] = x[3]; R[9] = -x[3];
    R[2] = x[2]; R[6] = -x[2]; R[3] = x[1]; R[7] = x[1];
    su2_near_identity(u12 + 4, eps, x);
    S[0] = x[0]; S[16] = x[0]; S[1] = x[3]; S[17] = -x[3];
    S[4] = x[2]; S[12] = -x[2]; S[5] = x[1]; S[13] = x[1];
    su2_near_identity(u12 + 8, eps, x);
    T[8] = x[0]; T[16] = x[0]; T[9] = x[3]; T[17] = -x[3];
    T[10] = x[2]; T[14] = -x[2]; T[11] = x[1]; T[15] = x[1];
    or_su3_mul(R, S, RS);
    or_su3_mul(RS, T, X);
}

/* X = exp(i eps H), H = sum_a h_a lambda_a (Gell-Mann), h_a = 2u_a - 1.
 * H -> -H is equally likely, so the proposal is symmetric. */
void or_rand_su3_exph(const double *u8, double eps, double *X)
{
    double h[8], Q[18];
    for (int a = 0; a < 8; a++) h[a] = 2.0 * u8[a] - 1.0;
    const double r3 = 0.57735026918962576451; /* 1/sqrt(3) */
    Q[RE(0, 0)] = eps * (h[2] + h[7] * r3);
    Q[RE(1, 1)] = eps * (-h[2] + h[7] * r3);
    Q[RE(2, 2)] = eps * (-2.0 * h[7] * r3);
    Q[IM(0, 0)] = Q[IM(1, 1)] = Q[IM(2, 2)] = 0.0;
    Q[RE(0, 1)] = eps * h[0];
    Q[IM(0, 1)] = -eps * h[1];
    Q[RE(0, 2)] = eps * h[3];
    Q[IM(0, 2)] = -eps * h[4];
    Q[RE(1, 2)] = eps * h[5];
    Q[IM(1, 2)] = -eps * h[6];
    Q[RE(1, 0)] = Q[RE(0, 1)];
    Q[IM(1, 0)] = -Q[IM(0, 1)];
    Q[RE(2, 0)] = Q[RE(0, 2)];
    Q[IM(2, 0)] = -Q[IM(0, 2)];
    Q[RE(2, 1)] = Q[RE(1, 2)];
    Q[IM(2, 1)] = -Q[IM(1, 2)];
    or_su3_exp(Q, X);
}

int or_uniforms_per_hit(int candidate) { return candidate == 0 ? 13 : 9; }

/* Haar-random SU(3): two complex Gaussian rows (Box-Muller on Philox
 * uniforms), Gram-Schmidt, third row = conj(row0 x row1) so det = 1.
 * Counter {link, 0xFFFFFFFF, 0, blk}, key {seed_lo, seed_hi ^ 0x48414152}. */
void or_hot_start(double *links, const int dims[4], uint64_t seed)
{
    geom g = make_geom(dims);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ 0x48414152u};
    for (long l = 0; l < 4 * g.V; l++) {
        uint32_t r[12];
        for (uint32_t b = 0; b < 3; b++) {
            uint32_t ctr[4] = {(uint32_t)l, 0xFFFFFFFFu, 0u, b};
            or_philox4x32_10(ctr, key, r + 4 * b);
        }
        double gs[12];
        for (int k = 0; k < 6; k++) {
            double u1 = ((double)r[2 * k] + 1.0) * (1.0 / 4294967296.0);
            double u2 = u32_to_unit(r[2 * k + 1]);
            double rad = sqrt(-2.0 * log(u1));
            gs[2 * k] = rad * cos(6.283185307179586477 * u2);
            gs[2 * k + 1] = rad * sin(6.283185307179586477 * u2);
        }
        double ar[3], ai[3], br[3], bi[3];
        for (int j = 0; j < 3; j++) {
            ar[j] = gs[2 * j];
            ai[j] = gs[2 * j + 1];
            br[j] = gs[6 + 2 * j];
            bi[j] = gs[6 + 2 * j + 1];
        }
        double na = sqrt(ar[0] * ar[0] + ai[0] * ai[0] + ar[1] * ar[1] + ai[1] * ai[1] + ar[2] * ar[2] + ai[2] * ai[2]);
        for (int j = 0; j < 3; j++) {
            ar[j] /= na;
            ai[j] /= na;
        }
        double pr = 0, pi = 0; /* <a|b> = sum conj(a_j) b_j */
        for (int j = 0; j < 3; j++) {
            pr += ar[j] * br[j] + ai[j] * bi[j];
            pi += ar[j] * bi[j] - ai[j] * br[j];
        }
        for (int j = 0; j < 3; j++) {
            br[j] -= pr * ar[j] - pi * ai[j];
            bi[j] -= pr * ai[j] + pi * ar[j];
        }
        double nb = sqrt(br[0] * br[0] + bi[0] * bi[0] + br[1] * br[1] + bi[1] * bi[1] + br[2] * br[2] + bi[2] * bi[2]);
        for (int j = 0; j < 3; j++) {
            br[j] /= nb;
            bi[j] /= nb;
        }
        double *u = links + l * 18;
        for (int j = 0; j < 3; j++) {
            u[RE(0, j)] = ar[j];
            u[IM(0, j)] = ai[j];
            u[RE(1, j)] = br[j];
            u[IM(1, j)] = bi[j];
        }
        for (int j = 0; j < 3; j++) {
            int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            /* (a x b)_j = a_j1 b_j2 - a_j2 b_j1 ; row2 = conj */
            double cr = (ar[j1] * br[j2] - ai[j1] * bi[j2]) - (ar[j2] * br[j1] - ai[j2] * bi[j1]);
            double ci = (ar[j1] * bi[j2] + ai[j1] * br[j2]) - (ar[j2] * bi[j1] + ai[j2] * br[j1]);
            u[RE(2, j)] = cr;
            u[IM(2, j)] = -ci;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* checkerboard Metropolis                                                  */
/* ------------------------------------------------------------------------ */

/* Per-link update of GaugeFieldFactory::updateLink (src/Apps/gaugefieldfactory.cpp:205-220)
 * with the intended Delta S of PureGauge::compute (src/Actions/puregauge.cpp:28):
 *   U' = X U ; dS = -(beta/3) Re Tr[(U' - U) omega] ; accept iff exp(-dS) > u.
 * The sweep is a checkerboard: sub-sweeps (mu = 0..3) x (parity 0, 1), all
 * links of one (mu, parity) are independent.  Uniforms per hit come either
 * from predrawn[((sweep*4V + link)*nhit + hit)*nu + k] (test mode) or from
 * Philox with counter {link, sweep0+sweep, hit, k/4}, key {seed_lo, seed_hi}.
 * candidate 0: R*S*T of src/Math/random.cpp:71-156 (12 uniforms + 1);
 * candidate 1: exp(i eps H) (8 uniforms + 1).  Returns 0, or -1 for odd extents. */
int or_metropolis(double *links, const int dims[4], double beta, double eps,
                  int nhit, int nsweeps, int sweep0, int candidate,
                  uint64_t seed, const double *predrawn, int64_t *accepted)
{
    for (int d = 0; d < 4; d++)
        if (dims[d] % 2) return -1;
    geom g = make_geom(dims);
    long V = g.V;
    int nu_per_hit = or_uniforms_per_hit(candidate);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t acc = 0;
    for (int sw = 0; sw < nsweeps; sw++) {
        for (int mu = 0; mu < 4; mu++) {
            for (int par = 0; par < 2; par++) {
                for (long s = 0; s < V; s++) {
                    long x = s % g.L[0], y = (s / g.stride[1]) % g.L[1];
                    long z = (s / g.stride[2]) % g.L[2], t = s / g.stride[3];
                    if ((int)((x + y + z + t) & 1) != par) continue;
                    double omega[18];
                    staple_at(links, &g, mu, s, omega);
                    double *U = links + ((long)mu * V + s) * 18;
                    long lk = (long)mu * V + s;
                    for (int hit = 0; hit < nhit; hit++) {
                        double u[16];
                        if (predrawn) {
                            const double *src = predrawn + (((long)sw * 4 * V + lk) * nhit + hit) * nu_per_hit;
                            for (int k = 0; k < nu_per_hit; k++) u[k] = src[k];
                        } else {
                            for (uint32_t b = 0; b < 4; b++) {
                                uint32_t ctr[4] = {(uint32_t)lk, (uint32_t)(sweep0 + sw), (uint32_t)hit, b};
                                uint32_t r[4];
                                or_philox4x32_10(ctr, key, r);
                                for (int k = 0; k < 4; k++) u[4 * b + k] = u32_to_unit(r[k]);
                            }
                        }
                        double X[18], Un[18], diff[18];
                        if (candidate == 0)
                            or_rand_su3_transf(u, eps, X);
                        else
                            or_rand_su3_exph(u, eps, X);
                        or_su3_mul(X, U, Un);
                        for (int k = 0; k < 18; k++) diff[k] = Un[k] - U[k];
                        double dS = -beta / 3.0 * or_su3_retr_mul(diff, omega);
                        if (exp(-dS) > u[nu_per_hit - 1]) {
                            memcpy(U, Un, sizeof Un);
                            acc++;
                        }
                    }
                }
            }
        }
    }
    if (accepted) *accepted = acc;
    return 0;
}
