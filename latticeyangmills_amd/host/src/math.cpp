// Host SU(3) algebra (no device code: the lattice-wide hot path runs in the
// HIP kernels behind include/lym.h).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>

#include "Math/lattice.h"
#include "Math/su3.h"

SU3 &SU3::operator*=(const SU3 &o) noexcept
{
    SU3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = 0, im = 0;
            for (int k = 0; k < 3; k++) {
                const double ar = mat[2 * (3 * i + k)], ai = mat[2 * (3 * i + k) + 1];
                const double br = o.mat[2 * (3 * k + j)], bi = o.mat[2 * (3 * k + j) + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            r.mat[2 * (3 * i + j)] = re;
            r.mat[2 * (3 * i + j) + 1] = im;
        }
    return *this = r;
}

void SU3::setSU3Identity()
{
    mat.fill(0.0);
    mat[0] = mat[8] = mat[16] = 1.0;
}

void SU3::setSU3Zero() { mat.fill(0.0); }

void SU3::printSU3() const
{
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) printf("( %2.5g, %2.5g )\t", mat[2 * (3 * i + j)], mat[2 * (3 * i + j) + 1]);
        printf("\n");
    }
    printf("\n");
}

SU3 operator~(const SU3 &a)
{
    SU3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            c.mat[2 * (3 * i + j)] = a.mat[2 * (3 * j + i)];
            c.mat[2 * (3 * i + j) + 1] = -a.mat[2 * (3 * j + i) + 1];
        }
    return c;
}

double getMultSumTrace(const SU3 &a, const SU3 &b)
{
    double s = 0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            s += a.mat[2 * (3 * i + k)] * b.mat[2 * (3 * k + i)] - a.mat[2 * (3 * i + k) + 1] * b.mat[2 * (3 * k + i) + 1];
    return s;
}

// exp(iQ) for Hermitian traceless Q by the Cayley-Hamilton method of
// Morningstar and Peardon (Phys. Rev. D 69 (2004) 054501, eqs. 23-35):
// exp(iQ) = f0 + f1 Q + f2 Q^2 with f_j = h_j / (9u^2 - w^2), where
// c0 = det Q, c1 = Tr Q^2 / 2, theta = acos(|c0| / c0max), c0max = 2 (c1/3)^(3/2),
// u = sqrt(c1/3) cos(theta/3), w = sqrt(c1) sin(theta/3), xi0(w) = sin w / w and
//   h0 = (u^2 - w^2) e^{2iu} + e^{-iu} [8u^2 cos w + 2iu (3u^2 + w^2) xi0(w)]
//   h1 = 2u e^{2iu} - e^{-iu} [2u cos w - i (3u^2 - w^2) xi0(w)]
//   h2 = e^{2iu} - e^{-iu} [cos w + 3iu xi0(w)];
// for c0 < 0, f_j(c0) = (-1)^j conj f_j(-c0).  This is the algorithm of the
// reference's exp(SU3) (src/Math/su3.cpp:96-183), written here in complex
// arithmetic; Q = 0 gives the identity.
SU3 exp(const SU3 &Q)
{
    using C = std::complex<double>;
    C q[3][3], q2[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) q[i][j] = C(Q.mat[2 * (3 * i + j)], Q.mat[2 * (3 * i + j) + 1]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) q2[i][j] = q[i][0] * q[0][j] + q[i][1] * q[1][j] + q[i][2] * q[2][j];
    const double c1 = 0.5 * (q2[0][0] + q2[1][1] + q2[2][2]).real();
    SU3 r;
    r.setSU3Identity();
    if (!(c1 > 0.0)) return r;
    const C det = q[0][0] * (q[1][1] * q[2][2] - q[1][2] * q[2][1]) - q[0][1] * (q[1][0] * q[2][2] - q[1][2] * q[2][0]) +
                  q[0][2] * (q[1][0] * q[2][1] - q[1][1] * q[2][0]);
    const double c0 = det.real();
    const double c0max = 2.0 * std::pow(c1 / 3.0, 1.5);
    const double theta = std::acos(std::min(1.0, std::fabs(c0) / c0max));
    const double u = std::sqrt(c1 / 3.0) * std::cos(theta / 3.0);
    const double w = std::sqrt(c1) * std::sin(theta / 3.0);
    const double w2 = w * w, u2 = u * u;
    const double xi0 = std::fabs(w) < 0.05 ? 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0)) : std::sin(w) / w;
    const C I(0.0, 1.0), e2iu = std::exp(2.0 * I * u), emiu = std::exp(-I * u);
    const double cw = std::cos(w);
    C h[3];
    h[0] = (u2 - w2) * e2iu + emiu * (8.0 * u2 * cw + 2.0 * I * u * (3.0 * u2 + w2) * xi0);
    h[1] = 2.0 * u * e2iu - emiu * (2.0 * u * cw - I * (3.0 * u2 - w2) * xi0);
    h[2] = e2iu - emiu * (cw + 3.0 * I * u * xi0);
    C f[3];
    for (int j = 0; j < 3; j++) {
        f[j] = h[j] / (9.0 * u2 - w2);
        if (c0 < 0.0) f[j] = (j % 2 ? -1.0 : 1.0) * std::conj(f[j]);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const C e = (i == j ? f[0] : C(0.0)) + f[1] * q[i][j] + f[2] * q2[i][j];
            r.mat[2 * (3 * i + j)] = e.real();
            r.mat[2 * (3 * i + j) + 1] = e.imag();
        }
    return r;
}

void setToZero(Lattice<SU3> &l)
{
    for (auto &m : l.lattice) m.setSU3Zero();
}

void setToIdentity(Lattice<SU3> &l)
{
    for (auto &m : l.lattice) m.setSU3Identity();
}
