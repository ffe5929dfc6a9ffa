// SU(3) element on the host, the value type of the reference's API
// (include/Math/su3.h:53-103): 18 doubles, mat[2*(3i+j)] = Re U_ij,
// mat[2*(3i+j)+1] = Im U_ij.  Arithmetic here is for host-side use of the
// API (user code, per-link Metropolis Delta S); the lattice-wide hot path runs
// in the HIP kernels behind include/lym.h.
#pragma once
#include <array>

struct SU3 {
    std::array<double, 18> mat;

    SU3() noexcept {}
    explicit SU3(double value) noexcept { mat.fill(value); }

    SU3 &operator+=(const SU3 &o) noexcept
    {
        for (int i = 0; i < 18; i++) mat[i] += o.mat[i];
        return *this;
    }
    SU3 &operator-=(const SU3 &o) noexcept
    {
        for (int i = 0; i < 18; i++) mat[i] -= o.mat[i];
        return *this;
    }
    SU3 &operator*=(const SU3 &o) noexcept;  // matrix product, *this = *this * o
    SU3 &operator+=(double s) noexcept
    {
        for (auto &x : mat) x += s;
        return *this;
    }
    SU3 &operator-=(double s) noexcept
    {
        for (auto &x : mat) x -= s;
        return *this;
    }
    SU3 &operator*=(double s) noexcept
    {
        for (auto &x : mat) x *= s;
        return *this;
    }

    void setSU3Identity();
    void setSU3Zero();
    double realTrace() const { return mat[0] + mat[8] + mat[16]; }
    double imagTrace() const { return mat[1] + mat[9] + mat[17]; }
    void printSU3() const;
};

inline SU3 operator+(SU3 a, const SU3 &b) { return a += b; }
inline SU3 operator-(SU3 a, const SU3 &b) { return a -= b; }
inline SU3 operator*(SU3 a, const SU3 &b) { return a *= b; }
inline SU3 operator+(SU3 a, double s) { return a += s; }
inline SU3 operator-(SU3 a, double s) { return a -= s; }
inline SU3 operator*(SU3 a, double s) { return a *= s; }

SU3 operator~(const SU3 &a);                         // Hermitian conjugate
SU3 exp(const SU3 &Q);                               // exp(iQ), Q Hermitian traceless
double getMultSumTrace(const SU3 &a, const SU3 &b);  // Re Tr(a b)
