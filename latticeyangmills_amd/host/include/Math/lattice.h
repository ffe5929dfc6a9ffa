// Site container of the reference's API (include/Math/lattice.h:61-229):
// V values, site index x + Lx*(y + Ly*(z + Lz*t)).  Whole-lattice element-wise
// operators are kept for host-side use; the MPI halo shift of the reference is
// not part of this host API (the kernels index periodic neighbours directly).
#pragma once
#include <array>
#include <vector>

#include "Math/su3.h"

template <typename T> class Lattice {
public:
    Lattice() {}
    explicit Lattice(std::array<int, 4> s) { allocate(s); }
    void allocate(std::array<int, 4> s)
    {
        size = s;
        sites = s[0] * s[1] * s[2] * s[3];
        lattice.resize(sites);
    }

    std::vector<T> lattice;
    std::array<int, 4> size{{0, 0, 0, 0}};
    int sites = 0;

    T &at(int x, int y, int z, int t) { return lattice[x + size[0] * (y + size[1] * (z + size[2] * t))]; }
    const T &at(int x, int y, int z, int t) const { return lattice[x + size[0] * (y + size[1] * (z + size[2] * t))]; }
    T &at(int i) { return lattice[i]; }
    const T &at(int i) const { return lattice[i]; }

    Lattice &operator+=(const Lattice &o) noexcept
    {
        for (int i = 0; i < sites; i++) lattice[i] += o.lattice[i];
        return *this;
    }
    Lattice &operator-=(const Lattice &o) noexcept
    {
        for (int i = 0; i < sites; i++) lattice[i] -= o.lattice[i];
        return *this;
    }
    Lattice &operator*=(const Lattice &o) noexcept
    {
        for (int i = 0; i < sites; i++) lattice[i] *= o.lattice[i];
        return *this;
    }
    Lattice &operator*=(double s) noexcept
    {
        for (int i = 0; i < sites; i++) lattice[i] *= s;
        return *this;
    }
    friend Lattice operator+(Lattice a, const Lattice &b) { return a += b; }
    friend Lattice operator-(Lattice a, const Lattice &b) { return a -= b; }
    friend Lattice operator*(Lattice a, const Lattice &b) { return a *= b; }
    friend Lattice operator*(Lattice a, double s) { return a *= s; }
};

template <typename T> Lattice<T> adj(const Lattice<T> &a)
{
    Lattice<T> r(a.size);
    for (int i = 0; i < a.sites; i++) r.at(i) = ~a.at(i);
    return r;
}

template <typename T> T sum(const Lattice<T> &a)
{
    T r(0.0);
    for (int i = 0; i < a.sites; i++) r += a.at(i);
    return r;
}

// Site-wise exp(iQ) of a lattice of Hermitian traceless matrices
// (include/Math/lattice.h:231-245 of the reference).
template <typename T> Lattice<T> exp(const Lattice<T> &a)
{
    Lattice<T> r(a.size);
    for (int i = 0; i < a.sites; i++) r.at(i) = exp(a.at(i));
    return r;
}

// Site-wise Re Tr and Im Tr (include/Math/lattice.h:298-324 of the reference).
inline Lattice<double> realTrace(const Lattice<SU3> &a)
{
    Lattice<double> r(a.size);
    for (int i = 0; i < a.sites; i++) r.at(i) = a.at(i).realTrace();
    return r;
}
inline Lattice<double> imagTrace(const Lattice<SU3> &a)
{
    Lattice<double> r(a.size);
    for (int i = 0; i < a.sites; i++) r.at(i) = a.at(i).imagTrace();
    return r;
}

void setToZero(Lattice<SU3> &l);
void setToIdentity(Lattice<SU3> &l);
