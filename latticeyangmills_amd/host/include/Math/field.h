// Field<T, N> / GluonField of the reference's API (include/Math/field.h:52-76):
// one Lattice per direction.  The GluonField's four Lattice<SU3> buffers are
// exactly the by_mu[4] arrays of the C ABI (include/lym.h).
#pragma once
#include <array>

#include "Math/lattice.h"

template <typename T, int N> class Field {
public:
    explicit Field(std::array<int, 4> size) : m_size(size), m_dimensions(N)
    {
        for (auto &l : m_lats) l.allocate(size);
    }
    Lattice<T> &operator[](int mu) { return m_lats[mu]; }
    const Lattice<T> &operator[](int mu) const { return m_lats[mu]; }

    std::array<int, 4> m_size;
    int m_dimensions;

private:
    std::array<Lattice<T>, N> m_lats;
};

typedef Field<SU3, 4> GluonField;
