// Device-side SU(3) algebra, su(3) exponential, lattice geometry, Philox and
// compensated reductions for gfx950.
//
// Link storage in HBM is structure-of-arrays by complex element:
//   field[(mu*9 + e)*V + s]   (e = 3*i + j, one complex per slot)
// so a wavefront of 64 consecutive sites reads 64 consecutive complexes
// (1 KiB FP64 / 512 B FP32) per element.  The reference keeps each link as
// 18 interleaved doubles per site (include/Math/su3.h:63); conversion happens
// once at upload/download.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lym {

// ---------------------------------------------------------------- types ---

template <typename R> struct cplx {
    R re, im;
};

template <typename R> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <typename R> using vec2_t = typename vec2<R>::type;

// 3x3 complex matrix, row-major, e[3*i + j]
template <typename R> struct su3 {
    cplx<R> e[9];
};

// Hermitian traceless 3x3 stored as its 8 real su(3) coordinates:
// diagonal d0, d1 (d2 = -d0 - d1), upper off-diagonal o01, o02, o12.
template <typename R> struct herm8 {
    R d0, d1;
    cplx<R> o01, o02, o12;
};

// Anti-Hermitian traceless 3x3: diagonal i*a0, i*a1, i*a2 (a0+a1+a2 = 0 up
// to rounding), upper off-diagonal o01, o02, o12, lower = -conj(upper).
template <typename R> struct ah8 {
    R a0, a1, a2;
    cplx<R> o01, o02, o12;
};

#define LYM_DI __device__ __forceinline__

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename R> LYM_DI cplx<R> cmk(R a, R b) { return cplx<R>{a, b}; }
template <typename R> LYM_DI cplx<R> conj(cplx<R> a) { return cplx<R>{a.re, -a.im}; }

// acc + a*b
template <typename R> LYM_DI cplx<R> cfma(cplx<R> a, cplx<R> b, cplx<R> acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}
// acc + a*conj(b)
template <typename R> LYM_DI cplx<R> cfma_nc(cplx<R> a, cplx<R> b, cplx<R> acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.im, b.re, acc.im);
    acc.im = fma(-a.re, b.im, acc.im);
    return acc;
}
// acc + conj(a)*b
template <typename R> LYM_DI cplx<R> cfma_cn(cplx<R> a, cplx<R> b, cplx<R> acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
    return acc;
}
// acc + conj(a)*conj(b)
template <typename R> LYM_DI cplx<R> cfma_cc(cplx<R> a, cplx<R> b, cplx<R> acc)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(-a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
    return acc;
}

template <typename R> LYM_DI su3<R> zero3()
{
    su3<R> r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.e[k] = cplx<R>{R(0), R(0)};
    return r;
}

// r = a * b
template <typename R> LYM_DI su3<R> mul_nn(const su3<R> &a, const su3<R> &b)
{
    su3<R> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{a.e[3 * i].re * b.e[j].re, a.e[3 * i].re * b.e[j].im};
            s.re = fma(-a.e[3 * i].im, b.e[j].im, s.re);
            s.im = fma(a.e[3 * i].im, b.e[j].re, s.im);
            s = cfma(a.e[3 * i + 1], b.e[3 + j], s);
            s = cfma(a.e[3 * i + 2], b.e[6 + j], s);
            r.e[3 * i + j] = s;
        }
    return r;
}
// r = a * b^+
template <typename R> LYM_DI su3<R> mul_na(const su3<R> &a, const su3<R> &b)
{
    su3<R> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_nc(a.e[3 * i + k], b.e[3 * j + k], s);
            r.e[3 * i + j] = s;
        }
    return r;
}
// r = a^+ * b
template <typename R> LYM_DI su3<R> mul_an(const su3<R> &a, const su3<R> &b)
{
    su3<R> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_cn(a.e[3 * k + i], b.e[3 * k + j], s);
            r.e[3 * i + j] = s;
        }
    return r;
}
// acc += a * b^+
template <typename R> LYM_DI void acc_na(su3<R> &acc, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s = acc.e[3 * i + j];
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_nc(a.e[3 * i + k], b.e[3 * j + k], s);
            acc.e[3 * i + j] = s;
        }
}
// acc += a^+ * b
template <typename R> LYM_DI void acc_an(su3<R> &acc, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s = acc.e[3 * i + j];
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_cn(a.e[3 * k + i], b.e[3 * k + j], s);
            acc.e[3 * i + j] = s;
        }
}
// Two-row products.  For SU(3) matrices the third row of a product follows
// from the first two (row2 = conj(row0 x row1)), so a chain of products of
// SU(3) links needs rows 0 and 1 of each factor only on the left.
// Rows 0, 1 of a * b^+ (a: rows 0, 1 read; b: all rows)
template <typename R> LYM_DI void mul_na_rows01(su3<R> &r, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_nc(a.e[3 * i + k], b.e[3 * j + k], s);
            r.e[3 * i + j] = s;
        }
}
// Rows 0, 1 of a * b (a: rows 0, 1 read; b: all rows)
template <typename R> LYM_DI void mul_nn_rows01(su3<R> &r, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma(a.e[3 * i + k], b.e[3 * k + j], s);
            r.e[3 * i + j] = s;
        }
}
// Columns 0, 1 of a * b, i.e. rows 0, 1 of (a * b)^+ after conjugation
// (a: all rows; b: columns 0, 1 read); stored as r.e[3*i + j] for j < 2.
template <typename R> LYM_DI void mul_nn_cols01(su3<R> &r, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma(a.e[3 * i + k], b.e[3 * k + j], s);
            r.e[3 * i + j] = s;
        }
}
// Rows 0, 1 of a^+ * b with only columns 0, 1 of a read.
template <typename R> LYM_DI void mul_an_rows01(su3<R> &r, const su3<R> &a, const su3<R> &b)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<R> s{R(0), R(0)};
#pragma unroll
            for (int k = 0; k < 3; k++) s = cfma_cn(a.e[3 * k + i], b.e[3 * k + j], s);
            r.e[3 * i + j] = s;
        }
}
// Row 2 of an SU(3) matrix from its rows 0, 1: row2 = conj(row0 x row1).
template <typename R> LYM_DI void complete_row2(su3<R> &m)
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int k = (j + 1) % 3, l = (j + 2) % 3;
        const cplx<R> a = m.e[k], b = m.e[3 + l], c = m.e[l], d = m.e[3 + k];
        R re = a.re * b.re;
        re = fma(-a.im, b.im, re);
        re = fma(-c.re, d.re, re);
        re = fma(c.im, d.im, re);
        R im = a.re * b.im;
        im = fma(a.im, b.re, im);
        im = fma(-c.re, d.im, im);
        im = fma(-c.im, d.re, im);
        m.e[6 + j] = cplx<R>{re, -im};
    }
}
// acc += m for m in SU(3) given by rows 0, 1: row 2 = conj(row0 x row1).
template <typename R> LYM_DI void acc_su3_rows01(su3<R> &acc, const su3<R> &m)
{
#pragma unroll
    for (int j = 0; j < 6; j++) {
        acc.e[j].re += m.e[j].re;
        acc.e[j].im += m.e[j].im;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int k = (j + 1) % 3, l = (j + 2) % 3;
        // acc row 2 += conj(m0k m1l - m0l m1k), accumulated into acc directly
        const cplx<R> a = m.e[k], b = m.e[3 + l], c = m.e[l], d = m.e[3 + k];
        R re = fma(a.re, b.re, acc.e[6 + j].re);
        re = fma(-a.im, b.im, re);
        re = fma(-c.re, d.re, re);
        re = fma(c.im, d.im, re);
        R im = fma(-a.re, b.im, acc.e[6 + j].im);
        im = fma(-a.im, b.re, im);
        im = fma(c.re, d.im, im);
        im = fma(c.im, d.re, im);
        acc.e[6 + j] = cplx<R>{re, im};
    }
}

template <typename R> LYM_DI su3<R> adj(const su3<R> &a)
{
    su3<R> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.e[3 * i + j] = conj(a.e[3 * j + i]);
    return r;
}
template <typename R> LYM_DI R retrace(const su3<R> &a) { return a.e[0].re + a.e[4].re + a.e[8].re; }

// Re Tr(a * b)
template <typename R> LYM_DI R retr_mul(const su3<R> &a, const su3<R> &b)
{
    R s = R(0);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s = fma(a.e[3 * i + k].re, b.e[3 * k + i].re, s);
            s = fma(-a.e[3 * i + k].im, b.e[3 * k + i].im, s);
        }
    return s;
}

template <typename To, typename From> LYM_DI su3<To> convert(const su3<From> &a)
{
    su3<To> r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.e[k] = cplx<To>{To(a.e[k].re), To(a.e[k].im)};
    return r;
}

// ------------------------------------------------------------ HBM access ---

// Link and accumulator accesses go through buffer resources (scalar
// registers): the element planes of a direction are taken in groups of
// kPlanesPerRsrc<S> (6 in FP32: rows 0 and 1 of a link; 3 in FP64: one row),
// one descriptor per group; the element's offset in its group k*V*size is the
// scalar soffset and the site's byte offset the one vector offset shared by
// all elements.  No per-element address arithmetic runs in the vector ALU,
// and an access outside the group reads 0 instead of faulting.  Requires
// 48 * V < 2^31 bytes, i.e. V < 44.7 M sites (checked when a lattice is
// created), so that every size and offset is a non-negative 32-bit integer.
// The two-row FP32 flow (the mixed default) reads and writes a link through
// one descriptor; FP64 takes three, which also leaves its observables and
// stage kernels fewer live registers than one descriptor per direction did.
constexpr int kBufferDword3 = 0x00020000;  // raw buffer, 32-bit data format (gfx9 family)

template <typename T> LYM_DI __amdgpu_buffer_rsrc_t make_rsrc(const T *base, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base), (short)0, int(bytes), kBufferDword3);
}
template <typename S> constexpr int kPlanesPerRsrc = sizeof(S) == 4 ? 6 : 3;

// AUX is the cache policy of the access: 0, or kStreamAux (the non-temporal
// bit) for data that this kernel does not read again: the links a stage
// writes and the RK accumulator.  That leaves the XCD's L2 to the neighbour
// links the other workgroups still read (measured: -4 % per mixed stage and
// -7 % per FP64 stage at 32^3 x 64).
constexpr int kStreamAux = 2;

template <typename T, int AUX = 0> LYM_DI T buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), AUX));
    else if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), int(soff), AUX));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), int(soff), AUX));
}

template <typename T, int AUX = 0> LYM_DI void buf_store(T v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(voff), int(soff), AUX);
    else if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, int(voff), int(soff), AUX);
    else
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, int(voff), int(soff), AUX);
}

// Elements 0 .. NE-1 of the link of site s: plane[k*V + s], the planes taken
// in groups of kPlanesPerRsrc<S> per descriptor.
template <typename S, int NE, int AUX = 0>
LYM_DI void load_elems(const vec2_t<S> *__restrict__ plane, long V, unsigned s, vec2_t<S> out[NE])
{
    using F2 = vec2_t<S>;
    constexpr int G = kPlanesPerRsrc<S>;
    const unsigned off = s * unsigned(sizeof(F2)), es = unsigned(V * sizeof(F2));
#pragma unroll
    for (int k0 = 0; k0 < NE; k0 += G) {
        const int n = NE - k0 < G ? NE - k0 : G;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(plane + long(k0) * V, long(n) * V * long(sizeof(F2)));
#pragma unroll
        for (int k = 0; k < n; k++) out[k0 + k] = buf_load<F2, AUX>(r, off, unsigned(k) * es);
    }
}
template <typename S, int NE, int AUX = 0>
LYM_DI void store_elems(vec2_t<S> *__restrict__ plane, long V, unsigned s, const vec2_t<S> in[NE])
{
    using F2 = vec2_t<S>;
    constexpr int G = kPlanesPerRsrc<S>;
    const unsigned off = s * unsigned(sizeof(F2)), es = unsigned(V * sizeof(F2));
#pragma unroll
    for (int k0 = 0; k0 < NE; k0 += G) {
        const int n = NE - k0 < G ? NE - k0 : G;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(plane + long(k0) * V, long(n) * V * long(sizeof(F2)));
#pragma unroll
        for (int k = 0; k < n; k++) buf_store<F2, AUX>(in[k0 + k], r, off, unsigned(k) * es);
    }
}

// One link of a direction plane: 9 complexes at plane[k*V + s], k = 0..8.
template <typename R, typename S>
LYM_DI su3<R> load_link(const vec2_t<S> *__restrict__ plane, long V, unsigned s)
{
    vec2_t<S> v[9];
    load_elems<S, 9>(plane, V, s, v);
    su3<R> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = cplx<R>{R(v[k].x), R(v[k].y)};
    return m;
}

// FP32 links (the mixed path) hold row 2 = conj(row0 x row1) formed in FP32
// from their stored rows 0, 1: store_link writes it so whatever row 2 of m
// is, and the upload kernel derives it too.  The two-row form (below) is then
// a lossless compression of an FP32 field: a reader sees the same values
// whichever form the field is in, so the choice of form changes no result.
// FP64 links are stored as they are (the FP64 path multiplies out all rows).
template <typename S, typename R, int AUX = 0>
LYM_DI void store_link(vec2_t<S> *__restrict__ plane, long V, unsigned s, const su3<R> &m)
{
    su3<S> t;
#pragma unroll
    for (int k = 0; k < 9; k++) t.e[k] = cplx<S>{S(m.e[k].re), S(m.e[k].im)};
    if constexpr (sizeof(S) == 4) complete_row2(t);
    vec2_t<S> v[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        v[k].x = t.e[k].re;
        v[k].y = t.e[k].im;
    }
    store_elems<S, 9, AUX>(plane, V, s, v);
}

// Links in two-row form (the mixed flow, chosen per lattice shape): a
// direction plane holds the first two rows only, plane[k*V + s] for
// k = 0..5, and row 2 = conj(row0 x row1) is rebuilt on load in the storage
// precision S, exactly as store_link forms the row 2 it stores for FP32
// links, then converted to R.  A third fewer bytes per link read and
// written; where a product uses rows 0, 1 of a link only, the rebuild is dead
// code and drops out.
template <int ROWS> constexpr int plane_elems() { return ROWS == 2 ? 6 : 9; }

template <typename R, typename S, int ROWS = 3>
LYM_DI su3<R> load_link_r(const vec2_t<S> *__restrict__ plane, long V, unsigned s)
{
    if constexpr (ROWS == 3) {
        return load_link<R, S>(plane, V, s);
    } else {
        vec2_t<S> v[6];
        load_elems<S, 6>(plane, V, s, v);
        su3<S> m;
#pragma unroll
        for (int k = 0; k < 6; k++) m.e[k] = cplx<S>{v[k].x, v[k].y};
        complete_row2(m);
        return convert<R>(m);
    }
}

template <typename S, typename R, int AUX = 0, int ROWS = 3>
LYM_DI void store_link_r(vec2_t<S> *__restrict__ plane, long V, unsigned s, const su3<R> &m)
{
    if constexpr (ROWS == 3) {
        store_link<S, R, AUX>(plane, V, s, m);
    } else {
        vec2_t<S> v[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            v[k].x = S(m.e[k].re);
            v[k].y = S(m.e[k].im);
        }
        store_elems<S, 6, AUX>(plane, V, s, v);
    }
}

// ------------------------------------------------------------- geometry ---

// Lexicographic site index s = x + Lx*(y + Ly*(z + Lz*t)), periodic
// boundaries in every direction (one-rank semantics of the reference's
// halo shift, include/Math/lattice.h:360-594).
struct Geom {
    int L[4];
    int stride[4];
    long V;
};

// Neighbour offsets of one site: fwd[d] = index(x + e_d) - s, bwd[d] = index(x - e_d) - s.
struct Nbr {
    int fwd[4], bwd[4];
};

// Site indices fit in 32 bits (V < 2^31 / (9 * sizeof(complex)), checked when
// a lattice is created), so the coordinates come from unsigned 32-bit
// division rather than the far longer 64-bit sequence.
// From the site's coordinates c = (x, y, z, t):
LYM_DI Nbr neighbours_at(const Geom g, const int c[4])
{
    Nbr n;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        n.fwd[d] = (c[d] == g.L[d] - 1) ? -(g.L[d] - 1) * g.stride[d] : g.stride[d];
        n.bwd[d] = (c[d] == 0) ? (g.L[d] - 1) * g.stride[d] : -g.stride[d];
    }
    return n;
}
// From the site index (three divisions per lane):
LYM_DI Nbr neighbours(const Geom g, long s)
{
    int c[4];
    uint32_t r = uint32_t(s);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const uint32_t L = uint32_t(g.L[d]);
        c[d] = int(r % L);
        r /= L;
    }
    c[3] = int(r);
    return neighbours_at(g, c);
}

// ------------------------------------------------- su(3) exponential map ---

// Q^2 for Hermitian traceless Q (Hermitian result: diagonal real parts in
// q2.e[0,4,8].re, upper triangle in e[1], e[2], e[5]; lower filled as conj).
LYM_DI su3<double> herm_square(const herm8<double> &q)
{
    double d2 = -q.d0 - q.d1;
    double n01 = fma(q.o01.re, q.o01.re, q.o01.im * q.o01.im);
    double n02 = fma(q.o02.re, q.o02.re, q.o02.im * q.o02.im);
    double n12 = fma(q.o12.re, q.o12.re, q.o12.im * q.o12.im);
    su3<double> r;
    r.e[0] = cplx<double>{fma(q.d0, q.d0, n01 + n02), 0.0};
    r.e[4] = cplx<double>{fma(q.d1, q.d1, n01 + n12), 0.0};
    r.e[8] = cplx<double>{fma(d2, d2, n02 + n12), 0.0};
    // (Q^2)_01 = (d0 + d1) o01 + o02 conj(o12)
    cplx<double> t01{(q.d0 + q.d1) * q.o01.re, (q.d0 + q.d1) * q.o01.im};
    t01 = cfma_nc(q.o02, q.o12, t01);
    // (Q^2)_02 = (d0 + d2) o02 + o01 o12
    cplx<double> t02{(q.d0 + d2) * q.o02.re, (q.d0 + d2) * q.o02.im};
    t02 = cfma(q.o01, q.o12, t02);
    // (Q^2)_12 = (d1 + d2) o12 + conj(o01) o02
    cplx<double> t12{(q.d1 + d2) * q.o12.re, (q.d1 + d2) * q.o12.im};
    t12 = cfma_cn(q.o01, q.o02, t12);
    r.e[1] = t01;
    r.e[2] = t02;
    r.e[5] = t12;
    r.e[3] = conj(t01);
    r.e[6] = conj(t02);
    r.e[7] = conj(t12);
    return r;
}

// E = f0 + f1 Q + f2 Q^2 for Hermitian traceless Q (d2 = -d0 - d1), Q^2 from herm_square
LYM_DI su3<double> exp_assemble(const herm8<double> &q, const su3<double> &q2, double d2, cplx<double> f0,
                                cplx<double> f1, cplx<double> f2)
{
    su3<double> r;
    const double dg[3] = {q.d0, q.d1, d2};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double q2ii = q2.e[4 * i].re;
        r.e[4 * i] = cplx<double>{fma(f2.re, q2ii, fma(f1.re, dg[i], f0.re)), fma(f2.im, q2ii, fma(f1.im, dg[i], f0.im))};
    }
    r.e[1] = cfma(f2, q2.e[1], cfma(f1, q.o01, cplx<double>{0.0, 0.0}));
    r.e[2] = cfma(f2, q2.e[2], cfma(f1, q.o02, cplx<double>{0.0, 0.0}));
    r.e[5] = cfma(f2, q2.e[5], cfma(f1, q.o12, cplx<double>{0.0, 0.0}));
    r.e[3] = cfma(f2, q2.e[3], cfma(f1, conj(q.o01), cplx<double>{0.0, 0.0}));
    r.e[6] = cfma(f2, q2.e[6], cfma(f1, conj(q.o02), cplx<double>{0.0, 0.0}));
    r.e[7] = cfma(f2, q2.e[7], cfma(f1, conj(q.o12), cplx<double>{0.0, 0.0}));
    return r;
}

// exp(iQ) by its power series reduced with Cayley-Hamilton: for traceless Q,
// Q^3 = c1 Q + c0 1 (c1 = Tr Q^2 / 2, c0 = det Q), so Q^n = a_n + b_n Q + d_n Q^2
// with a_{n+1} = c0 d_n, b_{n+1} = a_n + c1 d_n, d_{n+1} = b_n, and
// f_j = sum_n i^n coef_{j,n} / n!.  The spectral radius is at most
// sqrt(4 c1 / 3) <= 0.817 for c1 <= 0.5, so the terms n = 0..18 kept leave a
// truncation error below 2e-19 (the first omitted term, 0.817^19 / 19!); the
// result agrees with the Morningstar-Peardon form to rounding (4.6e-16 against
// a Pade matrix exponential at c1 = 0.5) and needs no transcendental function.
constexpr double kSeriesC1Max = 0.5;
constexpr int kSeriesTerms = 18;
// Short tier: c1 <= 0.004 (spectral radius <= 0.073) needs 9 terms for a
// truncation error below 2e-18.  The Wilson flow at eps <= 0.02 is almost
// always there.  Each lane decides from its own c1: lanes in the tier take
// the short sum of terms 0..9 (short_series), lanes above it the recursion
// over terms 0..18 (a wave whose lanes are all in the tier skips it), so a
// link's result depends on its own Q alone and not on which sites share its
// wave.
constexpr double kShortSeriesC1Max = 0.004;
constexpr int kShortSeriesTerms = 9;

// Running state of the series: Q^n = a + b Q + d Q^2, inv_fact = 1/n!, and the
// partial sums of f0, f1, f2 over the terms added so far.
struct SeriesState {
    double a, b, d, inv_fact;
    double f0r, f0i, f1r, f1i, f2r, f2i;
};

// adds the terms n = N0 .. N1 (in order) to the sums
template <int N0, int N1> LYM_DI void series_terms(SeriesState &st, double c0, double c1)
{
#pragma unroll
    for (int n = N0; n <= N1; n++) {
        if (n > 0) st.inv_fact *= 1.0 / double(n);
        const double sg = (n & 2) ? -1.0 : 1.0;  // i^n = 1, i, -1, -i: real for even n
        const double w = sg * st.inv_fact;
        if (n & 1) {
            st.f0i = fma(st.a, w, st.f0i);
            st.f1i = fma(st.b, w, st.f1i);
            st.f2i = fma(st.d, w, st.f2i);
        } else {
            st.f0r = fma(st.a, w, st.f0r);
            st.f1r = fma(st.b, w, st.f1r);
            st.f2r = fma(st.d, w, st.f2r);
        }
        const double an = c0 * st.d, bn = fma(c1, st.d, st.a);
        st.d = st.b;
        st.a = an;
        st.b = bn;
    }
}

// The short tier, terms 0 .. N (N = 6 or 9), written out: with Q^n = a_n + b_n Q
// + d_n Q^2, (a, b, d)_n = (1,0,0), (0,1,0), (0,0,1), (c0,c1,0), (0,c0,c1),
// (c0 c1, c1^2, c0), (c0^2, 2 c0 c1, c1^2), (c0 c1^2, c0^2 + c1^3, 2 c0 c1),
// (2 c0^2 c1, 3 c0 c1^2, c0^2 + c1^3), (c0^3 + c0 c1^3, 3 c0^2 c1 + c1^4, 3 c0 c1^2)
// for n = 0 .. 9, and f_j = sum_n i^n / n! times the j-th of them, grouped as
// polynomials in c1 (about a third of the operations of the recursion, whose
// sums it equals to rounding).
template <int N> LYM_DI void short_series(double c0, double c1, cplx<double> &f0, cplx<double> &f1, cplx<double> &f2)
{
    static_assert(N == 6 || N == 9, "short series of 6 or 9 terms");
    const double c0sq = c0 * c0;
    if constexpr (N == 6) {
        f0 = cplx<double>{fma(c0sq, -1.0 / 720.0, 1.0), c0 * fma(c1, 1.0 / 120.0, -1.0 / 6.0)};
        f1 = cplx<double>{c0 * fma(c1, -1.0 / 360.0, 1.0 / 24.0), fma(c1, fma(c1, 1.0 / 120.0, -1.0 / 6.0), 1.0)};
        f2 = cplx<double>{fma(c1, fma(c1, -1.0 / 720.0, 1.0 / 24.0), -0.5), c0 * (1.0 / 120.0)};
    } else {
        const double p1 = fma(c1, fma(c1, fma(c1, 1.0 / 362880.0, -1.0 / 5040.0), 1.0 / 120.0), -1.0 / 6.0);
        f0 = cplx<double>{fma(c0sq, fma(c1, 2.0 / 40320.0, -1.0 / 720.0), 1.0), c0 * fma(c0sq, 1.0 / 362880.0, p1)};
        f1 = cplx<double>{c0 * fma(c1, fma(c1, 3.0 / 40320.0, -1.0 / 360.0), 1.0 / 24.0),
                          fma(c0sq, fma(c1, 3.0 / 362880.0, -1.0 / 5040.0), fma(c1, p1, 1.0))};
        f2 = cplx<double>{fma(c0sq, 1.0 / 40320.0, fma(c1, fma(c1, fma(c1, 1.0 / 40320.0, -1.0 / 720.0), 1.0 / 24.0), -0.5)),
                          c0 * fma(c1, fma(c1, 3.0 / 362880.0, -2.0 / 5040.0), 1.0 / 120.0)};
    }
}

// Links stored in FP32 (the mixed path) carry a rounding of 6e-8, so there the
// tiers stop earlier: terms 0..6 for c1 <= 0.004 (truncation below 2.2e-12)
// and 0..14 up to c1 = 0.5 (below 4e-14).
constexpr int kShortSeriesTermsF32 = 6, kSeriesTermsF32 = 14;

template <int N_SHORT = kShortSeriesTerms, int N_LONG = kSeriesTerms>
LYM_DI su3<double> exp_iq_series(const herm8<double> &q, const su3<double> &q2, double d2, double c1)
{
    cplx<double> p = cfma(q.o01, q.o12, cplx<double>{0.0, 0.0});
    const double re3 = fma(p.re, q.o02.re, p.im * q.o02.im);
    const double n01 = fma(q.o01.re, q.o01.re, q.o01.im * q.o01.im);
    const double n02 = fma(q.o02.re, q.o02.re, q.o02.im * q.o02.im);
    const double n12 = fma(q.o12.re, q.o12.re, q.o12.im * q.o12.im);
    // c0 = det Q = d0 d1 d2 + 2 Re(o01 o12 conj(o02)) - d0 |o12|^2 - d1 |o02|^2 - d2 |o01|^2
    const double c0 = fma(-d2, n01, fma(-q.d1, n02, fma(-q.d0, n12, fma(q.d0 * q.d1, d2, 2.0 * re3))));
    cplx<double> f0, f1, f2;
    if (c1 > kShortSeriesC1Max) {  // long tier: the recursion over terms 0 .. N_LONG
        SeriesState st{1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        series_terms<0, N_LONG>(st, c0, c1);
        f0 = cplx<double>{st.f0r, st.f0i};
        f1 = cplx<double>{st.f1r, st.f1i};
        f2 = cplx<double>{st.f2r, st.f2i};
    } else {
        short_series<N_SHORT>(c0, c1, f0, f1, f2);
    }
    return exp_assemble(q, q2, d2, f0, f1, f2);
}

// exp(iQ) by the series alone, for Q known to have c1 = Tr Q^2 / 2 <= kSeriesC1Max
// (Metropolis candidates exp(i eps H) with 8 eps^2 <= kSeriesC1Max): no
// transcendental branch, and so far fewer registers than exp_iq.
LYM_DI su3<double> exp_iq_small(const herm8<double> &q)
{
    const double d2 = -q.d0 - q.d1;
    su3<double> q2 = herm_square(q);
    const double c1 = 0.5 * (q2.e[0].re + q2.e[4].re + q2.e[8].re);
    return exp_iq_series(q, q2, d2, c1);
}

// exp(iQ) for Hermitian traceless Q: Cayley-Hamilton form of Morningstar &
// Peardon (hep-lat/0311018), the algorithm of the reference's exp(SU3)
// (src/Math/su3.cpp:96-183): c0 = det Q = Tr Q^3 / 3, c1 = Tr Q^2 / 2,
// theta = acos(c0/c0max), u, w, xi0(w) (series for |w| < 0.05), h0..h2,
// the c0 < 0 symmetry, f_j = h_j / (9u^2 - w^2), exp(iQ) = f0 + f1 Q + f2 Q^2.
// Q == 0 gives the identity (the reference divides 0/0 there), and c0/c0max
// is clamped to 1 against rounding.  Small Q (c1 <= kSeriesC1Max) takes the
// equivalent power series above.
template <int N_SHORT = kShortSeriesTerms, int N_LONG = kSeriesTerms> LYM_DI su3<double> exp_iq(const herm8<double> &q)
{
    const double d2 = -q.d0 - q.d1;
    su3<double> q2 = herm_square(q);
    const double c1raw = 0.5 * (q2.e[0].re + q2.e[4].re + q2.e[8].re);
    if (c1raw <= kSeriesC1Max) return exp_iq_series<N_SHORT, N_LONG>(q, q2, d2, c1raw);
    const bool tiny = !(c1raw > 1e-40);  // |Q| < 1e-20: exp(iQ) = 1 + iQ to double precision
    const double c1 = tiny ? 1.0 : c1raw;
    // det Q = d0 d1 d2 + 2 Re(o01 o12 conj(o02)) - d0|o12|^2 - d1|o02|^2 - d2|o01|^2
    cplx<double> p = cfma(q.o01, q.o12, cplx<double>{0.0, 0.0});
    double re3 = p.re * q.o02.re + p.im * q.o02.im;
    double n01 = q.o01.re * q.o01.re + q.o01.im * q.o01.im;
    double n02 = q.o02.re * q.o02.re + q.o02.im * q.o02.im;
    double n12 = q.o12.re * q.o12.re + q.o12.im * q.o12.im;
    double c0 = q.d0 * q.d1 * d2 + 2.0 * re3 - q.d0 * n12 - q.d1 * n02 - d2 * n01;
    bool c0neg = c0 < 0.0;
    c0 = fabs(c0);
    const double inv_sqrt3 = 0.57735026918962576451;
    double c1sqrt = sqrt(c1);
    double c0max = 2.0 * c1 * c1sqrt * inv_sqrt3 / 3.0;
    double ratio = fmin(c0 / c0max, 1.0);
    double theta = acos(ratio);
    double st, ct;
    sincos(theta * (1.0 / 3.0), &st, &ct);
    double w = c1sqrt * st;
    double u = c1sqrt * inv_sqrt3 * ct;
    double u2 = u * u, w2 = w * w;
    double sw, cw;
    sincos(w, &sw, &cw);
    double xi = (fabs(w) < 0.05) ? 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0)) : sw / w;
    double su, cu;
    sincos(u, &su, &cu);
    double cmu = cu, smu = -su;  // e^{-iu}
    double c2u = cmu * cmu - smu * smu;
    double s2u = -2.0 * cmu * smu;  // e^{2iu}
    double cwcmu = cw * cmu, cwsmu = cw * smu, xicmu = xi * cmu, xismu = xi * smu;
    double h0r = (u2 - w2) * c2u + 8.0 * u2 * cwcmu - 2.0 * u * (3.0 * u2 + w2) * xismu;
    double h0i = (u2 - w2) * s2u + 8.0 * u2 * cwsmu + 2.0 * u * (3.0 * u2 + w2) * xicmu;
    double h1r = 2.0 * u * c2u - 2.0 * u * cwcmu - (3.0 * u2 - w2) * xismu;
    double h1i = 2.0 * u * s2u - 2.0 * u * cwsmu + (3.0 * u2 - w2) * xicmu;
    double h2r = c2u - cwcmu + 3.0 * u * xismu;
    double h2i = s2u - cwsmu - 3.0 * u * xicmu;
    if (c0neg) {  // f_j(-c0) = (-1)^j conj f_j(c0)
        h0i = -h0i;
        h1r = -h1r;
        h2i = -h2i;
    }
    double nrm = 1.0 / (9.0 * u2 - w2);
    cplx<double> f0{h0r * nrm, h0i * nrm}, f1{h1r * nrm, h1i * nrm}, f2{h2r * nrm, h2i * nrm};
    if (tiny) {
        f0 = cplx<double>{1.0, 0.0};
        f1 = cplx<double>{0.0, 1.0};
        f2 = cplx<double>{0.0, 0.0};
    }
    return exp_assemble(q, q2, d2, f0, f1, f2);
}

// Z = (i/2) [X - X^+]_traceless for X = U * Omega
// (src/Actions/puregauge.cpp:62-79, there formed as M = Omega^+ U^+, D = M^+ - M).
// In the precision T of the result: FP64 on the FP64 path; the mixed path
// keeps Z and the RK accumulator in FP32 (its storage precision) and moves to
// FP64 at the exponential.
template <typename R, typename T = double> LYM_DI herm8<T> force_from_x(const su3<R> &X)
{
    T i0 = T(X.e[0].im), i1 = T(X.e[4].im), i2 = T(X.e[8].im);
    T tr = (i0 + i1 + i2) * T(1.0 / 3.0);
    herm8<T> z;
    z.d0 = -(i0 - tr);
    z.d1 = -(i1 - tr);
    // Z_ij = (i/2)(X_ij - conj(X_ji)) for i < j
    auto off = [](cplx<R> a, cplx<R> b) {
        return cplx<T>{T(-0.5) * (T(a.im) + T(b.im)), T(0.5) * (T(a.re) - T(b.re))};
    };
    z.o01 = off(X.e[1], X.e[3]);
    z.o02 = off(X.e[2], X.e[6]);
    z.o12 = off(X.e[5], X.e[7]);
    return z;
}

// Entry (i, j) of a * b, and its imaginary part alone, in the operation order
// of mul_nn (so each equals the corresponding entry of mul_nn(a, b) bit for bit).
template <typename R> LYM_DI cplx<R> mul_entry(const su3<R> &a, const su3<R> &b, int i, int j)
{
    cplx<R> s{a.e[3 * i].re * b.e[j].re, a.e[3 * i].re * b.e[j].im};
    s.re = fma(-a.e[3 * i].im, b.e[j].im, s.re);
    s.im = fma(a.e[3 * i].im, b.e[j].re, s.im);
    s = cfma(a.e[3 * i + 1], b.e[3 + j], s);
    return cfma(a.e[3 * i + 2], b.e[6 + j], s);
}
template <typename R> LYM_DI R mul_entry_im(const su3<R> &a, const su3<R> &b, int i, int j)
{
    R im = a.e[3 * i].re * b.e[j].im;
    im = fma(a.e[3 * i].im, b.e[j].re, im);
#pragma unroll
    for (int k = 1; k < 3; k++) {
        im = fma(a.e[3 * i + k].re, b.e[3 * k + j].im, im);
        im = fma(a.e[3 * i + k].im, b.e[3 * k + j].re, im);
    }
    return im;
}

// Z = force_from_x(U * Omega) from the entries of X = U * Omega it reads: the
// imaginary parts of the diagonal and the off-diagonal pairs (X_ij, X_ji),
// 90 instead of 216 arithmetic operations, the same values.
template <typename R, typename T = double> LYM_DI herm8<T> force_from_product(const su3<R> &U, const su3<R> &om)
{
    su3<R> X;
#pragma unroll
    for (int i = 0; i < 3; i++) X.e[4 * i] = cplx<R>{R(0), mul_entry_im(U, om, i, i)};
    X.e[1] = mul_entry(U, om, 0, 1);
    X.e[3] = mul_entry(U, om, 1, 0);
    X.e[2] = mul_entry(U, om, 0, 2);
    X.e[6] = mul_entry(U, om, 2, 0);
    X.e[5] = mul_entry(U, om, 1, 2);
    X.e[7] = mul_entry(U, om, 2, 1);
    return force_from_x<R, T>(X);
}

template <typename R> LYM_DI herm8<R> h8_axpby(R a, const herm8<R> &x, R b, const herm8<R> &y)
{
    herm8<R> r;
    r.d0 = a * x.d0 + b * y.d0;
    r.d1 = a * x.d1 + b * y.d1;
    r.o01 = cplx<R>{a * x.o01.re + b * y.o01.re, a * x.o01.im + b * y.o01.im};
    r.o02 = cplx<R>{a * x.o02.re + b * y.o02.re, a * x.o02.im + b * y.o02.im};
    r.o12 = cplx<R>{a * x.o12.re + b * y.o12.re, a * x.o12.im + b * y.o12.im};
    return r;
}
// a * x in FP64 (the generator handed to the exponential)
template <typename R> LYM_DI herm8<double> h8_scale(double a, const herm8<R> &x)
{
    herm8<double> r;
    r.d0 = a * double(x.d0);
    r.d1 = a * double(x.d1);
    r.o01 = cplx<double>{a * double(x.o01.re), a * double(x.o01.im)};
    r.o02 = cplx<double>{a * double(x.o02.re), a * double(x.o02.im)};
    r.o12 = cplx<double>{a * double(x.o12.re), a * double(x.o12.im)};
    return r;
}

// RK accumulator in HBM: acc[(mu*8 + k)*V + s], k = d0, d1, o01.re, o01.im,
// o02.re, o02.im, o12.re, o12.im (buffer addressing as load_link)
template <typename S, int AUX = 0> LYM_DI herm8<S> load_h8(const S *__restrict__ a, long V, unsigned s)
{
    const unsigned off = s * unsigned(sizeof(S)), es = unsigned(V * sizeof(S));
    constexpr int G = sizeof(S) == 4 ? 8 : 4;  // planes per descriptor: 8 V sizeof(S) < 2^31 bytes in FP32
    S v[8];
#pragma unroll
    for (int h = 0; h < 8; h += G) {
        const __amdgpu_buffer_rsrc_t r = make_rsrc(a + long(h) * V, G * V * long(sizeof(S)));
#pragma unroll
        for (int k = 0; k < G; k++) v[h + k] = buf_load<S, AUX>(r, off, unsigned(k) * es);
    }
    herm8<S> z;
    z.d0 = v[0];
    z.d1 = v[1];
    z.o01 = cplx<S>{v[2], v[3]};
    z.o02 = cplx<S>{v[4], v[5]};
    z.o12 = cplx<S>{v[6], v[7]};
    return z;
}
template <typename S, int AUX = 0> LYM_DI void store_h8(S *__restrict__ a, long V, unsigned s, const herm8<S> &z)
{
    const unsigned off = s * unsigned(sizeof(S)), es = unsigned(V * sizeof(S));
    constexpr int G = sizeof(S) == 4 ? 8 : 4;
    const S v[8] = {z.d0, z.d1, z.o01.re, z.o01.im, z.o02.re, z.o02.im, z.o12.re, z.o12.im};
#pragma unroll
    for (int h = 0; h < 8; h += G) {
        const __amdgpu_buffer_rsrc_t r = make_rsrc(a + long(h) * V, G * V * long(sizeof(S)));
#pragma unroll
        for (int k = 0; k < G; k++) buf_store<S, AUX>(v[h + k], r, off, unsigned(k) * es);
    }
}

// 1/sqrt(n) in FP64.  STORED_FP32: the result feeds a link stored in FP32,
// so the hardware estimate (v_rsq_f64, relative error 5.2e-8 measured on
// gfx950) refined by one Newton step (4.1e-15) is ample, seven orders below
// the storage rounding; otherwise rsqrt() and one more Newton step (2e-16).
template <bool STORED_FP32> LYM_DI double rsqrt_refined(double n)
{
    const double r = STORED_FP32 ? __builtin_amdgcn_rsq(n) : rsqrt(n);
    return r * fma(-0.5 * n, r * r, 1.5);
}

// Project a near-unitary matrix back onto SU(3): Gram-Schmidt on rows 0, 1,
// row 2 = conj(row0 x row1) (so det = 1).  STORED_FP32 as rsqrt_refined.
template <bool STORED_FP32 = false> LYM_DI void reunitarize(su3<double> &m)
{
    double n0 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) n0 = fma(m.e[j].im, m.e[j].im, fma(m.e[j].re, m.e[j].re, n0));
    const double r0 = rsqrt_refined<STORED_FP32>(n0);
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[j] = cplx<double>{m.e[j].re * r0, m.e[j].im * r0};
    cplx<double> p{0.0, 0.0};  // <row0|row1>
#pragma unroll
    for (int j = 0; j < 3; j++) p = cfma_cn(m.e[j], m.e[3 + j], p);
    const cplx<double> mp{-p.re, -p.im};
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[3 + j] = cfma(mp, m.e[j], m.e[3 + j]);  // row1 -= <row0|row1> row0
    double n1 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) n1 = fma(m.e[3 + j].im, m.e[3 + j].im, fma(m.e[3 + j].re, m.e[3 + j].re, n1));
    const double r1 = rsqrt_refined<STORED_FP32>(n1);
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[3 + j] = cplx<double>{m.e[3 + j].re * r1, m.e[3 + j].im * r1};
#pragma unroll
    for (int j = 0; j < 3; j++) {  // row2 = conj(row0 x row1)
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const cplx<double> a = m.e[j2];
        cplx<double> c = cfma(m.e[j1], m.e[3 + j2], cplx<double>{0.0, 0.0});
        c = cfma(cplx<double>{-a.re, -a.im}, m.e[3 + j1], c);
        m.e[6 + j] = conj(c);
    }
}

// ----------------------------------------------------------------- Philox ---

// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
LYM_DI uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32x32 -> 64-bit product gives both halves (v_mad_u64_u32)
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x, p1 = uint64_t(0xCD9E8D57u) * c.z;
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
LYM_DI double u32_unit(uint32_t r) { return double(r) * (1.0 / 4294967296.0); }

// ------------------------------------------------- compensated reductions ---

// (s, c) with value s + c; combined by Knuth's TwoSum so no rounding error of
// the partial sums is lost (Kahan-Babuska compensation in tree form).
struct ksum {
    double s, c;
};
LYM_DI ksum kadd(ksum a, ksum b)
{
    double s = a.s + b.s;
    double bp = s - a.s;
    double err = (a.s - (s - bp)) + (b.s - bp);
    return ksum{s, a.c + b.c + err};
}
LYM_DI ksum kshfl_xor(ksum a, int m)
{
    return ksum{__shfl_xor(a.s, m, 64), __shfl_xor(a.c, m, 64)};
}
LYM_DI ksum wave_ksum(ksum a)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a = kadd(a, kshfl_xor(a, m));
    return a;
}

}  // namespace lym
