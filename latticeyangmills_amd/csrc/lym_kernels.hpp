// Per-site device routines and the kernels of the single-lattice path.
//
// Thread mapping of the flow stage: a 256-thread workgroup covers 64
// consecutive sites, wave w = direction mu = w.  The four waves of a
// workgroup therefore read overlapping neighbour links and share them in the
// CU's L1 / the XCD's L2; each wave's loads of one matrix element are 64
// consecutive complexes.
#pragma once
#include "lym_device.hpp"

namespace lym {

constexpr int kSitesPerBlock = 64;

// Workgroup -> run of 64 sites (see lym_sched.hpp): workgroup b runs on XCD
// b % 8 and takes entry b / 8 of that XCD's row of the table; no table means
// workgroup b takes run b.  Returns -1 for padding entries.
struct Sched {
    const int *order;
    int per_xcd;
};
LYM_DI long sched_run(const Sched sc, unsigned b)
{
    if (!sc.order) return long(b);
    return sc.order[(b & 7u) * unsigned(sc.per_xcd) + (b >> 3)];
}

template <typename S> using F2 = vec2_t<S>;

// ---------------------------------------------------------- link sources ---

// A link read of the staple is tagged at compile time with its offset from
// the site being updated: DX, DY in {-1, 0, 1} along x and y, ZT = 1 when
// the offset has a z or t part, and FLIP = 1 when the offset is one lattice
// step (the other checkerboard parity) rather than zero or two.  A link
// source decides from the tag alone whether the read comes from LDS, so no
// lane ever branches on it.

// Links of a field read straight from HBM (through L1/L2).  ROWS: the
// field's storage form, three rows or two (load_link_r).
template <typename S, int ROWS = 3> struct GlobalLinks {
    static constexpr int rows = ROWS;
    const F2<S> *f;
    long V;
    template <typename R, int DX, int DY, int ZT, int FLIP = 0> LYM_DI su3<R> get(int dir, long site) const
    {
        return load_link_r<R, S, ROWS>(f + (long)dir * plane_elems<ROWS>() * V, V, unsigned(site));
    }
};

template <typename R, typename S> LYM_DI su3<R> lds_link(const F2<S> *p, int stride)
{
    su3<R> m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const F2<S> v = p[k * stride];
        m.e[k] = cplx<R>{R(v.x), R(v.y)};
    }
    return m;
}

// The four links of the run's 64 sites staged in LDS:
// lds[(dir*9 + k)*64 + (site - site0)], consecutive lanes at consecutive
// addresses (no bank conflicts).  Reads at the site itself, and along x when
// XROWS (runs hold whole x-rows, Lx | 64), come from LDS.  Per link update
// 3 of the 19 staple reads are at the site itself and, with XROWS, another 3
// are along x.
template <typename S, bool XROWS, int ROWS = 3> struct RunLinks {
    static constexpr int rows = ROWS;
    const F2<S> *f;
    long V;
    long site0;
    const F2<S> *lds;
    template <typename R, int DX, int DY, int ZT, int FLIP = 0> LYM_DI su3<R> get(int dir, long site) const
    {
        if constexpr (ZT == 0 && DY == 0 && (DX == 0 || XROWS)) {
            unsigned i = unsigned(site - site0);
            asm volatile("" : "+v"(i));
            return lds_link<R, S>(lds + dir * 9 * kSitesPerBlock + i, kSitesPerBlock);
        } else
            return load_link_r<R, S, ROWS>(f + (long)dir * plane_elems<ROWS>() * V, V, unsigned(site));
    }
};

// Slab staging for runs of R = 64 / Lx whole x-rows (Lx | 64, Lx <= 32, and
// Ly a multiple of R so that a run never crosses a z-plane): the run's rows
// and one halo row in y (periodic) are staged in LDS, per direction the halo
// row that staple reads use most.  No staple reads U_y one row up (y0 + R),
// and U_x, U_z, U_t one row down (y0 - 1) only once per site (the lower
// staple of plane (mu, y) for mu = x, z, t), against two or three reads of
// the row up; so direction y keeps rows -1 .. R-1 and the others rows 0 .. R,
// and the reads one row down of x, z and t go to HBM.  Direction d's element
// planes are kSlabSites sites apart, row r at (r - slab_row0(d)) * Lx:
//   slab[(d*9 + k)*kSlabSites + (r - slab_row0(d))*Lx + x].
// 27.6 KB per workgroup in FP32 instead of 36.9 KB with both halo rows of
// every direction, so five workgroups share a CU instead of four (the FP32
// slab kernel is compiled for the registers of five waves).  Every
// other read whose offset has no z or t part comes from LDS: 9.75 of the 19
// reads per link update on average, against 7 with XROWS alone.
constexpr int kSlabSites = 96;  // (R + 1) * Lx = 64 + Lx <= 96
constexpr int slab_row0(int d) { return d == 1 ? -1 : 0; }
constexpr bool slab_staged(int d, int dy) { return d == 1 ? dy <= 0 : dy >= 0; }

// Whether a lattice takes the slab staging: Lx | 64, Lx <= 32, Ly % (64 / Lx) == 0.
inline bool slab_fits(const int dims[4])
{
    const int lx = dims[0];
    return lx >= 2 && lx <= 32 && kSitesPerBlock % lx == 0 && dims[1] % (kSitesPerBlock / lx) == 0;
}
template <typename S, int ROWS = 3> struct SlabLinks {
    static constexpr int rows = ROWS;
    const F2<S> *f;
    long V;
    const F2<S> *slab;
    int lx_mask, row, x;  // Lx - 1; this lane's row in the run and x coordinate
    // dir is a constant at every call site once inlined, so the choice of
    // LDS or HBM folds away.
    template <typename R, int DX, int DY, int ZT, int FLIP = 0> LYM_DI su3<R> get(int dir, long site) const
    {
        if constexpr (ZT == 0) {
            if (slab_staged(dir, DY)) {
                int li = (row - slab_row0(dir) + DY) * (lx_mask + 1) + ((x + DX) & lx_mask);
                asm volatile("" : "+v"(li));
                return lds_link<R, S>(slab + dir * 9 * kSlabSites + li, kSlabSites);
            }
        }
        return load_link_r<R, S, ROWS>(f + (long)dir * plane_elems<ROWS>() * V, V, unsigned(site));
    }
};

// Links in the even-odd order of the Metropolis sweeps: the link of site s,
// of parity P(s) = (x + y + z + t) & 1, sits at P(s) V/2 + s/2.  With Lx even,
// s/2 is the rank of s among the sites of its parity in lexicographic order,
// so a wave that updates 64 consecutive sites of one parity reads 64
// consecutive complexes per matrix element, for its own links and for every
// neighbour; in lexicographic order it would use every second one.
template <typename S> struct EOLinks {
    const F2<S> *f;
    long V;
    unsigned half;  // V / 2
    unsigned par;   // parity of the site being updated
    template <typename R, int DX, int DY, int ZT, int FLIP> LYM_DI su3<R> get(int dir, long site) const
    {
        const unsigned base = (par ^ unsigned(FLIP)) ? half : 0u;
        return load_link<R, S>(f + (long)dir * 9 * V, V, base + (unsigned(site) >> 1));
    }
};

// Tile staging for lattices whose x-rows do not divide a run (Lx = 48, 80,
// ...): the 64 sites of a workgroup are a kTileX x kTileY = 16 x 4 block of
// one (z, t) plane, and the block with a one-site halo ring in x and y
// (periodic) is staged as tile[(dir*9 + k)*kTileSites + (ly + 1)*(kTileX + 2)
// + lx + 1], ly = -1 .. 4, lx = -1 .. 16: 108 sites.  Every read whose offset
// has no z or t part comes from LDS, as with the slab.  A tile row of one
// element is a whole 128-byte line in FP32; narrower tiles (8 x 8 for
// Lx = 24) were measured slower than staging own links only.
constexpr int kTileX = 16, kTileY = kSitesPerBlock / kTileX;
constexpr int kTileSites = 108;  // (kTileX + 2) * (kTileY + 2)
inline bool tile_fits(const int dims[4]) { return dims[0] % kTileX == 0 && dims[1] % kTileY == 0; }
template <typename S, int ROWS = 3> struct TileLinks {
    static constexpr int rows = ROWS;
    const F2<S> *f;
    long V;
    const F2<S> *tile;
    int self;  // this lane's slot (ly + 1) * (kTileX + 2) + lx + 1
    template <typename R, int DX, int DY, int ZT, int FLIP = 0> LYM_DI su3<R> get(int dir, long site) const
    {
        if constexpr (ZT == 0) {
            int base = self;
            asm volatile("" : "+v"(base));
            return lds_link<R, S>(tile + dir * 9 * kTileSites + base + DY * (kTileX + 2) + DX, kTileSites);
        } else
            return load_link_r<R, S, ROWS>(f + (long)dir * plane_elems<ROWS>() * V, V, unsigned(site));
    }
};

// ---------------------------------------------------------------- staple ---

// The two staples of plane (MU, NU) for link U_MU(x):
//   U_NU(x+MU) U_MU(x+NU)^+ U_NU(x)^+  +  U_NU(x+MU-NU)^+ U_MU(x-NU)^+ U_NU(x-NU)
// In FP32 (the mixed path, whose links are reunitarised after every stage)
// each staple is formed from its first two rows and completed by unitarity:
// 186 instead of 216 arithmetic operations per staple, with a deviation of
// the order of the links' own distance from SU(3) (FP32 rounding).  The FP64
// path multiplies out all three rows as the reference does.
template <typename R, int MU, int NU, typename Src>
LYM_DI void staple_plane(const Src &src, long s, const Nbr &n, su3<R> &om)
{
    // offset tags (unit vectors of MU and NU split into x, y and z/t parts)
    constexpr int MX = MU == 0, MY = MU == 1, MZ = MU >= 2, NX = NU == 0, NY = NU == 1, NZ = NU >= 2;
    const long y = s + n.bwd[NU];  // x - NU
    if constexpr (sizeof(R) == 4) {
        {
            su3<R> a = src.template get<R, MX, MY, MZ, 1>(NU, s + n.fwd[MU]);
            su3<R> b = src.template get<R, NX, NY, NZ, 1>(MU, s + n.fwd[NU]);
            su3<R> t, u;
            mul_na_rows01(t, a, b);
            su3<R> c = src.template get<R, 0, 0, 0, 0>(NU, s);
            mul_na_rows01(u, t, c);
            acc_su3_rows01(om, u);
        }
        {
            su3<R> a = src.template get<R, -NX, -NY, NZ, 1>(MU, y);
            su3<R> b = src.template get<R, MX - NX, MY - NY, MZ | NZ, 0>(NU, y + n.fwd[MU]);
            su3<R> t, u;
            mul_nn_cols01(t, a, b);  // (a b)^+ rows 0, 1 need columns 0, 1 of a b
            su3<R> c = src.template get<R, -NX, -NY, NZ, 1>(NU, y);
            mul_an_rows01(u, t, c);
            acc_su3_rows01(om, u);
        }
        return;
    }
    {
        su3<R> a = src.template get<R, MX, MY, MZ, 1>(NU, s + n.fwd[MU]);
        su3<R> b = src.template get<R, NX, NY, NZ, 1>(MU, s + n.fwd[NU]);
        su3<R> t = mul_na(a, b);
        su3<R> c = src.template get<R, 0, 0, 0, 0>(NU, s);
        acc_na(om, t, c);
    }
    {
        su3<R> a = src.template get<R, -NX, -NY, NZ, 1>(MU, y);
        su3<R> b = src.template get<R, MX - NX, MY - NY, MZ | NZ, 0>(NU, y + n.fwd[MU]);
        su3<R> t = mul_nn(a, b);
        su3<R> c = src.template get<R, -NX, -NY, NZ, 1>(NU, y);
        acc_an(om, t, c);
    }
}

// Omega_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                              + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) ]
// (src/Actions/puregauge.cpp:36-54), so that U_mu(x) Omega_mu(x) is the sum
// of the six plaquettes through the link.
template <typename R, int MU, typename Src>
LYM_DI su3<R> staple_from(const Src &src, long s, const Nbr &n)
{
    su3<R> om = zero3<R>();
    if constexpr (MU != 0) staple_plane<R, MU, 0>(src, s, n, om);
    if constexpr (MU != 1) staple_plane<R, MU, 1>(src, s, n, om);
    if constexpr (MU != 2) staple_plane<R, MU, 2>(src, s, n, om);
    if constexpr (MU != 3) staple_plane<R, MU, 3>(src, s, n, om);
    return om;
}

template <typename R, typename S, int MU>
LYM_DI su3<R> staple(const F2<S> *__restrict__ f, long V, long s, const Nbr &n)
{
    return staple_from<R, MU>(GlobalLinks<S>{f, V}, s, n);
}

// ---------------------------------------------------- flow: one RK stage ---

// One Runge-Kutta-Munthe-Kaas stage for one link (src/Apps/wilsonflow.cpp:141-172):
//   STAGE 1: acc = Z(W0)                     ; W1 = exp(i eps/4 acc) W0
//   STAGE 2: acc = 8/9 Z(W1) - 17/36 acc     ; W2 = exp(i eps acc) W1
//   STAGE 3: q   = 3/4 Z(W2) - acc           ; W3 = exp(i eps q) W2
// Links are read from win and written to wout (all four directions of a
// stage see the same input field, as the reference computes all four Z
// before any update).  Staple and force run in the storage precision S;
// the RK combination, the exponential and exp*U run in FP64.  REUNIT
// projects the result back onto SU(3) before it is stored (mixed path).
// n: the neighbour offsets of site s.
template <typename S, int STAGE, bool REUNIT, int MU, typename Src>
LYM_DI void flow_link(const Src &src, F2<S> *__restrict__ wout, S *__restrict__ acc, const Geom g, long s,
                      const Nbr &n, double eps)
{
    const long V = g.V;
    su3<S> om = staple_from<S, MU>(src, s, n);
    su3<S> U = src.template get<S, 0, 0, 0>(MU, s);
    // Z of X = U Omega and the RK combination in the storage precision S
    const herm8<S> z = force_from_product<S, S>(U, om);
    S *am = acc + (long)MU * 8 * V;
    herm8<double> q;
    if constexpr (STAGE == 1) {
        store_h8<S, kStreamAux>(am, V, unsigned(s), z);
        q = h8_scale(eps * 0.25, z);
    } else if constexpr (STAGE == 2) {
        herm8<S> a = load_h8<S, kStreamAux>(am, V, unsigned(s));
        a = h8_axpby(S(8.0 / 9.0), z, S(-17.0 / 36.0), a);
        store_h8<S, kStreamAux>(am, V, unsigned(s), a);
        q = h8_scale(eps, a);
    } else {
        const herm8<S> a = load_h8<S, kStreamAux>(am, V, unsigned(s));
        q = h8_scale(eps, h8_axpby(S(0.75), z, S(-1.0), a));
    }
    su3<double> E = (sizeof(S) == 4) ? exp_iq<kShortSeriesTermsF32, kSeriesTermsF32>(q) : exp_iq(q);
    su3<double> W = mul_nn(E, convert<double>(U));
    if constexpr (REUNIT) reunitarize<sizeof(S) == 4>(W);
    constexpr int ROWS = Src::rows;  // the output field takes the input's form
    store_link_r<S, double, kStreamAux, ROWS>(wout + (long)MU * plane_elems<ROWS>() * V, V, unsigned(s), W);
}

// One 64-site run, all four directions (wave = direction).  LDS selects what
// is staged in LDS before the staples (each wave stages its own direction):
//   0  nothing
//   1  the run's own links (RunLinks, reads at the site itself)
//   2  the same, runs holding whole x-rows (reads along x too)
//   3  the run's rows plus one halo row on each side in y (SlabLinks, every
//      read without a z or t part)
//   5  run r is tile r instead (TileLinks: a 16 x 4 block of one (z, t)
//      plane plus a halo ring in x and y; every read without a z or t part)
// All 256 threads reach the barrier, then lanes past the end of the lattice
// drop out.  The staging buffer lds (the kernel's LDS array, at least
// stage_lds_elems<LDS>() complexes) is free again when the function returns
// (trailing barrier), so a persistent caller may loop, and a kernel that
// picks the mode per workgroup passes every mode the same array.
template <int LDS> constexpr int stage_lds_elems()
{
    return LDS == 0 ? 1 : LDS == 3 ? 4 * 9 * kSlabSites : LDS == 5 ? 4 * 9 * kTileSites : 4 * 9 * kSitesPerBlock;
}
// Waves per SIMD the register allocation of a stage kernel must allow: the
// FP32 slab kernel fits five workgroups' LDS on a CU, so it is held to the
// registers of five waves; the others take what they need.
template <typename S, int LDS> constexpr int stage_min_waves() { return sizeof(S) == 4 && (LDS == 3 || LDS == 5) ? 5 : 1; }

// tid: the thread's index in the workgroup (threadIdx.x; a persistent caller
// passes it through an opaque copy, so that values derived from it are formed
// per run and not held in registers across its loop).
template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS = 3>
LYM_DI void flow_block(const F2<S> *__restrict__ win, F2<S> *__restrict__ wout, S *__restrict__ acc,
                       const Geom g, long site0, double eps, F2<S> *__restrict__ lds, unsigned tid = threadIdx.x)
{
    const int mu = __builtin_amdgcn_readfirstlane(int(tid >> 6));  // uniform over the wave
    const int lane = tid & 63;
    if constexpr (LDS == 5) {
        // Tiles cover the lattice exactly (checked on the host), so every lane
        // is a site.  Tile r = tx + ntx (ty + nty zt), zt = z + Lz t: the same
        // (z, t) plane as run r, so the run schedules keep their locality.
        // Wave mu stages direction mu: lane l its own link, and lanes below
        // 2 (TX + 2) + 2 TY = 44 one halo site each (rows -1 and TY, then
        // columns -1 and TX).
        constexpr int TX = kTileX, TY = kTileY, W = TX + 2;
        F2<S> *const tile = lds;
        const int Lx = g.L[0], Ly = g.L[1];
        const uint32_t r = uint32_t(site0 / kSitesPerBlock), ntx = uint32_t(Lx / TX), nty = uint32_t(Ly / TY);
        const uint32_t q = r / ntx, tx = r - q * ntx, zt = q / nty, ty = q - zt * nty;
        const int x0 = int(tx) * TX, y0 = int(ty) * TY;
        const long pb = (long)zt * Lx * Ly;
        const int lx = lane % TX, ly = lane / TX;
        const long s = pb + x0 + lx + (long)Lx * (y0 + ly);
        const F2<S> *plane = win + (long)mu * plane_elems<ROWS>() * g.V;
        F2<S> *dst = tile + mu * 9 * kTileSites;
        const int self = (ly + 1) * W + lx + 1;
        {
            const su3<S> u = load_link_r<S, S, ROWS>(plane, g.V, unsigned(s));
#pragma unroll
            for (int k = 0; k < 9; k++) {
                dst[k * kTileSites + self].x = u.e[k].re;
                dst[k * kTileSites + self].y = u.e[k].im;
            }
        }
        if (lane < 2 * W + 2 * TY) {
            int hx, hy;
            if (lane < W) {
                hx = lane - 1;
                hy = -1;
            } else if (lane < 2 * W) {
                hx = lane - W - 1;
                hy = TY;
            } else if (lane < 2 * W + TY) {
                hx = -1;
                hy = lane - 2 * W;
            } else {
                hx = TX;
                hy = lane - 2 * W - TY;
            }
            int gx = x0 + hx, gy = y0 + hy;
            gx += (gx < 0) ? Lx : 0;
            gx -= (gx >= Lx) ? Lx : 0;
            gy += (gy < 0) ? Ly : 0;
            gy -= (gy >= Ly) ? Ly : 0;
            const su3<S> u = load_link_r<S, S, ROWS>(plane, g.V, unsigned(pb + gx + (long)Lx * gy));
            const int li = (hy + 1) * W + hx + 1;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                dst[k * kTileSites + li].x = u.e[k].re;
                dst[k * kTileSites + li].y = u.e[k].im;
            }
        }
        __syncthreads();
        const TileLinks<S, ROWS> src{win, g.V, tile, self};
        // zt is uniform over the workgroup: z and t cost no per-lane division
        const int c[4] = {x0 + lx, y0 + ly, int(zt % uint32_t(g.L[2])), int(zt / uint32_t(g.L[2]))};
        const Nbr n = neighbours_at(g, c);
        switch (mu) {
        case 0: flow_link<S, STAGE, REUNIT, 0>(src, wout, acc, g, s, n, eps); break;
        case 1: flow_link<S, STAGE, REUNIT, 1>(src, wout, acc, g, s, n, eps); break;
        case 2: flow_link<S, STAGE, REUNIT, 2>(src, wout, acc, g, s, n, eps); break;
        default: flow_link<S, STAGE, REUNIT, 3>(src, wout, acc, g, s, n, eps); break;
        }
        __syncthreads();
        return;
    }
    const long s = site0 + lane;
    const bool active = s < g.V;
    if constexpr (LDS == 3) {
        // Slab: runs are whole x-rows and never cross a z-plane, so every lane
        // is a lattice site (checked on the host).  Wave mu stages direction mu:
        // the run's rows 0 .. R-1, and of the rows y0 - 1 and y0 + R (periodic
        // in y) those that direction keeps (SlabLinks).
        F2<S> *const slab = lds;
        const int Lx = g.L[0], Ly = g.L[1], R = kSitesPerBlock / Lx;
        const uint32_t rows0 = uint32_t(site0) / uint32_t(Lx), zt = rows0 / uint32_t(Ly);  // uniform
        const int y0 = int(rows0 - zt * uint32_t(Ly));
        const long lo = (y0 > 0) ? site0 - Lx : site0 - Lx + long(Ly) * Lx;
        const long hi = (y0 + R < Ly) ? site0 + kSitesPerBlock : site0 + kSitesPerBlock - long(Ly) * Lx;
        const F2<S> *plane = win + (long)mu * plane_elems<ROWS>() * g.V;
        F2<S> *dst = slab + mu * 9 * kSlabSites;
        const int r0 = slab_row0(mu);
        {
            const su3<S> u = load_link_r<S, S, ROWS>(plane, g.V, unsigned(s));
            const int li = -r0 * Lx + lane;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                dst[k * kSlabSites + li].x = u.e[k].re;
                dst[k * kSlabSites + li].y = u.e[k].im;
            }
        }
        if (lane < Lx) {  // the halo row: y0 - 1 for direction y, y0 + R for the others
            const long hs = (mu == 1 ? lo : hi) + lane;
            const int li = (mu == 1 ? 0 : R * Lx) + lane;
            const su3<S> u = load_link_r<S, S, ROWS>(plane, g.V, unsigned(hs));
#pragma unroll
            for (int k = 0; k < 9; k++) {
                dst[k * kSlabSites + li].x = u.e[k].re;
                dst[k * kSlabSites + li].y = u.e[k].im;
            }
        }
        __syncthreads();
        const SlabLinks<S, ROWS> src{win, g.V, slab, Lx - 1, lane >> __builtin_ctz(unsigned(Lx)), lane & (Lx - 1)};
        const int c[4] = {src.x, y0 + src.row, int(zt % uint32_t(g.L[2])), int(zt / uint32_t(g.L[2]))};
        const Nbr n = neighbours_at(g, c);
        switch (mu) {
        case 0: flow_link<S, STAGE, REUNIT, 0>(src, wout, acc, g, s, n, eps); break;
        case 1: flow_link<S, STAGE, REUNIT, 1>(src, wout, acc, g, s, n, eps); break;
        case 2: flow_link<S, STAGE, REUNIT, 2>(src, wout, acc, g, s, n, eps); break;
        default: flow_link<S, STAGE, REUNIT, 3>(src, wout, acc, g, s, n, eps); break;
        }
        __syncthreads();
    } else if constexpr (LDS != 0) {
        if (active) {
            const su3<S> u = load_link_r<S, S, ROWS>(win + (long)mu * plane_elems<ROWS>() * g.V, g.V, unsigned(s));
#pragma unroll
            for (int k = 0; k < 9; k++) {
                F2<S> v;
                v.x = u.e[k].re;
                v.y = u.e[k].im;
                lds[(mu * 9 + k) * kSitesPerBlock + lane] = v;
            }
        }
        __syncthreads();
        if (active) {
            const RunLinks<S, LDS == 2, ROWS> src{win, g.V, site0, lds};
            const Nbr n = neighbours(g, s);
            switch (mu) {
            case 0: flow_link<S, STAGE, REUNIT, 0>(src, wout, acc, g, s, n, eps); break;
            case 1: flow_link<S, STAGE, REUNIT, 1>(src, wout, acc, g, s, n, eps); break;
            case 2: flow_link<S, STAGE, REUNIT, 2>(src, wout, acc, g, s, n, eps); break;
            default: flow_link<S, STAGE, REUNIT, 3>(src, wout, acc, g, s, n, eps); break;
            }
        }
        __syncthreads();
    } else {
        if (!active) return;
        const GlobalLinks<S, ROWS> src{win, g.V};
        const Nbr n = neighbours(g, s);
        switch (mu) {
        case 0: flow_link<S, STAGE, REUNIT, 0>(src, wout, acc, g, s, n, eps); break;
        case 1: flow_link<S, STAGE, REUNIT, 1>(src, wout, acc, g, s, n, eps); break;
        case 2: flow_link<S, STAGE, REUNIT, 2>(src, wout, acc, g, s, n, eps); break;
        default: flow_link<S, STAGE, REUNIT, 3>(src, wout, acc, g, s, n, eps); break;
        }
    }
}

template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS = 3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_min_waves<S, LDS>())))
void k_flow_stage(const F2<S> *__restrict__ win, F2<S> *__restrict__ wout, S *__restrict__ acc, Geom g, Sched sc,
                  double eps)
{
    const long run = sched_run(sc, blockIdx.x);  // uniform over the workgroup
    if (run < 0) return;
    __shared__ F2<S> lds[stage_lds_elems<LDS>()];
    flow_block<S, STAGE, REUNIT, LDS, ROWS>(win, wout, acc, g, run * kSitesPerBlock, eps, lds);
}

// ----------------------------------------------------------- observables ---

// Clover of plane (MU, NU) at site s (src/Observables/superobs.cpp:40-106):
// clover = P_{mu nu}(x) + [(-mu,-nu) leaf] - [(+mu,-nu) leaf] - [(-mu,+nu) leaf]
// with each leaf formed as in the reference; returns C = [clover - clover^+]
// made traceless (8 reals), and Re Tr P_{mu nu}(x) in *plaq.  In FP32 (mixed
// path) each leaf, a product of four SU(3) links, is formed from its first two
// rows and completed by unitarity (240 instead of 324 operations per leaf);
// FP64 multiplies out all three rows as the reference does.
template <typename R, typename S, int MU, int NU, int ROWS = 3>
LYM_DI ah8<R> clover(const F2<S> *__restrict__ f, long V, long s, const Nbr &n, R *plaq)
{
    const F2<S> *Fm = f + (long)MU * plane_elems<ROWS>() * V;
    const F2<S> *Fn = f + (long)NU * plane_elems<ROWS>() * V;
    auto load_link = [](const F2<S> *plane, long V, long site) { return load_link_r<R, S, ROWS>(plane, V, unsigned(site)); };
    constexpr bool TWO_ROW = sizeof(R) == 4;
    auto add = [](su3<R> &cl, const su3<R> &t, R sign) {  // sign = +-1: one rounding, as cl +- t
#pragma unroll
        for (int k = 0; k < 9; k++)
            cl.e[k] = cplx<R>{cl.e[k].re + sign * t.e[k].re, cl.e[k].im + sign * t.e[k].im};
    };
    su3<R> cl;
    {  // P = ((U_mu(x) U_nu(x+mu)) U_mu(x+nu)^+) U_nu(x)^+
        if constexpr (TWO_ROW) {
            su3<R> t, u;
            mul_nn_rows01(t, load_link(Fm, V, s), load_link(Fn, V, s + n.fwd[MU]));
            mul_na_rows01(u, t, load_link(Fm, V, s + n.fwd[NU]));
            mul_na_rows01(cl, u, load_link(Fn, V, s));
            complete_row2(cl);
        } else {
            su3<R> t = mul_nn(load_link(Fm, V, s), load_link(Fn, V, s + n.fwd[MU]));
            t = mul_na(t, load_link(Fm, V, s + n.fwd[NU]));
            cl = mul_na(t, load_link(Fn, V, s));
        }
        *plaq = retrace(cl);
    }
    {  // y = x-mu-nu: ((U_mu(y+nu)^+ U_nu(y)^+) U_mu(y)) U_nu(y+mu) = (U_nu(y) U_mu(x-mu))^+ U_mu(y) U_nu(x-nu)
        const long y = s + n.bwd[MU] + n.bwd[NU];
        su3<R> t;
        if constexpr (TWO_ROW) {
            su3<R> p, u;
            mul_nn_cols01(p, load_link(Fn, V, y), load_link(Fm, V, s + n.bwd[MU]));
            mul_an_rows01(u, p, load_link(Fm, V, y));
            mul_nn_rows01(t, u, load_link(Fn, V, s + n.bwd[NU]));
            complete_row2(t);
        } else {
            t = mul_nn(load_link(Fn, V, y), load_link(Fm, V, s + n.bwd[MU]));
            t = mul_an(t, load_link(Fm, V, y));
            t = mul_nn(t, load_link(Fn, V, s + n.bwd[NU]));
        }
        add(cl, t, R(1));
    }
    {  // y = x-nu: ((U_mu(x) U_nu(x+mu-nu)^+) U_mu(y)^+) U_nu(y)
        const long y = s + n.bwd[NU];
        su3<R> t;
        if constexpr (TWO_ROW) {
            su3<R> p, u;
            mul_na_rows01(p, load_link(Fm, V, s), load_link(Fn, V, s + n.fwd[MU] + n.bwd[NU]));
            mul_na_rows01(u, p, load_link(Fm, V, y));
            mul_nn_rows01(t, u, load_link(Fn, V, y));
            complete_row2(t);
        } else {
            t = mul_na(load_link(Fm, V, s), load_link(Fn, V, s + n.fwd[MU] + n.bwd[NU]));
            t = mul_na(t, load_link(Fm, V, y));
            t = mul_nn(t, load_link(Fn, V, y));
        }
        add(cl, t, R(-1));
    }
    {  // y = x-mu: ((U_mu(y)^+ U_nu(y)) U_mu(x-mu+nu)) U_nu(x)^+
        const long y = s + n.bwd[MU];
        su3<R> t;
        if constexpr (TWO_ROW) {
            su3<R> p, u;
            mul_an_rows01(p, load_link(Fm, V, y), load_link(Fn, V, y));
            mul_nn_rows01(u, p, load_link(Fm, V, y + n.fwd[NU]));
            mul_na_rows01(t, u, load_link(Fn, V, s));
            complete_row2(t);
        } else {
            t = mul_an(load_link(Fm, V, y), load_link(Fn, V, y));
            t = mul_nn(t, load_link(Fm, V, y + n.fwd[NU]));
            t = mul_na(t, load_link(Fn, V, s));
        }
        add(cl, t, R(-1));
    }
    // C = clover - clover^+ (anti-Hermitian), made traceless
    ah8<R> C;
    const R i0 = R(2) * cl.e[0].im, i1 = R(2) * cl.e[4].im, i2 = R(2) * cl.e[8].im;
    const R tr = (i0 + i1 + i2) * R(1.0 / 3.0);
    C.a0 = i0 - tr;
    C.a1 = i1 - tr;
    C.a2 = i2 - tr;
    C.o01 = cplx<R>{cl.e[1].re - cl.e[3].re, cl.e[1].im + cl.e[3].im};
    C.o02 = cplx<R>{cl.e[2].re - cl.e[6].re, cl.e[2].im + cl.e[6].im};
    C.o12 = cplx<R>{cl.e[5].re - cl.e[7].re, cl.e[5].im + cl.e[7].im};
    return C;
}

// Re Tr(A B) for anti-Hermitian A, B given by diagonal imaginary parts a_i and
// upper off-diagonals (A_ji = -conj(A_ij)):  -sum_i a_i b_i - 2 sum_{i<j} Re(A_ij conj(B_ij)).
// Re Tr(A B^+) = -Re Tr(A B).
template <typename R> LYM_DI R retr_ah(const ah8<R> &A, const ah8<R> &B)
{
    R d = A.a0 * B.a0 + A.a1 * B.a1 + A.a2 * B.a2;
    R o = A.o01.re * B.o01.re + A.o01.im * B.o01.im + A.o02.re * B.o02.re + A.o02.im * B.o02.im +
          A.o12.re * B.o12.re + A.o12.im * B.o12.im;
    return -d - R(2) * o;
}

// Contributions of one plane pair at one site (src/Observables/superobs.cpp:40-124):
// PAIR 0 = planes (01, 23), 1 = (02, 13), 2 = (03, 12).  out[0] = Re Tr P over
// both planes, out[1] = Re Tr(C C) over both planes (the reference's raw
// energy terms, superobs.cpp:109), out[2] = this pair's term of
// q(x) = -[ReTr(C01 C23) + ReTr(C02 C13^+) + ReTr(C03 C12)] (superobs.cpp:113-124).
template <typename R, typename S, int PAIR, int ROWS = 3>
LYM_DI void pair_observables(const F2<S> *__restrict__ f, const Geom g, long s, double out[3])
{
    constexpr int M1 = 0, N1 = PAIR + 1;
    constexpr int M2 = (PAIR == 2) ? 1 : (PAIR == 0 ? 2 : 1), N2 = (PAIR == 0) ? 3 : (PAIR == 1 ? 3 : 2);
    const long V = g.V;
    Nbr n = neighbours(g, s);
    R p1, p2;
    ah8<R> A = clover<R, S, M1, N1, ROWS>(f, V, s, n, &p1);
    ah8<R> B = clover<R, S, M2, N2, ROWS>(f, V, s, n, &p2);
    out[0] = double(p1) + double(p2);
    out[1] = double(retr_ah(A, A)) + double(retr_ah(B, B));
    const double t = double(retr_ah(A, B));
    out[2] = (PAIR == 1) ? t : -t;  // -Re Tr(C02 C13^+) = +Re Tr(C02 C13)
}

// Block reduction of 4 compensated sums; thread 0 returns the block total.
template <int NT> LYM_DI void block_ksum4(ksum v[4], ksum *out)
{
    __shared__ ksum sh[NT / 64][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = wave_ksum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) sh[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            ksum t = sh[0][k];
            for (int i = 1; i < NT / 64; i++) t = kadd(t, sh[i][k]);
            out[k] = t;
        }
    }
}

// 192-thread workgroups: 64 consecutive sites x 3 plane pairs (wave = pair).
// Per-block compensated partials of {sum Re Tr P, sum Re Tr C^2, sum q,
// sum |q terms|}.
constexpr int kObsSites = 64;

template <typename S, int ROWS = 3>
__global__ __launch_bounds__(192) void k_observables(const F2<S> *__restrict__ f, Geom g, Sched sc,
                                                     ksum *__restrict__ partials)
{
    const long run = sched_run(sc, blockIdx.x);
    const long s = run * kObsSites + (threadIdx.x & 63);
    const int pair = threadIdx.x >> 6;
    ksum v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (run >= 0 && s < g.V) {
        double o[3];
        if (pair == 0)
            pair_observables<S, S, 0, ROWS>(f, g, s, o);
        else if (pair == 1)
            pair_observables<S, S, 1, ROWS>(f, g, s, o);
        else
            pair_observables<S, S, 2, ROWS>(f, g, s, o);
        v[0].s = o[0];
        v[1].s = o[1];
        v[2].s = o[2];
        v[3].s = fabs(o[2]);
    }
    block_ksum4<192>(v, partials + 4 * blockIdx.x);
}

// Sum nblk block partials into dst[0..3] (one workgroup of 1024 threads).
constexpr int kReduceThreads = 1024;
static __global__ __launch_bounds__(kReduceThreads) void k_reduce_partials(const ksum *__restrict__ partials, int nblk,
                                                                    double *__restrict__ dst)
{
    ksum v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int i = threadIdx.x; i < nblk; i += kReduceThreads)
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = kadd(v[k], partials[4 * i + k]);
    __shared__ ksum tot[4];
    block_ksum4<kReduceThreads>(v, tot);
    __syncthreads();
    if (threadIdx.x < 4) dst[threadIdx.x] = tot[threadIdx.x].s + tot[threadIdx.x].c;
}

// ------------------------------------------------ layout and start fields ---

// streaming copy b = a (grid-stride), the HBM calibration kernel
static __global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = a[i];
}

// clears counters with a kernel (a memset node misbehaves on graph replays)
static __global__ void k_zero_u64(unsigned long long *p, int n)
{
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}

// Link health: out[0] = max over links of max_ij |(U U^+ - 1)_ij| and
// |det U - 1| as the bit pattern of a non-negative double (which orders like
// its value, so an unsigned 64-bit max is a max of the values); out[1] =
// links with a non-finite entry.  out[] must be zero before the launch.
template <typename S>
static __global__ __launch_bounds__(256) void k_check_links(const F2<S> *__restrict__ f, long V,
                                                           unsigned long long *__restrict__ out)
{
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;  // link mu*V + s (4V < 2^31)
    double dev = 0.0;
    bool bad = false;
    if (l < uint32_t(4 * V)) {
        const uint32_t mu = l / uint32_t(V), s = l - mu * uint32_t(V);
        const su3<double> u = load_link<double, S>(f + long(mu) * 9 * V, V, s);
#pragma unroll
        for (int k = 0; k < 9; k++) bad = bad || !isfinite(u.e[k].re) || !isfinite(u.e[k].im);
        const su3<double> p = mul_na(u, u);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double re = p.e[3 * i + j].re - (i == j ? 1.0 : 0.0), im = p.e[3 * i + j].im;
                dev = fmax(dev, sqrt(re * re + im * im));
            }
        // det U = u0 . (u1 x u2) over rows
        cplx<double> det{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int k = (j + 1) % 3, m = (j + 2) % 3;
            cplx<double> c = cfma(u.e[3 + k], u.e[6 + m], cplx<double>{0.0, 0.0});
            c = cfma(cplx<double>{-u.e[3 + m].re, -u.e[3 + m].im}, u.e[6 + k], c);
            det = cfma(u.e[j], c, det);
        }
        dev = fmax(dev, sqrt((det.re - 1.0) * (det.re - 1.0) + det.im * det.im));
        if (bad) dev = 0.0;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) dev = fmax(dev, __shfl_xor(dev, m, 64));
    const unsigned long long nbad = __popcll(__ballot(bad));
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, (unsigned long long)__double_as_longlong(dev));
        if (nbad) atomicAdd(out + 1, nbad);
    }
}

// Three-row field -> two-row field (the first two rows of every link);
// blockIdx.y = direction.
template <typename S>
__global__ __launch_bounds__(256) void k_to_rows2(const F2<S> *__restrict__ in, F2<S> *__restrict__ out, long V)
{
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    if (s >= V) return;
    const long mu = blockIdx.y;
#pragma unroll
    for (int k = 0; k < 6; k++) out[(mu * 6 + k) * V + s] = in[(mu * 9 + k) * V + s];
}

// Two-row field -> three-row field: row 2 is the one load_link_r rebuilds
// and store_link stores, so the conversion loses nothing.
template <typename S>
__global__ __launch_bounds__(256) void k_from_rows2(const F2<S> *__restrict__ in, F2<S> *__restrict__ out, long V)
{
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    if (s >= V) return;
    const long mu = blockIdx.y;
    const su3<S> m = load_link_r<S, S, 2>(in + mu * 6 * V, V, unsigned(s));
    store_link<S, S>(out + mu * 9 * V, V, unsigned(s), m);
}

// reference layout (V x 18 doubles for one mu) -> SoA plane of direction mu.
// FP64 fields take every element as it is; FP32 fields take rows 0, 1
// rounded and derive row 2 from them (store_link).
template <typename S>
__global__ void k_aos_to_soa(const double *__restrict__ aos, F2<S> *__restrict__ plane, long V)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    if constexpr (sizeof(S) == 4) {
        su3<double> m;
#pragma unroll
        for (int k = 0; k < 9; k++) m.e[k] = cplx<double>{k < 6 ? aos[s * 18 + 2 * k] : 0.0, k < 6 ? aos[s * 18 + 2 * k + 1] : 0.0};
        store_link<S, double>(plane, V, unsigned(s), m);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            F2<S> v;
            v.x = S(aos[s * 18 + 2 * k]);
            v.y = S(aos[s * 18 + 2 * k + 1]);
            plane[k * V + s] = v;
        }
    }
}

// reference layout (V x 18 doubles for one mu) -> two-row plane of direction
// mu (rows 0, 1 rounded to S; the row 2 a reader rebuilds is the one an FP32
// three-row field would hold)
template <typename S>
__global__ void k_aos_to_rows2(const double *__restrict__ aos, F2<S> *__restrict__ plane, long V)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    su3<double> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = cplx<double>{k < 6 ? aos[s * 18 + 2 * k] : 0.0, k < 6 ? aos[s * 18 + 2 * k + 1] : 0.0};
    store_link_r<S, double, 0, 2>(plane, V, unsigned(s), m);
}

template <typename S>
__global__ void k_soa_to_aos(const F2<S> *__restrict__ plane, double *__restrict__ aos, long V)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        F2<S> v = plane[k * V + s];
        aos[s * 18 + 2 * k] = double(v.x);
        aos[s * 18 + 2 * k + 1] = double(v.y);
    }
}

// all links = 1 (the reference's setToIdentity, src/Math/lattice.cpp:33-41)
template <typename S> __global__ void k_cold(F2<S> *__restrict__ f, long V)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= 4 * V) return;
    const long mu = l / V, s = l % V;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        F2<S> v;
        v.x = (k == 0 || k == 4 || k == 8) ? S(1) : S(0);
        v.y = S(0);
        f[(mu * 9 + k) * V + s] = v;
    }
}

// Haar-random SU(3) per link: two complex Gaussian rows (Box-Muller on
// Philox uniforms, counter {link, 0xFFFFFFFF, 0, blk}, key {seed_lo,
// seed_hi ^ 0x48414152}), Gram-Schmidt, row 2 = conj(row0 x row1).
LYM_DI su3<double> haar_su3(uint32_t link, uint64_t seed)
{
    uint2 key = make_uint2(uint32_t(seed), uint32_t(seed >> 32) ^ 0x48414152u);
    uint32_t r[12];
#pragma unroll
    for (uint32_t b = 0; b < 3; b++) {
        uint4 o = philox4x32_10(make_uint4(link, 0xFFFFFFFFu, 0u, b), key);
        r[4 * b] = o.x;
        r[4 * b + 1] = o.y;
        r[4 * b + 2] = o.z;
        r[4 * b + 3] = o.w;
    }
    double gs[12];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double u1 = (double(r[2 * k]) + 1.0) * (1.0 / 4294967296.0);
        double u2 = u32_unit(r[2 * k + 1]);
        double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(6.283185307179586477 * u2, &sn, &cs);
        gs[2 * k] = rad * cs;
        gs[2 * k + 1] = rad * sn;
    }
    su3<double> m;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        m.e[j] = cplx<double>{gs[2 * j], gs[2 * j + 1]};
        m.e[3 + j] = cplx<double>{gs[6 + 2 * j], gs[6 + 2 * j + 1]};
    }
    reunitarize(m);
    return m;
}

template <typename S> __global__ void k_hot(F2<S> *__restrict__ f, long V, uint64_t seed)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= 4 * V) return;
    const long mu = l / V, s = l % V;
    su3<double> m = haar_su3(uint32_t(l), seed);
    store_link<S, double>(f + mu * 9 * V, V, s, m);
}

// Staple (what = 0) or force Z (what = 1) of direction mu in the reference's
// per-site layout (V x 18 doubles): the results of PureGauge::computeStaples
// and PureGauge::computeDerivative (src/Actions/puregauge.cpp:36-79).
template <typename S, int MU>
LYM_DI void export_force_site(const F2<S> *__restrict__ f, const Geom g, long s, int what, double *__restrict__ out)
{
    Nbr n = neighbours(g, s);
    su3<double> om = staple<double, S, MU>(f, g.V, s, n);
    double *o = out + s * 18;
    if (what == 0) {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            o[2 * k] = om.e[k].re;
            o[2 * k + 1] = om.e[k].im;
        }
        return;
    }
    su3<double> U = load_link<double, S>(f + (long)MU * 9 * g.V, g.V, s);
    herm8<double> z = force_from_product(U, om);
    const double d2 = -z.d0 - z.d1;
    o[0] = z.d0;  o[1] = 0.0;
    o[8] = z.d1;  o[9] = 0.0;
    o[16] = d2;   o[17] = 0.0;
    o[2] = z.o01.re;  o[3] = z.o01.im;  o[6] = z.o01.re;   o[7] = -z.o01.im;
    o[4] = z.o02.re;  o[5] = z.o02.im;  o[12] = z.o02.re;  o[13] = -z.o02.im;
    o[10] = z.o12.re; o[11] = z.o12.im; o[14] = z.o12.re;  o[15] = -z.o12.im;
}

template <typename S>
__global__ __launch_bounds__(256) void k_export_force(const F2<S> *__restrict__ f, Geom g, int mu, int what,
                                                      double *__restrict__ out)
{
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    if (s >= g.V) return;
    switch (mu) {
    case 0: export_force_site<S, 0>(f, g, s, what, out); break;
    case 1: export_force_site<S, 1>(f, g, s, what, out); break;
    case 2: export_force_site<S, 2>(f, g, s, what, out); break;
    default: export_force_site<S, 3>(f, g, s, what, out); break;
    }
}

// ------------------------------------------------ checkerboard Metropolis ---

// SU(2) subgroup element of randSU3Transf (src/Math/random.cpp:81-102).
LYM_DI void su2_near_identity(const double *u4, double eps, double x[4])
{
    double norm = 0.0;
    x[0] = u4[0] - 0.5;
#pragma unroll
    for (int i = 1; i < 4; i++) {
        x[i] = u4[i] - 0.5;
        norm += x[i] * x[i];
    }
    norm = sqrt(norm);
    double x0 = sqrt(1.0 - eps * eps);
    x[0] = (x[0] < 0) ? -x0 : x0;
#pragma unroll
    for (int i = 1; i < 4; i++) x[i] = x[i] * eps / norm;
}

// X = R S T (src/Math/random.cpp:71-156), SU(2) blocks
// [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 - i x3]] in (0,1), (0,2), (1,2).
LYM_DI su3<double> cand_su2_product(const double *u, double eps)
{
    double x[4];
    su3<double> R = zero3<double>(), Sm = zero3<double>(), T = zero3<double>();
    R.e[0].re = R.e[4].re = R.e[8].re = 1.0;
    Sm.e[0].re = Sm.e[4].re = Sm.e[8].re = 1.0;
    T.e[0].re = T.e[4].re = T.e[8].re = 1.0;
    su2_near_identity(u, eps, x);
    R.e[0] = cplx<double>{x[0], x[3]};
    R.e[4] = cplx<double>{x[0], -x[3]};
    R.e[1] = cplx<double>{x[2], x[1]};
    R.e[3] = cplx<double>{-x[2], x[1]};
    su2_near_identity(u + 4, eps, x);
    Sm.e[0] = cplx<double>{x[0], x[3]};
    Sm.e[8] = cplx<double>{x[0], -x[3]};
    Sm.e[2] = cplx<double>{x[2], x[1]};
    Sm.e[6] = cplx<double>{-x[2], x[1]};
    su2_near_identity(u + 8, eps, x);
    T.e[4] = cplx<double>{x[0], x[3]};
    T.e[8] = cplx<double>{x[0], -x[3]};
    T.e[5] = cplx<double>{x[2], x[1]};
    T.e[7] = cplx<double>{-x[2], x[1]};
    return mul_nn(mul_nn(R, Sm), T);
}

// X = exp(i eps H), H = sum_a (2u_a - 1) lambda_a (Gell-Mann matrices).
// Tr (eps H)^2 / 2 = eps^2 sum_a h_a^2 < 8 eps^2, so for 8 eps^2 <= kSeriesC1Max
// (eps <= 0.25) SMALL takes the series-only exponential.
template <bool SMALL> LYM_DI su3<double> cand_exph(const double *u, double eps)
{
    double h[8];
#pragma unroll
    for (int a = 0; a < 8; a++) h[a] = 2.0 * u[a] - 1.0;
    const double r3 = 0.57735026918962576451;
    herm8<double> q;
    q.d0 = eps * (h[2] + h[7] * r3);
    q.d1 = eps * (-h[2] + h[7] * r3);
    q.o01 = cplx<double>{eps * h[0], -eps * h[1]};
    q.o02 = cplx<double>{eps * h[3], -eps * h[4]};
    q.o12 = cplx<double>{eps * h[5], -eps * h[6]};
    if constexpr (SMALL)
        return exp_iq_small(q);
    else
        return exp_iq(q);
}

struct MetroParams {
    double beta, eps;
    int nhit, candidate, nu;  // nu = uniforms per hit (13 or 9)
    uint32_t sweep;           // global sweep index (Philox counter)
    int parity;
    uint64_t seed;
    const double *predrawn;  // test mode: this sweep's block, or nullptr
};

// Metropolis update of link (mu, s): nhit hits of U' = X U with
// dS = -(beta/3) Re Tr[(U' - U) Omega], accepted iff exp(-dS) > u
// (src/Apps/gaugefieldfactory.cpp:205-220, src/Actions/puregauge.cpp:28).
// The field f is in even-odd order (EOLinks); s is the lexicographic index
// (it keys the random numbers) and i its rank within its parity.
// Arithmetic in FP64; FP32 storage is reunitarised before the store.
// CAND (0: R S T product, 13 uniforms per hit; 1: exp(i eps H), 9; 2: the
// same with 8 eps^2 <= kSeriesC1Max, series-only exponential) and
// PREDRAWN (uniforms read from an array instead of drawn from Philox) are
// compile-time, so each kernel carries the registers of one candidate only.
template <typename S, int MU, int CAND, bool PREDRAWN>
LYM_DI int metro_link(F2<S> *__restrict__ f, const Geom g, long s, long i, const MetroParams &p)
{
    constexpr int NU = (CAND == 0) ? 13 : 9;
    constexpr int BLOCKS = (NU + 3) / 4;  // Philox outputs 4 uniforms per call
    const long V = g.V;
    const unsigned half = unsigned(V / 2), own = unsigned(p.parity) * half + unsigned(i);
    Nbr n = neighbours(g, s);
    su3<double> om = staple_from<double, MU>(EOLinks<S>{f, V, half, unsigned(p.parity)}, s, n);
    F2<S> *Fm = f + (long)MU * 9 * V;
    su3<double> U = load_link<double, S>(Fm, V, own);
    const uint32_t lk = uint32_t(MU * V + s);
    const uint2 key = make_uint2(uint32_t(p.seed), uint32_t(p.seed >> 32));
    int acc = 0;
    for (int hit = 0; hit < p.nhit; hit++) {
        double u[4 * BLOCKS];
        if constexpr (PREDRAWN) {
            const double *src = p.predrawn + ((long)lk * p.nhit + hit) * NU;
#pragma unroll
            for (int k = 0; k < NU; k++) u[k] = src[k];
        } else {
#pragma unroll
            for (uint32_t b = 0; b < BLOCKS; b++) {
                uint4 o = philox4x32_10(make_uint4(lk, p.sweep, uint32_t(hit), b), key);
                u[4 * b] = u32_unit(o.x);
                u[4 * b + 1] = u32_unit(o.y);
                u[4 * b + 2] = u32_unit(o.z);
                u[4 * b + 3] = u32_unit(o.w);
            }
        }
        su3<double> X;
        if constexpr (CAND == 0)
            X = cand_su2_product(u, p.eps);
        else
            X = cand_exph<CAND == 2>(u, p.eps);
        su3<double> Un = mul_nn(X, U);
        su3<double> d;
#pragma unroll
        for (int k = 0; k < 9; k++) d.e[k] = cplx<double>{Un.e[k].re - U.e[k].re, Un.e[k].im - U.e[k].im};
        double dS = -p.beta / 3.0 * retr_mul(d, om);
        if (exp(-dS) > u[NU - 1]) {
            U = Un;
            acc++;
        }
    }
    if constexpr (sizeof(S) == 4) reunitarize(U);
    store_link<S, double>(Fm, V, own, U);
    return acc;
}

// metro_link for a direction known only at run time (uniform over the wave)
template <typename S, int CAND, bool PREDRAWN>
LYM_DI int metro_link_mu(F2<S> *__restrict__ f, const Geom g, int mu, long s, long i, const MetroParams &p)
{
    switch (mu) {
    case 0: return metro_link<S, 0, CAND, PREDRAWN>(f, g, s, i, p);
    case 1: return metro_link<S, 1, CAND, PREDRAWN>(f, g, s, i, p);
    case 2: return metro_link<S, 2, CAND, PREDRAWN>(f, g, s, i, p);
    default: return metro_link<S, 3, CAND, PREDRAWN>(f, g, s, i, p);
    }
}

// site index of the i-th site of the given parity (requires even extents);
// 32-bit arithmetic, as site indices fit in 32 bits
LYM_DI long parity_site(const Geom g, long i, int parity)
{
    const uint32_t half = uint32_t(g.L[0]) / 2u, L1 = uint32_t(g.L[1]), L2 = uint32_t(g.L[2]);
    const uint32_t r = uint32_t(i) / half, xh = uint32_t(i) - r * half;
    const uint32_t r1 = r / L1, y = r - r1 * L1;
    const uint32_t t = r1 / L2, z = r1 - t * L2;
    const uint32_t x = 2u * xh + ((uint32_t(parity) + y + z + t) & 1u);
    return long(x + uint32_t(g.L[0]) * r);
}

// Reorders the 36 (direction, element) planes of a link field between the
// lexicographic order and the even-odd order of EOLinks; one thread per
// complex, thread j of a plane handles even-odd position j.
template <typename S, bool TO_EO>
__global__ __launch_bounds__(256) void k_reorder_eo(const F2<S> *__restrict__ in, F2<S> *__restrict__ out, Geom g)
{
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= g.V) return;
    const long half = g.V / 2, base = (long)blockIdx.y * g.V;
    const int par = j >= half;
    const long s = parity_site(g, j - par * half, par);
    if constexpr (TO_EO)
        out[base + j] = in[base + s];
    else
        out[base + s] = in[base + j];
}

// One Metropolis sub-sweep (direction mu, one parity) over a field in
// even-odd order; block q = sites [256 q, 256 q + 256) of the parity.
// Workgroups are dispatched round-robin over the 8 XCDs, so workgroup b runs
// on XCD b % 8 and takes slot (b % 8) * per_xcd + b / 8 of the block order:
// order[slot] from the host's column table (metro_column_schedule: each XCD
// marches (z, y) columns along t, so the neighbouring time slices stay in its
// L2), or the slot itself when order is null (each XCD one contiguous range).
// The grid is 8 * per_xcd workgroups; padding slots (-1) and blocks past the
// lattice return at once.
template <typename S, int CAND, bool PREDRAWN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CAND == 1 ? 1 : 2))) void k_metropolis(
    F2<S> *__restrict__ f, Geom g, int mu, MetroParams p, const int *__restrict__ order, unsigned per_xcd,
    unsigned long long *__restrict__ accepted)
{
    const unsigned slot = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    const long blk = order ? long(order[slot]) : long(slot);  // -1: padding
    if (blk < 0) return;
    const long i = blk * 256 + threadIdx.x;
    int acc = 0;
    if (i < g.V / 2) acc = metro_link_mu<S, CAND, PREDRAWN>(f, g, mu, parity_site(g, i, p.parity), i, p);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(accepted, (unsigned long long)acc);
}

}  // namespace lym
