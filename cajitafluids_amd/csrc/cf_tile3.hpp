// cf_tile3.hpp -- the 3-D PCG kernels: z-marching tiles with a software
// pipeline over planes.
//
// A block owns a 64 x 16 (x, y) tile and marches up a chunk of z planes.
// Each lane keeps its column's values at planes k-1, k, k+1 in registers (4
// rows per lane); the tile's plane-k values plus a one-cell x/y halo ring sit
// in LDS for the x and y neighbours.  Loads run two planes ahead of the
// stencil: while plane k is computed, plane k+1 (loaded one step earlier) is
// finished into LDS and plane k+2 is in flight, so a wave seldom waits on
// memory latency.  Every vector element is read once per chunk; halo cells
// come from the neighbouring tiles, which the XCD-ordered block mapping keeps
// in the same L2.
//
// Everything that depends on the lane's x position is computed once per
// block, everything that depends on a row once per block and wave (rows are
// wave-uniform), and everything that depends on the plane once per plane, so
// the per-cell work is the arithmetic itself.
#pragma once

#include "cf_stream.hpp"

namespace cfmi
{
namespace tile3
{
constexpr int TX = 64;          // lanes along x
constexpr int NW = kBlock / TX; // waves per block
constexpr int RY = 4;           // rows per lane
constexpr int TY = NW * RY;     // tile rows
constexpr int NHALO = 2 * TX + 2 * TY;
constexpr int NS = RY + 1;      // load slots per lane: RY own cells + 1 halo
static_assert(NHALO <= kBlock, "one halo cell per lane at most");
} // namespace tile3

struct Tile3Grid
{
    int tiles_x, tiles_y, nchunk, kz, work;
};

inline Tile3Grid tile3_grid(const Geo& g)
{
    using namespace tile3;
    Tile3Grid t;
    t.tiles_x = (g.n[0] + TX - 1) / TX;
    t.tiles_y = (g.n[1] + TY - 1) / TY;
    t.kz = g.n[2] / 16 > 16 ? g.n[2] / 16 : 16;
    t.nchunk = (g.n[2] + t.kz - 1) / t.kz;
    t.work = t.tiles_x * t.tiles_y * t.nchunk;
    return t;
}

inline int tile3_blocks(const Geo& g)
{
    const Tile3Grid t = tile3_grid(g);
    return (t.work + 7) / 8 * 8;
}

// Maps a global index that may lie one cell outside the domain to the
// stored cell it stands for: wrapped on a periodic axis; -1 (a zero ghost)
// otherwise, and for anything further out.
__device__ __forceinline__ int wrap_or_ghost(const Geo& g, int d, int gi)
{
    const int n = g.n[d];
    if (gi >= 0 && gi < n)
        return gi;
    if (g.periodic[d] && (gi == -1 || gi == n))
        return gi < 0 ? n - 1 : 0;
    return -1;
}

// Jacobi diagonal inverse for a per-lane wall count (at most 6 walls).
template <class T>
__device__ __forceinline__ T dinv_sel(const Geo& g, int cnt)
{
    T v = dinv_of<T>(g, 0);
#pragma unroll
    for (int q = 1; q <= 6; ++q)
        v = cnt == q ? dinv_of<T>(g, q) : v;
    return v;
}

// Whether the -/+ coupling of an owned cell with global index gi along axis
// d is cut by a wall (BoundaryConditions.hpp:57-97: gi <= min, gi > max - 1).
__device__ __forceinline__ bool wall_lo(const Geo& g, int d, int gi)
{
    return !g.periodic[d] && gi <= 0;
}
__device__ __forceinline__ bool wall_hi(const Geo& g, int d, int gi)
{
    return !g.periodic[d] && gi > g.n[d] - 2;
}

// The block's place in the grid and its lanes' load slots.  Slot q < RY is
// the lane's own cell in row w*RY + q; slot RY is the lane's halo cell (if
// any).  colb is the linear index of the slot's (x, y) column at local
// z = 0, or -1 for a zero ghost.
struct Tile3Ctx
{
    bool active;
    int i0, j0, k0, k1;
    int lane, w;
    int64_t colb[tile3::NS];
    int hwxy;           // x + y wall count of the halo slot
    int hrow, hcol;     // LDS position of the halo slot (hrow < 0: none)
    // own cells
    int wx;             // x wall count of the lane's column
    bool xwall_m, xwall_p;
    int wy[tile3::RY];  // y wall count of each row (wave-uniform)
    bool ywall_m[tile3::RY], ywall_p[tile3::RY];
    bool valid[tile3::RY];
};

__device__ __forceinline__ Tile3Ctx tile3_ctx(const Geo& g, const Tile3Grid& t)
{
    using namespace tile3;
    Tile3Ctx c;
    const int L = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    c.active = L < t.work;
    const int tx = L % t.tiles_x;
    const int rest = L / t.tiles_x;
    const int ty = rest % t.tiles_y;
    const int ch = rest / t.tiles_y;
    c.i0 = tx * TX;
    c.j0 = ty * TY;
    c.k0 = ch * t.kz;
    c.k1 = min(c.k0 + t.kz, g.n[2]);
    c.lane = threadIdx.x & (TX - 1);
    c.w = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    auto col = [&](int gi, int gj) -> int64_t {
        const int a = wrap_or_ghost(g, 0, gi), b = wrap_or_ghost(g, 1, gj);
        if (a < 0 || b < 0)
            return -1;
        return g.sx + (g.h + a) + g.px * (int64_t)(g.h + b);
    };
    const int gi = c.i0 + c.lane;
    const int a = wrap_or_ghost(g, 0, gi);
    c.wx = a < 0 ? 0 : wall_count(g, 0, a);
    c.xwall_m = wall_lo(g, 0, gi);
    c.xwall_p = wall_hi(g, 0, gi);
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int gj = c.j0 + c.w * RY + q;
        const int b = wrap_or_ghost(g, 1, gj);
        c.colb[q] = col(gi, gj);
        c.wy[q] = b < 0 ? 0 : wall_count(g, 1, b);
        c.ywall_m[q] = wall_lo(g, 1, gj);
        c.ywall_p[q] = wall_hi(g, 1, gj);
        c.valid[q] = gi < g.n[0] && gj < g.n[1];
    }
    const int hh = threadIdx.x;
    int gih = 0, gjh = 0;
    c.hrow = -1;
    c.hcol = 0;
    if (hh < TX) {
        c.hrow = 0; c.hcol = hh + 1; gih = c.i0 + hh; gjh = c.j0 - 1;
    } else if (hh < 2 * TX) {
        c.hrow = TY + 1; c.hcol = hh - TX + 1; gih = c.i0 + hh - TX; gjh = c.j0 + TY;
    } else if (hh < 2 * TX + TY) {
        c.hrow = hh - 2 * TX + 1; c.hcol = 0; gih = c.i0 - 1; gjh = c.j0 + hh - 2 * TX;
    } else if (hh < NHALO) {
        c.hrow = hh - 2 * TX - TY + 1; c.hcol = TX + 1; gih = c.i0 + TX;
        gjh = c.j0 + hh - 2 * TX - TY;
    }
    c.colb[RY] = c.hrow >= 0 ? col(gih, gjh) : -1;
    c.hwxy = 0;
    if (c.colb[RY] >= 0)
        c.hwxy = wall_count(g, 0, wrap_or_ghost(g, 0, gih)) +
                 wall_count(g, 1, wrap_or_ghost(g, 1, gjh));
    return c;
}

// Plane-uniform data of global plane gk: offset of the plane (-1 for a zero
// ghost plane), its z wall count, and its z couplings.
struct Plane3Meta
{
    int64_t off;
    int wz;
    bool zwall_m, zwall_p;
};

__device__ __forceinline__ Plane3Meta plane3_meta(const Geo& g, int gk)
{
    Plane3Meta m;
    const int zc = wrap_or_ghost(g, 2, gk);
    m.off = zc < 0 ? -1 : g.px * g.py * (int64_t)(g.h + zc);
    m.wz = zc < 0 ? 0 : wall_count(g, 2, zc);
    m.zwall_m = wall_lo(g, 2, gk);
    m.zwall_p = wall_hi(g, 2, gk);
    return m;
}

// Raw operands of one plane: a = r (direction) or p (update), b = old
// direction (direction kernel only), both 0 at zero ghosts.
template <class T, bool DIR>
struct Plane3
{
    T a[tile3::NS];
    T b[tile3::NS];
};

template <class T, bool DIR>
__device__ __forceinline__ void tile3_load(const Tile3Ctx& c, const Plane3Meta& m,
                                           const T* __restrict__ va,
                                           const T* __restrict__ vb, bool need_b,
                                           Plane3<T, DIR>& P)
{
#pragma unroll
    for (int s = 0; s < tile3::NS; ++s) {
        const bool ok = c.colb[s] >= 0 && m.off >= 0;
        const int64_t at = c.colb[s] + m.off;
        P.a[s] = ok ? va[at] : (T)0;
        if constexpr (DIR)
            P.b[s] = (ok && need_b) ? vb[at] : (T)0;
    }
}

// Value of slot s: the new direction fma(beta, p_old, D^-1 r) (D^-1 r
// alone on the first iteration) for DIR, else the loaded p.
template <class T, bool DIR>
__device__ __forceinline__ T tile3_value(const Geo& g, const Tile3Ctx& c,
                                         const Plane3Meta& m,
                                         const Plane3<T, DIR>& P, int s,
                                         bool first, T beta)
{
    if constexpr (!DIR) {
        return P.a[s];
    } else {
        T dinv;
        if (s < tile3::RY)
            dinv = dinv_pick<T>(g, c.wy[s] + m.wz, c.wx); // uniform y, z part
        else
            dinv = dinv_sel<T>(g, c.hwxy + m.wz);
        const T z = dinv * P.a[s];
        return first ? z : fma_t<T>(beta, P.b[s], z);
    }
}

// Writes the plane's values into LDS plane pl and own[]; the halo slot
// fills its ring cell.
template <class T, bool DIR>
__device__ __forceinline__ void tile3_finish(const Geo& g, const Tile3Ctx& c,
                                             const Plane3Meta& m,
                                             const Plane3<T, DIR>& P, bool first,
                                             T beta, T (*pl)[tile3::TX + 2],
                                             T (&own)[tile3::RY])
{
    using namespace tile3;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        own[q] = tile3_value<T, DIR>(g, c, m, P, q, first, beta);
        pl[c.w * RY + q + 1][c.lane + 1] = own[q];
    }
    if (c.hrow >= 0)
        pl[c.hrow][c.hcol] = tile3_value<T, DIR>(g, c, m, P, RY, first, beta);
}

// y = A p at own cell (row q) of a plane: centre pc, z neighbours pm / pp,
// x and y neighbours from the LDS plane; reference stencil order
// {0}, -x, +x, -y, +y, -z, +z with zero couplings across walls.
template <class T>
__device__ __forceinline__ T tile3_apply(const Geo& g, const Tile3Ctx& c,
                                         const Plane3Meta& m,
                                         const T (*pl)[tile3::TX + 2], int q,
                                         T pc, T pm, T pp)
{
    const T s = offdiag<T>(g);
    const int jj = c.w * tile3::RY + q;
    const int ln = c.lane;
    T y = diag_pick<T>(g, c.wy[q] + m.wz, c.wx) * pc;
    y = fma_t<T>(c.xwall_m ? (T)0 : s, pl[jj + 1][ln], y);
    y = fma_t<T>(c.xwall_p ? (T)0 : s, pl[jj + 1][ln + 2], y);
    y = fma_t<T>(c.ywall_m[q] ? (T)0 : s, pl[jj][ln + 1], y);
    y = fma_t<T>(c.ywall_p[q] ? (T)0 : s, pl[jj + 2][ln + 1], y);
    y = fma_t<T>(m.zwall_m ? (T)0 : s, pm, y);
    y = fma_t<T>(m.zwall_p ? (T)0 : s, pp, y);
    return y;
}

// The z march shared by the direction (DIR) and update kernels; epi.cell()
// receives every owned cell's centre value and A p.
template <class T, bool DIR, class Epi>
__device__ __forceinline__ void tile3_march(const Geo& g, const Tile3Grid& tg,
                                            const T* __restrict__ va,
                                            const T* __restrict__ vb, bool first,
                                            T beta, T (*pl)[tile3::TY + 2][tile3::TX + 2],
                                            Epi&& epi)
{
    using namespace tile3;
    const Tile3Ctx c = tile3_ctx(g, tg);
    if (!c.active)
        return;
    const bool need_b = !first;
    Plane3<T, DIR> P0, P1;
    T pm[RY], pc[RY], pp[RY];
    // prologue: plane k0-1 (own column only), plane k0 into LDS, k0+1 in flight
    Plane3Meta mm = plane3_meta(g, c.k0 - 1);
    tile3_load<T, DIR>(c, mm, va, vb, need_b, P0);
#pragma unroll
    for (int q = 0; q < RY; ++q)
        pm[q] = tile3_value<T, DIR>(g, c, mm, P0, q, first, beta);
    Plane3Meta mc = plane3_meta(g, c.k0);
    tile3_load<T, DIR>(c, mc, va, vb, need_b, P1);
    tile3_finish<T, DIR>(g, c, mc, P1, first, beta, pl[0], pc);
    Plane3Meta mn = plane3_meta(g, c.k0 + 1);
    tile3_load<T, DIR>(c, mn, va, vb, need_b, P0);
    epi.prefetch(c, mc, c.k0);
    epi.rotate();
    __syncthreads();
    int cur = 0;
    for (int k = c.k0; k < c.k1; ++k) {
        // P0 holds plane k+1 (loaded last step); start plane k+2
        const Plane3Meta mnn = plane3_meta(g, k + 2);
        if (k + 2 <= c.k1)
            tile3_load<T, DIR>(c, mnn, va, vb, need_b, P1);
        tile3_finish<T, DIR>(g, c, mn, P0, first, beta, pl[cur ^ 1], pp);
        epi.prefetch(c, mn, k + 1);
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            if (c.valid[q]) {
                const T y = tile3_apply<T>(g, c, mc, pl[cur], q, pc[q], pm[q], pp[q]);
                epi.cell(c, mc, q, pc[q], y);
            }
            pm[q] = pc[q];
            pc[q] = pp[q];
        }
        __syncthreads();
        cur ^= 1;
        P0 = P1;
        mc = mn;
        mn = mnn;
        epi.rotate();
    }
}

} // namespace cfmi
