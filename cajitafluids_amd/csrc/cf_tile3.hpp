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

// The block's place in the grid and its lanes' load slots.  Slot q < RY is
// the lane's own cell in row w*RY + q; slot RY is the lane's halo cell (if
// any).  colb is the linear index of the slot's (x, y) column at local
// z = 0, or -1 for a zero ghost; wx / wxy are its x and x+y wall counts.
struct Tile3Ctx
{
    bool active;
    int i0, j0, k0, k1;
    int lane, w;
    int64_t colb[tile3::NS];
    int wxy[tile3::NS];
    int wx;          // own cells' x wall count
    int hrow, hcol;  // LDS position of the halo slot (hrow < 0: none)
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
    auto col = [&](int gi, int gj, int& wxy) -> int64_t {
        const int a = wrap_or_ghost(g, 0, gi), b = wrap_or_ghost(g, 1, gj);
        if (a < 0 || b < 0) {
            wxy = 0;
            return -1;
        }
        wxy = wall_count(g, 0, a) + wall_count(g, 1, b);
        return g.sx + (g.h + a) + g.px * (int64_t)(g.h + b);
    };
    const int gi = c.i0 + c.lane;
    const int a = wrap_or_ghost(g, 0, gi);
    c.wx = a < 0 ? 0 : wall_count(g, 0, a);
#pragma unroll
    for (int q = 0; q < RY; ++q)
        c.colb[q] = col(gi, c.j0 + c.w * RY + q, c.wxy[q]);
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
    c.colb[RY] = c.hrow >= 0 ? col(gih, gjh, c.wxy[RY]) : -1;
    if (c.hrow < 0)
        c.wxy[RY] = 0;
    return c;
}

// Linear index of slot s at global plane gk, or -1 for a zero ghost.
__device__ __forceinline__ int64_t tile3_at(const Geo& g, const Tile3Ctx& c,
                                            int s, int gk)
{
    const int zc = wrap_or_ghost(g, 2, gk);
    if (c.colb[s] < 0 || zc < 0)
        return -1;
    return c.colb[s] + g.px * g.py * (int64_t)(g.h + zc);
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
__device__ __forceinline__ void tile3_load(const Geo& g, const Tile3Ctx& c,
                                           int gk, const T* __restrict__ va,
                                           const T* __restrict__ vb, bool need_b,
                                           Plane3<T, DIR>& P)
{
#pragma unroll
    for (int s = 0; s < tile3::NS; ++s) {
        const int64_t at = tile3_at(g, c, s, gk);
        P.a[s] = at >= 0 ? va[at] : (T)0;
        if constexpr (DIR)
            P.b[s] = (at >= 0 && need_b) ? vb[at] : (T)0;
    }
}

// Value of slot s at plane gk: the new direction fma(beta, p_old, D^-1 r)
// (D^-1 r alone on the first iteration) for DIR, else the loaded p.
template <class T, bool DIR>
__device__ __forceinline__ T tile3_value(const Geo& g, const Tile3Ctx& c,
                                         const Plane3<T, DIR>& P, int s, int gk,
                                         bool first, T beta)
{
    if constexpr (!DIR) {
        return P.a[s];
    } else {
        const int zc = wrap_or_ghost(g, 2, gk);
        const int wz = zc < 0 ? 0 : wall_count(g, 2, zc);
        T dinv;
        if (s < tile3::RY) {
            // own cell: the y and z parts are wave-uniform
            const int gj = c.j0 + c.w * tile3::RY + s;
            const int b = wrap_or_ghost(g, 1, gj);
            const int wy = b < 0 ? 0 : wall_count(g, 1, b);
            dinv = dinv_pick<T>(g, wy + wz, c.wx);
        } else {
            dinv = dinv_sel<T>(g, c.wxy[s] + wz);
        }
        const T z = dinv * P.a[s];
        return first ? z : fma_t<T>(beta, P.b[s], z);
    }
}

// Writes plane gk's values into LDS plane pl and own[]; the halo slot fills
// its ring cell.
template <class T, bool DIR>
__device__ __forceinline__ void tile3_finish(const Geo& g, const Tile3Ctx& c,
                                             const Plane3<T, DIR>& P, int gk,
                                             bool first, T beta,
                                             T (*pl)[tile3::TX + 2],
                                             T (&own)[tile3::RY])
{
    using namespace tile3;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        own[q] = tile3_value<T, DIR>(g, c, P, q, gk, first, beta);
        pl[c.w * RY + q + 1][c.lane + 1] = own[q];
    }
    if (c.hrow >= 0)
        pl[c.hrow][c.hcol] = tile3_value<T, DIR>(g, c, P, RY, gk, first, beta);
}

// y = A p at own cell (row jj) of plane k: centre pc, z neighbours pm / pp,
// x and y neighbours from the LDS plane; reference stencil order
// {0}, -x, +x, -y, +y, -z, +z with zero couplings across walls.
template <class T>
__device__ __forceinline__ T tile3_apply(const Geo& g, const T (*pl)[tile3::TX + 2],
                                         int lane, int jj, int gi, int gj, int gk,
                                         T pc, T pm, T pp)
{
    const T s = offdiag<T>(g);
    auto cf = [&](int d, int side, int gidx) -> T {
        if (g.periodic[d])
            return s;
        const bool wall = side == 0 ? gidx <= 0 : gidx > g.n[d] - 2;
        return wall ? (T)0 : s;
    };
    T y = diag_pick<T>(g, wall_count(g, 1, gj) + wall_count(g, 2, gk),
                       wall_count(g, 0, gi)) *
          pc;
    y = fma_t<T>(cf(0, 0, gi), pl[jj + 1][lane], y);
    y = fma_t<T>(cf(0, 1, gi), pl[jj + 1][lane + 2], y);
    y = fma_t<T>(cf(1, 0, gj), pl[jj][lane + 1], y);
    y = fma_t<T>(cf(1, 1, gj), pl[jj + 2][lane + 1], y);
    y = fma_t<T>(cf(2, 0, gk), pm, y);
    y = fma_t<T>(cf(2, 1, gk), pp, y);
    return y;
}

// Direction + SpMV: p_new = fma(beta, p_old, D^-1 r); q = A p_new;
// partial p_new.q.  UPDATE (DIR = false): q = A p; x += alpha p;
// r -= alpha q; partials r.r, r.(D^-1 r).
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
    tile3_load<T, DIR>(g, c, c.k0 - 1, va, vb, need_b, P0);
#pragma unroll
    for (int q = 0; q < RY; ++q)
        pm[q] = tile3_value<T, DIR>(g, c, P0, q, c.k0 - 1, first, beta);
    tile3_load<T, DIR>(g, c, c.k0, va, vb, need_b, P1);
    tile3_finish<T, DIR>(g, c, P1, c.k0, first, beta, pl[0], pc);
    tile3_load<T, DIR>(g, c, c.k0 + 1, va, vb, need_b, P0);
    epi.prefetch(c, c.k0);
    epi.rotate();
    __syncthreads();
    int cur = 0;
    for (int k = c.k0; k < c.k1; ++k) {
        // P0 holds plane k+1 (loaded last step); start plane k+2
        if (k + 2 <= c.k1)
            tile3_load<T, DIR>(g, c, k + 2, va, vb, need_b, P1);
        tile3_finish<T, DIR>(g, c, P0, k + 1, first, beta, pl[cur ^ 1], pp);
        epi.prefetch(c, k + 1);
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            const int jj = c.w * RY + q;
            const int gi = c.i0 + c.lane;
            const int gj = c.j0 + jj;
            if (gi < g.n[0] && gj < g.n[1]) {
                const T y = tile3_apply<T>(g, pl[cur], c.lane, jj, gi, gj, k,
                                           pc[q], pm[q], pp[q]);
                epi.cell(c, q, gi, gj, k, pc[q], y);
            }
            pm[q] = pc[q];
            pc[q] = pp[q];
        }
        __syncthreads();
        cur ^= 1;
        P0 = P1;
        epi.rotate();
    }
}

} // namespace cfmi
