// cf_tile3.hpp -- the 3-D PCG kernels: z-marching tiles with a software
// pipeline over planes.
//
// A block owns a (64 CPL) x 16 (x, y) tile and marches up a chunk of z
// planes.  Each lane owns CPL consecutive x cells in each of 4 rows (CPL = 2:
// one 8-byte (FP32) or 16-byte (FP64) access per row, half the address and
// memory instructions per cell) and keeps its cells' values at planes k-1,
// k, k+1 in registers; the tile's plane-k values plus a one-cell x/y halo
// ring sit in LDS for the x and y neighbours (a cell's x neighbour inside
// the lane's own group is read from registers).  Loads run two planes ahead
// of the stencil: while plane k is computed, plane k+1 (loaded one step
// earlier) is finished into LDS and plane k+2 is in flight, so a wave seldom
// waits on memory latency.  Every vector element is read once per chunk;
// halo cells come from the neighbouring tiles, which the XCD-ordered block
// mapping keeps in the same L2.
//
// Loads that feed the pipeline are issued on every path (a lane with
// nothing to read loads a cell known to be in bounds and discards it): the
// compiler counts vector memory operations per path, and one conditional
// load makes it wait for every load in flight, the next plane's included,
// before touching the current one.  The loaded registers are kept raw and
// masked where they are used: a select right after the load may be placed
// at the end of the loop body, where it waits for the loads just as well.
// Groups of two cells (CPL = 2, FP32) read aligned pairs on every path too,
// with the periodic-x wrap compiled in only where x is periodic (PX).
//
// Everything that depends on the lane's x position is computed once per
// block, everything that depends on a row once per block and wave (rows are
// wave-uniform), and everything that depends on the plane once per plane, so
// the per-cell work is the arithmetic itself.
//
// The value marched through the stencil is chosen by MODE: the loaded
// vector (0), the PCG direction fma(beta, p_old, D^-1 r) (1), or the loaded
// vector plus beta times its parent's value on a grid coarsened by 2 along
// every axis (2, the multigrid prolongation; CPL = 1).
#pragma once

#include "cf_stream.hpp"

namespace cfmi
{
// Lanes per block of the tile kernels (4 waves: 16-row tiles).  With 512
// (8 waves, 32-row tiles: half the y-halo rows per cell) the FP64 step was
// the same (284.9 against 285.1 ms at 256^3) and the FP32-refined one 1%
// slower: the halo rows' reads hit the caches.
#ifndef CF_TILE_BLOCK
#define CF_TILE_BLOCK 256
#endif
constexpr int kTileBlock = CF_TILE_BLOCK;

namespace tile3
{
constexpr int LX = 64;              // lanes along x
constexpr int NW = kTileBlock / LX; // waves per block
constexpr int RY = 4;           // rows per lane
constexpr int TY = NW * RY;     // tile rows
constexpr int NS = RY + 1;      // load slots per lane: RY own rows + 1 halo
// halo slots: the two y-halo rows (one lane per CPL cells) and the two
// x-halo columns (one lane per cell)
constexpr int NHALO = 2 * LX + 2 * TY;
static_assert(NHALO <= kTileBlock, "one halo slot per lane at most");
template <int CPL>
constexpr int tx()
{
    return LX * CPL; // tile cells along x
}
// LDS plane layout: a row holds the tile's cells from column pcol() on,
// with the x-halo cells just left and right of them.  With two cells per
// lane the first cell sits at column 2, so each lane's pair, and the pairs
// beside and above / below it, are 8-byte aligned: every LDS access of the
// stencil is one conflict-free 8-byte access per lane.
template <int CPL>
constexpr int pcol()
{
    return CPL;
}
template <int CPL>
constexpr int prow()
{
    return tx<CPL>() + 2 * pcol<CPL>();
}
} // namespace tile3

struct Tile3Grid
{
    int tiles_x, tiles_y, nchunk, kz, work;
};

// Parent grid of MODE 2: value(i, j, k) = fma(alpha, xc[base + i/2 + sy j/2 +
// sz k/2], a(i, j, k)) (the parent array compact or padded); the march
// receives alpha as its `beta`.
template <class TP>
struct Tile3ParentT
{
    const TP* xc; // the parent level's values (FP64, or an FP32 level's)
    int64_t base, sy, sz;
    double alpha;
};
using Tile3Parent = Tile3ParentT<double>;

// An epilogue that declares `static constexpr bool kRaw = true` receives the
// raw loaded operand of its own rows (epi.take_raw(plane)) just before each
// prefetch of the same plane.
template <class E, class = void>
struct epi_takes_raw : std::false_type
{
};
template <class E>
struct epi_takes_raw<E, std::void_t<decltype(E::kRaw)>> : std::bool_constant<E::kRaw>
{
};

// An epilogue with kLate = true loads its data of a plane itself, in
// late(c): the march calls it after the prologue and after every step's
// barrier, for the plane last passed to prefetch (which then only records
// the plane).  One buffer then holds that data instead of two.
template <class E, class = void>
struct epi_loads_late : std::false_type
{
};
template <class E>
struct epi_loads_late<E, std::void_t<decltype(E::kLate)>> : std::bool_constant<E::kLate>
{
};

template <int CPL>
inline Tile3Grid tile3_grid(const Geo& g)
{
    using namespace tile3;
    Tile3Grid t;
    t.tiles_x = (g.n[0] + tx<CPL>() - 1) / tx<CPL>();
    t.tiles_y = (g.n[1] + TY - 1) / TY;
    // z chunks: as many blocks as the one-cell-per-lane tiling
    const int zc = 16 * CPL;
    t.kz = g.n[2] / zc > 16 / CPL ? g.n[2] / zc : 16 / CPL;
    t.nchunk = (g.n[2] + t.kz - 1) / t.kz;
    t.work = t.tiles_x * t.tiles_y * t.nchunk;
    return t;
}

template <int CPL>
inline int tile3_blocks(const Geo& g)
{
    const Tile3Grid t = tile3_grid<CPL>(g);
    return (t.work + 7) / 8 * 8;
}

// v[at] where ok, else v[safe] (an index known to be in bounds), for a
// caller that masks or ignores the !ok lanes' value where it uses it.
template <class T>
__device__ __forceinline__ T load_any(const T* __restrict__ v, bool ok, int64_t at,
                                      int64_t safe)
{
    return v[ok ? at : safe];
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

// CPL consecutive elements as one vector access (8 or 16 bytes for CPL = 2;
// the element index of a lane's first cell is even, so it is aligned).
template <class T, int CPL>
struct Vec
{
};
template <class T>
struct Vec<T, 1>
{
    using type = T;
};
template <>
struct Vec<float, 2>
{
    using type = float2;
};
template <>
struct Vec<double, 2>
{
    using type = double2;
};

template <class T, int CPL>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&v)[CPL])
{
    if constexpr (CPL == 1) {
        v[0] = *p;
    } else {
        using V = typename Vec<T, 2>::type;
        const V w = *reinterpret_cast<const V*>(p);
        v[0] = w.x;
        v[1] = w.y;
    }
}

template <class T, int CPL>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&v)[CPL])
{
    if constexpr (CPL == 1) {
        *p = v[0];
    } else {
        using V = typename Vec<T, 2>::type;
        V w;
        w.x = v[0];
        w.y = v[1];
        *reinterpret_cast<V*>(p) = w;
    }
}

// Stores the valid cells of a lane's group (all of them as one access when
// every cell is valid; only the last tile of a row has partial groups).
template <class T, int CPL>
__device__ __forceinline__ void vstore_valid(T* __restrict__ p, const T (&v)[CPL],
                                             const bool (&ok)[CPL])
{
    if constexpr (CPL == 1) {
        if (ok[0])
            *p = v[0];
    } else {
        if (ok[0] && ok[1])
            vstore<T, 2>(p, v);
        else if (ok[0])
            p[0] = v[0];
    }
}

// The block's place in the grid and its lanes' load slots.  Slot q < RY is
// the lane's own CPL cells in row w*RY + q; slot RY is the lane's halo slot:
// CPL cells of a y-halo row (hkind 1) or one cell of an x-halo column
// (hkind 2).  colb is the linear index of the slot's first (x, y) column at
// local z = 0, or -1 for a zero ghost.
template <int CPL>
struct Tile3Ctx
{
    bool active;
    int i0, j0, k0, k1;
    int lane, w;
    int64_t colb[tile3::NS];
    int64_t safe;                       // an owned cell: the address of discarded loads
    int sgi[tile3::NS], sgj[tile3::NS]; // (wrapped) x, y of each slot's first cell
    // a group whose last cell is the row's end, on a periodic x axis: its
    // second cell (one past the end) is the row's first cell (CPL = 2)
    bool split, hsplit;
    int wrapback;         // n0 - 1
    int hkind;            // 0: no halo slot, 1: y-halo cells, 2: x-halo cell
    int hrow, hcol;       // LDS position of the halo slot's first cell
    int hwxy[CPL];        // x + y wall count of the halo slot's cells
    // own cells
    int wx[CPL];          // x wall count of the lane's columns
    bool xwall_m[CPL], xwall_p[CPL];
    int wy[tile3::RY];    // y wall count of each row (wave-uniform)
    bool ywall_m[tile3::RY], ywall_p[tile3::RY];
    bool valid[tile3::RY][CPL];
    // the march's coefficient table in LDS (element type of the march):
    // diag[k] at k, dinv[k] at 8 + k, k the cell's SOLID wall count
    const void* tab;
};

// A cell's diagonal and Jacobi inverse for the wall count ubase + wx
// (ubase wave-uniform, wx per lane).  FP32 marches read the march's LDS
// table: one LDS read with a per-lane address (the same address in most
// lanes), where uniform reads of the table and selects cost these
// issue-bound kernels scalar registers (spilled) and vector instructions.
// FP64 marches keep the uniform reads and selects: the table read raised
// the FP64 update's registers past 128 (four waves per SIMD).
template <class T, int DIR>
constexpr bool tile3_tab()
{
    return std::is_same<T, float>::value || DIR == 1;
}
template <class T, int DIR, int CPL>
__device__ __forceinline__ T ctx_diag(const Geo& g, const Tile3Ctx<CPL>& c, int ubase, int wx)
{
    if constexpr (tile3_tab<T, DIR>())
        return static_cast<const T*>(c.tab)[ubase + wx];
    else
        return diag_pick<T>(g, ubase, wx);
}
template <class T, int DIR, int CPL>
__device__ __forceinline__ T ctx_dinv(const Geo& g, const Tile3Ctx<CPL>& c, int ubase, int wx)
{
    if constexpr (tile3_tab<T, DIR>())
        return static_cast<const T*>(c.tab)[8 + ubase + wx];
    else
        return dinv_pick<T>(g, ubase, wx);
}

template <int CPL>
__device__ __forceinline__ Tile3Ctx<CPL> tile3_ctx(const Geo& g, const Tile3Grid& t)
{
    using namespace tile3;
    Tile3Ctx<CPL> c;
    const int L = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    c.active = L < t.work;
    const int tx_ = L % t.tiles_x;
    const int rest = L / t.tiles_x;
    const int ty = rest % t.tiles_y;
    const int ch = rest / t.tiles_y;
    c.i0 = tx_ * tx<CPL>();
    c.j0 = ty * TY;
    c.k0 = ch * t.kz;
    c.k1 = min(c.k0 + t.kz, g.n[2]);
    c.lane = threadIdx.x & (LX - 1);
    c.w = __builtin_amdgcn_readfirstlane(threadIdx.x / LX);
    c.safe = g.idx(g.h, g.h, g.h);
    auto col = [&](int gi, int gj) -> int64_t {
        const int a = wrap_or_ghost(g, 0, gi), b = wrap_or_ghost(g, 1, gj);
        if (a < 0 || b < 0)
            return -1;
        return g.sx + (g.h + a) + g.px * (int64_t)(g.h + b);
    };
    const int gi0 = c.i0 + c.lane * CPL;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
        const int gi = gi0 + j;
        const int a = wrap_or_ghost(g, 0, gi);
        c.wx[j] = a < 0 ? 0 : wall_count(g, 0, a);
        c.xwall_m[j] = wall_lo(g, 0, gi);
        c.xwall_p[j] = wall_hi(g, 0, gi);
    }
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int gj = c.j0 + c.w * RY + q;
        const int b = wrap_or_ghost(g, 1, gj);
        // cells just past the domain stand in for their periodic images
        // (or zero ghosts) as neighbours of the last own cells
        c.colb[q] = col(gi0, gj);
        c.sgi[q] = wrap_or_ghost(g, 0, gi0);
        c.sgj[q] = b;
        c.wy[q] = b < 0 ? 0 : wall_count(g, 1, b);
        c.ywall_m[q] = wall_lo(g, 1, gj);
        c.ywall_p[q] = wall_hi(g, 1, gj);
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            c.valid[q][j] = gi0 + j < g.n[0] && gj < g.n[1];
    }
    c.split = CPL > 1 && g.periodic[0] && gi0 + 1 == g.n[0];
    c.wrapback = g.n[0] - 1;
    // halo slot
    const int hh = threadIdx.x;
    c.hsplit = false;
    c.hkind = 0;
    c.hrow = -1;
    c.hcol = 0;
    c.colb[RY] = -1;
    c.sgi[RY] = c.sgj[RY] = 0;
#pragma unroll
    for (int j = 0; j < CPL; ++j)
        c.hwxy[j] = 0;
    if (hh < 2 * LX) {
        // y-halo rows: the CPL cells above or below this lane's column group
        const int l2 = hh < LX ? hh : hh - LX;
        const int gj = hh < LX ? c.j0 - 1 : c.j0 + TY;
        const int gi = c.i0 + l2 * CPL;
        c.hkind = 1;
        c.hrow = hh < LX ? 0 : TY + 1;
        c.hcol = l2 * CPL + pcol<CPL>();
        c.colb[RY] = col(gi, gj);
        c.sgi[RY] = wrap_or_ghost(g, 0, gi);
        c.sgj[RY] = wrap_or_ghost(g, 1, gj);
        c.hsplit = CPL > 1 && g.periodic[0] && gi + 1 == g.n[0];
        const int b = wrap_or_ghost(g, 1, gj);
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const int a = wrap_or_ghost(g, 0, gi + j);
            c.hwxy[j] = (a < 0 ? 0 : wall_count(g, 0, a)) + (b < 0 ? 0 : wall_count(g, 1, b));
        }
    } else if (hh < NHALO) {
        // x-halo columns: one cell left or right of the tile
        const int left = hh < 2 * LX + TY;
        const int r = left ? hh - 2 * LX : hh - 2 * LX - TY;
        const int gi = left ? c.i0 - 1 : c.i0 + tx<CPL>();
        const int gj = c.j0 + r;
        c.hkind = 2;
        c.hrow = r + 1;
        c.hcol = left ? pcol<CPL>() - 1 : tx<CPL>() + pcol<CPL>();
        c.colb[RY] = col(gi, gj);
        c.sgi[RY] = wrap_or_ghost(g, 0, gi);
        c.sgj[RY] = wrap_or_ghost(g, 1, gj);
        if (c.colb[RY] >= 0)
            c.hwxy[0] = wall_count(g, 0, wrap_or_ghost(g, 0, gi)) +
                        wall_count(g, 1, wrap_or_ghost(g, 1, gj));
    }
    return c;
}

// Plane-uniform data of global plane gk: offset of the plane (-1 for a zero
// ghost plane), its z wall count, and its z couplings.
struct Plane3Meta
{
    int64_t off;
    int wz;
    int zc; // wrapped z, or -1
    bool zwall_m, zwall_p;
};

__device__ __forceinline__ Plane3Meta plane3_meta(const Geo& g, int gk)
{
    Plane3Meta m;
    const int zc = wrap_or_ghost(g, 2, gk);
    m.off = zc < 0 ? -1 : g.px * g.py * (int64_t)(g.h + zc);
    m.zc = zc;
    m.wz = zc < 0 ? 0 : wall_count(g, 2, zc);
    m.zwall_m = wall_lo(g, 2, gk);
    m.zwall_p = wall_hi(g, 2, gk);
    return m;
}

// Raw operands of one plane: a = r (direction) or p (update), b = old
// direction (MODE 1) or the parent's value (MODE 2), all 0 at zero ghosts.
template <class T, int MODE, int CPL>
struct Plane3
{
    T a[tile3::NS][CPL];
    T b[tile3::NS][CPL];
};

// TA: the element type of va (an FP32 operand widened to T on load, one
// cell per lane only).
template <class T, int MODE, int CPL, bool PX = false, class TA = T, class TP = double>
__device__ __forceinline__ void tile3_load(const Tile3Ctx<CPL>& c, const Plane3Meta& m,
                                           const TA* __restrict__ va,
                                           const T* __restrict__ vb, bool need_b,
                                           const Tile3ParentT<TP>& par,
                                           Plane3<T, MODE, CPL>& P)
{
    static_assert(std::is_same<TA, T>::value || CPL == 1, "a widened operand: one cell per lane");
    constexpr bool DIR = MODE == 1;
#pragma unroll
    for (int s = 0; s < tile3::NS; ++s) {
        const bool ok = c.colb[s] >= 0 && m.off >= 0;
        const int64_t at = c.colb[s] + m.off;
        // the halo slot of an x-halo lane is one cell; every other slot is a
        // lane's group of CPL cells
        const bool single = CPL > 1 && s == tile3::RY && c.hkind == 2;
        const bool split = CPL > 1 && (s < tile3::RY ? c.split : c.hsplit);
        if constexpr (MODE == 2 && CPL == 1) {
            const int64_t pat = par.base + (c.sgi[s] >> 1) + par.sy * (c.sgj[s] >> 1) +
                                par.sz * (m.zc >> 1);
            P.a[s][0] = (T)load_any<TA>(va, ok, at, c.safe);
            // an FP32 level rounds an FP64 parent's value
            P.b[s][0] = (T)load_any<TP>(par.xc, ok, pat, par.base);
            continue;
        } else if constexpr (CPL == 1) {
            // the direction's old value is read on the first iteration too
            // (and ignored there): a load on every path
            P.a[s][0] = (T)load_any<TA>(va, ok, at, c.safe);
            if constexpr (DIR)
                P.b[s][0] = load_any<T>(vb, ok, at, c.safe);
            continue;
        } else {
            // two cells per lane: one pair access per stream on every path.
            // Pairs start at even indices (rows start at multiples of 16), so
            // an x-halo lane reads its cell with the pair neighbour; a zero
            // ghost reads the block's safe pair (masked where used).  On a
            // periodic x axis (PX) the last group of an odd row takes its
            // second cell from the row's start, read on every path too.
            const int64_t ap = ok ? (at & ~(int64_t)1) : c.safe;
            const bool odd = (at & 1) != 0;
            T a2[2], b2[2];
            vload<T, 2>(va + ap, a2);
            if constexpr (DIR)
                vload<T, 2>(vb + ap, b2);
            P.a[s][0] = single && odd ? a2[1] : a2[0];
            P.a[s][1] = single ? (T)0 : a2[1];
            if constexpr (DIR) {
                P.b[s][0] = single && odd ? b2[1] : b2[0];
                P.b[s][1] = single ? (T)0 : b2[1];
            }
            if constexpr (PX) {
                // the row's first cell: n0 - 1 cells back from the last
                const int64_t aw = ok && split ? at - c.wrapback : c.safe;
                const T wa = va[aw];
                P.a[s][1] = split ? wa : P.a[s][1];
                if constexpr (DIR) {
                    const T wb = vb[aw];
                    P.b[s][1] = split ? wb : P.b[s][1];
                }
            }
            if constexpr (MODE == 2) {
                // the parents of the slot's cells (the parent level is FP64;
                // an FP32 level rounds its values): the second cell is the
                // row's first (split), or a zero ghost past a non-periodic
                // row's end, or absent (an x-halo slot)
                // A pair starts at an even x, so both cells have one parent
                // unless the second is the wrapped first cell (PX).
                const int x0 = c.sgi[s];
                const bool ghost1 = single || (!split && x0 + 1 > c.wrapback);
                const int64_t prow =
                    par.base + par.sy * (c.sgj[s] >> 1) + par.sz * (m.zc >> 1);
                P.b[s][0] = (T)load_any<TP>(par.xc, ok, prow + (x0 >> 1), par.base);
                T b1 = P.b[s][0];
                if constexpr (PX) {
                    const T bw = (T)load_any<TP>(par.xc, ok && split, prow, par.base);
                    b1 = split ? bw : b1;
                }
                P.b[s][1] = ghost1 ? (T)0 : b1;
            }
        }
    }
}

// Value of cell j of slot s: the new direction fma(beta, p_old, D^-1 r)
// (D^-1 r alone on the first iteration) for MODE 1, fma(beta, parent, a) for MODE 2,
// else the loaded vector.
template <class T, int MODE, int CPL>
__device__ __forceinline__ T tile3_value(const Geo& g, const Tile3Ctx<CPL>& c,
                                         const Plane3Meta& m,
                                         const Plane3<T, MODE, CPL>& P, int s, int j,
                                         bool first, T beta)
{
    T v;
    if constexpr (MODE == 0) {
        v = P.a[s][j];
    } else if constexpr (MODE == 2) {
        v = fma_t<T>(beta, P.b[s][j], P.a[s][j]);
    } else {
        T dinv;
        if (s < tile3::RY)
            dinv = ctx_dinv<T, MODE>(g, c, c.wy[s] + m.wz, c.wx[j]); // uniform y, z part
        else
            dinv = ctx_dinv<T, MODE>(g, c, m.wz, c.hwxy[j]);
        const T z = dinv * P.a[s][j];
        v = first ? z : fma_t<T>(beta, P.b[s][j], z);
    }
    // loads are raw: a zero ghost slot read some owned cell
    return c.colb[s] >= 0 && m.off >= 0 ? v : (T)0;
}

// Writes the plane's values into LDS plane pl and own[]; the halo slot
// fills its ring cells.
template <class T, int DIR, int CPL>
__device__ __forceinline__ void tile3_finish(const Geo& g, const Tile3Ctx<CPL>& c,
                                             const Plane3Meta& m,
                                             const Plane3<T, DIR, CPL>& P, bool first,
                                             T beta, T (*pl)[tile3::prow<CPL>()],
                                             T (&own)[tile3::RY][CPL])
{
    using namespace tile3;
#pragma unroll
    for (int q = 0; q < RY; ++q) {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            own[q][j] = tile3_value<T, DIR, CPL>(g, c, m, P, q, j, first, beta);
        // one aligned access per lane (a pair with two cells per lane)
        vstore<T, CPL>(&pl[c.w * RY + q + 1][c.lane * CPL + pcol<CPL>()], own[q]);
    }
    if constexpr (CPL == 1) {
        if (c.hkind != 0)
            pl[c.hrow][c.hcol] = tile3_value<T, DIR, CPL>(g, c, m, P, RY, 0, first, beta);
    } else {
        if (c.hkind == 1) {
            T hv[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                hv[j] = tile3_value<T, DIR, CPL>(g, c, m, P, RY, j, first, beta);
            vstore<T, CPL>(&pl[c.hrow][c.hcol], hv);
        } else if (c.hkind == 2) {
            pl[c.hrow][c.hcol] = tile3_value<T, DIR, CPL>(g, c, m, P, RY, 0, first, beta);
        }
    }
}

// y = A p at own cell (row q, cell j) of a plane: centre pc[j], z neighbours
// pm / pp, x neighbours from the lane's group or the LDS plane, y neighbours
// from the LDS plane; reference stencil order {0}, -x, +x, -y, +y, -z, +z
// with zero couplings across walls.
// The x and y neighbours of a lane's cells in own row q, from the LDS
// plane (the x neighbours inside a lane's pair come from registers).  With
// two cells per lane: the pairs left, right, above and below the lane's
// pair, four aligned 8-byte reads.
template <class T, int CPL>
struct RowNb
{
    T xm[CPL], xp[CPL], ym[CPL], yp[CPL];
};

template <class T, int CPL>
__device__ __forceinline__ RowNb<T, CPL> tile3_neighbours(const Tile3Ctx<CPL>& c,
                                                          const T (*pl)[tile3::prow<CPL>()],
                                                          int q, const T (&pc)[CPL])
{
    using namespace tile3;
    RowNb<T, CPL> nb;
    const int jj = c.w * RY + q;
    const int c0 = c.lane * CPL + pcol<CPL>(); // LDS column of the lane's first cell
    if constexpr (CPL == 1) {
        nb.xm[0] = pl[jj + 1][c0 - 1];
        nb.xp[0] = pl[jj + 1][c0 + 1];
        nb.ym[0] = pl[jj][c0];
        nb.yp[0] = pl[jj + 2][c0];
    } else {
        T L[2], R[2], U[2], D[2];
        vload<T, 2>(&pl[jj + 1][c0 - 2], L);
        vload<T, 2>(&pl[jj + 1][c0 + 2], R);
        vload<T, 2>(&pl[jj][c0], U);
        vload<T, 2>(&pl[jj + 2][c0], D);
        nb.xm[0] = L[1];
        nb.xp[0] = pc[1];
        nb.xm[1] = pc[0];
        nb.xp[1] = R[0];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            nb.ym[j] = U[j];
            nb.yp[j] = D[j];
        }
    }
    return nb;
}

template <class T, int DIR, int CPL>
__device__ __forceinline__ T tile3_apply(const Geo& g, const Tile3Ctx<CPL>& c,
                                         const Plane3Meta& m, const RowNb<T, CPL>& nb, int q,
                                         int j, const T (&pc)[CPL], T pm, T pp)
{
    const T s = offdiag<T>(g);
    T y = ctx_diag<T, DIR>(g, c, c.wy[q] + m.wz, c.wx[j]) * pc[j];
    y = fma_t<T>(c.xwall_m[j] ? (T)0 : s, nb.xm[j], y);
    y = fma_t<T>(c.xwall_p[j] ? (T)0 : s, nb.xp[j], y);
    y = fma_t<T>(c.ywall_m[q] ? (T)0 : s, nb.ym[j], y);
    y = fma_t<T>(c.ywall_p[q] ? (T)0 : s, nb.yp[j], y);
    y = fma_t<T>(m.zwall_m ? (T)0 : s, pm, y);
    y = fma_t<T>(m.zwall_p ? (T)0 : s, pp, y);
    return y;
}

// The z march shared by the direction (MODE 1), update (MODE 0) and
// multigrid kernels; epi.cells<B>() receives each own row's centre values and
// A applied to the values.
//
// With one cell per lane the loop is unrolled by two with ping-pong
// buffers, so no register copy of a plane's loads (which would wait for
// them) happens before their use: at step s (plane k = k0 + s; with REV,
// k1 - 1 - s, and "k+1" below means the next plane down) the march's
// plane k+1 is in P[s & 1], plane k+2 loads into P[(s + 1) & 1], and the
// epilogue's data of plane k sits in its buffer s & 1 while plane k+1's
// loads into (s + 1) & 1.  With two cells per lane (FP32) the second set of
// live buffers costs more registers than the overlap gains: that loop keeps
// one buffer pair and copies (P0 = P1, epi.shift()) after each plane.
// Epilogue interface: prefetch<B>(c, m, k) loads plane k's data into buffer
// B (k outside [k0, k1): values never used), take_raw<B>(P) takes the marched operand
// instead, cells<B>() uses buffer B, shift() copies buffer 1 into 0 (CPL = 2
// only), rotate() runs after every step's barrier; late(c) as described at
// epi_loads_late.
//
// REV marches the chunk downwards (k1-1 to k0).  A kernel that follows one
// marching the other way then starts on the planes its predecessor touched
// last, which are still in the Infinity Cache.  The stencil sees the same
// operands in the same order either way.
template <class T, int DIR, int CPL, bool PX = false, bool REV = false, class TA = T,
          class Epi, class TP = double>
__device__ __forceinline__ void tile3_march(
    const Geo& g, const Tile3Grid& tg, const TA* __restrict__ va,
    const T* __restrict__ vb, bool first, T beta,
    T (*pl)[tile3::TY + 2][tile3::prow<CPL>()], Epi&& epi,
    const Tile3ParentT<TP> par = Tile3ParentT<TP>{nullptr, 0, 0, 0, 0.0})
{
    using namespace tile3;
    using E = std::decay_t<Epi>;
    Tile3Ctx<CPL> c = tile3_ctx<CPL>(g, tg);
    if (!c.active)
        return; // the whole block (active depends on the block only)
    c.tab = nullptr;
    if constexpr (tile3_tab<T, DIR>()) {
        __shared__ T tab[16];
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (threadIdx.x == q) {
                tab[q] = diag_of<T>(g, q);
                tab[8 + q] = dinv_of<T>(g, q);
            }
        __syncthreads();
        c.tab = tab;
    }
    const bool need_b = DIR == 2 || !first;
    // MODE 1 reads vb on every path (first iteration included, where its
    // values are ignored); a caller without an old direction gets va's (a
    // caller with a widened va passes its vb)
    if constexpr (std::is_same<TA, T>::value)
        if (vb == nullptr)
            vb = va;
    Plane3<T, DIR, CPL> P0, P1;
    // pm: the plane behind the march, pc: the current one, pp: the one ahead
    T pm[RY][CPL], pc[RY][CPL], pp[RY][CPL];
    constexpr int D = REV ? -1 : 1;
    const int kbeg = REV ? c.k1 - 1 : c.k0;
    // the plane two steps ahead, clamped to the plane just past the chunk
    auto ahead2 = [&](int k) { return REV ? max(k - 2, c.k0 - 1) : min(k + 2, c.k1); };
    // prologue: the plane behind kbeg (own column only), kbeg into LDS, the
    // plane ahead in flight
    Plane3Meta mm = plane3_meta(g, kbeg - D);
    tile3_load<T, DIR, CPL, PX, TA, TP>(c, mm, va, vb, need_b, par, P0);
#pragma unroll
    for (int q = 0; q < RY; ++q)
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            pm[q][j] = tile3_value<T, DIR, CPL>(g, c, mm, P0, q, j, first, beta);
    Plane3Meta mc = plane3_meta(g, kbeg);
    tile3_load<T, DIR, CPL, PX, TA, TP>(c, mc, va, vb, need_b, par, P1);
    tile3_finish<T, DIR, CPL>(g, c, mc, P1, first, beta, pl[0], pc);
    Plane3Meta mn = plane3_meta(g, kbeg + D);
    tile3_load<T, DIR, CPL, PX, TA, TP>(c, mn, va, vb, need_b, par, P0);
    if constexpr (epi_takes_raw<E>::value)
        epi.template take_raw<0>(P1);
    epi.template prefetch<0>(c, mc, kbeg);
    if constexpr (epi_loads_late<E>::value)
        epi.late(c);
    epi.rotate();
    __syncthreads();
    auto step = [&](int k, auto parity) {
        constexpr int S = decltype(parity)::value;
        Plane3<T, DIR, CPL>& Pf = S == 0 ? P0 : P1; // plane k+D
        Plane3<T, DIR, CPL>& Pl = S == 0 ? P1 : P0; // plane k+2D, started here
        // the last step loads the plane past the chunk again (an L2 hit,
        // never used) rather than branching around the loads
        const Plane3Meta mnn = plane3_meta(g, ahead2(k));
        tile3_load<T, DIR, CPL, PX, TA, TP>(c, mnn, va, vb, need_b, par, Pl);
        tile3_finish<T, DIR, CPL>(g, c, mn, Pf, first, beta, pl[S ^ 1], pp);
        if constexpr (epi_takes_raw<E>::value)
            epi.template take_raw<S ^ 1>(Pf);
        epi.template prefetch<S ^ 1>(c, mn, k + D);
#pragma unroll
        for (int q = 0; q < RY; ++q) {
            // a group with no valid cell has nothing to compute or store
            if (c.valid[q][0]) {
                const RowNb<T, CPL> nb = tile3_neighbours<T, CPL>(c, pl[S], q, pc[q]);
                T y[CPL];
#pragma unroll
                for (int j = 0; j < CPL; ++j)
                    y[j] = tile3_apply<T, DIR, CPL>(g, c, mc, nb, q, j, pc[q],
                                               REV ? pp[q][j] : pm[q][j],
                                               REV ? pm[q][j] : pp[q][j]);
                epi.template cells<S>(c, mc, q, pc[q], y);
            }
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                pm[q][j] = pc[q][j];
                pc[q][j] = pp[q][j];
            }
        }
        __syncthreads();
        if constexpr (epi_loads_late<E>::value)
            epi.late(c);
        mc = mn;
        mn = mnn;
        epi.rotate();
    };
    if constexpr (CPL == 1) {
        int k = kbeg, left = c.k1 - c.k0;
        for (; left >= 2; left -= 2, k += 2 * D) {
            step(k, std::integral_constant<int, 0>{});
            step(k + D, std::integral_constant<int, 1>{});
        }
        if (left > 0)
            step(k, std::integral_constant<int, 0>{});
    } else {
        int cur = 0;
        for (int s = 0, k = kbeg; s < c.k1 - c.k0; ++s, k += D) {
            const Plane3Meta mnn = plane3_meta(g, ahead2(k));
            tile3_load<T, DIR, CPL, PX, TA, TP>(c, mnn, va, vb, need_b, par, P1);
            tile3_finish<T, DIR, CPL>(g, c, mn, P0, first, beta, pl[cur ^ 1], pp);
            if constexpr (epi_takes_raw<E>::value)
                epi.template take_raw<1>(P0);
            epi.template prefetch<1>(c, mn, k + D);
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                if (c.valid[q][0]) {
                    const RowNb<T, CPL> nb = tile3_neighbours<T, CPL>(c, pl[cur], q, pc[q]);
                    T y[CPL];
#pragma unroll
                    for (int j = 0; j < CPL; ++j)
                        y[j] = tile3_apply<T, DIR, CPL>(g, c, mc, nb, q, j, pc[q],
                                                   REV ? pp[q][j] : pm[q][j],
                                                   REV ? pm[q][j] : pp[q][j]);
                    epi.template cells<0>(c, mc, q, pc[q], y);
                }
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    pm[q][j] = pc[q][j];
                    pc[q][j] = pp[q][j];
                }
            }
            __syncthreads();
            if constexpr (epi_loads_late<E>::value)
                epi.late(c);
            cur ^= 1;
            P0 = P1;
            mc = mn;
            mn = mnn;
            epi.shift();
            epi.rotate();
        }
    }
}

} // namespace cfmi
