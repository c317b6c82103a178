// cf_mg.hip -- kernels of the multigrid preconditioner (cf_mg.hpp).
//
// Arithmetic order, shared with the oracle's restatement so that the
// V-cycle is bit-identical: A x = diag x, then fma over the -x, +x, -y, +y,
// -z, +z couplings (a missing neighbour contributes fma(0, 0, .)); a Jacobi
// sweep is fma(w, b - A x, x); restriction sums the children's residuals in
// z, y, x order starting from 0; prolongation adds the parent's value.
//
// Sweeps are fused with what precedes them: the first sweep from zero reads
// x = w b at each stencil point from b, and the first sweep after the coarse
// correction reads fma(alpha, xc[parent], x) at each point, so neither x = w b
// nor the prolongated x is written out.  The residual is formed inside the
// restriction, and the last level-0 sweep also reduces r.z for the PCG.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "cf_mg.hpp"
#include "cf_solver.hpp"
#include "cf_stream.hpp"
#include "cf_tile3.hpp"

namespace cfmi
{

// ---- level 0: padded layout, matrix-free operator ---------------------------------

// table entry by total wall count without dynamic indexing of the argument
__device__ __forceinline__ double tab_sel(const MgTab& t, int cnt)
{
    double v = t.v[0];
#pragma unroll
    for (int q = 1; q <= 6; ++q)
        v = cnt == q ? t.v[q] : v;
    return v;
}

__device__ __forceinline__ int wrap_nbr(int periodic, int n, int i, int dir)
{
    const int m = i + dir;
    if (periodic)
        return m < 0 ? n - 1 : (m >= n ? 0 : m);
    return m;
}

// x = w b (w rounded to T on an FP32 level)
template <int DIM, class T = double>
__global__ __launch_bounds__(256) void k_mg0_first(Geo g, MgTab wd, const T* __restrict__ b,
                                                   T* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const int ub = wall_count(g, 1, j) + wall_count(g, 2, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            x[c] = (T)tab_sel(wd, ub + wall_count(g, 0, i)) * b[c];
        }
    });
}

// b32 = (float) r over the owned cells: the FP32 level 0's right-hand side
__global__ __launch_bounds__(256) void k_mg_cast32(Geo g, const double* __restrict__ r,
                                                   float* __restrict__ b32, const PcgState* st)
{
    if (st->done)
        return;
    for_owned_rows(g, [&](int64_t base, int, int) {
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock)
            b32[base + i] = (float)r[base + i];
    });
}

// One Jacobi sweep xo = fma(w, b - A x, x) on a padded level (row kernel,
// one pass over memory), x from SRC (the parent of a cell at pm); RZ adds
// r.z.
template <int DIM, int SRC, bool RZ>
__global__ __launch_bounds__(256) void k_mg0_smooth(Geo g, MgTab wd, const double* __restrict__ xi,
                                                    const double* __restrict__ b, MgParentMap pm,
                                                    double alpha, const double* __restrict__ xc,
                                                    double* __restrict__ xo, PcgState* st,
                                                    RedBuf rb)
{
    if (st->done)
        return;
    double hi[1] = {0.0}, lo[1] = {0.0};
    const double s = offdiag<double>(g);
    // the next level coarsens every axis with more than 2 cells
    const bool cx = g.n[0] > 2, cy = g.n[1] > 2, cz = DIM > 2 && g.n[2] > 2;
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        // level-0 coordinates of the y / z neighbour rows
        const int jn[2] = {wrap_nbr(g.periodic[1], g.n[1], j, -1),
                           wrap_nbr(g.periodic[1], g.n[1], j, +1)};
        const int kn[2] = {DIM > 2 ? wrap_nbr(g.periodic[2], g.n[2], k, -1) : 0,
                           DIM > 2 ? wrap_nbr(g.periodic[2], g.n[2], k, +1) : 0};
        auto parent = [&](int i0, int j0, int k0) -> int64_t {
            const int P0 = cx ? i0 >> 1 : i0, P1 = cy ? j0 >> 1 : j0, P2 = cz ? k0 >> 1 : k0;
            return pm.base + P0 + pm.sy * P1 + pm.sz * P2;
        };
        // value at a stencil point (0 at a ghost across a wall)
        auto val = [&](int64_t at, bool wall, int cnt, int i0, int j0, int k0) -> double {
            if constexpr (SRC == MG_PLAIN) {
                return xi[at];
            } else if constexpr (SRC == MG_FIRST) {
                return wall ? 0.0 : tab_sel(wd, cnt) * b[at];
            } else {
                return wall ? 0.0 : fma(alpha, xc[parent(i0, j0, k0)], xi[at]);
            }
        };
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            const int cnt = rs.cnt + lx.cnt;
            const double vc = val(c, false, cnt, i, j, k);
            double y = diag_pick<double>(g, rs.cnt, lx.cnt) * vc;
            y = fma(lx.wall[0] ? 0.0 : s,
                    val(c + lx.off[0], lx.wall[0], rs.cnt + lx.ncnt[0], i + (int)lx.off[0], j, k), y);
            y = fma(lx.wall[1] ? 0.0 : s,
                    val(c + lx.off[1], lx.wall[1], rs.cnt + lx.ncnt[1], i + (int)lx.off[1], j, k), y);
            y = fma(rs.wall[1][0] ? 0.0 : s,
                    val(c + rs.off[1][0], rs.wall[1][0], rs.ncnt[1][0] + lx.cnt, i, jn[0], k), y);
            y = fma(rs.wall[1][1] ? 0.0 : s,
                    val(c + rs.off[1][1], rs.wall[1][1], rs.ncnt[1][1] + lx.cnt, i, jn[1], k), y);
            if constexpr (DIM > 2) {
                y = fma(rs.wall[2][0] ? 0.0 : s,
                        val(c + rs.off[2][0], rs.wall[2][0], rs.ncnt[2][0] + lx.cnt, i, j, kn[0]),
                        y);
                y = fma(rs.wall[2][1] ? 0.0 : s,
                        val(c + rs.off[2][1], rs.wall[2][1], rs.ncnt[2][1] + lx.cnt, i, j, kn[1]),
                        y);
            }
            const double bc = b[c];
            const double out = fma(tab_sel(wd, cnt), bc - y, vc);
            xo[c] = out;
            if constexpr (RZ)
                dd_add_prod(hi[0], lo[0], bc, out);
        }
    });
    if constexpr (RZ) {
        grid_reduce<1>(hi, lo, rb, [&](double* sm) {
            if (st->iter > 0)
                st->beta = sm[0] / st->rho;
            st->rho = sm[0];
        });
    }
}

// ---- tiled levels, 3-D: z-marching tiles (cf_tile3.hpp) ----------------------------
// Level 0 and the coarse levels whose cells are all full 2^l boxes (a padded
// layout, like level 0, with that level's Geo: couplings -(4^l s), diagonal
// 4^l (6 - SOLID walls) s) when every axis is coarsened further.  The tile
// grid has an even z chunk so that the two z children of a parent are
// marched by one block; the tile's x origin and rows are even, so the x
// children of a parent are an even lane and the next one, and its y
// children two rows of one lane.

// Sweep epilogue: xo = fma(w, b - y, x) at each own cell, w = omega / diag
// from gw's Jacobi table; RZ also accumulates r.z = b . xo.  RAW: b is the
// marched operand (the first sweep), taken from the tile's loads.  T is the
// level's vector type.  On an FP32 level the RZ sweep (level 0's last) reads
// the FP64 right-hand side r (TB = double): the sweep computes in FP32 from
// (float) r, and r.z uses r itself and (double) z.
template <class T, bool RZ>
using MgRhsT = std::conditional_t<std::is_same<T, float>::value && RZ, double, T>;

template <class T, int CPL, bool RZ, bool RAW>
struct MgSweepEpi
{
    using TB = MgRhsT<T, RZ>;
    static constexpr bool kRaw = RAW;
    const Geo& gw;
    const TB* __restrict__ b;
    T* __restrict__ xo;
    double hi[1] = {0.0}, lo[1] = {0.0};
    TB bb[2][tile3::RY][CPL]; // b of a plane, alternating with the march's steps
    template <int B, class PL>
    __device__ void take_raw(const PL& P)
    {
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q)
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                bb[B][q][j] = (TB)P.a[q][j];
    }
    template <int B>
    __device__ void prefetch(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int k)
    {
        if constexpr (!RAW) {
#pragma unroll
            for (int q = 0; q < tile3::RY; ++q) {
                const bool ok = k < c.k1 && c.valid[q][0];
                if constexpr (CPL == 1) {
                    bb[B][q][0] = load_any<TB>(b, ok, c.colb[q] + m.off, c.safe); // raw: read on valid rows only
                } else {
                    // an aligned pair on every path (raw, as above)
                    const int64_t ap = ok ? c.colb[q] + m.off : c.safe;
                    vload<TB, CPL>(b + ap, bb[B][q]);
                }
            }
        }
    }
    __device__ void shift()
    {
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q)
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                bb[0][q][j] = bb[1][q][j];
    }
    __device__ void rotate() {}
    template <int B>
    __device__ void cells(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int q, const T (&pc)[CPL],
                          const T (&y)[CPL])
    {
        T out[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const T w = dinv_pick<T>(gw, c.wy[q] + m.wz, c.wx[j]);
            out[j] = fma_t<T>(w, (T)bb[B][q][j] - y[j], pc[j]);
        }
        if constexpr (CPL == 1)
            xo[c.colb[q] + m.off] = out[0];
        else
            vstore_valid<T, CPL>(xo + c.colb[q] + m.off, out, c.valid[q]);
        if constexpr (RZ) {
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                if (c.valid[q][j])
                    dd_acc_prod<false>(hi[0], lo[0], (double)bb[B][q][j], (double)out[j]);
        }
    }
};

// MODE: 1 = first sweep from zero (values w b, read from b through gw's
// Jacobi table), 0 = plain, 2 = after the coarse correction.  CPL, PX: the
// tile shape (two cells per lane for an FP32 level, with its periodic-x
// variant).
template <class T, int MODE, bool RZ, int CPL = 1, bool PX = false, class TP = double>
__global__ __launch_bounds__(kTileBlock) void k_mg_smooth3(Geo g, Geo gw, Tile3Grid tg,
                                                    const T* __restrict__ xi,
                                                    const MgRhsT<T, RZ>* __restrict__ b,
                                                    Tile3ParentT<TP> par, T* __restrict__ xo,
                                                    PcgState* st, RedBuf rb)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ __align__(16) T pl[2][TY + 2][prow<CPL>()];
    MgSweepEpi<T, CPL, RZ, MODE == 1> epi{gw, b, xo};
    if constexpr (MODE == 1) {
        static_assert(std::is_same<MgRhsT<T, RZ>, T>::value, "the first sweep marches b itself");
        tile3_march<T, 1, CPL, PX>(gw, tg, b, nullptr, true, (T)0, pl, epi);
    } else {
        tile3_march<T, MODE, CPL, PX>(g, tg, xi, nullptr, false, (T)par.alpha, pl, epi, par);
    }
    if constexpr (RZ) {
        dd_norm(epi.hi[0], epi.lo[0]);
        grid_reduce<1, kTileBlock>(epi.hi, epi.lo, rb, [&](double* sm) {
            if (st->iter > 0)
                st->beta = sm[0] / st->rho;
            st->rho = sm[0];
        });
    }
}

// Residual-restriction epilogue: r = b - A x per own cell; after each plane
// an even lane adds its own and the next lane's residuals of each row pair
// (children in x, y order), and after the second plane of a z pair (or the
// last plane) writes the parent.  rotate() runs on every lane of an active
// block, so the lane exchange is not under divergent control.
// TO: the parent level's type (the sums, in FP64, are rounded to an FP32
// parent level).
template <class T, int CPL = 1, class TO = double>
struct MgRestrictEpi
{
    const Geo& g;
    const T* __restrict__ b;
    TO* __restrict__ bc;
    MgParentMap pm;
    const Tile3Ctx<CPL>* ctx = nullptr;
    // b of a plane, alternating with the march's steps; the residuals of
    // an FP32 level are summed as doubles
    T bb[2][tile3::RY][CPL], r[tile3::RY][CPL] = {};
    double sum[tile3::RY / 2] = {};
    int knext = -1, kcur = -1;
    template <int B>
    __device__ void prefetch(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int k)
    {
        ctx = &c;
        knext = k;
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q) {
            const bool ok = k < c.k1 && c.valid[q][0];
            if constexpr (CPL == 1) {
                bb[B][q][0] = load_any<T>(b, ok, c.colb[q] + m.off, c.safe); // raw: read on valid rows only
            } else {
                const int64_t ap = ok ? c.colb[q] + m.off : c.safe;
                vload<T, CPL>(b + ap, bb[B][q]);
            }
        }
    }
    __device__ void shift()
    {
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q)
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                bb[0][q][j] = bb[1][q][j];
    }
    template <int B>
    __device__ void cells(const Tile3Ctx<CPL>&, const Plane3Meta&, int q, const T (&)[CPL],
                          const T (&y)[CPL])
    {
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            r[q][j] = bb[B][q][j] - y[j];
    }
    // the children of a parent in z, y, x order: the x pair is an even lane
    // and the next one (CPL = 1) or one lane's two cells (CPL = 2)
    __device__ void rotate()
    {
        const Tile3Ctx<CPL>& c = *ctx;
        if (kcur >= 0) {
            const int gi = c.i0 + c.lane * CPL;
            const bool lead = CPL == 2 || (c.lane & 1) == 0;
#pragma unroll
            for (int p = 0; p < tile3::RY / 2; ++p) {
                double a0, a1, b0, b1;
                bool x1;
                if constexpr (CPL == 1) {
                    a0 = (double)r[2 * p][0];
                    a1 = (double)r[2 * p + 1][0];
                    b0 = __shfl_down(a0, 1);
                    b1 = __shfl_down(a1, 1);
                    x1 = gi + 1 < g.n[0];
                } else {
                    a0 = (double)r[2 * p][0];
                    b0 = (double)r[2 * p][1];
                    a1 = (double)r[2 * p + 1][0];
                    b1 = (double)r[2 * p + 1][1];
                    x1 = gi + 1 < g.n[0];
                }
                if (lead && c.valid[2 * p][0]) {
                    sum[p] = sum[p] + a0;
                    if (x1)
                        sum[p] = sum[p] + b0;
                    if (c.valid[2 * p + 1][0]) {
                        sum[p] = sum[p] + a1;
                        if (x1)
                            sum[p] = sum[p] + b1;
                    }
                }
            }
            if ((kcur & 1) || kcur == g.n[2] - 1) {
#pragma unroll
                for (int p = 0; p < tile3::RY / 2; ++p) {
                    if (lead && c.valid[2 * p][0]) {
                        const int gj = c.j0 + c.w * tile3::RY + 2 * p;
                        bc[pm.base + (gi >> 1) + pm.sy * (gj >> 1) + pm.sz * (kcur >> 1)] =
                            (TO)sum[p];
                    }
                    sum[p] = 0.0;
                }
            }
        }
        kcur = knext;
    }
};

template <class T, int CPL = 1, bool PX = false, class TO = double>
__global__ __launch_bounds__(kTileBlock) void k_mg_resrestrict3(Geo g, Tile3Grid tg, MgParentMap pm,
                                                         const T* __restrict__ x,
                                                         const T* __restrict__ b,
                                                         TO* __restrict__ bc,
                                                         const PcgState* st)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ __align__(16) T pl[2][TY + 2][prow<CPL>()];
    MgRestrictEpi<T, CPL, TO> epi{g, b, bc, pm};
    tile3_march<T, 0, CPL, PX>(g, tg, x, nullptr, false, (T)0, pl, epi);
}

// children of parent index P along an axis of n cells
__device__ __forceinline__ void children(bool coarsen, int P, int n, int& lo, int& hi)
{
    if (coarsen) {
        lo = 2 * P;
        hi = 2 * P + 1 < n ? 2 * P + 1 : 2 * P;
    } else {
        lo = hi = P;
    }
}

__device__ __forceinline__ void unflatten(int id, const int (&n)[3], int (&c)[3])
{
    const unsigned u = (unsigned)id;
    c[0] = (int)(u % (unsigned)n[0]);
    const unsigned r = u / (unsigned)n[0];
    c[1] = (int)(r % (unsigned)n[1]);
    c[2] = (int)(r / (unsigned)n[1]);
}

__device__ __forceinline__ int flat(const int (&n)[3], int i, int j, int k)
{
    return i + n[0] * (j + n[1] * k);
}

#define CF_GRID_STRIDE(id, count) \
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < (count); id += gridDim.x * blockDim.x)

// b_c[parent] = sum over children (z, y, x order) of (b - A x) on a padded
// level; one thread per parent of C (written at pm).
template <int DIM>
__global__ __launch_bounds__(256) void k_mg0_resrestrict(Geo g, MgLev C, MgParentMap pm,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ b,
                                                         double* __restrict__ bc,
                                                         const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, C.ncell)
    {
        int P[3];
        unflatten(id, C.n, P);
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d)
            children(d < g.dim && g.n[d] > 2, P[d], d < g.dim ? g.n[d] : 1, lo[d], hi[d]);
        double sum = 0.0;
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j) {
                const RowStencil rs = row_stencil(g, j, k);
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const int64_t c = g.idx(g.h + i, g.h + j, DIM == 3 ? g.h + k : 0);
                    const LaneX lx = lane_x(g, i);
                    sum = sum + (b[c] - apply_A<DIM, double>(g, rs, lx, x, c));
                }
            }
        bc[pm.base + P[0] + pm.sy * P[1] + pm.sz * P[2]] = sum;
    }
}

// ---- compact levels: matrix-free Galerkin operators -------------------------------

// level-0 cells of the box of coarse index P along axis d
__device__ __forceinline__ int agg_ext(const MgLev& L, int d, int P)
{
    const int lo = P << L.shift[d];
    const int hi = min(lo + (1 << L.shift[d]), L.n0[d]);
    return hi - lo;
}

// Face counts of cell c: area[d] (faces to a neighbouring box along d) and K
// (boundary faces that are not SOLID walls).
__device__ __forceinline__ void lev_counts(const MgLev& L, const int (&c)[3], int64_t (&area)[3],
                                           int64_t& K)
{
    int64_t e[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d < L.dim)
            e[d] = agg_ext(L, d, c[d]);
    K = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        area[d] = d == 0 ? e[1] * e[2] : (d == 1 ? e[0] * e[2] : e[0] * e[1]);
        if (d >= L.dim)
            continue;
        const bool lo_cut = !L.periodic[d] && c[d] == 0 && L.solid[d][0];
        const bool hi_cut = !L.periodic[d] && c[d] == L.n[d] - 1 && L.solid[d][1];
        K += (lo_cut ? 0 : area[d]) + (hi_cut ? 0 : area[d]);
    }
}

__device__ __forceinline__ double lev_w(const MgLev& L, const int (&c)[3])
{
    int64_t area[3], K;
    lev_counts(L, c, area, K);
    return L.omega / ((double)K * L.s);
}

// neighbour coordinates along d (dir -1 / +1) with the periodic wrap;
// false outside a non-periodic axis
__device__ __forceinline__ bool lev_nbr(const MgLev& L, const int (&c)[3], int d, int dir,
                                        int (&m)[3])
{
    m[0] = c[0];
    m[1] = c[1];
    m[2] = c[2];
    m[d] += dir;
    if (m[d] < 0 || m[d] >= L.n[d]) {
        if (!L.periodic[d])
            return false;
        m[d] = m[d] < 0 ? L.n[d] - 1 : 0;
    }
    return true;
}

// y = A x at cell c of a compact level, x(m) from val(coords, index); also
// returns the cell's Jacobi weight.
template <class V>
__device__ __forceinline__ double lev_apply(const MgLev& L, const int (&c)[3], double xc, V&& val,
                                            double& w)
{
    int64_t area[3], K;
    lev_counts(L, c, area, K);
    const double diag = (double)K * L.s;
    w = L.omega / diag;
    double acc = diag * xc;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= L.dim)
            break;
        const double ca = -((double)area[d] * L.s);
        int m[3];
        const bool hm = lev_nbr(L, c, d, -1, m);
        const double xm = hm ? val(m, flat(L.n, m[0], m[1], m[2])) : 0.0;
        acc = fma(hm ? ca : 0.0, xm, acc);
        const bool hp = lev_nbr(L, c, d, +1, m);
        const double xp = hp ? val(m, flat(L.n, m[0], m[1], m[2])) : 0.0;
        acc = fma(hp ? ca : 0.0, xp, acc);
    }
    return acc;
}

__device__ __forceinline__ int lev_parent(const MgLev& F, const MgLev& C, const int (&f)[3])
{
    const int P0 = F.coarsen[0] ? f[0] >> 1 : f[0];
    const int P1 = F.coarsen[1] ? f[1] >> 1 : f[1];
    const int P2 = F.coarsen[2] ? f[2] >> 1 : f[2];
    return flat(C.n, P0, P1, P2);
}

// The per-cell bodies, shared by the grid kernels and k_mg_tail.
__device__ __forceinline__ void mg_first_cell(const MgLev& L, int id, const double* __restrict__ b,
                                              double* __restrict__ x)
{
    int c[3];
    unflatten(id, L.n, c);
    x[id] = lev_w(L, c) * b[id];
}

template <int SRC>
__device__ __forceinline__ void mg_smooth_cell(const MgLev& L, int id, const double* __restrict__ xi,
                                               const double* __restrict__ b, const MgLev& C,
                                               const double* __restrict__ xc,
                                               double* __restrict__ xo)
{
    int c[3];
    unflatten(id, L.n, c);
    auto val = [&](const int (&m)[3], int mid) -> double {
        if constexpr (SRC == MG_PLAIN)
            return xi[mid];
        else if constexpr (SRC == MG_FIRST)
            return lev_w(L, m) * b[mid];
        else
            return fma(L.alpha, xc[lev_parent(L, C, m)], xi[mid]);
    };
    const double vc = val(c, id);
    double w;
    const double y = lev_apply(L, c, vc, val, w);
    xo[id] = fma(w, b[id] - y, vc);
}

__device__ __forceinline__ void mg_resrestrict_cell(const MgLev& F, const MgLev& C, int id,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ bc)
{
    int P[3];
    unflatten(id, C.n, P);
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d)
        children(F.coarsen[d] != 0, P[d], F.n[d], lo[d], hi[d]);
    auto val = [&](const int (&)[3], int mid) -> double { return x[mid]; };
    double sum = 0.0;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const int c[3] = {i, j, k};
                const int fi = flat(F.n, i, j, k);
                double w;
                sum = sum + (b[fi] - lev_apply(F, c, x[fi], val, w));
            }
    bc[id] = sum;
}

__global__ __launch_bounds__(256) void k_mg_first(MgLev L, const double* __restrict__ b,
                                                  double* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, L.ncell) { mg_first_cell(L, id, b, x); }
}

template <int SRC>
__global__ __launch_bounds__(256) void k_mg_smooth(MgLev L, const double* __restrict__ xi,
                                                   const double* __restrict__ b, MgLev C,
                                                   const double* __restrict__ xc,
                                                   double* __restrict__ xo, const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, L.ncell) { mg_smooth_cell<SRC>(L, id, xi, b, C, xc, xo); }
}

__global__ __launch_bounds__(256) void k_mg_resrestrict(MgLev F, MgLev C,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ bc,
                                                        const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, C.ncell) { mg_resrestrict_cell(F, C, id, x, b, bc); }
}

// The coarsest level (every axis at most 2 cells, so at most 8 cells) by
// one wave: lane i holds cell i's x and exchanges neighbour values by
// shuffles, so its sweeps need no barrier.  Same arithmetic as
// mg_first_cell / mg_smooth_cell (a neighbour's w b is the value that lane
// computes itself).  The result goes to out.
__device__ __forceinline__ void mg_coarsest_wave(const MgLev& L, const double* __restrict__ b,
                                                 double* __restrict__ out, int sweeps)
{
    const int lane = threadIdx.x;
    const bool on = lane < L.ncell;
    int c[3];
    unflatten(on ? lane : 0, L.n, c);
    int64_t area[3], K;
    lev_counts(L, c, area, K);
    const double diag = (double)K * L.s;
    const double w = L.omega / diag;
    double cm[3] = {0.0, 0.0, 0.0}, cp[3] = {0.0, 0.0, 0.0};
    int lm[3] = {lane, lane, lane}, lp[3] = {lane, lane, lane};
    bool hm[3] = {false, false, false}, hp[3] = {false, false, false};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= L.dim)
            break;
        const double ca = -((double)area[d] * L.s);
        int m[3];
        hm[d] = lev_nbr(L, c, d, -1, m);
        if (hm[d])
            lm[d] = flat(L.n, m[0], m[1], m[2]);
        hp[d] = lev_nbr(L, c, d, +1, m);
        if (hp[d])
            lp[d] = flat(L.n, m[0], m[1], m[2]);
        cm[d] = hm[d] ? ca : 0.0;
        cp[d] = hp[d] ? ca : 0.0;
    }
    const double bv = on ? b[lane] : 0.0;
    double x = w * bv;
    for (int it = 1; it < sweeps; ++it) {
        double acc = diag * x;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= L.dim)
                break;
            const double xm = __shfl(x, lm[d]);
            const double xp = __shfl(x, lp[d]);
            acc = fma(cm[d], hm[d] ? xm : 0.0, acc);
            acc = fma(cp[d], hp[d] ? xp : 0.0, acc);
        }
        x = fma(w, bv - acc, x);
    }
    if (on)
        out[lane] = x;
}

// ---- the small levels in one workgroup ------------------------------------------------
// Same cycle as the host runs on the large levels (Multigrid::vcycle), one
// phase per sweep, __syncthreads between phases.  Every vector of these
// levels lives in LDS (level l: X, T, B of ncell each from element
// T->off[l]); the kernel reads level l0's right-hand side from global memory
// and writes its cycle result to T->res[l0].
#define CF_BLOCK_STRIDE(id, count) for (int id = threadIdx.x; id < (count); id += blockDim.x)

__global__ __launch_bounds__(kMgTailThreads) void k_mg_tail(const MgTail* __restrict__ T,
                                                            const PcgState* st)
{
    extern __shared__ double lds[];
    if (st->done)
        return;
    const int l0 = T->l0, last = T->last, nu = T->nu, coarse = T->coarse;
    auto X = [&](int l) { return lds + T->off[l]; };
    auto Tv = [&](int l) { return lds + T->off[l] + T->lev[l].ncell; };
    auto Bv = [&](int l) { return lds + T->off[l] + 2 * T->lev[l].ncell; };
    auto res = [&](int l) { return (mg_level_swaps(l == last, nu, coarse) & 1) ? Tv(l) : X(l); };
    {
        const double* bg = T->B[l0];
        double* b = Bv(l0);
        CF_BLOCK_STRIDE(id, T->lev[l0].ncell) { b[id] = bg[id]; }
        __syncthreads();
    }
    // down: pre-smoothing and restriction; the coarsest level in wave 0
    for (int l = l0; l <= last; ++l) {
        const MgLev L = T->lev[l];
        const MgLev& C = L; // unused parent level of the non-prolong sweeps
        const double* b = Bv(l);
        if (l == last) {
            if (threadIdx.x < kWave)
                mg_coarsest_wave(L, b, res(l), coarse);
            __syncthreads();
            break;
        }
        double* x = X(l);
        double* t = Tv(l);
        const int pre = nu;
        if (pre == 1) {
            CF_BLOCK_STRIDE(id, L.ncell) { mg_first_cell(L, id, b, x); }
        } else {
            CF_BLOCK_STRIDE(id, L.ncell) { mg_smooth_cell<MG_FIRST>(L, id, nullptr, b, C, nullptr, x); }
        }
        __syncthreads();
        for (int it = 2; it < pre; ++it) {
            CF_BLOCK_STRIDE(id, L.ncell) { mg_smooth_cell<MG_PLAIN>(L, id, x, b, C, nullptr, t); }
            __syncthreads();
            double* tmp = x;
            x = t;
            t = tmp;
        }
        const MgLev Cn = T->lev[l + 1];
        double* bn = Bv(l + 1);
        CF_BLOCK_STRIDE(id, Cn.ncell) { mg_resrestrict_cell(L, Cn, id, x, b, bn); }
        __syncthreads();
    }
    // up: prolongation fused into the first post-sweep
    for (int l = last - 1; l >= l0; --l) {
        const MgLev L = T->lev[l];
        const MgLev Cn = T->lev[l + 1];
        const double* b = Bv(l);
        // the pre-smoothing result: X after an even number of swaps
        const bool odd = (nu >= 2 ? nu - 2 : 0) & 1;
        double* x = odd ? Tv(l) : X(l);
        double* t = odd ? X(l) : Tv(l);
        const double* xc = res(l + 1);
        CF_BLOCK_STRIDE(id, L.ncell) { mg_smooth_cell<MG_PROLONG>(L, id, x, b, Cn, xc, t); }
        __syncthreads();
        double* tmp = x;
        x = t;
        t = tmp;
        for (int it = 1; it < nu; ++it) {
            CF_BLOCK_STRIDE(id, L.ncell) { mg_smooth_cell<MG_PLAIN>(L, id, x, b, L, nullptr, t); }
            __syncthreads();
            tmp = x;
            x = t;
            t = tmp;
        }
    }
    {
        const double* x = res(l0);
        double* xg = T->res[l0];
        CF_BLOCK_STRIDE(id, T->lev[l0].ncell) { xg[id] = x[id]; }
    }
}

// ---- MG-PCG scalars ----------------------------------------------------------------

__global__ __launch_bounds__(256) void k_mgpcg_init(Geo g, const double* __restrict__ b,
                                                    double* __restrict__ x,
                                                    double* __restrict__ r, PcgState* st,
                                                    double tol, int maxit, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        const double bc = b[c];
        r[c] = bc;
        x[c] = 0.0;
        dd_add_prod(hi[0], lo[0], bc, bc);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) {
        const double norm = sqrt(s[0]);
        st->bnorm = norm;
        st->tol = tol;
        st->maxit = maxit;
        st->iter = 0;
        st->rho = 0.0;
        st->alpha = 0.0;
        st->beta = 0.0;
        st->alpha_lag = 0.0;
        const bool zero = norm == 0.0;
        const bool bad = !isfinite(norm); // NaN or Inf in b: stop at once
        st->done = zero || bad ? 1 : (maxit <= 0 ? 1 : 0);
        st->converged = zero ? 1 : 0;
        st->res = zero ? 0.0 : (bad ? norm : 1.0);
        st->mg = 1;
        st->nonfinite = bad ? 1 : 0;
    });
}

__global__ __launch_bounds__(256) void k_mgpcg_rz(Geo g, const double* __restrict__ r,
                                                  const double* __restrict__ z, PcgState* st,
                                                  RedBuf rb)
{
    if (st->done)
        return;
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int, int, int) { dd_add_prod(hi[0], lo[0], r[c], z[c]); });
    grid_reduce<1>(hi, lo, rb, [&](double* s) {
        if (st->iter > 0)
            st->beta = s[0] / st->rho;
        st->rho = s[0];
    });
}

// ---- launchers -------------------------------------------------------------------------

static dim3 compact_grid(int64_t n)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    return dim3((unsigned)(b < kRedBlocks ? (b > 0 ? b : 1) : kRedBlocks));
}

#define CF_MG0_LAUNCH(kernel, ...)                                                      \
    do {                                                                                \
        if (g.dim == 2)                                                                 \
            hipLaunchKernelGGL(kernel<2>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, \
                               __VA_ARGS__);                                            \
        else                                                                            \
            hipLaunchKernelGGL(kernel<3>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, \
                               __VA_ARGS__);                                            \
    } while (0)

void launch_mg0_first(const Geo& g, const MgTab& wd0, const double* b, double* x,
                      const PcgState* st, hipStream_t s)
{
    CF_MG0_LAUNCH(k_mg0_first, g, wd0, b, x, st);
}

template <int DIM>
static void mg0_smooth_dim(const Geo& g, const MgTab& wd0, int src, bool rz, const double* xi,
                           const double* b, MgParentMap pm, double alpha, const double* xc,
                           double* xo, PcgState* st, RedBuf rb, hipStream_t s)
{
    const dim3 grid(stream_blocks(g)), blk(kBlock);
#define CF_MG0_SMOOTH(SRC, RZ)                                                                   \
    hipLaunchKernelGGL((k_mg0_smooth<DIM, SRC, RZ>), grid, blk, 0, s, g, wd0, xi, b, pm, alpha, \
                       xc, xo, st, rb)
    if (src == MG_FIRST) {
        if (rz)
            CF_MG0_SMOOTH(MG_FIRST, true);
        else
            CF_MG0_SMOOTH(MG_FIRST, false);
    } else if (src == MG_PLAIN) {
        if (rz)
            CF_MG0_SMOOTH(MG_PLAIN, true);
        else
            CF_MG0_SMOOTH(MG_PLAIN, false);
    } else {
        if (rz)
            CF_MG0_SMOOTH(MG_PROLONG, true);
        else
            CF_MG0_SMOOTH(MG_PROLONG, false);
    }
#undef CF_MG0_SMOOTH
}

void launch_mg0_smooth(const Geo& g, const MgTab& wd0, int src, bool rz, const double* xi,
                       const double* b, MgParentMap pm, double alpha, const double* xc,
                       double* xo, PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        mg0_smooth_dim<2>(g, wd0, src, rz, xi, b, pm, alpha, xc, xo, st, rb, s);
    else
        mg0_smooth_dim<3>(g, wd0, src, rz, xi, b, pm, alpha, xc, xo, st, rb, s);
}

// the tile grid of the tiled multigrid kernels: an even z chunk, shortened
// on small levels until there are 512 blocks of 256 lanes (whose few tiles would
// otherwise leave most CUs idle while each block marches its planes one
// after the other; each halving adds two halo planes per chunk, and at 256
// blocks the V-cycle took longer: 693 against 681 us at 256^3, 1024 blocks
// 682 against 678 at 512)
template <int CPL = 1>
static Tile3Grid mg_tile_grid(const Geo& g)
{
    Tile3Grid t = tile3_grid<CPL>(g);
    t.kz += t.kz & 1;
    while (t.kz > 2 && t.tiles_x * t.tiles_y * ((g.n[2] + t.kz - 1) / t.kz) * kTileBlock <
                           512 * 256)
        t.kz = (t.kz / 2 + 1) & ~1;
    t.nchunk = (g.n[2] + t.kz - 1) / t.kz;
    t.work = t.tiles_x * t.tiles_y * t.nchunk;
    return t;
}

bool mg_tiled3(const Geo& g)
{
    return g.dim == 3 && g.n[0] > 2 && g.n[1] > 2 && g.n[2] > 2;
}

void launch_mg_smooth3(const Geo& g, const Geo& gw, int src, bool rz, const double* xi,
                       const double* b, const double* xc, MgParentMap pm, double alpha,
                       double* xo, PcgState* st, RedBuf rb, hipStream_t s)
{
    const Tile3Grid tg = mg_tile_grid(g);
    const dim3 grid((tg.work + 7) / 8 * 8), blk(kTileBlock);
    const Tile3Parent par{xc, pm.base, pm.sy, pm.sz, alpha};
#define CF_MG_SMOOTH3(M, RZ)                                                                \
    hipLaunchKernelGGL((k_mg_smooth3<double, M, RZ>), grid, blk, 0, s, g, gw, tg, xi, b, par, \
                       xo, st, rb)
    if (src == MG_FIRST) {
        if (rz)
            CF_MG_SMOOTH3(1, true);
        else
            CF_MG_SMOOTH3(1, false);
    } else if (src == MG_PLAIN) {
        if (rz)
            CF_MG_SMOOTH3(0, true);
        else
            CF_MG_SMOOTH3(0, false);
    } else {
        if (rz)
            CF_MG_SMOOTH3(2, true);
        else
            CF_MG_SMOOTH3(2, false);
    }
#undef CF_MG_SMOOTH3
}

// The FP32 level 0 marches two cells per lane (8-byte FP32 accesses, as the
// FP32 PCG tiles), with the periodic-x variant on a periodic x axis; the
// sweep that reads the FP64 r and reduces r.z keeps one cell per lane
// (256^3, rocprof: 54.1 against 56.2 us; the others gain 1.4-3.7 us).  An
// FP32 level 1 marches one cell per lane (at 128^3 two would leave 512
// blocks), with an FP64 (level 2) or FP32 (level 1 for level 0) parent.
template <int CPL, bool PX, class TP>
static void mg_smooth3_f32(const Geo& g, const Geo& gw, int src, bool rz, const float* xi,
                           const float* b32, const double* r64, const Tile3ParentT<TP>& par,
                           float* xo, PcgState* st, RedBuf rb, hipStream_t s)
{
    const dim3 blk(kTileBlock);
    if (rz) {
        const Tile3Grid tg = mg_tile_grid<1>(g);
        const dim3 grid((tg.work + 7) / 8 * 8);
        if (src == MG_PLAIN)
            hipLaunchKernelGGL((k_mg_smooth3<float, 0, true, 1, false, TP>), grid, blk, 0, s, g,
                               gw, tg, xi, r64, par, xo, st, rb);
        else
            hipLaunchKernelGGL((k_mg_smooth3<float, 2, true, 1, false, TP>), grid, blk, 0, s, g,
                               gw, tg, xi, r64, par, xo, st, rb);
        return;
    }
    const Tile3Grid tg = mg_tile_grid<CPL>(g);
    const dim3 grid((tg.work + 7) / 8 * 8);
#define CF_MG_SMOOTH3F(M)                                                                     \
    hipLaunchKernelGGL((k_mg_smooth3<float, M, false, CPL, PX, TP>), grid, blk, 0, s, g, gw, tg, \
                       xi, b32, par, xo, st, rb)
    if (src == MG_FIRST)
        CF_MG_SMOOTH3F(1);
    else if (src == MG_PLAIN)
        CF_MG_SMOOTH3F(0);
    else
        CF_MG_SMOOTH3F(2);
#undef CF_MG_SMOOTH3F
}

void launch_mg_smooth3_f32(const Geo& g, const Geo& gw, int src, bool rz, const float* xi,
                           const float* b32, const double* r64, const double* xc,
                           const float* xc32, MgParentMap pm, double alpha, float* xo,
                           PcgState* st, RedBuf rb, hipStream_t s, int cpl)
{
    if (rz && src == MG_FIRST)
        throw CfError(CF_EINVAL, "the FP32 first sweep does not reduce r.z");
    if (cpl != 1 && cpl != 2)
        throw CfError(CF_EINVAL, "cells per lane of an FP32 multigrid level: 1 or 2");
    if (xc32) {
        const Tile3ParentT<float> par{xc32, pm.base, pm.sy, pm.sz, alpha};
        if (cpl == 1)
            mg_smooth3_f32<1, false>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
        else if (g.periodic[0])
            mg_smooth3_f32<2, true>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
        else
            mg_smooth3_f32<2, false>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
        return;
    }
    const Tile3Parent par{xc, pm.base, pm.sy, pm.sz, alpha};
    if (cpl == 1)
        mg_smooth3_f32<1, false>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
    else if (g.periodic[0])
        mg_smooth3_f32<2, true>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
    else
        mg_smooth3_f32<2, false>(g, gw, src, rz, xi, b32, r64, par, xo, st, rb, s);
}

void launch_mg_resrestrict3(const Geo& g, const double* x, const double* b, double* bc,
                            MgParentMap pm, const PcgState* st, hipStream_t s)
{
    const Tile3Grid tg = mg_tile_grid(g);
    hipLaunchKernelGGL(k_mg_resrestrict3<double>, dim3((tg.work + 7) / 8 * 8), dim3(kTileBlock), 0,
                       s, g, tg, pm, x, b, bc, st);
}

template <int CPL, bool PX, class TO>
static void mg_resrestrict3_f32(const Geo& g, const float* x, const float* b, TO* bc,
                                MgParentMap pm, const PcgState* st, hipStream_t s)
{
    const Tile3Grid tg = mg_tile_grid<CPL>(g);
    hipLaunchKernelGGL((k_mg_resrestrict3<float, CPL, PX, TO>), dim3((tg.work + 7) / 8 * 8),
                       dim3(kTileBlock), 0, s, g, tg, pm, x, b, bc, st);
}

void launch_mg_resrestrict3_f32(const Geo& g, const float* x, const float* b, double* bc,
                                float* bc32, MgParentMap pm, const PcgState* st, hipStream_t s,
                                int cpl)
{
    if (cpl != 1 && cpl != 2)
        throw CfError(CF_EINVAL, "cells per lane of an FP32 multigrid level: 1 or 2");
    if (bc32) {
        if (cpl == 1)
            mg_resrestrict3_f32<1, false>(g, x, b, bc32, pm, st, s);
        else if (g.periodic[0])
            mg_resrestrict3_f32<2, true>(g, x, b, bc32, pm, st, s);
        else
            mg_resrestrict3_f32<2, false>(g, x, b, bc32, pm, st, s);
        return;
    }
    if (cpl == 1)
        mg_resrestrict3_f32<1, false>(g, x, b, bc, pm, st, s);
    else if (g.periodic[0])
        mg_resrestrict3_f32<2, true>(g, x, b, bc, pm, st, s);
    else
        mg_resrestrict3_f32<2, false>(g, x, b, bc, pm, st, s);
}

void launch_mg0_first_f32(const Geo& g, const MgTab& wd0, const float* b, float* x,
                          const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL((k_mg0_first<3, float>), dim3(stream_blocks(g)), dim3(kBlock), 0, s, g,
                       wd0, b, x, st);
}

void launch_mg_cast32(const Geo& g, const double* r, float* b32, const PcgState* st,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_cast32, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g, r, b32, st);
}

void launch_mg0_resrestrict(const Geo& g, const MgLev& c1, MgParentMap pm, const double* x,
                            const double* b, double* b1, const PcgState* st, hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_mg0_resrestrict<2>, compact_grid(c1.ncell), dim3(kBlock), 0, s, g,
                           c1, pm, x, b, b1, st);
    else
        hipLaunchKernelGGL(k_mg0_resrestrict<3>, compact_grid(c1.ncell), dim3(kBlock), 0, s, g,
                           c1, pm, x, b, b1, st);
}

void launch_mg_first(const MgLev& L, const double* b, double* x, const PcgState* st,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_first, compact_grid(L.ncell), dim3(kBlock), 0, s, L, b, x, st);
}

void launch_mg_smooth(const MgLev& L, int src, const double* xi, const double* b,
                      const MgLev& C, const double* xc, double* xo, const PcgState* st,
                      hipStream_t s)
{
    const dim3 grid = compact_grid(L.ncell), blk(kBlock);
    if (src == MG_FIRST)
        hipLaunchKernelGGL(k_mg_smooth<MG_FIRST>, grid, blk, 0, s, L, xi, b, C, xc, xo, st);
    else if (src == MG_PLAIN)
        hipLaunchKernelGGL(k_mg_smooth<MG_PLAIN>, grid, blk, 0, s, L, xi, b, C, xc, xo, st);
    else
        hipLaunchKernelGGL(k_mg_smooth<MG_PROLONG>, grid, blk, 0, s, L, xi, b, C, xc, xo, st);
}

void launch_mg_resrestrict(const MgLev& F, const MgLev& C, const double* x, const double* b,
                           double* bc, const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_resrestrict, compact_grid(C.ncell), dim3(kBlock), 0, s, F, C, x, b,
                       bc, st);
}

void launch_mg_tail(const MgTail* t_dev, size_t lds_bytes, const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kMgTailThreads), lds_bytes, s, t_dev, st);
}

void launch_mgpcg_init(const Geo& g, const double* b, double* x, double* r, PcgState* st,
                       double tol, int maxit, RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_mgpcg_init, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g, b, x, r,
                       st, tol, maxit, rb);
}

void launch_mgpcg_rz(const Geo& g, const double* r, const double* z, PcgState* st,
                     RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_mgpcg_rz, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g, r, z, st,
                       rb);
}

} // namespace cfmi
