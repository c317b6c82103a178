// cf_pcg.hip -- matrix-free Jacobi-PCG (FP64, and FP32 inside an FP64
// defect-correction loop) for the pressure solve of
// src/VelocityCorrector.hpp:266-282.
//
// Control flow (this build's statement of Cajita's ReferenceConjugateGradient,
// called at src/VelocityCorrector.hpp:276), from x = 0:
//   r = b; rho = r.(D^-1 r); bnorm = ||b||            (bnorm = 0: done)
//   iteration k:
//     p = D^-1 r                      (k = 1)
//     p = fma(beta, p, D^-1 r)        (k > 1)
//     q = A p; alpha = rho / p.q
//     x = fma(alpha, p, x); r = fma(-alpha, q, r)
//     stop if ||r|| / bnorm < tol, or k = maxit
//     beta = r.(D^-1 r) / rho; rho = r.(D^-1 r)
//
// One iteration is two streaming kernels:
//   dir_spmv: p_new = fma(beta, p_old, D^-1 r) at the cell, and at each
//             stencil neighbour recomputed from r and p_old there; q = A p_new;
//             p_new.q -> alpha.                reads r, p_old; writes p_new
//   update:   q = A p_new again (same expression, same bits); x += alpha p_new;
//             r -= alpha q; r.r, r.(D^-1 r) -> convergence test, beta, rho.
//                                              reads x, p_new, r; writes x, r
// 8 vector words per cell per iteration instead of the 11 of the
// SpMV / update / direction split.  p is double-buffered because dir_spmv
// reads the old direction around each cell while writing the new one.  The
// last block of each kernel to finish computes the scalars on the device;
// every kernel returns at once when the device-side `done` flag is set, so
// the host queues iterations in batches and polls once per batch.
#include <hip/hip_runtime.h>

#include "cf_stream.hpp"

namespace cfmi
{

// ---- PCG initialisation ------------------------------------------------------
__device__ __forceinline__ void init_state(PcgState* st, double rz, double nn,
                                           double tol, int maxit)
{
    const double norm = sqrt(nn);
    st->bnorm = norm;
    st->tol = tol;
    st->maxit = maxit;
    st->iter = 0;
    st->rho = rz;
    st->alpha = 0.0;
    st->beta = 0.0;
    const bool zero = norm == 0.0;
    st->done = zero ? 1 : (maxit <= 0 ? 1 : 0);
    st->converged = zero ? 1 : 0;
    st->res = zero ? 0.0 : 1.0;
}

__global__ __launch_bounds__(256) void k_pcg_init64(
    Geo g, const double* __restrict__ b, double* __restrict__ x,
    double* __restrict__ r, PcgState* st, double tol, int maxit, RedBuf rb)
{
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const int ryz = wall_count(g, 1, j) + wall_count(g, 2, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const double bc = b[c];
            const double z = dinv_pick<double>(g, ryz, wall_count(g, 0, i)) * bc;
            r[c] = bc; // r = b - A 0 = b
            x[c] = 0.0;
            dd_add_prod(hi[0], lo[0], bc, z);
            dd_add_prod(hi[1], lo[1], bc, bc);
        }
    });
    grid_reduce<2>(hi, lo, rb,
                   [&](double* s) { init_state(st, s[0], s[1], tol, maxit); });
}

__global__ __launch_bounds__(256) void k_pcg_init32(
    Geo g, const float* __restrict__ r, float* __restrict__ d, PcgState* st,
    double tol, int maxit, RedBuf rb)
{
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const int ryz = wall_count(g, 1, j) + wall_count(g, 2, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const float rc = r[c];
            const float z = dinv_pick<float>(g, ryz, wall_count(g, 0, i)) * rc;
            d[c] = 0.0f;
            dd_add_prod(hi[0], lo[0], (double)rc, (double)z);
            dd_add_prod(hi[1], lo[1], (double)rc, (double)rc);
        }
    });
    grid_reduce<2>(hi, lo, rb,
                   [&](double* s) { init_state(st, s[0], s[1], tol, maxit); });
}

// ---- one PCG iteration --------------------------------------------------------
template <int DIM, class T>
__global__ __launch_bounds__(256) void k_pcg_dir_spmv(
    Geo g, const T* __restrict__ r, const T* __restrict__ p_old,
    T* __restrict__ p_new, PcgState* st, RedBuf rb)
{
    if (st->done)
        return;
    const bool first = st->iter == 0;
    const T beta = (T)st->beta;
    const T s = offdiag<T>(g);
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            // new direction at cell `at` whose wall count is ubase + wx
            auto pdir = [&](int64_t at, int ubase, int wx) -> T {
                const T z = dinv_pick<T>(g, ubase, wx) * r[at];
                return first ? z : fma_t<T>(beta, p_old[at], z);
            };
            const T pc = pdir(c, rs.cnt, lx.cnt);
            T y = diag_pick<T>(g, rs.cnt, lx.cnt) * pc;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const T pn = pdir(c + lx.off[side], rs.cnt, lx.ncnt[side]);
                y = fma_t<T>(lx.wall[side] ? (T)0 : s, lx.wall[side] ? (T)0 : pn, y);
            }
#pragma unroll
            for (int d = 1; d < DIM; ++d) {
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const T pn = pdir(c + rs.off[d][side], rs.ncnt[d][side], lx.cnt);
                    y = fma_t<T>(rs.wall[d][side] ? (T)0 : s,
                                 rs.wall[d][side] ? (T)0 : pn, y);
                }
            }
            p_new[c] = pc;
            dd_add_prod(hi[0], lo[0], (double)pc, (double)y);
        }
    });
    grid_reduce<1>(hi, lo, rb, [&](double* sm) { st->alpha = st->rho / sm[0]; });
}

template <int DIM, class T>
__global__ __launch_bounds__(256) void k_pcg_update(
    Geo g, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
    PcgState* st, RedBuf rb)
{
    if (st->done)
        return;
    const T alpha = (T)st->alpha;
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            const T q = apply_A<DIM, T>(g, rs, lx, p, c);
            x[c] = fma_t<T>(alpha, p[c], x[c]);
            const T rn = fma_t<T>(-alpha, q, r[c]);
            r[c] = rn;
            const T z = dinv_pick<T>(g, rs.cnt, lx.cnt) * rn;
            dd_add_prod(hi[0], lo[0], (double)rn, (double)rn);
            dd_add_prod(hi[1], lo[1], (double)rn, (double)z);
        }
    });
    grid_reduce<2>(hi, lo, rb, [&](double* sm) {
        const int it = st->iter + 1;
        st->iter = it;
        const double res = sqrt(sm[0]) / st->bnorm;
        st->res = res;
        if (res < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (it >= st->maxit) {
            st->done = 1;
        } else {
            st->beta = sm[1] / st->rho;
            st->rho = sm[1];
        }
    });
}

// ---- 3-D: z-marching tiles ------------------------------------------------------
// A block owns a 64 x 16 (x, y) tile and marches up a chunk of z planes.  Each
// lane keeps its column's values at planes k-1, k, k+1 in registers (4 rows
// per lane), and the plane k values of the tile plus a one-cell halo sit in
// LDS for the x and y neighbours.  Every vector element is loaded once per
// chunk (halo cells come from the neighbouring tiles, which the XCD-ordered
// block mapping keeps in the same L2).
namespace tile3
{
constexpr int TX = 64;          // lanes along x
constexpr int NW = kBlock / TX; // waves per block
constexpr int RY = 4;           // rows per lane
constexpr int TY = NW * RY;     // tile rows
constexpr int NHALO = 2 * TX + 2 * TY;
} // namespace tile3

struct Tile3Grid
{
    int tiles_x, tiles_y, nchunk, kz, work;
};

static Tile3Grid tile3_grid(const Geo& g)
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

static int tile3_blocks(const Geo& g)
{
    const Tile3Grid t = tile3_grid(g);
    return (t.work + 7) / 8 * 8;
}

// Block -> (tile x, tile y, chunk): XCD x (= blockIdx mod 8) gets a
// contiguous range of the x-fastest, then y, then chunk order.
struct Tile3Pos
{
    int i0, j0, k0, k1;
    bool active;
};

__device__ __forceinline__ Tile3Pos tile3_pos(const Geo& g, Tile3Grid t)
{
    using namespace tile3;
    const int L = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    Tile3Pos p;
    p.active = L < t.work;
    const int tx = L % t.tiles_x;
    const int rest = L / t.tiles_x;
    const int ty = rest % t.tiles_y;
    const int ch = rest / t.tiles_y;
    p.i0 = tx * TX;
    p.j0 = ty * TY;
    p.k0 = ch * t.kz;
    p.k1 = min(p.k0 + t.kz, g.n[2]);
    return p;
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

// New search direction at global cell (gi, gj, gk) (0 at a zero ghost).
// UNIFORM_YZ: gj and gk are the same across the wave, so their part of the
// wall count and the diagonal lookup stay scalar.
template <class T, bool UNIFORM_YZ>
__device__ __forceinline__ T pdir3(const Geo& g, const T* __restrict__ r,
                                   const T* __restrict__ p_old, bool first,
                                   T beta, int gi, int gj, int gk)
{
    const int a = wrap_or_ghost(g, 0, gi), b = wrap_or_ghost(g, 1, gj),
              c = wrap_or_ghost(g, 2, gk);
    if (a < 0 || b < 0 || c < 0)
        return (T)0;
    const int64_t at = g.idx(g.h + a, g.h + b, g.h + c);
    T dinv;
    if constexpr (UNIFORM_YZ)
        dinv = dinv_pick<T>(g, wall_count(g, 1, b) + wall_count(g, 2, c),
                            wall_count(g, 0, a));
    else
        dinv = dinv_sel<T>(g, wall_count(g, 0, a) + wall_count(g, 1, b) +
                                  wall_count(g, 2, c));
    const T z = dinv * r[at];
    return first ? z : fma_t<T>(beta, p_old[at], z);
}

// Stored value at global cell (gi, gj, gk) (0 at a zero ghost).
template <class T>
__device__ __forceinline__ T load3(const Geo& g, const T* __restrict__ p, int gi,
                                   int gj, int gk)
{
    const int a = wrap_or_ghost(g, 0, gi), b = wrap_or_ghost(g, 1, gj),
              c = wrap_or_ghost(g, 2, gk);
    if (a < 0 || b < 0 || c < 0)
        return (T)0;
    return p[g.idx(g.h + a, g.h + b, g.h + c)];
}

// Fills LDS plane `pl` (rows 0..TY+1, cols 0..TX+1) for the tile's cells
// (vo, rows wave-uniform) and its x/y halo ring (vh) at plane gk; own[] gets
// the lane's own cells.
template <class T, class VO, class VH>
__device__ __forceinline__ void tile3_fill(T (*pl)[tile3::TX + 2], int i0, int j0,
                                           int gk, T (&own)[tile3::RY], VO&& vo,
                                           VH&& v)
{
    using namespace tile3;
    const int lane = threadIdx.x & (TX - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int jj = w * RY + q;
        own[q] = vo(i0 + lane, j0 + jj, gk);
        pl[jj + 1][lane + 1] = own[q];
    }
    for (int hh = threadIdx.x; hh < NHALO; hh += kBlock) {
        int row, col, gi, gj;
        if (hh < TX) {
            row = 0; col = hh + 1; gi = i0 + hh; gj = j0 - 1;
        } else if (hh < 2 * TX) {
            row = TY + 1; col = hh - TX + 1; gi = i0 + hh - TX; gj = j0 + TY;
        } else if (hh < 2 * TX + TY) {
            row = hh - 2 * TX + 1; col = 0; gi = i0 - 1; gj = j0 + hh - 2 * TX;
        } else {
            row = hh - 2 * TX - TY + 1; col = TX + 1; gi = i0 + TX;
            gj = j0 + hh - 2 * TX - TY;
        }
        pl[row][col] = v(gi, gj, gk);
    }
}

// y = A p at own cell (lane, row jj) of plane k: centre pc, z neighbours
// pm / pp, x and y neighbours from the LDS plane; reference stencil order.
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

template <class T>
__global__ __launch_bounds__(256) void k_pcg_dir_spmv3(
    Geo g, Tile3Grid tg, const T* __restrict__ r, const T* __restrict__ p_old,
    T* __restrict__ p_new, PcgState* st, RedBuf rb)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ T pl[2][TY + 2][TX + 2];
    const bool first = st->iter == 0;
    const T beta = (T)st->beta;
    double hi[1] = {0.0}, lo[1] = {0.0};
    const Tile3Pos tp = tile3_pos(g, tg);
    if (tp.active) {
        const int lane = threadIdx.x & (TX - 1);
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
        const int gi = tp.i0 + lane;
        auto po = [&](int a, int b, int c) {
            return pdir3<T, true>(g, r, p_old, first, beta, a, b, c);
        };
        auto ph = [&](int a, int b, int c) {
            return pdir3<T, false>(g, r, p_old, first, beta, a, b, c);
        };
        T pm[RY], pc[RY], pp[RY];
#pragma unroll
        for (int q = 0; q < RY; ++q)
            pm[q] = po(gi, tp.j0 + w * RY + q, tp.k0 - 1);
        tile3_fill<T>(pl[0], tp.i0, tp.j0, tp.k0, pc, po, ph);
        __syncthreads();
        int cur = 0;
        for (int k = tp.k0; k < tp.k1; ++k) {
            tile3_fill<T>(pl[cur ^ 1], tp.i0, tp.j0, k + 1, pp, po, ph);
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                const int jj = w * RY + q;
                const int gj = tp.j0 + jj;
                if (gi < g.n[0] && gj < g.n[1]) {
                    const T y = tile3_apply<T>(g, pl[cur], lane, jj, gi, gj, k,
                                               pc[q], pm[q], pp[q]);
                    p_new[g.idx(g.h + gi, g.h + gj, g.h + k)] = pc[q];
                    dd_add_prod(hi[0], lo[0], (double)pc[q], (double)y);
                }
                pm[q] = pc[q];
                pc[q] = pp[q];
            }
            __syncthreads();
            cur ^= 1;
        }
    }
    grid_reduce<1>(hi, lo, rb, [&](double* sm) { st->alpha = st->rho / sm[0]; });
}

template <class T>
__global__ __launch_bounds__(256) void k_pcg_update3(
    Geo g, Tile3Grid tg, T* __restrict__ x, T* __restrict__ r,
    const T* __restrict__ p, PcgState* st, RedBuf rb)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ T pl[2][TY + 2][TX + 2];
    const T alpha = (T)st->alpha;
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    const Tile3Pos tp = tile3_pos(g, tg);
    if (tp.active) {
        const int lane = threadIdx.x & (TX - 1);
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
        const int gi = tp.i0 + lane;
        auto pv = [&](int a, int b, int c) { return load3<T>(g, p, a, b, c); };
        T pm[RY], pc[RY], pp[RY];
#pragma unroll
        for (int q = 0; q < RY; ++q)
            pm[q] = pv(gi, tp.j0 + w * RY + q, tp.k0 - 1);
        tile3_fill<T>(pl[0], tp.i0, tp.j0, tp.k0, pc, pv, pv);
        __syncthreads();
        int cur = 0;
        for (int k = tp.k0; k < tp.k1; ++k) {
            tile3_fill<T>(pl[cur ^ 1], tp.i0, tp.j0, k + 1, pp, pv, pv);
#pragma unroll
            for (int q = 0; q < RY; ++q) {
                const int jj = w * RY + q;
                const int gj = tp.j0 + jj;
                if (gi < g.n[0] && gj < g.n[1]) {
                    const T qv = tile3_apply<T>(g, pl[cur], lane, jj, gi, gj, k,
                                                pc[q], pm[q], pp[q]);
                    const int64_t c = g.idx(g.h + gi, g.h + gj, g.h + k);
                    x[c] = fma_t<T>(alpha, pc[q], x[c]);
                    const T rn = fma_t<T>(-alpha, qv, r[c]);
                    r[c] = rn;
                    const T z = dinv_pick<T>(g,
                                             wall_count(g, 1, gj) +
                                                 wall_count(g, 2, k),
                                             wall_count(g, 0, gi)) *
                                rn;
                    dd_add_prod(hi[0], lo[0], (double)rn, (double)rn);
                    dd_add_prod(hi[1], lo[1], (double)rn, (double)z);
                }
                pm[q] = pc[q];
                pc[q] = pp[q];
            }
            __syncthreads();
            cur ^= 1;
        }
    }
    grid_reduce<2>(hi, lo, rb, [&](double* sm) {
        const int it = st->iter + 1;
        st->iter = it;
        const double res = sqrt(sm[0]) / st->bnorm;
        st->res = res;
        if (res < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (it >= st->maxit) {
            st->done = 1;
        } else {
            st->beta = sm[1] / st->rho;
            st->rho = sm[1];
        }
    });
}

// ---- defect correction ---------------------------------------------------------
//   x = 0; r64 = b; bnorm = ||b||; res = ||r64|| / bnorm
//   pass: r32 = (float) r64; inner FP32 PCG d ~ A^-1 r32;
//         x = x + (double) d; r64 = b - A x; res = ||r64|| / bnorm
__global__ __launch_bounds__(256) void k_refine_init(
    Geo g, const double* __restrict__ b, double* __restrict__ x64,
    float* __restrict__ r32, RefineState* rst, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        const double bc = b[c];
        x64[c] = 0.0;
        r32[c] = (float)bc;
        dd_add_prod(hi[0], lo[0], bc, bc);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) {
        const double bnorm = sqrt(s[0]);
        rst->bnorm = bnorm;
        rst->res = bnorm == 0.0 ? 0.0 : sqrt(s[0]) / bnorm;
    });
}

__global__ __launch_bounds__(256) void k_refine_add(Geo g,
                                                     double* __restrict__ x64,
                                                     const float* __restrict__ d)
{
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        x64[c] = x64[c] + (double)d[c];
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_refine_residual(
    Geo g, const double* __restrict__ b, const double* __restrict__ x64,
    float* __restrict__ r32, RefineState* rst, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            const double y = apply_A<DIM, double>(g, rs, lx, x64, c);
            const double r = b[c] - y;
            r32[c] = (float)r;
            dd_add_prod(hi[0], lo[0], r, r);
        }
    });
    grid_reduce<1>(hi, lo, rb,
                   [&](double* s) { rst->res = sqrt(s[0]) / rst->bnorm; });
}

// ---- RHS mean removal --------------------------------------------------------
__global__ __launch_bounds__(256) void k_rhs_sum(Geo g, const double* rhs,
                                                  double* out, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        dd_add(hi[0], lo[0], rhs[c], 0.0);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) { out[0] = s[0]; });
}

__global__ __launch_bounds__(256) void k_rhs_shift(Geo g, double* rhs,
                                                    const double* sum)
{
    const double cnt = (double)g.n[0] * (double)g.n[1] * (double)g.n[2];
    const double mean = sum[0] / cnt;
    for_owned_cells(g, [&](int64_t c, int, int, int) { rhs[c] = rhs[c] - mean; });
}

// ---- launchers -------------------------------------------------------------------
#define CF_STREAM_LAUNCH(kernel, ...)                                         \
    hipLaunchKernelGGL(kernel, dim3(kRedBlocks), dim3(kBlock), 0, s, __VA_ARGS__)

void launch_rhs_sum(const Geo& g, const double* rhs, double* out_sum,
                    RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_rhs_sum, g, rhs, out_sum, rb);
}

void launch_rhs_shift(const Geo& g, double* rhs, const double* sum,
                      hipStream_t s)
{
    CF_STREAM_LAUNCH(k_rhs_shift, g, rhs, sum);
}

void launch_pcg_init64(const Geo& g, const double* b, double* x, double* r,
                       PcgState* st, double tol, int maxit, RedBuf rb,
                       hipStream_t s)
{
    CF_STREAM_LAUNCH(k_pcg_init64, g, b, x, r, st, tol, maxit, rb);
}

void launch_pcg_init32(const Geo& g, const float* r, float* d, PcgState* st,
                       double tol, int maxit, RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_pcg_init32, g, r, d, st, tol, maxit, rb);
}

template <class T>
void launch_pcg_dir_spmv(const Geo& g, const T* r, const T* p_old, T* p_new,
                         PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        CF_STREAM_LAUNCH((k_pcg_dir_spmv<2, T>), g, r, p_old, p_new, st, rb);
    else
        hipLaunchKernelGGL(k_pcg_dir_spmv3<T>, dim3(tile3_blocks(g)),
                           dim3(kBlock), 0, s, g, tile3_grid(g), r, p_old,
                           p_new, st, rb);
}

template <class T>
void launch_pcg_update(const Geo& g, T* x, T* r, const T* p, PcgState* st,
                       RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        CF_STREAM_LAUNCH((k_pcg_update<2, T>), g, x, r, p, st, rb);
    else
        hipLaunchKernelGGL(k_pcg_update3<T>, dim3(tile3_blocks(g)), dim3(kBlock),
                           0, s, g, tile3_grid(g), x, r, p, st, rb);
}

int reduction_blocks_max(const Geo& g)
{
    const int t = g.dim == 3 ? tile3_blocks(g) : 0;
    return t > kRedBlocks ? t : kRedBlocks;
}

template void launch_pcg_dir_spmv<double>(const Geo&, const double*,
                                          const double*, double*, PcgState*,
                                          RedBuf, hipStream_t);
template void launch_pcg_dir_spmv<float>(const Geo&, const float*, const float*,
                                         float*, PcgState*, RedBuf, hipStream_t);
template void launch_pcg_update<double>(const Geo&, double*, double*,
                                        const double*, PcgState*, RedBuf,
                                        hipStream_t);
template void launch_pcg_update<float>(const Geo&, float*, float*, const float*,
                                       PcgState*, RedBuf, hipStream_t);

void launch_refine_init(const Geo& g, const double* b, double* x64,
                        float* r32, RefineState* rs, RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_refine_init, g, b, x64, r32, rs, rb);
}

void launch_refine_add(const Geo& g, double* x64, const float* d32,
                       hipStream_t s)
{
    CF_STREAM_LAUNCH(k_refine_add, g, x64, d32);
}

void launch_refine_residual(const Geo& g, const double* b, const double* x64,
                            float* r32, RefineState* rs, RedBuf rb,
                            hipStream_t s)
{
    if (g.dim == 2)
        CF_STREAM_LAUNCH(k_refine_residual<2>, g, b, x64, r32, rs, rb);
    else
        CF_STREAM_LAUNCH(k_refine_residual<3>, g, b, x64, r32, rs, rb);
}

} // namespace cfmi
