// cf_adjoint.hip -- reverse-mode kernels of one step (advection, inputs,
// projection).  The reference has no adjoint; these are the transposes of
// the forward kernels in cf_kernels.hip, driven by Solver::adjointRun.
//
// Ghost handling.  The forward gathers give every ghost the value of the
// owned entry at the periodically wrapped index (wall ghosts are constant
// zeros), so the transpose of "gather, then read" is "write to the wrapped
// owned index, drop writes that land on a wall ghost".  The adjoint kernels
// do that wrap inline and never hold anything in ghost entries.
//
// Advection transpose.  Each owned point re-traces its RK3 path (the same
// device functions as the forward, so the trace is bit-identical), then
// scatters the adjoint of its Next value through the field sample at the
// trace and back through the three velocity samples with the derivatives of
// the B-spline weights.  Scatters are hardware FP64 atomics
// (global_atomic_add_f64); their order is not fixed, so the gradient is
// reproducible to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include "cf_adjoint.hpp"
#include "cf_stream.hpp"

namespace cfmi
{

// ---- spline weights with their derivatives --------------------------------------
// w as spline_weights<ORDER>; dw = d w / d p (the logical coordinate's
// derivative folded in); zero where the coordinate was clamped.
template <int ORDER>
__device__ __forceinline__ void spline_wd(const Geo& g, int e, int d, double p,
                                          int& s0, double (&w)[ORDER + 1],
                                          double (&dw)[ORDER + 1])
{
    const double xi_raw = (p - g.x0[e][d]) * g.rdx[e][d];
    const double lo = ORDER == 1 ? 0.0 : 1.0;
    const double hi = (double)(g.ext[e][d] - (ORDER == 1 ? 2 : 3));
    const double xi = fmin(fmax(xi_raw, lo), hi);
    const bool clamped = !(xi_raw >= lo && xi_raw <= hi);
    spline_weights<ORDER>(xi, s0, w);
    const double t = xi - (double)(int)xi;
    const double r = clamped ? 0.0 : g.rdx[e][d];
    if constexpr (ORDER == 1) {
        dw[0] = -r;
        dw[1] = r;
    } else {
        const double omt = 1.0 - t;
        dw[0] = (-0.5 * omt * omt) * r;
        dw[1] = (1.5 * t * t - 2.0 * t) * r;
        dw[2] = ((-1.5 * t * t + t) + 0.5) * r;
        dw[3] = (0.5 * t * t) * r;
    }
}

// Sample weights of entity e at p, per axis (axes >= DIM: one tap, weight 1).
template <int DIM, int ORDER>
struct Taps
{
    int s0[3];
    double w[3][ORDER + 1];
    double dw[3][ORDER + 1];
};

template <int DIM, int ORDER>
__device__ __forceinline__ void make_taps(const Geo& g, int e, const double (&p)[3],
                                          Taps<DIM, ORDER>& t)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        spline_wd<ORDER>(g, e, d, p[d], t.s0[d], t.w[d], t.dw[d]);
    if constexpr (DIM == 2)
        t.s0[2] = 0;
}

// a[i] for a loop index the compiler does not unroll, as a select chain, so
// that a stays in registers.
template <class T, int N>
__device__ __forceinline__ T pick(const T (&a)[N], int i)
{
    T r = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k)
        r = i == k ? a[k] : r;
    return r;
}

// d/dp of the sample of f.  3-D: BATCH loads the taps of all z planes
// before they are used (in flight together); otherwise one plane at a time,
// which the kernels with cubic velocity samples need to stay within the
// register file.
template <int DIM, int ORDER, bool BATCH = true>
__device__ __forceinline__ void sample_grad(const Geo& g, const Taps<DIM, ORDER>& t,
                                            const double* __restrict__ f,
                                            double (&gr)[3])
{
    gr[0] = gr[1] = gr[2] = 0.0;
    if constexpr (DIM == 2) {
#pragma unroll
        for (int b = 0; b < ORDER + 1; ++b) {
            const double* row = f + g.idx(t.s0[0], t.s0[1] + b, 0);
            double sx = 0.0, sv = 0.0;
#pragma unroll
            for (int a = 0; a < ORDER + 1; ++a) {
                const double fv = row[a];
                sx = fma(t.dw[0][a], fv, sx);
                sv = fma(t.w[0][a], fv, sv);
            }
            gr[0] = fma(t.w[1][b], sx, gr[0]);
            gr[1] = fma(t.dw[1][b], sv, gr[1]);
        }
    } else if constexpr (!BATCH) {
        const int64_t pz = g.px * g.py;
#pragma unroll 1
        for (int c = 0; c < ORDER + 1; ++c) {
            double gx = 0.0, gy = 0.0, vv = 0.0;
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b) {
                const double* row = f + g.idx(t.s0[0], t.s0[1] + b, t.s0[2]) + c * pz;
                double sx = 0.0, sv = 0.0;
#pragma unroll
                for (int a = 0; a < ORDER + 1; ++a) {
                    const double fv = row[a];
                    sx = fma(t.dw[0][a], fv, sx);
                    sv = fma(t.w[0][a], fv, sv);
                }
                gx = fma(t.w[1][b], sx, gx);
                gy = fma(t.dw[1][b], sv, gy);
                vv = fma(t.w[1][b], sv, vv);
            }
            const double wz = pick(t.w[2], c);
            gr[0] = fma(wz, gx, gr[0]);
            gr[1] = fma(wz, gy, gr[1]);
            gr[2] = fma(pick(t.dw[2], c), vv, gr[2]);
        }
    } else {
        const int64_t pz = g.px * g.py;
        const double* base = f + g.idx(t.s0[0], t.s0[1], t.s0[2]);
        // each z plane's taps are loaded before any is used, so that they
        // are in flight together (the planes are unrolled: constant indices
        // keep the weights in registers)
#pragma unroll
        for (int c = 0; c < ORDER + 1; ++c) {
            double fv[ORDER + 1][ORDER + 1];
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b)
#pragma unroll
                for (int a = 0; a < ORDER + 1; ++a)
                    fv[b][a] = base[c * pz + b * g.px + a];
            double gx = 0.0, gy = 0.0, vv = 0.0;
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b) {
                double sx = 0.0, sv = 0.0;
#pragma unroll
                for (int a = 0; a < ORDER + 1; ++a) {
                    sx = fma(t.dw[0][a], fv[b][a], sx);
                    sv = fma(t.w[0][a], fv[b][a], sv);
                }
                gx = fma(t.w[1][b], sx, gx);
                gy = fma(t.dw[1][b], sv, gy);
                vv = fma(t.w[1][b], sv, vv);
            }
            gr[0] = fma(t.w[2][c], gx, gr[0]);
            gr[1] = fma(t.w[2][c], gy, gr[1]);
            gr[2] = fma(t.dw[2][c], vv, gr[2]);
        }
    }
}

// Owned index of local index l of entity e along axis d: periodic axes wrap,
// a wall ghost returns -1 (a constant: its adjoint is dropped).
__device__ __forceinline__ int owned_index(const Geo& g, int e, int d, int l)
{
    const int lo = g.h, hi = g.own_hi[e][d];
    if (l >= lo && l < hi)
        return l;
    if (!g.periodic[d])
        return -1;
    while (l < lo)
        l += g.n[d];
    while (l >= hi)
        l -= g.n[d];
    return l;
}

// bar[taps] += weight * val, with the ghost transpose folded in.
template <int DIM, int ORDER>
__device__ __forceinline__ void scatter(const Geo& g, int e, double* __restrict__ bar,
                                        const Taps<DIM, ORDER>& t, double val)
{
    int ix[ORDER + 1], iy[ORDER + 1], iz[ORDER + 1];
#pragma unroll
    for (int a = 0; a < ORDER + 1; ++a) {
        ix[a] = owned_index(g, e, 0, t.s0[0] + a);
        iy[a] = owned_index(g, e, 1, t.s0[1] + a);
        iz[a] = DIM == 3 ? owned_index(g, e, 2, t.s0[2] + a) : 0;
    }
    constexpr int NZ = DIM == 3 ? ORDER + 1 : 1;
#pragma unroll 1
    for (int c = 0; c < NZ; ++c) {
        const int kz = pick(iz, c);
        if (kz < 0)
            continue;
        const double wz = DIM == 3 ? pick(t.w[2], c) * val : val;
#pragma unroll
        for (int b = 0; b < ORDER + 1; ++b) {
            if (iy[b] < 0)
                continue;
            const double wyz = t.w[1][b] * wz;
#pragma unroll
            for (int a = 0; a < ORDER + 1; ++a) {
                if (ix[a] < 0)
                    continue;
                atomicAdd(bar + g.idx(ix[a], iy[b], kz), t.w[0][a] * wyz);
            }
        }
    }
}

// Jacobian rows J[d] = grad v_d of the velocity at x.
template <int DIM, int VO>
__device__ __forceinline__ void velocity_jac(const Geo& g, const Fields& cur,
                                             const double (&x)[3], double (&J)[3][3])
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        Taps<DIM, VO> t;
        make_taps<DIM, VO>(g, d + 1, x, t);
        sample_grad<DIM, VO, VO == 1>(g, t, cur.e[d + 1], J[d]);
    }
}

// Scatter policies: Direct adds every tap to global memory; Windowed adds
// the taps of a sample whose whole tap box lies in the block's LDS window
// there (ds_add_f64) and the others to global memory.
struct DirectScatter
{
    const Geo& g;
    const Fields& bar;
    template <int DIM, int ORDER>
    __device__ __forceinline__ void add(int e, const Taps<DIM, ORDER>& t, double val) const
    {
        scatter<DIM, ORDER>(g, e, bar.e[e], t, val);
    }
};

// The LDS windows of k_adj_advect_win: per array, the raw local indices
// [slab origin - M, slab end + M) of a 32 x 4 x 2 slab of lattice points.
// The WZ planes form a ring along z: plane z sits in slot (z - zbase) mod WZ.
namespace adjwin
{
constexpr int TX = 32, TY = 4, TZ = 2, M = 3;
constexpr int WX = TX + 2 * M, WY = TY + 2 * M, WZ = TZ + 2 * M;
constexpr int WP = WX * WY;      // cells of one window plane
constexpr int WN = WP * WZ;      // 3040 doubles per array, 97 KB for four
static_assert((WZ & (WZ - 1)) == 0, "the z ring indexes with a mask");
// the Makefile's ADJ_WIN_LDS: check_lds.py fails the build unless the
// compiled kernels hold exactly this much LDS
static_assert(4 * WN * sizeof(double) == 97280, "LDS windows of k_adj_advect_win");
} // namespace adjwin

struct WindowedScatter
{
    const Geo& g;
    const Fields& bar;
    double (*win)[adjwin::WN];
    int wlo[3]; // the window's lowest raw local index per axis
    int zbase;  // the plane in ring slot 0 (z - zbase >= 0 for every window plane)
    template <int DIM, int ORDER>
    __device__ __forceinline__ void add(int e, const Taps<DIM, ORDER>& t, double val) const
    {
        using namespace adjwin;
        const bool in = t.s0[0] >= wlo[0] && t.s0[0] + ORDER < wlo[0] + WX &&
                        t.s0[1] >= wlo[1] && t.s0[1] + ORDER < wlo[1] + WY &&
                        t.s0[2] >= wlo[2] && t.s0[2] + ORDER < wlo[2] + WZ;
        if (!in) {
            scatter<DIM, ORDER>(g, e, bar.e[e], t, val);
            return;
        }
        double* w = win[e] + (t.s0[0] - wlo[0]) + WX * (t.s0[1] - wlo[1]);
        const int zs = t.s0[2] - zbase;
#pragma unroll
        for (int c = 0; c < ORDER + 1; ++c) {
            const double wz = t.w[2][c] * val;
            double* wp = w + WP * ((zs + c) & (WZ - 1));
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b) {
                const double wyz = t.w[1][b] * wz;
#pragma unroll
                for (int a = 0; a < ORDER + 1; ++a)
                    atomicAdd(wp + a + WX * b, t.w[0][a] * wyz);
            }
        }
    }
};

// Transpose of the velocity sample at x: bar_{d+1}[taps] += w * vb[d].  The
// taps are recomputed from x rather than kept, to hold register use down.
template <int DIM, int VO, class Sc>
__device__ __forceinline__ void velocity_scatter(const Geo& g, const Sc& sc,
                                                 const double (&x)[3],
                                                 const double (&vb)[3])
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        Taps<DIM, VO> t;
        make_taps<DIM, VO>(g, d + 1, x, t);
        sc.template add<DIM, VO>(d + 1, t, vb[d]);
    }
}

template <int DIM, int FO, int VO, class Sc>
__device__ __forceinline__ void advect_point_adj(const Geo& g, int e, const Fields& cur,
                                                 const double (&x0)[3], double ob,
                                                 const Sc& sc)
{
    const double dt = g.dt;
    const bool quirk = (g.quirks & CF_QUIRK_RK3_V0) != 0;
    double v0[3], v1[3], v2[3];
    double x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
    // the forward trace, bit-identical to k_advect's
    velocity<DIM, VO>(g, cur.e[1], cur.e[2], cur.e[3], x0, v0);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x1[d] = x0[d] - 0.5 * dt * v0[d];
    velocity<DIM, VO>(g, cur.e[1], cur.e[2], cur.e[3], x1, v1);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x2[d] = x0[d] - 0.75 * dt * (quirk ? v0[d] : v1[d]);
    velocity<DIM, VO>(g, cur.e[1], cur.e[2], cur.e[3], x2, v2);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        tr[d] = x0[d] - dt * ((2.0 / 9.0) * v0[d] + (3.0 / 9.0) * v1[d] +
                              (4.0 / 9.0) * v2[d]);
    // the field sample at the trace
    double gf[3];
    {
        Taps<DIM, FO> tf;
        make_taps<DIM, FO>(g, e, tr, tf);
        sc.template add<DIM, FO>(e, tf, ob);
        sample_grad<DIM, FO, VO == 1>(g, tf, cur.e[e], gf);
    }
    // back through the RK3 combination (x0 is fixed: no Jacobian there)
    double v0b[3] = {0, 0, 0}, v1b[3] = {0, 0, 0}, v2b[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double trb = ob * gf[d];
        v0b[d] = (-dt * (2.0 / 9.0)) * trb;
        v1b[d] = (-dt * (3.0 / 9.0)) * trb;
        v2b[d] = (-dt * (4.0 / 9.0)) * trb;
    }
    {
        double J[3][3];
        velocity_jac<DIM, VO>(g, cur, x2, J);
        velocity_scatter<DIM, VO>(g, sc, x2, v2b);
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            double x2b = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                x2b = fma(J[d][m], v2b[d], x2b);
            if (quirk)
                v0b[m] = fma(-0.75 * dt, x2b, v0b[m]);
            else
                v1b[m] = fma(-0.75 * dt, x2b, v1b[m]);
        }
    }
    {
        double J[3][3];
        velocity_jac<DIM, VO>(g, cur, x1, J);
        velocity_scatter<DIM, VO>(g, sc, x1, v1b);
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            double x1b = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                x1b = fma(J[d][m], v1b[d], x1b);
            v0b[m] = fma(-0.5 * dt, x1b, v0b[m]);
        }
    }
    velocity_scatter<DIM, VO>(g, sc, x0, v0b);
}

// One lane per index of the union of the owned boxes, as k_advect.
template <int DIM, int FO, int VO>
__global__ __launch_bounds__(256) void k_adj_advect(Geo g, Fields cur, Fields obar,
                                                    Fields bar)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
#pragma unroll 1
    for (int e = 0; e <= DIM; ++e) {
        bool own = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        const double ob = obar.e[e][c];
        if (ob == 0.0)
            continue;
        double x0[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            x0[d] = coord(g, e, d, l[d]);
        advect_point_adj<DIM, FO, VO>(g, e, cur, x0, ob, DirectScatter{g, bar});
    }
}

// 3-D: a block marches a column of 32 x 4 lattice points through `slabs`
// slabs of 2 planes.  Each slab's points accumulate their scatters in LDS
// windows (one per array, the slab plus M cells a side); the windows' planes
// form a ring along z, so when the march moves on by a slab, the TZ planes
// below the new window are complete: each non-zero cell of them is added to
// its owned index once and the slots are cleared for the planes entering at
// the top.  A point's traces stay within about one cell at the clamped time
// step, so nearly every sample lands in the window; the global atomics drop
// from ~136 per point to the window cells a slab leaves behind (2 planes of
// 38 x 10 per 256 points, plus the last window of the column).
template <int FO, int VO>
__global__ __launch_bounds__(256) void k_adj_advect_win(Geo g, Fields cur, Fields obar,
                                                        Fields bar, int slabs)
{
    using namespace adjwin;
    __shared__ double win[4][WN];
    const int t = threadIdx.x + TX * (threadIdx.y + TY * threadIdx.z);
    for (int q = t; q < 4 * WN; q += TX * TY * TZ)
        win[q / WN][q % WN] = 0.0;
    const int i0 = g.h + blockIdx.x * TX, j0 = g.h + blockIdx.y * TY;
    const int nslab = (g.uni_hi[2] - g.own_lo(2) + TZ - 1) / TZ;
    const int kb = g.own_lo(2) + (int)blockIdx.z * slabs * TZ;
    const int kend = g.own_lo(2) + min(nslab, ((int)blockIdx.z + 1) * slabs) * TZ;
    const int zbase = kb - M;
    __syncthreads();
#pragma unroll 1
    for (int k0 = kb; k0 < kend; k0 += TZ) {
        const WindowedScatter sc{g, bar, win, {i0 - M, j0 - M, k0 - M}, zbase};
        const int l[3] = {i0 + (int)threadIdx.x, j0 + (int)threadIdx.y, k0 + (int)threadIdx.z};
        const int64_t c = g.idx(l[0], l[1], l[2]);
#pragma unroll 1
        for (int e = 0; e <= 3; ++e) {
            const bool own = l[0] < g.own_hi[e][0] && l[1] < g.own_hi[e][1] &&
                             l[2] < g.own_hi[e][2];
            if (!own)
                continue;
            const double ob = obar.e[e][c];
            if (ob == 0.0)
                continue;
            double x0[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                x0[d] = coord(g, e, d, l[d]);
            advect_point_adj<3, FO, VO>(g, e, cur, x0, ob, sc);
        }
        __syncthreads();
        // flush the planes [k0 - M, k0 - M + np): the TZ lowest, which no
        // later slab's window holds, or after the last slab the whole window;
        // each cell goes to its owned index (wrapped; wall ghosts dropped)
        // and its slot is cleared
        const int np = k0 + TZ >= kend ? WZ : TZ;
#pragma unroll 1
        for (int a = 0; a < 4; ++a) {
            for (int q = t; q < np * WP; q += TX * TY * TZ) {
                const int z = k0 - M + q / WP, r = q % WP;
                double* cell = &win[a][WP * ((z - zbase) & (WZ - 1)) + r];
                const double v = *cell;
                if (v == 0.0)
                    continue;
                *cell = 0.0;
                const int ox = owned_index(g, a, 0, sc.wlo[0] + r % WX);
                const int oy = owned_index(g, a, 1, sc.wlo[1] + r / WX);
                const int oz = owned_index(g, a, 2, z);
                if (ox < 0 || oy < 0 || oz < 0)
                    continue;
                atomicAdd(bar.e[a] + g.idx(ox, oy, oz), v);
            }
        }
        __syncthreads();
    }
}

static dim3 union_grid(const Geo& g, dim3 blk)
{
    const int ex = g.uni_hi[0] - g.h, ey = g.uni_hi[1] - g.h;
    const int ez = g.dim == 3 ? g.uni_hi[2] - g.h : 1;
    return dim3((ex + blk.x - 1) / blk.x, (ey + blk.y - 1) / blk.y, ez);
}

template <int DIM, int FO, int VO>
static void adj_advect_t(const Geo& g, Fields cur, Fields obar, Fields bar,
                         hipStream_t s)
{
    if constexpr (DIM == 3) {
        using namespace adjwin;
        const dim3 blk(TX, TY, TZ);
        const int ex = g.uni_hi[0] - g.h, ey = g.uni_hi[1] - g.h;
        const int nslab = (g.uni_hi[2] - g.own_lo(2) + TZ - 1) / TZ;
        const int cols = ((ex + TX - 1) / TX) * ((ey + TY - 1) / TY);
        // columns of up to 16 slabs, shorter while that leaves fewer than
        // 2048 blocks (8 per CU)
        int slabs = 16;
        while (slabs > 1 && cols * ((nslab + slabs - 1) / slabs) < 2048)
            slabs /= 2;
        const dim3 grid((ex + TX - 1) / TX, (ey + TY - 1) / TY, (nslab + slabs - 1) / slabs);
        hipLaunchKernelGGL((k_adj_advect_win<FO, VO>), grid, blk, 0, s, g, cur, obar, bar,
                           slabs);
    } else {
        dim3 blk(64, 4, 1);
        hipLaunchKernelGGL((k_adj_advect<DIM, FO, VO>), union_grid(g, blk), blk, 0, s,
                           g, cur, obar, bar);
    }
}

void launch_adj_advect(const Geo& g, Fields cur, Fields obar, Fields bar,
                       hipStream_t s)
{
    const int key = g.dim * 100 + g.field_order * 10 + g.vel_order;
    switch (key) {
    case 211: adj_advect_t<2, 1, 1>(g, cur, obar, bar, s); break;
    case 231: adj_advect_t<2, 3, 1>(g, cur, obar, bar, s); break;
    case 213: adj_advect_t<2, 1, 3>(g, cur, obar, bar, s); break;
    case 233: adj_advect_t<2, 3, 3>(g, cur, obar, bar, s); break;
    case 311: adj_advect_t<3, 1, 1>(g, cur, obar, bar, s); break;
    case 331: adj_advect_t<3, 3, 1>(g, cur, obar, bar, s); break;
    case 313: adj_advect_t<3, 1, 3>(g, cur, obar, bar, s); break;
    case 333: adj_advect_t<3, 3, 3>(g, cur, obar, bar, s); break;
    default: break;
    }
}

// ---- inputs^T: Solver::_addInputs (src/Solver.hpp:181-263) reversed --------------
// pre: the values the inputs saw.  A face zeroed by a SOLID wall passes no
// adjoint; a value the inflow overwrote passes its adjoint to dL/dU_d
// (block sums, one atomic per block and component); an overwritten quantity
// passes nothing.
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_inputs(Geo g, Fields pre, Fields bar,
                                                    double* __restrict__ dvel)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e <= DIM; ++e) {
        bool own = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        double x[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            x[d] = coord(g, e, d, l[d]);
        const bool in = in_source(g, x);
        const double val = pre.e[e][c];
        if (e == 0) {
            if (in && val < g.Q)
                bar.e[0][c] = 0.0;
            continue;
        }
        const int d = e - 1;
        double b = bar.e[e][c];
        if (!g.periodic[d]) {
            const int gl = l[d] - g.h;
            if ((gl <= 0 && g.solid[d][0]) || (gl > g.n[d] - 1 && g.solid[d][1]))
                b = 0.0;
        }
        if (in && fabs(val) < fabs(g.U[d])) {
            acc[d] += b;
            b = 0.0;
        }
        bar.e[e][c] = b;
    }
    // block sums of the inflow-velocity adjoint
    __shared__ double sh[3][4];
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d)
        for (int off = kWave / 2; off > 0; off >>= 1)
            acc[d] += __shfl_xor(acc[d], off, kWave);
    if (lane == 0)
        for (int d = 0; d < 3; ++d)
            sh[d][threadIdx.y] = acc[d];
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0)
        for (int d = 0; d < DIM; ++d) {
            const double s = (sh[d][0] + sh[d][1]) + (sh[d][2] + sh[d][3]);
            if (s != 0.0)
                atomicAdd(dvel + d, s);
        }
}

void launch_adj_inputs(const Geo& g, Fields pre, Fields bar, double* dvel,
                       hipStream_t s)
{
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_adj_inputs<2>, union_grid(g, blk), blk, 0, s, g, pre,
                           bar, dvel);
    else
        hipLaunchKernelGGL(k_adj_inputs<3>, union_grid(g, blk), blk, 0, s, g, pre,
                           bar, dvel);
}

// ---- projection^T: VelocityCorrector::_applyPressure / _buildRHS reversed --------
// 1. The wall rule applied after the face updates (with CF_QUIRK_PROJ_BC_U the
//    y/z loops' rule lands on u, at the y/z-face positions, as in k_project).
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_mask(Geo g, Fields bar)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int e = d + 1;
        if (g.periodic[d])
            continue;
        bool own = true;
#pragma unroll
        for (int m = 0; m < DIM; ++m)
            own = own && l[m] < g.own_hi[e][m];
        if (!own)
            continue;
        const int gl = l[d] - g.h;
        if ((gl <= 0 && g.solid[d][0]) || (gl > g.n[d] - 1 && g.solid[d][1])) {
            double* bc = (d > 0 && (g.quirks & CF_QUIRK_PROJ_BC_U)) ? bar.e[1] : bar.e[e];
            bc[c] = 0.0;
        }
    }
}

// 2. y = -s G^T bar over the owned cells (the pressure gather folded in), and
//    its sum for the consistency shift.
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_gradT(Geo g, Fields bar,
                                                   double* __restrict__ y,
                                                   double* __restrict__ ysum, RedBuf rb)
{
    const int64_t stride[3] = {1, g.px, g.px * g.py};
    const double ns = -g.proj_scale;
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const int gi[3] = {i, j, k};
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            // faces c (this cell's low face) and c + e_d (its high face)
            const int64_t up = (g.periodic[d] && gi[d] == g.n[d] - 1)
                                   ? c - (int64_t)(g.n[d] - 1) * stride[d]
                                   : c + stride[d];
            t = t + (bar.e[d + 1][c] - bar.e[d + 1][up]);
        }
        const double v = ns * t;
        y[c] = v;
        dd_add(hi[0], lo[0], v, 0.0);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) { ysum[0] = s[0]; });
}

// 3. y -= diag(A) * sum(y) / sum(diag(A)) (singular operator only).
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_shift(Geo g, double* __restrict__ y,
                                                   const double* __restrict__ ysum,
                                                   double diag_sum)
{
    const double mu = ysum[0] / diag_sum;
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const int gi[3] = {i, j, k};
        y[c] = y[c] - g.diag[solid_count(g, gi)] * mu;
    });
}

// 4. bar += -(1/dx) D^T z over the owned faces (the divergence gather folded
//    in): face f of axis d gets z of its low cell minus z of its high cell.
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_divT(Geo g, const double* __restrict__ z,
                                                  Fields bar)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
    const int64_t stride[3] = {1, g.px, g.px * g.py};
    const double ns = -g.div_scale;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int e = d + 1;
        bool own = true;
#pragma unroll
        for (int m = 0; m < DIM; ++m)
            own = own && l[m] < g.own_hi[e][m];
        if (!own)
            continue;
        const int gl = l[d] - g.h;
        double zl = 0.0, zu = 0.0;
        if (gl > 0)
            zl = z[c - stride[d]];
        else if (g.periodic[d])
            zl = z[c + (int64_t)(g.n[d] - 1) * stride[d]];
        if (gl < g.n[d])
            zu = z[c];
        bar.e[e][c] = fma(ns, zl - zu, bar.e[e][c]);
    }
}

void launch_adj_mask(const Geo& g, Fields bar, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_adj_mask<2>, union_grid(g, blk), blk, 0, s, g, bar);
    else
        hipLaunchKernelGGL(k_adj_mask<3>, union_grid(g, blk), blk, 0, s, g, bar);
}

void launch_adj_gradT(const Geo& g, Fields bar, double* y, double* ysum, RedBuf rb,
                      hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_adj_gradT<2>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g,
                           bar, y, ysum, rb);
    else
        hipLaunchKernelGGL(k_adj_gradT<3>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g,
                           bar, y, ysum, rb);
}

void launch_adj_shift(const Geo& g, double* y, const double* ysum, double diag_sum,
                      hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_adj_shift<2>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g,
                           y, ysum, diag_sum);
    else
        hipLaunchKernelGGL(k_adj_shift<3>, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g,
                           y, ysum, diag_sum);
}

void launch_adj_divT(const Geo& g, const double* z, Fields bar, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_adj_divT<2>, union_grid(g, blk), blk, 0, s, g, z, bar);
    else
        hipLaunchKernelGGL(k_adj_divT<3>, union_grid(g, blk), blk, 0, s, g, z, bar);
}

} // namespace cfmi
