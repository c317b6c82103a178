// cf_kernels.hip -- HIP kernels of the CajitaFluids hot path for gfx950:
// advection, inputs, ghost fill, divergence and projection.  The PCG lives
// in cf_pcg.hip; the streaming-kernel building blocks in cf_stream.hpp.
#include <hip/hip_runtime.h>

#include "cf_stream.hpp"

namespace cfmi
{

// ---- advection: TimeIntegrator::advect for every entity in one launch ------
// (src/TimeIntegrator.hpp:81-116): each lane owns one (i, j, k) of the union
// of the owned boxes and advects every entity owning that index.
//
// Occupancy: the 3-D kernels with linear velocity sampling keep the 64
// cubic taps of a field sample in flight and would take 146 VGPRs (3 waves
// per SIMD); capping them at 128 VGPRs (4 waves, a 36-byte spill) runs the
// 256^3 vortex tank's advection in 3.0 ms instead of 4.15 (64 x 4 x 1
// blocks).  Higher caps spill more than they gain (5 waves: 4.2 ms, 6:
// 5.2 ms).
template <int DIM, int VORDER>
constexpr int advect_min_waves()
{
    return DIM == 3 && VORDER == 1 ? 4 : 1;
}

template <int DIM, int FORDER, int VORDER>
__global__ __launch_bounds__(256, (advect_min_waves<DIM, VORDER>())) void k_advect(Geo g,
                                                                                 Fields cur,
                                                                                 Fields next)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z * blockDim.z + threadIdx.z;
    const int l[3] = {i, j, k};
    const double* u = cur.e[1];
    const double* v = cur.e[2];
    const double* w = cur.e[3];
#pragma unroll
    for (int e = 0; e <= DIM; ++e) {
        bool own = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        double x0[3] = {0, 0, 0}, tr[3];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            x0[d] = coord(g, e, d, l[d]);
        rk3<DIM, VORDER>(g, u, v, w, x0, tr);
        next.e[e][g.idx(i, j, k)] =
            interpolate<DIM, FORDER>(g, e, cur.e[e], tr);
    }
}

static dim3 union_grid(const Geo& g, dim3 blk)
{
    const int ex = g.uni_hi[0] - g.h, ey = g.uni_hi[1] - g.h;
    const int ez = g.dim == 3 ? g.uni_hi[2] - g.h : 1;
    return dim3((ex + blk.x - 1) / blk.x, (ey + blk.y - 1) / blk.y, (ez + blk.z - 1) / blk.z);
}

// 3-D blocks of 32 x 4 x 2 points: two z planes share the cubic stencils in
// the CU's caches (2.88 against 3.0 ms for 64 x 4 x 1 at 256^3).
template <int DIM, int FORDER, int VORDER>
static void advect_t(const Geo& g, Fields cur, Fields next, hipStream_t s)
{
    const dim3 blk = DIM == 3 ? dim3(32, 4, 2) : dim3(64, 4, 1);
    hipLaunchKernelGGL((k_advect<DIM, FORDER, VORDER>), union_grid(g, blk), blk,
                       0, s, g, cur, next);
}

void launch_advect(const Geo& g, Fields cur, Fields next, hipStream_t s)
{
    const int key = g.dim * 100 + g.field_order * 10 + g.vel_order;
    switch (key) {
    case 211: advect_t<2, 1, 1>(g, cur, next, s); break;
    case 231: advect_t<2, 3, 1>(g, cur, next, s); break;
    case 213: advect_t<2, 1, 3>(g, cur, next, s); break;
    case 233: advect_t<2, 3, 3>(g, cur, next, s); break;
    case 311: advect_t<3, 1, 1>(g, cur, next, s); break;
    case 331: advect_t<3, 3, 1>(g, cur, next, s); break;
    case 313: advect_t<3, 1, 3>(g, cur, next, s); break;
    case 333: advect_t<3, 3, 3>(g, cur, next, s); break;
    default: break;
    }
}

// ---- inputs: Solver::_addInputs (src/Solver.hpp:181-263) -----------------
// The rule at one owned entity of kind e, local index l: the source
// (InflowSource.hpp:33-78), the body force (BodyForce.hpp:44-60), then the
// wall (BoundaryConditions.hpp:102-129), in the reference's order.
template <int DIM>
__device__ __forceinline__ void input_rule(const Geo& g, int e, const int (&l)[3],
                                           double* __restrict__ a)
{
    const int64_t c = g.idx(l[0], l[1], l[2]);
    double x[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x[d] = coord(g, e, d, l[d]);
    double val = a[c];
    const bool in = in_source(g, x);
    if (e == 0) {
        if (in && val < g.Q)
            val = g.Q;
    } else {
        const int d = e - 1;
        const double U = g.U[d];
        if (in && fabs(val) < fabs(U))
            val = U;
        val = val + g.force[d] * g.dt;
        if (!g.periodic[d]) {
            const int gl = l[d] - g.h;
            if (gl <= 0 && g.solid[d][0])
                val = 0;
            if (gl > g.n[d] - 1 && g.solid[d][1])
                val = 0;
        }
    }
    a[c] = val;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_inputs(Geo g, Fields cur)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
#pragma unroll
    for (int e = 0; e <= DIM; ++e) {
        bool own = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (own)
            input_rule<DIM>(g, e, l, cur.e[e]);
    }
}

// With no body force the rule changes an entity only inside the inflow
// region or on a SOLID wall (elsewhere it adds 0 * dt, which leaves every
// value but a -0.0 as it is), so the inputs visit only boxes of local
// indices that cover those: per entity, the inflow region's bounding box
// widened by one index, and each SOLID wall plane of a face entity.  The
// rule is idempotent, so a cell in two boxes gets the same value twice.
struct InputBoxes
{
    int n;
    int e[kMaxInputBoxes];
    int lo[kMaxInputBoxes][3], hi[kMaxInputBoxes][3];
    int64_t start[kMaxInputBoxes + 1]; // prefix sums of the box volumes
};

template <int DIM>
__global__ __launch_bounds__(256) void k_inputs_boxes(Geo g, Fields cur, InputBoxes bx)
{
    const int64_t total = bx.start[bx.n];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int b = 0;
        while (t >= bx.start[b + 1])
            ++b;
        const int64_t u = t - bx.start[b];
        const int nx = bx.hi[b][0] - bx.lo[b][0], ny = bx.hi[b][1] - bx.lo[b][1];
        const int l[3] = {bx.lo[b][0] + (int)(u % nx), bx.lo[b][1] + (int)((u / nx) % ny),
                          bx.lo[b][2] + (int)(u / ((int64_t)nx * ny))};
        input_rule<DIM>(g, bx.e[b], l, cur.e[bx.e[b]]);
    }
}

static InputBoxes input_boxes(const Geo& g)
{
    InputBoxes bx;
    bx.n = 0;
    auto add = [&](int e, int lo[3], int hi[3]) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = max(lo[d], g.own_lo(d));
            hi[d] = min(hi[d], g.own_hi[e][d]);
            if (hi[d] <= lo[d])
                return;
        }
        bx.e[bx.n] = e;
        for (int d = 0; d < 3; ++d) {
            bx.lo[bx.n][d] = lo[d];
            bx.hi[bx.n][d] = hi[d];
        }
        ++bx.n;
    };
    if (g.inflow_shape != CF_INFLOW_NONE) {
        for (int e = 0; e <= g.dim; ++e) {
            int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
            for (int d = 0; d < g.dim; ++d) {
                double a, b;
                if (g.inflow_shape == CF_INFLOW_BOX) {
                    a = g.in_lo[d];
                    b = g.in_hi[d];
                } else {
                    const double r = sqrt(g.in_r2);
                    a = g.in_c[d] - r;
                    b = g.in_c[d] + r;
                }
                // coordinate(l) = x0 + l dx: widened by one index each way
                lo[d] = (int)floor((a - g.x0[e][d]) / g.dx) - 1;
                hi[d] = (int)ceil((b - g.x0[e][d]) / g.dx) + 2;
            }
            add(e, lo, hi);
        }
    }
    for (int e = 1; e <= g.dim; ++e) {
        const int d = e - 1;
        if (g.periodic[d])
            continue;
        for (int side = 0; side < 2; ++side) {
            if (!g.solid[d][side])
                continue;
            int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
            for (int m = 0; m < g.dim; ++m) {
                lo[m] = g.own_lo(m);
                hi[m] = g.own_hi[e][m];
            }
            lo[d] = side == 0 ? g.h : g.h + g.n[d];
            hi[d] = lo[d] + 1;
            add(e, lo, hi);
        }
    }
    bx.start[0] = 0;
    for (int b = 0; b < bx.n; ++b)
        bx.start[b + 1] = bx.start[b] + (int64_t)(bx.hi[b][0] - bx.lo[b][0]) *
                                             (bx.hi[b][1] - bx.lo[b][1]) *
                                             (bx.hi[b][2] - bx.lo[b][2]);
    return bx;
}

bool inputs_are_sparse(const Geo& g)
{
    return g.force[0] == 0.0 && g.force[1] == 0.0 && g.force[2] == 0.0;
}

void launch_inputs(const Geo& g, Fields cur, hipStream_t s)
{
    if (inputs_are_sparse(g)) {
        const InputBoxes bx = input_boxes(g);
        const int64_t total = bx.start[bx.n];
        if (total == 0)
            return;
        int64_t blocks = (total + 255) / 256;
        const unsigned nb = (unsigned)(blocks < 2048 ? blocks : 2048);
        if (g.dim == 2)
            hipLaunchKernelGGL(k_inputs_boxes<2>, dim3(nb), dim3(256), 0, s, g, cur, bx);
        else
            hipLaunchKernelGGL(k_inputs_boxes<3>, dim3(nb), dim3(256), 0, s, g, cur, bx);
        return;
    }
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_inputs<2>, union_grid(g, blk), blk, 0, s, g, cur);
    else
        hipLaunchKernelGGL(k_inputs<3>, union_grid(g, blk), blk, 0, s, g, cur);
}

// ---- ghost fill: single-rank Cajita halo gather ---------------------------
// Periodic axis d: ghost layers copy the opposite owned layers, over the
// full extent of the other axes (axes are filled in order, so corners come
// out right).  Wall axes have no neighbour and keep their zero ghosts.
__global__ __launch_bounds__(256) void k_ghost_fill(Geo g, int e, int d,
                                                     double* a)
{
    const int E0 = g.ext[e][0], E1 = g.ext[e][1], E2 = g.ext[e][2];
    const int Ed = g.ext[e][d];
    const int nlay = Ed - g.n[d]; // ghost layers along d
    int64_t total;
    int o1, o2;
    if (d == 0) { o1 = 1; o2 = 2; }
    else if (d == 1) { o1 = 0; o2 = 2; }
    else { o1 = 0; o2 = 1; }
    const int Ea = g.ext[e][o1], Eb = g.ext[e][o2];
    total = (int64_t)nlay * Ea * Eb;
    (void)E0; (void)E1; (void)E2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int ia = (int)(t % Ea);
        const int rest = (int)(t / Ea);
        const int ib = rest % Eb;
        const int lay = rest / Eb;
        const int ld = lay < g.h ? lay : g.n[d] + lay; // ghost index along d
        const int src = ld < g.h ? ld + g.n[d] : ld - g.n[d];
        int dst3[3], src3[3];
        dst3[d] = ld;
        src3[d] = src;
        dst3[o1] = src3[o1] = ia;
        dst3[o2] = src3[o2] = ib;
        a[g.idx(dst3[0], dst3[1], dst3[2])] = a[g.idx(src3[0], src3[1], src3[2])];
    }
}

void launch_ghost_fill(const Geo& g, int entity, double* a, hipStream_t s)
{
    for (int d = 0; d < g.dim; ++d) {
        if (!g.periodic[d])
            continue;
        hipLaunchKernelGGL(k_ghost_fill, dim3(1024), dim3(256), 0, s, g, entity,
                           d, a);
    }
}

// ---- divergence RHS: VelocityCorrector::_buildRHS (:182-212) --------------
template <int DIM>
__global__ __launch_bounds__(256) void k_divergence(Geo g, Fields cur,
                                                     double* __restrict__ rhs)
{
    const double* __restrict__ u = cur.e[1];
    const double* __restrict__ v = cur.e[2];
    const double* __restrict__ w = cur.e[3];
    const int64_t py = g.px, pz = g.px * g.py;
    const double ns = -g.div_scale;
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        double t = u[c + 1] - u[c];
        t = t + v[c + py];
        t = t - v[c];
        if constexpr (DIM == 3) {
            t = t + w[c + pz];
            t = t - w[c];
        }
        rhs[c] = ns * t;
    });
}

void launch_divergence(const Geo& g, Fields cur, double* rhs, hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_divergence<2>, dim3(stream_blocks(g)), dim3(kBlock),
                           0, s, g, cur, rhs);
    else
        hipLaunchKernelGGL(k_divergence<3>, dim3(stream_blocks(g)), dim3(kBlock),
                           0, s, g, cur, rhs);
}

// ---- projection: VelocityCorrector::_applyPressure (:214-264) -------------
template <int DIM>
__global__ __launch_bounds__(256) void k_project(Geo g, Fields cur,
                                                  const double* __restrict__ p)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
    const int64_t stride[3] = {1, g.px, g.px * g.py};
    const double scale = g.proj_scale;
    double* u = cur.e[1];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int e = d + 1;
        bool own = true;
#pragma unroll
        for (int m = 0; m < DIM; ++m)
            own = own && l[m] < g.own_hi[e][m];
        if (!own)
            continue;
        double* a = cur.e[e];
        const double dp = p[c] - p[c - stride[d]];
        a[c] = a[c] - scale * dp;
        if (!g.periodic[d]) {
            double* bc = (d > 0 && (g.quirks & CF_QUIRK_PROJ_BC_U)) ? u : a;
            const int gl = l[d] - g.h;
            if (gl <= 0 && g.solid[d][0])
                bc[c] = 0;
            if (gl > g.n[d] - 1 && g.solid[d][1])
                bc[c] = 0;
        }
    }
}

void launch_project(const Geo& g, Fields cur, const double* p, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_project<2>, union_grid(g, blk), blk, 0, s, g, cur, p);
    else
        hipLaunchKernelGGL(k_project<3>, union_grid(g, blk), blk, 0, s, g, cur, p);
}

// SiloWriter.hpp:141-197: the owned cells of q and, at each owned cell
// centre, the velocity from linear interpolation of each face component
// (interpolateVelocity<2, 1>), as compact arrays with x fastest; vel_out
// holds dim arrays back to back.
template <int DIM>
__global__ __launch_bounds__(256) void k_cell_output(Geo g, Fields cur, double* __restrict__ q_out,
                                                     double* __restrict__ vel_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1])
        return;
    const int l[3] = {g.h + i, g.h + j, g.own_lo(2) + k};
    double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x[d] = coord(g, CF_CELL, d, l[d]);
    const int64_t o = i + (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k);
    const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
    q_out[o] = cur.e[CF_CELL][g.idx(l[0], l[1], l[2])];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        vel_out[d * N + o] = interpolate<DIM, 1>(g, d + 1, cur.e[d + 1], x);
}

// Dense ghosted array (x fastest, the entity's ghosted extents) <-> the
// padded layout: the host transfers of cf_copy_in / cf_copy_out are one
// contiguous copy each, with this repacking on the device.
__global__ __launch_bounds__(256) void k_repack(Geo g, int e, double* __restrict__ padded,
                                                double* __restrict__ dense, int to_padded)
{
    const int E0 = g.ext[e][0], E1 = g.ext[e][1], E2 = g.ext[e][2];
    const int64_t total = (int64_t)E0 * E1 * E2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % E0);
        const int64_t r = t / E0;
        const int j = (int)(r % E1);
        const int k = (int)(r / E1);
        const int64_t c = g.idx(i, j, k);
        if (to_padded)
            padded[c] = dense[t];
        else
            dense[t] = padded[c];
    }
}

void launch_repack(const Geo& g, int entity, double* padded, double* dense, bool to_padded,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_repack, dim3(2048), dim3(256), 0, s, g, entity, padded, dense,
                       to_padded ? 1 : 0);
}

__global__ __launch_bounds__(256) void k_copy_padded(Geo g, int e, double* __restrict__ dst,
                                                    const double* __restrict__ src,
                                                    int owned_only)
{
    const int64_t plane = g.px * g.py;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.total;
         t += (int64_t)gridDim.x * blockDim.x) {
        bool own = true;
        if (owned_only) {
            // t = sx + i + px (j + py k); entries before sx or past the last
            // plane belong to no cell
            const int64_t u = t - g.sx;
            const int i = (int)(u % g.px), j = (int)((u / g.px) % g.py);
            const int64_t k = u / plane;
            own = u >= 0 && i >= g.h && i < g.own_hi[e][0] && j >= g.own_lo(1) &&
                  j < g.own_hi[e][1] && k >= g.own_lo(2) && k < g.own_hi[e][2];
        }
        dst[t] = own ? src[t] : 0.0;
    }
}

void launch_copy_padded(const Geo& g, int entity, double* dst, const double* src,
                        bool owned_only, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_padded, dim3(2048), dim3(256), 0, s, g, entity, dst, src,
                       owned_only ? 1 : 0);
}

void launch_cell_output(const Geo& g, Fields cur, double* q_out, double* vel_out, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    dim3 grid((g.n[0] + 63) / 64, (g.n[1] + 3) / 4, g.n[2]);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_cell_output<2>, grid, blk, 0, s, g, cur, q_out, vel_out);
    else
        hipLaunchKernelGGL(k_cell_output<3>, grid, blk, 0, s, g, cur, q_out, vel_out);
}

} // namespace cfmi
