// cf_mg.hip -- kernels of the multigrid preconditioner (cf_mg.hpp).
//
// Arithmetic order, shared with the oracle's restatement so that the
// V-cycle is bit-identical: A x = diag x, then fma over the -x, +x, -y, +y,
// -z, +z couplings (a missing neighbour contributes fma(0, 0, .)); a Jacobi
// sweep is fma(w, b - A x, x); restriction sums the children in z, y, x
// order starting from 0; prolongation adds the parent's value.
#include <hip/hip_runtime.h>

#include "cf_mg.hpp"
#include "cf_stream.hpp"

namespace cfmi
{

// ---- level 0: padded layout, matrix-free operator ---------------------------------

__device__ __forceinline__ double tab_pick(const MgTab& t, int ub, int wx)
{
    const double a = t.v[ub], b = t.v[ub + 1], c = t.v[ub + 2];
    return wx == 0 ? a : (wx == 1 ? b : c);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_mg0_first(Geo g, MgTab wd, const double* __restrict__ b,
                                                   double* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const int ub = wall_count(g, 1, j) + wall_count(g, 2, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            x[c] = tab_pick(wd, ub, wall_count(g, 0, i)) * b[c];
        }
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_mg0_sweep(Geo g, MgTab wd, const double* __restrict__ xi,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ xo, const PcgState* st)
{
    if (st->done)
        return;
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            const double y = apply_A<DIM, double>(g, rs, lx, xi, c);
            xo[c] = fma(tab_pick(wd, rs.cnt, lx.cnt), b[c] - y, xi[c]);
        }
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_mg0_residual(Geo g, const double* __restrict__ x,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ r,
                                                      const PcgState* st)
{
    if (st->done)
        return;
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            r[c] = b[c] - apply_A<DIM, double>(g, rs, lx, x, c);
        }
    });
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

__device__ __forceinline__ void unflatten(int64_t id, const int (&n)[3], int (&c)[3])
{
    c[0] = (int)(id % n[0]);
    const int64_t r = id / n[0];
    c[1] = (int)(r % n[1]);
    c[2] = (int)(r / n[1]);
}

__device__ __forceinline__ int64_t flat(const int (&n)[3], int i, int j, int k)
{
    return (int64_t)i + (int64_t)n[0] * ((int64_t)j + (int64_t)n[1] * k);
}

#define CF_GRID_STRIDE(id, count)                                                   \
    for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < (count); \
         id += (int64_t)gridDim.x * blockDim.x)

__global__ __launch_bounds__(256) void k_mg0_restrict(Geo g, MgLev C, const double* __restrict__ r,
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
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i)
                    sum = sum + r[g.idx(g.h + i, g.h + j, g.dim == 3 ? g.h + k : 0)];
        bc[id] = sum;
    }
}

__global__ __launch_bounds__(256) void k_mg0_prolong(Geo g, MgLev C, const double* __restrict__ xc,
                                                     double* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const int f[3] = {i, j, k};
        int P[3];
        for (int d = 0; d < 3; ++d)
            P[d] = (d < g.dim && g.n[d] > 2) ? f[d] / 2 : f[d];
        x[c] = x[c] + xc[flat(C.n, P[0], P[1], P[2])];
    });
}

// ---- compact levels ------------------------------------------------------------------

// neighbour of cell c along d (dir -1 / +1) with the periodic wrap; -1 outside
__device__ __forceinline__ int64_t nbr(const MgLev& L, const int (&c)[3], int d, int dir)
{
    int m[3] = {c[0], c[1], c[2]};
    m[d] += dir;
    if (m[d] < 0 || m[d] >= L.n[d]) {
        if (!L.periodic[d])
            return -1;
        m[d] = m[d] < 0 ? L.n[d] - 1 : 0;
    }
    return flat(L.n, m[0], m[1], m[2]);
}

__device__ __forceinline__ double apply_lev(const MgLev& L, const double* __restrict__ x,
                                            int64_t id, const int (&c)[3])
{
    double acc = L.diag[id] * x[id];
    for (int d = 0; d < L.dim; ++d) {
        const int64_t m = nbr(L, c, d, -1), p = nbr(L, c, d, +1);
        const double cm = m >= 0 ? L.cp[d][m] : 0.0;
        const double xm = m >= 0 ? x[m] : 0.0;
        acc = fma(cm, xm, acc);
        const double cpl = p >= 0 ? L.cp[d][id] : 0.0;
        const double xp = p >= 0 ? x[p] : 0.0;
        acc = fma(cpl, xp, acc);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_mg_first(MgLev L, const double* __restrict__ b,
                                                  double* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, L.ncell) { x[id] = L.w[id] * b[id]; }
}

__global__ __launch_bounds__(256) void k_mg_sweep(MgLev L, const double* __restrict__ xi,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ xo, const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, L.ncell)
    {
        int c[3];
        unflatten(id, L.n, c);
        const double y = apply_lev(L, xi, id, c);
        xo[id] = fma(L.w[id], b[id] - y, xi[id]);
    }
}

__global__ __launch_bounds__(256) void k_mg_residual_restrict(MgLev F, MgLev C,
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
            children(F.coarsen[d] != 0, P[d], F.n[d], lo[d], hi[d]);
        double sum = 0.0;
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    const int64_t fi = flat(F.n, i, j, k);
                    sum = sum + (b[fi] - apply_lev(F, x, fi, c));
                }
        bc[id] = sum;
    }
}

__global__ __launch_bounds__(256) void k_mg_prolong(MgLev F, MgLev C, const double* __restrict__ xc,
                                                    double* __restrict__ x, const PcgState* st)
{
    if (st->done)
        return;
    CF_GRID_STRIDE(id, F.ncell)
    {
        int f[3];
        unflatten(id, F.n, f);
        int P[3];
        for (int d = 0; d < 3; ++d)
            P[d] = F.coarsen[d] ? f[d] / 2 : f[d];
        x[id] = x[id] + xc[flat(C.n, P[0], P[1], P[2])];
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
        const bool zero = norm == 0.0;
        st->done = zero ? 1 : (maxit <= 0 ? 1 : 0);
        st->converged = zero ? 1 : 0;
        st->res = zero ? 0.0 : 1.0;
        st->mg = 1;
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

void launch_mg0_sweep(const Geo& g, const MgTab& wd0, const double* xi, const double* b,
                      double* xo, const PcgState* st, hipStream_t s)
{
    CF_MG0_LAUNCH(k_mg0_sweep, g, wd0, xi, b, xo, st);
}

void launch_mg0_residual(const Geo& g, const double* x, const double* b, double* r,
                         const PcgState* st, hipStream_t s)
{
    CF_MG0_LAUNCH(k_mg0_residual, g, x, b, r, st);
}

void launch_mg0_restrict(const Geo& g, const MgLev& c1, const double* r, double* b1,
                         const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg0_restrict, compact_grid(c1.ncell), dim3(kBlock), 0, s, g, c1, r,
                       b1, st);
}

void launch_mg0_prolong(const Geo& g, const MgLev& c1, const double* x1, double* x,
                        const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg0_prolong, dim3(stream_blocks(g)), dim3(kBlock), 0, s, g, c1, x1,
                       x, st);
}

void launch_mg_first(const MgLev& L, const double* b, double* x, const PcgState* st,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_first, compact_grid(L.ncell), dim3(kBlock), 0, s, L, b, x, st);
}

void launch_mg_sweep(const MgLev& L, const double* xi, const double* b, double* xo,
                     const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_sweep, compact_grid(L.ncell), dim3(kBlock), 0, s, L, xi, b, xo,
                       st);
}

void launch_mg_residual_restrict(const MgLev& F, const MgLev& C, const double* x,
                                 const double* b, double* bc, const PcgState* st,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_residual_restrict, compact_grid(C.ncell), dim3(kBlock), 0, s, F,
                       C, x, b, bc, st);
}

void launch_mg_prolong(const MgLev& F, const MgLev& C, const double* xc, double* x,
                       const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_prolong, compact_grid(F.ncell), dim3(kBlock), 0, s, F, C, xc, x,
                       st);
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
