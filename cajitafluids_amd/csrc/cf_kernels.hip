// cf_kernels.hip -- HIP kernels of the CajitaFluids hot path for gfx950.
//
// Streaming kernels (divergence, PCG) walk the owned cells as x-rows: a
// 256-lane block takes one row at a time, so every load and store of a wave
// is a contiguous 512-byte run.  Rows are split into 8 contiguous chunks, one
// per XCD (workgroups are dealt to the XCDs round-robin), so the y and z
// neighbour rows a block reads were fetched by its own XCD's L2 shortly
// before.  Reductions accumulate double-double per lane, reduce across the
// wave with xor shuffles and across the block through LDS, and the last block
// to arrive reduces the per-block partials and computes the PCG scalars on
// the device.
#include <hip/hip_runtime.h>

#include "cf_device.hpp"
#include "cf_kernels.hpp"

namespace cfmi
{

int stream_grid(const Geo& g)
{
    (void)g;
    return kRedBlocks;
}

// ---- owned-cell iteration ------------------------------------------------
template <class F>
__device__ __forceinline__ void for_owned_cells(const Geo& g, F&& f)
{
    const int n0 = g.n[0], n1 = g.n[1];
    const int nrows = n1 * g.n[2];
    const int G8 = gridDim.x >> 3;
    const int xcd = blockIdx.x & 7;
    const int m = blockIdx.x >> 3;
    const int r0 = (int)(((int64_t)nrows * xcd) >> 3);
    const int r1 = (int)(((int64_t)nrows * (xcd + 1)) >> 3);
    const int k0 = g.own_lo(2);
    for (int r = r0 + m; r < r1; r += G8) {
        const int j = r % n1;
        const int k = r / n1;
        const int64_t base = g.idx(g.h, g.h + j, k0 + k);
        for (int i = threadIdx.x; i < n0; i += blockDim.x)
            f(base + i, i, j, k);
    }
}

// ---- reductions ------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_reduce(double (&hi)[NV], double (&lo)[NV])
{
    __shared__ double sh[NV][2][kBlock / kWave];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double oh = __shfl_xor(hi[v], off, kWave);
            double ol = __shfl_xor(lo[v], off, kWave);
            dd_add(hi[v], lo[v], oh, ol);
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sh[v][0][wv] = hi[v];
            sh[v][1][wv] = lo[v];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            hi[v] = sh[v][0][0];
            lo[v] = sh[v][1][0];
            for (int w = 1; w < (int)(blockDim.x / kWave); ++w)
                dd_add(hi[v], lo[v], sh[v][0][w], sh[v][1][w]);
        }
    }
}

// Reduces hi/lo over the whole grid; the last block to arrive calls
// fin(sum) on thread 0.  Every block of the launch must call it.
template <int NV, class Fin>
__device__ __forceinline__ void grid_reduce(double (&hi)[NV], double (&lo)[NV],
                                            RedBuf rb, Fin&& fin)
{
    block_reduce<NV>(hi, lo);
    __shared__ int last;
    if (threadIdx.x == 0) {
        double* part = rb.partials + (size_t)blockIdx.x * NV * 2;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            part[2 * v] = hi[v];
            part[2 * v + 1] = lo[v];
        }
        __threadfence();
        unsigned t = atomicInc(rb.counter, gridDim.x - 1);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last)
        return;
    __threadfence();
    double h2[NV], l2[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        h2[v] = 0.0;
        l2[v] = 0.0;
    }
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        const double* part = rb.partials + (size_t)b * NV * 2;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            dd_add(h2[v], l2[v], part[2 * v], part[2 * v + 1]);
    }
    __syncthreads();
    block_reduce<NV>(h2, l2);
    if (threadIdx.x == 0) {
        double sum[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sum[v] = h2[v] + l2[v];
        fin(sum);
    }
}

// ---- matrix-free pressure operator ----------------------------------------
// y = A x at owned cell c with global index (i, j, k): the reference's stored
// stencil (VelocityCorrector.hpp:116-144 with BoundaryConditions.hpp:57-97)
// evaluated from the wall mask, in stencil order {0}, -x, +x, -y, +y, -z, +z.
// Wall couplings are 0 (their ghost operand is 0); periodic axes wrap.
template <int DIM, class T>
struct Stencil
{
    T diag, cm[3], cp[3];
    int64_t om[3], op[3];
};

template <int DIM, class T>
__device__ __forceinline__ Stencil<DIM, T> stencil_at(const Geo& g, int i,
                                                      int j, int k)
{
    Stencil<DIM, T> st;
    const int gi[3] = {i, j, k};
    const int64_t stride[3] = {1, g.px, g.px * g.py};
    const int cnt = solid_count(g, gi);
    const T s = std::is_same<T, float>::value ? (T)g.s32 : (T)g.s;
    st.diag = std::is_same<T, float>::value ? (T)g.diag32[cnt] : (T)g.diag[cnt];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const bool lo = gi[d] <= 0, hi = gi[d] > g.n[d] - 2;
        if (g.periodic[d]) {
            st.cm[d] = -s;
            st.cp[d] = -s;
            st.om[d] = (gi[d] == 0) ? (int64_t)(g.n[d] - 1) * stride[d] : -stride[d];
            st.op[d] = (gi[d] == g.n[d] - 1) ? -(int64_t)(g.n[d] - 1) * stride[d]
                                             : stride[d];
        } else {
            st.cm[d] = lo ? (T)0 : -s;
            st.cp[d] = hi ? (T)0 : -s;
            st.om[d] = -stride[d];
            st.op[d] = stride[d];
        }
    }
    return st;
}

template <int DIM>
__device__ __forceinline__ double apply_A(const Stencil<DIM, double>& st,
                                          const double* __restrict__ x,
                                          int64_t c)
{
    double y = st.diag * x[c];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        y = fma(st.cm[d], x[c + st.om[d]], y);
        y = fma(st.cp[d], x[c + st.op[d]], y);
    }
    return y;
}

template <int DIM>
__device__ __forceinline__ float apply_A32(const Stencil<DIM, float>& st,
                                           const float* __restrict__ x,
                                           int64_t c)
{
    float y = st.diag * x[c];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        y = fmaf(st.cm[d], x[c + st.om[d]], y);
        y = fmaf(st.cp[d], x[c + st.op[d]], y);
    }
    return y;
}

__device__ __forceinline__ double dinv_at(const Geo& g, int i, int j, int k)
{
    const int gi[3] = {i, j, k};
    return g.dinv[solid_count(g, gi)];
}

__device__ __forceinline__ float dinv32_at(const Geo& g, int i, int j, int k)
{
    const int gi[3] = {i, j, k};
    return g.dinv32[solid_count(g, gi)];
}

// ---- advection: TimeIntegrator::advect for every entity in one launch ------
// (src/TimeIntegrator.hpp:81-116): each lane owns one (i, j, k) of the union
// of the owned boxes and advects every entity owning that index.
template <int DIM, int FORDER, int VORDER>
__global__ __launch_bounds__(256) void k_advect(Geo g, Fields cur, Fields next)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
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
    return dim3((ex + blk.x - 1) / blk.x, (ey + blk.y - 1) / blk.y, ez);
}

template <int DIM, int FORDER, int VORDER>
static void advect_t(const Geo& g, Fields cur, Fields next, hipStream_t s)
{
    dim3 blk(64, 4, 1);
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
template <int DIM>
__global__ __launch_bounds__(256) void k_inputs(Geo g, Fields cur)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const int64_t c = g.idx(i, j, k);
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
        double* a = cur.e[e];
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
}

void launch_inputs(const Geo& g, Fields cur, hipStream_t s)
{
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
        hipLaunchKernelGGL(k_divergence<2>, dim3(stream_grid(g)), dim3(kBlock),
                           0, s, g, cur, rhs);
    else
        hipLaunchKernelGGL(k_divergence<3>, dim3(stream_grid(g)), dim3(kBlock),
                           0, s, g, cur, rhs);
}

__global__ __launch_bounds__(256) void k_rhs_sum(Geo g, const double* rhs,
                                                  double* out, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        dd_add(hi[0], lo[0], rhs[c], 0.0);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) { out[0] = s[0]; });
}

void launch_rhs_sum(const Geo& g, const double* rhs, double* out_sum,
                    RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rhs_sum, dim3(stream_grid(g)), dim3(kBlock), 0, s, g,
                       rhs, out_sum, rb);
}

__global__ __launch_bounds__(256) void k_rhs_shift(Geo g, double* rhs,
                                                    const double* sum)
{
    const double cnt = (double)g.n[0] * (double)g.n[1] * (double)g.n[2];
    const double mean = sum[0] / cnt;
    for_owned_cells(g, [&](int64_t c, int, int, int) { rhs[c] = rhs[c] - mean; });
}

void launch_rhs_shift(const Geo& g, double* rhs, const double* sum,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_rhs_shift, dim3(stream_grid(g)), dim3(kBlock), 0, s, g,
                       rhs, sum);
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

// ---- FP64 Jacobi-PCG -------------------------------------------------------
// Control flow (this build's statement of Cajita's ReferenceConjugateGradient,
// called at src/VelocityCorrector.hpp:276), from x = 0:
//   r = b; z = D^-1 r; p = z; rho = r.z; bnorm = ||b||  (bnorm 0: done)
//   loop: q = A p; alpha = rho / p.q
//         x = fma(alpha, p, x); r = fma(-alpha, q, r)
//         stop if ||r|| / bnorm < tol (or iter == maxit)
//         beta = r.(D^-1 r) / rho; rho = r.(D^-1 r); p = fma(beta, p, D^-1 r)
__global__ __launch_bounds__(256) void k_pcg_init(
    Geo g, const double* __restrict__ b, double* __restrict__ x,
    double* __restrict__ r, double* __restrict__ p, PcgState* st, double tol,
    int maxit, RedBuf rb)
{
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const double bc = b[c];
        const double z = dinv_at(g, i, j, k) * bc;
        r[c] = bc;
        p[c] = z;
        x[c] = 0.0;
        dd_add_prod(hi[0], lo[0], bc, z);
        dd_add_prod(hi[1], lo[1], bc, bc);
    });
    grid_reduce<2>(hi, lo, rb, [&](double* s) {
        const double bnorm = sqrt(s[1]);
        st->bnorm = bnorm;
        st->tol = tol;
        st->maxit = maxit;
        st->iter = 0;
        st->rho = s[0];
        st->alpha = 0.0;
        st->beta = 0.0;
        if (bnorm == 0.0) {
            st->done = 1;
            st->converged = 1;
            st->res = 0.0;
        } else {
            st->done = 0;
            st->converged = 0;
            st->res = 1.0;
        }
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_pcg_spmv_dot(Geo g,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ q,
                                                       PcgState* st, RedBuf rb)
{
    if (st->done)
        return;
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const Stencil<DIM, double> A = stencil_at<DIM, double>(g, i, j, k);
        const double y = apply_A<DIM>(A, p, c);
        q[c] = y;
        dd_add_prod(hi[0], lo[0], p[c], y);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) { st->alpha = st->rho / s[0]; });
}

__global__ __launch_bounds__(256) void k_pcg_update(
    Geo g, double* __restrict__ x, double* __restrict__ r,
    const double* __restrict__ p, const double* __restrict__ q, PcgState* st,
    RedBuf rb)
{
    if (st->done)
        return;
    const double alpha = st->alpha;
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        x[c] = fma(alpha, p[c], x[c]);
        const double rn = fma(-alpha, q[c], r[c]);
        r[c] = rn;
        const double z = dinv_at(g, i, j, k) * rn;
        dd_add_prod(hi[0], lo[0], rn, rn);
        dd_add_prod(hi[1], lo[1], rn, z);
    });
    grid_reduce<2>(hi, lo, rb, [&](double* s) {
        const int it = st->iter + 1;
        st->iter = it;
        const double res = sqrt(s[0]) / st->bnorm;
        st->res = res;
        if (res < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (it >= st->maxit) {
            st->done = 1;
        } else {
            st->beta = s[1] / st->rho;
            st->rho = s[1];
        }
    });
}

__global__ __launch_bounds__(256) void k_pcg_dir(Geo g,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ p,
                                                  const PcgState* st)
{
    if (st->done)
        return;
    const double beta = st->beta;
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const double z = dinv_at(g, i, j, k) * r[c];
        p[c] = fma(beta, p[c], z);
    });
}

void launch_pcg_init(const Geo& g, const double* b, double* x, double* r,
                     double* p, PcgState* st, double tol, int maxit,
                     RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_init, dim3(stream_grid(g)), dim3(kBlock), 0, s, g,
                       b, x, r, p, st, tol, maxit, rb);
}

void launch_pcg_spmv_dot(const Geo& g, const double* p, double* q,
                         PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_pcg_spmv_dot<2>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, p, q, st, rb);
    else
        hipLaunchKernelGGL(k_pcg_spmv_dot<3>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, p, q, st, rb);
}

void launch_pcg_update(const Geo& g, double* x, double* r, const double* p,
                       const double* q, PcgState* st, RedBuf rb,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_update, dim3(stream_grid(g)), dim3(kBlock), 0, s,
                       g, x, r, p, q, st, rb);
}

void launch_pcg_dir(const Geo& g, const double* r, double* p,
                    const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_dir, dim3(stream_grid(g)), dim3(kBlock), 0, s, g,
                       r, p, st);
}

// ---- FP32-inner / FP64-outer defect correction -----------------------------
//   x = 0; r64 = b; bnorm = ||b||; res = ||r64|| / bnorm
//   pass: r32 = (float) r64; inner FP32 PCG d ~ A^-1 r32 (FP64 dots/scalars);
//         x = x + (double) d; r64 = b - A x; res = ||r64|| / bnorm
__global__ __launch_bounds__(256) void k_refine_init(
    Geo g, const double* __restrict__ b, double* __restrict__ x64,
    float* __restrict__ r32, RefineState* rs, RedBuf rb)
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
        rs->bnorm = bnorm;
        rs->res = bnorm == 0.0 ? 0.0 : sqrt(s[0]) / bnorm;
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
    float* __restrict__ r32, RefineState* rs, RedBuf rb)
{
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const Stencil<DIM, double> A = stencil_at<DIM, double>(g, i, j, k);
        const double y = apply_A<DIM>(A, x64, c);
        const double r = b[c] - y;
        r32[c] = (float)r;
        dd_add_prod(hi[0], lo[0], r, r);
    });
    grid_reduce<1>(hi, lo, rb,
                   [&](double* s) { rs->res = sqrt(s[0]) / rs->bnorm; });
}

void launch_refine_init(const Geo& g, const double* b, double* x64,
                        float* r32, RefineState* rs, RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_refine_init, dim3(stream_grid(g)), dim3(kBlock), 0, s,
                       g, b, x64, r32, rs, rb);
}

void launch_refine_add(const Geo& g, double* x64, const float* d32,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_refine_add, dim3(stream_grid(g)), dim3(kBlock), 0, s,
                       g, x64, d32);
}

void launch_refine_residual(const Geo& g, const double* b, const double* x64,
                            float* r32, RefineState* rs, RedBuf rb,
                            hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_refine_residual<2>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, b, x64, r32, rs, rb);
    else
        hipLaunchKernelGGL(k_refine_residual<3>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, b, x64, r32, rs, rb);
}

// Inner FP32 Jacobi-PCG: FP32 vectors and stencil, FP64 dot products (the
// products of two floats are exact in double) and FP64 alpha/beta, rounded
// to float once for the vector updates.  Stops at ||r|| / ||r0|| < tol.
__global__ __launch_bounds__(256) void k_pcg32_init(
    Geo g, const float* __restrict__ r, float* __restrict__ d,
    float* __restrict__ p, PcgState* st, double tol, int maxit, RedBuf rb)
{
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const float rc = r[c];
        const float z = dinv32_at(g, i, j, k) * rc;
        d[c] = 0.0f;
        p[c] = z;
        dd_add_prod(hi[0], lo[0], (double)rc, (double)z);
        dd_add_prod(hi[1], lo[1], (double)rc, (double)rc);
    });
    grid_reduce<2>(hi, lo, rb, [&](double* s) {
        const double r0 = sqrt(s[1]);
        st->bnorm = r0;
        st->tol = tol;
        st->maxit = maxit;
        st->iter = 0;
        st->rho = s[0];
        st->converged = 0;
        st->res = 1.0;
        st->done = (r0 == 0.0) ? 1 : 0;
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_pcg32_spmv_dot(
    Geo g, const float* __restrict__ p, float* __restrict__ q, PcgState* st,
    RedBuf rb)
{
    if (st->done)
        return;
    double hi[1] = {0.0}, lo[1] = {0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const Stencil<DIM, float> A = stencil_at<DIM, float>(g, i, j, k);
        const float y = apply_A32<DIM>(A, p, c);
        q[c] = y;
        dd_add_prod(hi[0], lo[0], (double)p[c], (double)y);
    });
    grid_reduce<1>(hi, lo, rb, [&](double* s) { st->alpha = st->rho / s[0]; });
}

__global__ __launch_bounds__(256) void k_pcg32_update(
    Geo g, float* __restrict__ d, float* __restrict__ r,
    const float* __restrict__ p, const float* __restrict__ q, PcgState* st,
    RedBuf rb)
{
    if (st->done)
        return;
    const float a32 = (float)st->alpha;
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        d[c] = fmaf(a32, p[c], d[c]);
        const float rn = fmaf(-a32, q[c], r[c]);
        r[c] = rn;
        const float z = dinv32_at(g, i, j, k) * rn;
        dd_add_prod(hi[0], lo[0], (double)rn, (double)rn);
        dd_add_prod(hi[1], lo[1], (double)rn, (double)z);
    });
    grid_reduce<2>(hi, lo, rb, [&](double* s) {
        const int it = st->iter + 1;
        st->iter = it;
        const double res = sqrt(s[0]) / st->bnorm;
        st->res = res;
        if (res < st->tol) {
            st->done = 1;
            st->converged = 1;
        } else if (it >= st->maxit) {
            st->done = 1;
        } else {
            st->beta = s[1] / st->rho;
            st->rho = s[1];
        }
    });
}

__global__ __launch_bounds__(256) void k_pcg32_dir(Geo g,
                                                    const float* __restrict__ r,
                                                    float* __restrict__ p,
                                                    const PcgState* st)
{
    if (st->done)
        return;
    const float b32 = (float)st->beta;
    for_owned_cells(g, [&](int64_t c, int i, int j, int k) {
        const float z = dinv32_at(g, i, j, k) * r[c];
        p[c] = fmaf(b32, p[c], z);
    });
}

void launch_pcg32_init(const Geo& g, const float* r, float* d, float* p,
                       PcgState* st, double tol, int maxit, RedBuf rb,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg32_init, dim3(stream_grid(g)), dim3(kBlock), 0, s,
                       g, r, d, p, st, tol, maxit, rb);
}

void launch_pcg32_spmv_dot(const Geo& g, const float* p, float* q,
                           PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        hipLaunchKernelGGL(k_pcg32_spmv_dot<2>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, p, q, st, rb);
    else
        hipLaunchKernelGGL(k_pcg32_spmv_dot<3>, dim3(stream_grid(g)),
                           dim3(kBlock), 0, s, g, p, q, st, rb);
}

void launch_pcg32_update(const Geo& g, float* d, float* r, const float* p,
                         const float* q, PcgState* st, RedBuf rb,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg32_update, dim3(stream_grid(g)), dim3(kBlock), 0,
                       s, g, d, r, p, q, st, rb);
}

void launch_pcg32_dir(const Geo& g, const float* r, float* p,
                      const PcgState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg32_dir, dim3(stream_grid(g)), dim3(kBlock), 0, s,
                       g, r, p, st);
}

} // namespace cfmi
