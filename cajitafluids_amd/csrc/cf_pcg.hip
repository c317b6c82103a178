// cf_pcg.hip -- matrix-free Jacobi-PCG (FP64, and FP32 inside an FP64
// defect-correction loop) for the pressure solve of
// src/VelocityCorrector.hpp:266-282.
//
// Control flow (this library's statement of Cajita's ReferenceConjugateGradient,
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
// 3-D FP64 update takes x every second iteration (XUpdate: the skipped step
// and the current one are applied together, in order, from the two
// direction buffers), 7.5 words per iteration.  The
// last block of each kernel to finish computes the scalars on the device;
// every kernel returns at once when the device-side `done` flag is set, so
// the host queues iterations in batches and polls once per batch.
#include <hip/hip_runtime.h>

#include "cf_solver.hpp"
#include "cf_stream.hpp"
#include "cf_tile3.hpp"


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
    st->alpha_lag = 0.0;
    const bool zero = norm == 0.0;
    // a NaN or Inf in the right-hand side stops the solve before it starts
    const bool bad = !isfinite(norm);
    st->done = zero || bad ? 1 : (maxit <= 0 ? 1 : 0);
    st->converged = zero ? 1 : 0;
    st->res = zero ? 0.0 : (bad ? norm : 1.0);
    st->mg = 0;
    st->nonfinite = bad ? 1 : 0;
    st->stalled = 0;
}

// Scalars after an update: iteration count, relative residual, the stop
// test, and (Jacobi-PCG) beta and rho.  A residual that is not finite stops
// the solve at once (the stop test alone would read false and run to maxit).
__device__ __forceinline__ void finish_update(PcgState* st, double rr, double rz)
{
    const int it = st->iter + 1;
    st->iter = it;
    const double res = sqrt(rr) / st->bnorm;
    st->res = res;
    if (!isfinite(res)) {
        st->done = 1;
        st->nonfinite = 1;
    } else if (res < st->tol) {
        st->done = 1;
        st->converged = 1;
    } else if (it >= st->maxit) {
        st->done = 1;
    } else {
        if (!st->mg) { // MG-PCG: beta and rho follow the V-cycle
            st->beta = rz / st->rho;
            st->rho = rz;
        }
    }
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

// Relative FP32 reduction that ends a defect-correction pass started from a
// residual of norm rref: inner_tol, or less when half the final target
// (0.5 tol ||b||) is closer, so the last pass stops near the target instead
// of a full inner_tol past it.
__device__ __forceinline__ double pass_tol(double itol, double gtol,
                                           double bnorm, double rref)
{
    return fmax(itol, (0.5 * gtol * bnorm) / rref);
}

__global__ __launch_bounds__(256) void k_pcg_init32(
    Geo g, const float* __restrict__ r, float* __restrict__ d, PcgState* st,
    const RefineState* rs, double itol, double gtol, int maxit, RedBuf rb)
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
    grid_reduce<2>(hi, lo, rb, [&](double* s) {
        const double r0 = sqrt(s[1]);
        init_state(st, s[0], s[1], pass_tol(itol, gtol, rs->bnorm, r0), maxit);
    });
}

// ---- one PCG iteration --------------------------------------------------------
// The two halves of an iteration over this block's owned rows (the row
// partition of for_owned_rows), accumulating the block's double-double
// partials.
template <int DIM, class T>
__device__ __forceinline__ void dir_spmv_rows(const Geo& g, const T* __restrict__ r,
                                              const T* __restrict__ p_old,
                                              T* __restrict__ p_new, bool first, T beta,
                                              double (&hi)[1], double (&lo)[1])
{
    const T s = offdiag<T>(g);
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
}

template <int DIM, class T>
__device__ __forceinline__ void update_rows(const Geo& g, T* __restrict__ x, T* __restrict__ r,
                                            const T* __restrict__ p, T alpha, double (&hi)[2],
                                            double (&lo)[2])
{
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
}

template <int DIM, class T>
__global__ __launch_bounds__(256) void k_pcg_dir_spmv(
    Geo g, const T* __restrict__ r, const T* __restrict__ p_old,
    T* __restrict__ p_new, PcgState* st, RedBuf rb)
{
    if (st->done)
        return;
    double hi[1] = {0.0}, lo[1] = {0.0};
    dir_spmv_rows<DIM, T>(g, r, p_old, p_new, st->iter == 0, (T)st->beta, hi, lo);
    grid_reduce<1>(hi, lo, rb, [&](double* sm) { st->alpha = st->rho / sm[0]; });
}

template <int DIM, class T>
__global__ __launch_bounds__(256) void k_pcg_update(
    Geo g, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
    PcgState* st, RedBuf rb)
{
    if (st->done)
        return;
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    update_rows<DIM, T>(g, x, r, p, (T)st->alpha, hi, lo);
    grid_reduce<2>(hi, lo, rb, [&](double* sm) { finish_update(st, sm[0], sm[1]); });
}

// ---- whole solve in one cooperative launch (2-D grids) --------------------------------
// A 2-D streaming iteration at 128^2-512^2 is two launches whose time goes
// to start-up, one sweep per lane group and the grid reduction's arrival
// chain (about 7.5 us each at 128^2), not to bandwidth.  Here one
// cooperative launch of G workgroups runs all iterations:
//   - workgroup g owns the rows [g R, g R + R) as a strip; lane t of it owns
//     strip cells t + 256 m (m < CPT, 4 or 8, x fastest); r, x and q = A p of
//     those cells stay in registers;
//   - the direction lives in the strip's LDS window: its rows, one halo row
//     above and below, and a one-cell ring in x.  Ring cells of periodic
//     axes are copies of the wrapped cells; the others hold a zero whose
//     sign is that of the coupling s, so that fma(s, ring, y) adds
//     s * ring = +0, the term fma(0, 0, y) the streaming kernels add across a
//     non-periodic boundary (the same bits for every y, -0 included);
//   - a halo row's direction is computed by the strip itself, by the same
//     expression, from the neighbour strip's boundary r and old direction,
//     which that strip publishes with the reduction below;
//   - each of an iteration's two reductions: a workgroup reduction, the
//     workgroup's double-double partial published with agent-scope stores
//     (written through to memory), then an agent-scope store of a tag (the
//     reduction's sequence number) after the partial's stores completed;
//     every workgroup polls all tags, reads all partials and sums them in
//     one fixed order, so every workgroup holds the same alpha, beta and
//     stop decision.  No fence is issued (an agent-scope fence would write
//     back and invalidate the XCD's whole L2).
// A partial or halo slot is rewritten only after its readers have
// published the next reduction, which they do after reading it, so one
// buffer per slot suffices.  Every value is the streaming kernels'
// expression on the same operands, so x, r, the direction and the state
// come out with the same bits.  A poll that exceeds its bound sets
// PcgState::stalled and ends the launch (the host reports CF_EDEVICE).
constexpr int kCoopThreads = 256;
constexpr int kCoopCpt = 8;                         // cells per lane, at most
constexpr int kCoopCells = kCoopThreads * kCoopCpt; // cells per strip, at most
constexpr int kCoopMaxGroups = 256;
constexpr int kCoopWaves = kCoopThreads / kWave;
constexpr unsigned kCoopSpin = 1u << 22; // polls (about 1 us each) before a wait counts as stalled
// unsigned per tag: each strip's tag on its own 64-byte line, so that the
// polls of all workgroups spread over as many lines as there are strips
constexpr int kCoopTagStride = 16;

// Strip layout: R rows per strip, G strips; window row pitch e0 = n0 + 2
// (window index of strip-relative cell (i, jj), halo rows jj = -1 and
// jj = R: (i+1) + e0 (jj+1)).
struct CoopGrid
{
    int G, R, e0, nwin;
    int c00; // padded index of owned cell (0, 0)
    int px;
    int cpt; // cells per lane the kernel is built for: 4 or kCoopCpt
};

// Strips of whole rows, at most cpt cells per lane.  When strips of 4 cells
// per lane leave at most 16 workgroups (64 wave partials, one per lane of
// the gather), those: 8.1 against 9.0 us per iteration at 128^2, where the
// strip's own work is a good part of an iteration.  Otherwise as many rows
// as kCoopCells holds, the fewest workgroups: more strips cost more in the
// waits than they save in work (at 512^2, 256 strips of 2 rows took 14.1
// us per iteration against 13.05 for 128 strips of 4, and 64 strips of 8
// rows at 16 cells per lane 15.7; strips of 2 cells per lane took 3% longer
// at 128^2).
inline CoopGrid coop_grid(const Geo& g)
{
    CoopGrid cg;
    const int r4 = (kCoopThreads * 4) / g.n[0];
    if (r4 > 0 && ((g.n[1] + r4 - 1) / r4) * kCoopWaves <= kWave) {
        cg.cpt = 4;
        cg.R = r4;
    } else {
        cg.cpt = kCoopCpt;
        cg.R = kCoopCells / g.n[0];
    }
    if (cg.R > g.n[1])
        cg.R = g.n[1];
    cg.G = (g.n[1] + cg.R - 1) / cg.R;
    cg.e0 = g.n[0] + 2;
    cg.nwin = cg.e0 * (cg.R + 2);
    cg.c00 = (int)g.idx(g.h, g.h, g.own_lo(2));
    cg.px = (int)g.px;
    return cg;
}

// The exchange buffer: tags (B: p.q, C: r.r and r.z) and partials per
// strip, and each strip's boundary rows (r, then the old direction; top
// row, then bottom row).
struct CoopXch
{
    unsigned* tag;   // [2][G][kCoopTagStride]
    double* part;    // [G][kCoopWaves][8]: B hi, lo; C hi0, lo0, hi1, lo1
    double* halo;    // [G][2 rows][2 arrays][n0]
};

// An iteration's reduction in three steps:
//   coop_publish: each wave's double-double partial (wave butterfly) to its
//     slot with agent-scope stores (WP), or the workgroup's (thread 0 over
//     the wave sums, to slot strip); every lane's stores
//     of this step complete (the halo rows too) before thread 0 stores the
//     strip's tag;
//   coop_wait: wave 0 polls every strip's tag; a barrier tells the others;
//   coop_gather: wave 0 reads all wave partials and sums them in one fixed
//     order (the same in every workgroup), while the other waves may load
//     the halo rows.
// A lane of wave 0 polls strips lane + 64 u (u < kCoopPerLane) and reads
// partials lane + 64 u (u < PPL, the kernel's bucket for the partial
// count); the loads of a step are issued together.  Partial slots, one
// 64-byte line each: part[slot 8 + poff + 2 v + {0, 1}], slot = strip
// kCoopWaves + wave (WP) or strip.
constexpr int kCoopPerLane = kCoopMaxGroups / kWave;

template <int NV, bool WP>
__device__ __forceinline__ void coop_publish(double (&hi)[NV], double (&lo)[NV], double* part,
                                             int poff, unsigned* tag, unsigned seq,
                                             double (*red)[2][kCoopWaves])
{
#pragma unroll
    for (int v = 0; v < NV; ++v)
        dd_norm(hi[v], lo[v]);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double oh = __shfl_xor(hi[v], off, kWave);
            const double ol = __shfl_xor(lo[v], off, kWave);
            dd_add(hi[v], lo[v], oh, ol);
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    if (lane == 0) {
        if constexpr (WP) {
            double* mine = part + ((size_t)blockIdx.x * kCoopWaves + wv) * 8 + poff;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                __hip_atomic_store(&mine[2 * v], hi[v], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&mine[2 * v + 1], lo[v], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                red[v][0][wv] = hi[v];
                red[v][1][wv] = lo[v];
            }
        }
    }
    // this lane's stores complete before the tag (the waitcnt orders the
    // hardware; the signal fences keep the compiler from reordering)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        if constexpr (!WP) {
            // the workgroup's partial: slot blockIdx.x
            double* mine = part + (size_t)blockIdx.x * 8 + poff;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double h = red[v][0][0], l = red[v][1][0];
#pragma unroll
                for (int w = 1; w < kCoopWaves; ++w)
                    dd_add(h, l, red[v][0][w], red[v][1][w]);
                __hip_atomic_store(&mine[2 * v], h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&mine[2 * v + 1], l, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        __hip_atomic_store(&tag[(size_t)blockIdx.x * kCoopTagStride], seq, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ void coop_wait(const unsigned* tag, unsigned seq, int G, int* stall)
{
    if (threadIdx.x < kWave) {
        unsigned polls = 0;
        for (;;) {
            unsigned got[kCoopPerLane];
#pragma unroll
            for (int u = 0; u < kCoopPerLane; ++u) {
                const int q = threadIdx.x + kWave * u;
                got[u] = q < G ? __hip_atomic_load(&tag[(size_t)q * kCoopTagStride],
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : seq;
            }
            bool all = true;
#pragma unroll
            for (int u = 0; u < kCoopPerLane; ++u)
                all = all && got[u] >= seq;
            if (__all(all))
                break;
            if (++polls >= kCoopSpin) {
                if (threadIdx.x == 0)
                    *stall = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    // the loads that follow stay after the poll (compiler-only barrier)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
}

// in wave 0: the rounded totals in every lane
template <int NV, int PPL, bool WP>
__device__ __forceinline__ void coop_gather(const double* part, int poff, int G, double (&out)[NV])
{
    // wave partials (slot strip * kCoopWaves + wave), or one partial per
    // strip (slot strip)
    const int nparts = WP ? G * kCoopWaves : G;
    double ph[PPL][NV], pl[PPL][NV];
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        const int q = threadIdx.x + kWave * u;
        const double* pq = part + (size_t)(q < nparts ? q : 0) * 8 + poff;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            ph[u][v] = q < nparts ? __hip_atomic_load(&pq[2 * v], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0.0;
            pl[u][v] = q < nparts ? __hip_atomic_load(&pq[2 * v + 1], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0.0;
        }
    }
    double h2[NV], l2[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        h2[v] = ph[0][v];
        l2[v] = pl[0][v];
#pragma unroll
        for (int u = 1; u < PPL; ++u)
            dd_add(h2[v], l2[v], ph[u][v], pl[u][v]);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double oh = __shfl_xor(h2[v], off, kWave);
            const double ol = __shfl_xor(l2[v], off, kWave);
            dd_add(h2[v], l2[v], oh, ol);
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
        out[v] = h2[v] + l2[v];
}

// A lane's cells: slot word = window index (bits 0-15), SOLID wall count
// (bits 16-18), kCoopOwned (bit 20), and whether the cell is in the strip's
// top / bottom row (bits 21, 22); a slot past the strip's last cell carries
// the first cell's window index without kCoopOwned, and its stores and sums
// are masked.
constexpr int kCoopOwned = 1 << 20;
constexpr int kCoopTop = 1 << 21;
constexpr int kCoopBottom = 1 << 22;

template <int PPL, bool WP, int CPT>
__global__ __launch_bounds__(kCoopThreads) void k_pcg_solve_coop(
    Geo g, CoopGrid cg, double* __restrict__ x, double* __restrict__ r,
    double* __restrict__ pg, PcgState* st, CoopXch xc, int max_its)
{
    extern __shared__ double lds[];
    double* pw = lds;             // the direction window
    double* tab = lds + cg.nwin;  // diag[0..6], then dinv[0..6]
    double* hr = tab + 14;        // halo rows' r   [2][n0]
    double* hp = hr + 2 * g.n[0]; // halo rows' old direction [2][n0]
    __shared__ double red[2][2][kCoopWaves];
    __shared__ double sh_ab[2]; // alpha, beta
    __shared__ int sh_stop, sh_stall;
    // The state is read once, here; every workgroup keeps the same copy in
    // LDS (only thread 0 writes it), and workgroup 0 stores it back at the
    // end.  A read of *st after a write in this kernel could be served by
    // the scalar cache, which vector stores do not update.
    __shared__ PcgState ls;
    const int t = threadIdx.x;
    const int gid = blockIdx.x, G = cg.G;
    const int n0 = g.n[0], n1 = g.n[1], e0 = cg.e0;
    const int j0 = gid * cg.R;
    const int nrows = (j0 + cg.R < n1 ? cg.R : n1 - j0);
    const double s = offdiag<double>(g);
    const double ring0 = copysign(0.0, s);
    if (t == 0) {
        ls = *st;
        sh_stall = 0;
    }
    for (int e = t; e < cg.nwin; e += kCoopThreads)
        pw[e] = ring0;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if (t == q) {
            tab[q] = g.diag[q];
            tab[7 + q] = g.dinv[q];
        }
    __syncthreads();
    if (ls.done)
        return;
    bool first = ls.iter == 0;
    double beta = ls.beta;

    // halo rows: window rows 0 (global row j0 - 1) and nrows + 1 (row j0 +
    // nrows), when they exist; `src` the strip that owns them
    const bool py = g.periodic[1];
    const bool has_top = j0 > 0 || py, has_bot = j0 + nrows < n1 || py;
    const int jtop = j0 > 0 ? j0 - 1 : n1 - 1, jbot = j0 + nrows < n1 ? j0 + nrows : 0;
    const int gtop = gid > 0 ? gid - 1 : G - 1, gbot = gid + 1 < G ? gid + 1 : 0;
    // SOLID wall count of a cell of global row j (in 2-D, z adds its own
    // count for k = 0 as the streaming kernels do)
    auto cell_cnt = [&](int i, int j) -> int { return row_stencil(g, j, 0).cnt + lane_x(g, i).cnt; };
    // the two halo rows' r and old direction into LDS from the padded
    // arrays (start of a launch)
    if (!first)
        for (int e = t; e < 2 * n0; e += kCoopThreads) {
            const int side = e / n0, i = e % n0;
            if (side == 0 ? has_top : has_bot) {
                const int64_t c = cg.c00 + i + (int64_t)cg.px * (side == 0 ? jtop : jbot);
                hr[e] = r[c];
                hp[e] = pg[c];
            }
        }
    else
        for (int e = t; e < 2 * n0; e += kCoopThreads) {
            const int side = e / n0, i = e % n0;
            if (side == 0 ? has_top : has_bot)
                hr[e] = r[cg.c00 + i + (int64_t)cg.px * (side == 0 ? jtop : jbot)];
        }

    int info[CPT];
    double rv[CPT], qv[CPT], xv[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int L = t + kCoopThreads * m;
        const int i = L % n0, jj = L / n0;
        info[m] = 1 + e0; // window index of strip cell (0, 0)
        rv[m] = 0.0;
        qv[m] = 0.0;
        xv[m] = 0.0;
        if (jj < nrows) {
            const int li = (i + 1) + e0 * (jj + 1);
            info[m] = li | (cell_cnt(i, j0 + jj) << 16) | kCoopOwned |
                      (jj == 0 ? kCoopTop : 0) | (jj == nrows - 1 ? kCoopBottom : 0);
            const int64_t c = cg.c00 + i + (int64_t)cg.px * (j0 + jj);
            rv[m] = r[c];
            xv[m] = x[c];
            if (!first)
                pw[li] = pg[c];
        }
    }
    // padded index of a slot's cell
    auto padded = [&](int li) -> int64_t {
        const int jj = li / e0 - 1, i = li - e0 * (jj + 1) - 1;
        return cg.c00 + i + (int64_t)cg.px * (j0 + jj);
    };
    unsigned seq = 0;
    for (int its = 1;; ++its) {
        // direction: p = D^-1 r, or fma(beta, p, D^-1 r), for the strip's
        // cells (staged in qv, free until the stencil) and its halo rows
        if (first) {
#pragma unroll
            for (int m = 0; m < CPT; ++m)
                qv[m] = tab[7 + ((info[m] >> 16) & 7)] * rv[m];
        } else {
#pragma unroll
            for (int m = 0; m < CPT; ++m)
                qv[m] = fma(beta, pw[info[m] & 0xffff], tab[7 + ((info[m] >> 16) & 7)] * rv[m]);
        }
#pragma unroll
        for (int m = 0; m < CPT; ++m)
            if (info[m] & kCoopOwned)
                pw[info[m] & 0xffff] = qv[m];
        for (int e = t; e < 2 * n0; e += kCoopThreads) {
            const int side = e / n0, i = e % n0;
            if (side == 0 ? has_top : has_bot) {
                const double z = tab[7 + cell_cnt(i, side == 0 ? jtop : jbot)] * hr[e];
                pw[(i + 1) + (side == 0 ? 0 : e0 * (nrows + 1))] =
                    first ? z : fma(beta, hp[e], z);
            }
        }
        __syncthreads();
        if (g.periodic[0]) {
            for (int jw = t; jw < nrows + 2; jw += kCoopThreads) {
                double* row = pw + e0 * jw;
                row[0] = row[n0];
                row[n0 + 1] = row[1];
            }
            __syncthreads();
        }
        // q = A p in the stencil order {0}, -x, +x, -y, +y; p.q
        double h1[1] = {0.0}, l1[1] = {0.0};
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int im = info[m];
            const double* pp = pw + (im & 0xffff);
            const double pc = pp[0];
            double y = tab[(im >> 16) & 7] * pc;
            y = fma(s, pp[-1], y);
            y = fma(s, pp[1], y);
            y = fma(s, pp[-e0], y);
            y = fma(s, pp[e0], y);
            const bool own = im & kCoopOwned;
            qv[m] = own ? y : 0.0;
            dd_acc_prod(h1[0], l1[0], own ? pc : 0.0, qv[m]);
        }
        coop_publish<1, WP>(h1, l1, xc.part, 0, xc.tag, ++seq, red);
        coop_wait(xc.tag, seq, G, &sh_stall);
        if (sh_stall)
            break;
        if (t < kWave) {
            double pq[1];
            coop_gather<1, PPL, WP>(xc.part, 0, G, pq);
            if (t == 0) {
                const double a = ls.rho / pq[0];
                ls.alpha = a;
                sh_ab[0] = a;
            }
        }
        __syncthreads();
        const double alpha = sh_ab[0];
        // x = fma(alpha, p, x); r = fma(-alpha, q, r); r.r and r.(D^-1 r).
        // Past the strip's last cell q = r = 0, so rn = 0 and the sums are
        // unchanged.
        double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int im = info[m];
            xv[m] = fma(alpha, pw[im & 0xffff], xv[m]);
            const double rn = fma(-alpha, qv[m], rv[m]);
            rv[m] = rn;
            const double z = tab[7 + ((im >> 16) & 7)] * rn;
            dd_acc_prod(hi[0], lo[0], rn, rn);
            dd_acc_prod(hi[1], lo[1], rn, z);
            // boundary rows for the neighbour strips: the new r and this
            // iteration's direction
#pragma unroll
            for (int side = 0; side < 2; ++side)
                if (im & (side == 0 ? kCoopTop : kCoopBottom)) {
                    const int li = im & 0xffff;
                    const int i = li - e0 * (side == 0 ? 1 : nrows) - 1;
                    double* hb = xc.halo + ((size_t)gid * 2 + side) * 2 * n0;
                    __hip_atomic_store(&hb[i], rn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&hb[n0 + i], pw[li], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
        }
        coop_publish<2, WP>(hi, lo, xc.part, 2, xc.tag + (size_t)G * kCoopTagStride, ++seq, red);
        coop_wait(xc.tag + (size_t)G * kCoopTagStride, seq, G, &sh_stall);
        if (sh_stall)
            break;
        if (t < kWave) {
            double rr[2];
            coop_gather<2, PPL, WP>(xc.part, 2, G, rr);
            if (t == 0) {
                finish_update(&ls, rr[0], rr[1]);
                sh_stop = ls.done || its >= max_its;
                sh_ab[1] = ls.beta;
            }
        } else {
            // the other waves: the neighbours' boundary rows (published
            // before their tags)
            for (int e = t - kWave; e < 2 * n0; e += kCoopThreads - kWave) {
                const int side = e / n0, i = e % n0;
                if (side == 0 ? has_top : has_bot) {
                    // the top halo row is the bottom row of the strip above
                    const double* hb = xc.halo + ((size_t)(side == 0 ? gtop : gbot) * 2 +
                                                  (side == 0 ? 1 : 0)) * 2 * n0;
                    hr[e] = __hip_atomic_load(&hb[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    hp[e] = __hip_atomic_load(&hb[n0 + i], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        __syncthreads();
        if (sh_stop)
            break;
        beta = sh_ab[1];
        first = false;
    }
    // x, r and the direction back to memory (a later launch resumes from
    // them), and the state from workgroup 0
    if (sh_stall && t == 0)
        ls.stalled = 1;
    if (gid == 0 && t == 0)
        *st = ls;
#pragma unroll
    for (int m = 0; m < CPT; ++m)
        if (info[m] & kCoopOwned) {
            const int li = info[m] & 0xffff;
            const int64_t c = padded(li);
            x[c] = xv[m];
            r[c] = rv[m];
            pg[c] = pw[li];
        }
}

// ---- 3-D kernels (cf_tile3.hpp) ------------------------------------------------
// Products of two floats are exact in double, so the FP32 kernels skip the
// TwoProd error term.
template <class T>
constexpr bool exact_products()
{
    return std::is_same<T, float>::value;
}

// Cells per lane along x of the 3-D kernels: 2 for FP32 (8-byte accesses);
// CF_CPL64 for FP64.
#ifndef CF_CPL64
#define CF_CPL64 1
#endif
template <class T>
constexpr int cpl_of()
{
    return std::is_same<T, float>::value ? 2 : CF_CPL64;
}

template <class T, int CPL>
struct DirEpiT
{
    const Geo& g;
    T* __restrict__ p_new;
    double hi[1] = {0.0}, lo[1] = {0.0};
    template <int B>
    __device__ void prefetch(const Tile3Ctx<CPL>&, const Plane3Meta&, int)
    {
    }
    __device__ void shift() {}
    __device__ void rotate() {}
    template <int B>
    __device__ void cells(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int q,
                          const T (&pc)[CPL], const T (&y)[CPL])
    {
        vstore_valid<T, CPL>(p_new + c.colb[q] + m.off, pc, c.valid[q]);
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            if (c.valid[q][j])
                dd_acc_prod<exact_products<T>()>(hi[0], lo[0], (double)pc[j], (double)y[j]);
    }
};

// The update's epilogue; XM is its x mode (XUpdate).
//
// X_EACH and X_SKIP: x and r of the next plane load into the second of two
// buffers (prefetch) while the march works on the current one.  X_PAIR
// reads one more vector; it loads x, r and the previous direction of the
// next plane after each step's barrier instead (kLate), into one buffer,
// which keeps it at 4 waves per SIMD in FP32 (double-buffered it would take
// 143 VGPRs, 3 waves).
//
// R32: the update also writes (float) r, the FP32 right-hand side of a
// multigrid V-cycle whose finest level runs in FP32 (the V-cycle then needs
// no cast of its own).
template <class T, int CPL, int XM, bool R32 = false>
struct UpdEpiT
{
    const Geo& g;
    T* __restrict__ x;
    T* __restrict__ r;
    T alpha;
    const T* __restrict__ p_lag; // X_PAIR: the previous direction
    T alpha_lag;
    float* __restrict__ r32;     // R32: (float) r
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    static constexpr bool kReadX = XM != X_SKIP;
    static constexpr bool kLate = XM == X_PAIR;
    // x and r of a plane, in two buffers that alternate with the march's
    // steps (X_EACH, X_SKIP)
    T xb[2][tile3::RY][CPL], rb[2][tile3::RY][CPL];
    // kLate: x, r and p_lag of the current plane, and where the next one is
    T xs[tile3::RY][CPL], rs[tile3::RY][CPL], ls[tile3::RY][CPL];
    int next_k = 0;
    int64_t next_off = 0;

    __device__ void late(const Tile3Ctx<CPL>& c)
    {
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q) {
            // a group with a valid first cell lies inside its padded row
            const bool ok = next_k >= c.k0 && next_k < c.k1 && c.valid[q][0];
            const int64_t at = c.colb[q] + next_off;
            if constexpr (CPL == 1) {
                xs[q][0] = load_any<T>(x, ok, at, c.safe);
                ls[q][0] = load_any<T>(p_lag, ok, at, c.safe);
                rs[q][0] = load_any<T>(r, ok, at, c.safe);
                continue;
            }
            const int64_t ap = ok ? at : c.safe;
            vload<T, CPL>(x + ap, xs[q]);
            vload<T, CPL>(p_lag + ap, ls[q]);
            vload<T, CPL>(r + ap, rs[q]);
        }
    }
    template <int B>
    __device__ void prefetch(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int k)
    {
        if constexpr (kLate) { // late() loads it
            next_k = k;
            next_off = m.off;
            return;
        }
#pragma unroll
        for (int q = 0; q < tile3::RY; ++q) {
            // a group with a valid first cell lies inside its padded row
            const bool ok = k >= c.k0 && k < c.k1 && c.valid[q][0];
            const int64_t at = c.colb[q] + m.off;
            if constexpr (CPL == 1) {
                // raw: cells() reads only valid rows of planes below k1
                if constexpr (kReadX)
                    xb[B][q][0] = load_any<T>(x, ok, at, c.safe);
                rb[B][q][0] = load_any<T>(r, ok, at, c.safe);
                continue;
            }
            // an aligned pair on every path (raw, as above)
            const int64_t ap = ok ? at : c.safe;
            if constexpr (kReadX)
                vload<T, CPL>(x + ap, xb[B][q]);
            vload<T, CPL>(r + ap, rb[B][q]);
        }
    }
    __device__ void shift()
    {
        if constexpr (!kLate) {
#pragma unroll
            for (int q = 0; q < tile3::RY; ++q)
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    if constexpr (kReadX)
                        xb[0][q][j] = xb[1][q][j];
                    rb[0][q][j] = rb[1][q][j];
                }
        }
    }
    __device__ void rotate() {}
    template <int B>
    __device__ void cells(const Tile3Ctx<CPL>& c, const Plane3Meta& m, int q,
                          const T (&pc)[CPL], const T (&qv)[CPL])
    {
        const int64_t at = c.colb[q] + m.off;
        T ro[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j)
            ro[j] = fma_t<T>(-alpha, qv[j], kLate ? rs[q][j] : rb[B][q][j]);
        if constexpr (kReadX) {
            T xo[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                // X_PAIR: the skipped step first, then this one
                const T xl = kLate ? fma_t<T>(alpha_lag, ls[q][j], xs[q][j]) : xb[B][q][j];
                xo[j] = fma_t<T>(alpha, pc[j], xl);
            }
            vstore_valid<T, CPL>(x + at, xo, c.valid[q]);
        }
        vstore_valid<T, CPL>(r + at, ro, c.valid[q]);
        if constexpr (R32) {
            float rf[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                rf[j] = (float)ro[j];
            vstore_valid<float, CPL>(r32 + at, rf, c.valid[q]);
        }
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            if (!c.valid[q][j])
                continue;
            const T z = ctx_dinv<T, 0>(g, c, c.wy[q] + m.wz, c.wx[j]) * ro[j];
            dd_acc_prod<exact_products<T>()>(hi[0], lo[0], (double)ro[j], (double)ro[j]);
            dd_acc_prod<exact_products<T>()>(hi[1], lo[1], (double)ro[j], (double)z);
        }
    }
};

// TR: the element type of r (float: an FP32 multigrid V-cycle's z, widened
// on load; its values are the FP64 z's exactly).
template <class T, int CPL, bool PX, class TR = T>
__global__ __launch_bounds__(kTileBlock) void k_pcg_dir_spmv3(
    Geo g, Tile3Grid tg, const TR* __restrict__ r, const T* __restrict__ p_old,
    T* __restrict__ p_new, PcgState* st, RedBuf rb)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ __align__(16) T pl[2][TY + 2][prow<CPL>()];
    const bool first = st->iter == 0;
    const T beta = (T)st->beta;
    DirEpiT<T, CPL> epi{g, p_new};
    tile3_march<T, true, CPL, PX, false, TR>(g, tg, r, p_old, first, beta, pl, epi);
    dd_norm(epi.hi[0], epi.lo[0]);
    grid_reduce<1, kTileBlock>(epi.hi, epi.lo, rb,
                   [&](double* sm) { st->alpha = st->rho / sm[0]; });
}

template <class T, int CPL, bool PX, int XM, bool R32>
__global__ __launch_bounds__(kTileBlock) void k_pcg_update3(
    Geo g, Tile3Grid tg, T* __restrict__ x, T* __restrict__ r,
    const T* __restrict__ p, const T* __restrict__ p_lag, float* __restrict__ r32,
    PcgState* st, RedBuf rb)
{
    using namespace tile3;
    if (st->done)
        return;
    __shared__ __align__(16) T pl[2][TY + 2][prow<CPL>()];
    UpdEpiT<T, CPL, XM, R32> epi{g, x, r, (T)st->alpha, p_lag, (T)st->alpha_lag, r32};
    // downwards: the direction kernel marches up, so each kernel starts on
    // the planes the other touched last (Infinity Cache hits; FP64 256^3:
    // update 127 -> 121 us, direction 101 -> 95 us)
    tile3_march<T, false, CPL, PX, true>(g, tg, p, nullptr, false, (T)0, pl, epi);
    dd_norm(epi.hi[0], epi.lo[0]);
    dd_norm(epi.hi[1], epi.lo[1]);
    grid_reduce<2, kTileBlock>(epi.hi, epi.lo, rb, [&](double* sm) {
        if (XM == X_SKIP) // every block has read the state by now
            st->alpha_lag = st->alpha;
        finish_update(st, sm[0], sm[1]);
    });
}

// x = fma(alpha, p, x): the x step an X_SKIP update left out, when it was
// the last of a run (alpha is still that iteration's).
template <class T>
__global__ __launch_bounds__(256) void k_pcg_xfix(Geo g, T* __restrict__ x,
                                                  const T* __restrict__ p,
                                                  const PcgState* st)
{
    const T a = (T)st->alpha;
    for_owned_cells(g, [&](int64_t c, int, int, int) { x[c] = fma_t<T>(a, p[c], x[c]); });
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
                                                     float* __restrict__ d)
{
    for_owned_cells(g, [&](int64_t c, int, int, int) {
        x64[c] = x64[c] + (double)d[c];
        d[c] = 0.0f;
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

// ---- defect correction with the search direction carried across passes -----
//   x = 0; bnorm = ||b||; r32 = (float) b; d = 0; rref = ||r32||;
//   rho = r32.(D32^-1 r32); then one FP32 PCG run (the iteration kernels
//   above, which stop when ||r|| / rref < pass_tol).  At each pass boundary:
//   x = x + (double) d; d = 0; r64 = b - A x; res = ||r64|| / bnorm;
//   r32 = (float) r64; rref = ||r32||; beta = r32.(D32^-1 r32) / rho;
//   rho = r32.(D32^-1 r32); the PCG resumes with the same direction.
__global__ __launch_bounds__(256) void k_refine_init_carry(
    Geo g, const double* __restrict__ b, double* __restrict__ x64,
    float* __restrict__ r32, float* __restrict__ d32, RefineState* rst,
    PcgState* st, double itol, double gtol, int maxit, RedBuf rb)
{
    double hi[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const int ryz = wall_count(g, 1, j) + wall_count(g, 2, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const double bc = b[c];
            const float rc = (float)bc;
            const float z = dinv_pick<float>(g, ryz, wall_count(g, 0, i)) * rc;
            x64[c] = 0.0;
            r32[c] = rc;
            d32[c] = 0.0f;
            dd_add_prod(hi[0], lo[0], bc, bc);
            dd_add_prod(hi[1], lo[1], (double)rc, (double)rc);
            dd_add_prod(hi[2], lo[2], (double)rc, (double)z);
        }
    });
    grid_reduce<3>(hi, lo, rb, [&](double* s) {
        const double bnorm = sqrt(s[0]);
        const double rref = sqrt(s[1]);
        rst->bnorm = bnorm;
        rst->res = 1.0;
        st->bnorm = rref;
        st->tol = pass_tol(itol, gtol, bnorm, rref);
        st->maxit = maxit;
        st->iter = 0;
        st->rho = s[2];
        st->alpha = 0.0;
        st->beta = 0.0;
        st->alpha_lag = 0.0;
        st->mg = 0;
        // a non-finite ||b|| is reported by the host from rst->bnorm
        st->nonfinite = isfinite(bnorm) ? 0 : 1;
        st->done = bnorm == 0.0 || st->nonfinite ? 1 : (maxit <= 0 ? 1 : 0);
        st->converged = bnorm == 0.0 ? 1 : 0;
        if (bnorm == 0.0)
            rst->res = 0.0;
    });
}

template <int DIM>
__global__ __launch_bounds__(256) void k_refine_residual_carry(
    Geo g, const double* __restrict__ b, const double* __restrict__ x64,
    float* __restrict__ r32, RefineState* rst, PcgState* st, double itol,
    double gtol, RedBuf rb)
{
    double hi[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        const RowStencil rs = row_stencil(g, j, k);
        for (int i = threadIdx.x; i < g.n[0]; i += kBlock) {
            const int64_t c = base + i;
            const LaneX lx = lane_x(g, i);
            const double y = apply_A<DIM, double>(g, rs, lx, x64, c);
            const double r = b[c] - y;
            const float rc = (float)r;
            const float z = dinv_pick<float>(g, rs.cnt, lx.cnt) * rc;
            r32[c] = rc;
            dd_add_prod(hi[0], lo[0], r, r);
            dd_add_prod(hi[1], lo[1], (double)rc, (double)rc);
            dd_add_prod(hi[2], lo[2], (double)rc, (double)z);
        }
    });
    grid_reduce<3>(hi, lo, rb, [&](double* s) {
        rst->res = sqrt(s[0]) / rst->bnorm;
        const double rref = sqrt(s[1]);
        st->bnorm = rref;
        st->tol = pass_tol(itol, gtol, rst->bnorm, rref);
        st->beta = s[2] / st->rho;
        st->rho = s[2];
        st->done = 0;
        st->converged = 0;
    });
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
    hipLaunchKernelGGL(kernel, dim3(stream_blocks(g)), dim3(kBlock), 0, s,    \
                       __VA_ARGS__)

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
                       const RefineState* rs, double itol, double gtol,
                       int maxit, RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_pcg_init32, g, r, d, st, rs, itol, gtol, maxit, rb);
}

void launch_refine_init_carry(const Geo& g, const double* b, double* x64,
                              float* r32, float* d32, RefineState* rs,
                              PcgState* st, double itol, double gtol, int maxit,
                              RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_refine_init_carry, g, b, x64, r32, d32, rs, st, itol,
                     gtol, maxit, rb);
}

void launch_refine_residual_carry(const Geo& g, const double* b,
                                  const double* x64, float* r32,
                                  RefineState* rs, PcgState* st, double itol,
                                  double gtol, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2)
        CF_STREAM_LAUNCH(k_refine_residual_carry<2>, g, b, x64, r32, rs, st,
                         itol, gtol, rb);
    else
        CF_STREAM_LAUNCH(k_refine_residual_carry<3>, g, b, x64, r32, rs, st,
                         itol, gtol, rb);
}

// The two-cell (FP32) tile kernels come in a periodic-x instantiation,
// which also reads the wrapped first cell of an odd row, and a plain one.
template <class T, bool PX>
static void dir_spmv3_t(const Geo& g, const T* r, const T* p_old, T* p_new, PcgState* st,
                        RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL((k_pcg_dir_spmv3<T, cpl_of<T>(), PX>),
                       dim3(tile3_blocks<cpl_of<T>()>(g)), dim3(kTileBlock), 0, s, g,
                       tile3_grid<cpl_of<T>()>(g), r, p_old, p_new, st, rb);
}

template <class T, bool PX, int XM, bool R32>
static void update3_t(const Geo& g, T* x, T* r, const T* p, const T* p_lag, float* r32,
                      PcgState* st, RedBuf rb, hipStream_t s)
{
    hipLaunchKernelGGL((k_pcg_update3<T, cpl_of<T>(), PX, XM, R32>),
                       dim3(tile3_blocks<cpl_of<T>()>(g)), dim3(kTileBlock), 0, s, g,
                       tile3_grid<cpl_of<T>()>(g), x, r, p, p_lag, r32, st, rb);
}

template <class T, bool PX, bool R32>
static void update3_x(const Geo& g, T* x, T* r, const T* p, const T* p_lag, float* r32,
                      PcgState* st, RedBuf rb, hipStream_t s, int xmode)
{
    if (xmode == X_EACH) {
        update3_t<T, PX, X_EACH, R32>(g, x, r, p, p_lag, r32, st, rb, s);
        return;
    }
    if (xmode == X_SKIP) {
        update3_t<T, PX, X_SKIP, R32>(g, x, r, p, p_lag, r32, st, rb, s);
        return;
    }
    if (xmode == X_PAIR) {
        if (!p_lag || p_lag == p)
            throw CfError(CF_EINVAL, "X_PAIR update needs the previous direction buffer");
        update3_t<T, PX, X_PAIR, R32>(g, x, r, p, p_lag, r32, st, rb, s);
        return;
    }
    throw CfError(CF_EINVAL, "x mode not available for this PCG update");
}

void launch_pcg_dir_spmv_z32(const Geo& g, const float* z, const double* p_old, double* p_new,
                             PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim != 3)
        throw CfError(CF_EINVAL, "an FP32 z is a 3-D multigrid V-cycle's");
    hipLaunchKernelGGL((k_pcg_dir_spmv3<double, 1, false, float>), dim3(tile3_blocks<1>(g)),
                       dim3(kTileBlock), 0, s, g, tile3_grid<1>(g), z, p_old, p_new, st, rb);
}

template <class T>
void launch_pcg_dir_spmv(const Geo& g, const T* r, const T* p_old, T* p_new,
                         PcgState* st, RedBuf rb, hipStream_t s)
{
    if (g.dim == 2) {
        CF_STREAM_LAUNCH((k_pcg_dir_spmv<2, T>), g, r, p_old, p_new, st, rb);
        return;
    }
    if constexpr (cpl_of<T>() > 1)
        if (g.periodic[0]) {
            dir_spmv3_t<T, true>(g, r, p_old, p_new, st, rb, s);
            return;
        }
    dir_spmv3_t<T, false>(g, r, p_old, p_new, st, rb, s);
}

template <class T>
void launch_pcg_update(const Geo& g, T* x, T* r, const T* p, PcgState* st,
                       RedBuf rb, hipStream_t s, int xmode, const T* p_lag, float* r32)
{
    if (g.dim == 2) {
        if (xmode != X_EACH || r32)
            throw CfError(CF_EINVAL, "the 2-D PCG update takes x every iteration and "
                                     "writes no FP32 copy of r");
        CF_STREAM_LAUNCH((k_pcg_update<2, T>), g, x, r, p, st, rb);
        return;
    }
    if constexpr (cpl_of<T>() > 1) {
        if (r32)
            throw CfError(CF_EINVAL, "the FP32 PCG update writes no FP32 copy of r");
        if (g.periodic[0]) {
            update3_x<T, true, false>(g, x, r, p, p_lag, nullptr, st, rb, s, xmode);
            return;
        }
        update3_x<T, false, false>(g, x, r, p, p_lag, nullptr, st, rb, s, xmode);
    } else {
        if (r32)
            update3_x<T, false, true>(g, x, r, p, p_lag, r32, st, rb, s, xmode);
        else
            update3_x<T, false, false>(g, x, r, p, p_lag, nullptr, st, rb, s, xmode);
    }
}

template <class T>
void launch_pcg_xfix(const Geo& g, T* x, const T* p, const PcgState* st, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_pcg_xfix<T>, g, x, p, st);
}

int reduction_blocks_max(const Geo& g)
{
    const int t = g.dim == 3 ? tile3_blocks<1>(g) : 0; // the largest tile grid
    return t > kRedBlocks ? t : kRedBlocks;
}

bool pcg_coop_fits(const Geo& g)
{
    if (g.dim != 2 || g.n[0] > kCoopCells)
        return false;
    const CoopGrid cg = coop_grid(g);
    if (cg.G > kCoopMaxGroups)
        return false;
    // the window, the coefficient tables and the halo rows fit the LDS;
    // window indices fit the slot word's 16 bits
    const int64_t lds = ((int64_t)cg.nwin + 14 + 4 * g.n[0]) * (int64_t)sizeof(double);
    return cg.nwin < (1 << 16) && lds <= 64 * 1024;
}

size_t coop_xch_bytes(const Geo& g)
{
    const CoopGrid cg = coop_grid(g);
    return (size_t)cg.G * (8 * kCoopWaves + 4 * (size_t)g.n[0]) * sizeof(double) +
           2 * (size_t)cg.G * kCoopTagStride * sizeof(unsigned);
}

bool launch_pcg_solve_coop(const Geo& g, double* x, double* r, double* p, PcgState* st,
                           void* xch, int max_its, hipStream_t s)
{
    if (!pcg_coop_fits(g))
        throw CfError(CF_EINVAL, "grid does not fit the cooperative PCG");
    if (max_its < 1)
        throw CfError(CF_EINVAL, "the cooperative PCG runs at least one iteration per launch");
    const CoopGrid cg = coop_grid(g);
    CoopXch xc;
    xc.part = (double*)xch;
    xc.halo = xc.part + (size_t)cg.G * 8 * kCoopWaves;
    xc.tag = (unsigned*)(xc.halo + (size_t)cg.G * 4 * g.n[0]);
    // tags start at 0 in every launch; sequence numbers count from 1
    CF_HIP(hipMemsetAsync(xc.tag, 0, 2 * (size_t)cg.G * kCoopTagStride * sizeof(unsigned), s));
    const size_t lds = ((size_t)cg.nwin + 14 + 4 * (size_t)g.n[0]) * sizeof(double);
    Geo ga = g;
    CoopGrid cga = cg;
    double *xa = x, *ra = r, *pa = p;
    PcgState* sta = st;
    CoopXch xca = xc;
    int ma = max_its;
    void* args[] = {&ga, &cga, &xa, &ra, &pa, &sta, &xca, &ma};
    // Each wave publishes its own partial while the gather reads at most 4
    // per lane (up to 64 strips: one store-acknowledgement wait per
    // reduction); beyond that each workgroup publishes one partial, summed
    // over its waves (a second wait, but 4 times fewer reads).  Measured per
    // 2-D iteration: 8.6 us at 128^2 (wave partials, 8 cells per lane)
    // against 9.1, 12.8 us at 512^2 (workgroup partials) against 13.9.
    const int nwp = cg.G * kCoopWaves;
    const void* fn = cg.cpt == 4        ? (const void*)k_pcg_solve_coop<1, true, 4>
                     : nwp <= kWave     ? (const void*)k_pcg_solve_coop<1, true, kCoopCpt>
                     : nwp <= 4 * kWave ? (const void*)k_pcg_solve_coop<4, true, kCoopCpt>
                                        : (const void*)k_pcg_solve_coop<4, false, kCoopCpt>;
    static_assert(kCoopMaxGroups <= 4 * kWave, "gather bucket");
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(cg.G), dim3(kCoopThreads), args,
                                                    (unsigned)lds, s);
    if (e == hipErrorCooperativeLaunchTooLarge || e == hipErrorNotSupported) {
        (void)hipGetLastError(); // refused before anything ran
        return false;
    }
    CF_HIP(e);
    return true;
}

template void launch_pcg_dir_spmv<double>(const Geo&, const double*,
                                          const double*, double*, PcgState*,
                                          RedBuf, hipStream_t);
template void launch_pcg_dir_spmv<float>(const Geo&, const float*, const float*,
                                         float*, PcgState*, RedBuf, hipStream_t);
template void launch_pcg_update<double>(const Geo&, double*, double*,
                                        const double*, PcgState*, RedBuf,
                                        hipStream_t, int, const double*, float*);
template void launch_pcg_update<float>(const Geo&, float*, float*, const float*,
                                       PcgState*, RedBuf, hipStream_t, int, const float*,
                                       float*);
template void launch_pcg_xfix<double>(const Geo&, double*, const double*, const PcgState*,
                                      hipStream_t);
template void launch_pcg_xfix<float>(const Geo&, float*, const float*, const PcgState*,
                                     hipStream_t);

void launch_refine_init(const Geo& g, const double* b, double* x64,
                        float* r32, RefineState* rs, RedBuf rb, hipStream_t s)
{
    CF_STREAM_LAUNCH(k_refine_init, g, b, x64, r32, rs, rb);
}

void launch_refine_add(const Geo& g, double* x64, float* d32, hipStream_t s)
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
