// cf_stream.hpp -- device building blocks of the streaming kernels
// (divergence, PCG, defect correction): owned-cell row iteration, the
// fence-free grid reduction, and the matrix-free pressure stencil.
//
// Streaming kernels walk the owned cells as x-rows: a 256-lane block takes
// one row at a time, so every wave load and store is a contiguous run.  The
// rows are split into 8 contiguous chunks, one per XCD (workgroups are dealt
// to the XCDs round-robin), so the y and z neighbour rows a block reads were
// fetched into its own XCD's L2 shortly before.
//
// Everything that depends only on the row (j, k) -- neighbour-row offsets,
// wall couplings, the y/z part of each cell's wall count -- is computed once
// per row and is wave-uniform; only the x part varies per lane.  The
// stencil is therefore branch-free.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cf_device.hpp"
#include "cf_kernels.hpp"

namespace cfmi
{

// ---- grid reduction ----------------------------------------------------------
// Each block reduces its lanes' double-double sums (xor shuffles, then LDS),
// and thread 0 publishes the block's partial with agent-scope atomic stores
// (written through to memory, so every XCD sees them), waits for them to
// complete, and counts its arrival with agent-scope atomics (per group of
// blocks, then per group).  The block that arrives last reads all partials
// with agent-scope atomic loads (which
// read memory, not its own XCD's L2), reduces them and runs `fin` on thread
// 0.  No fence is issued: an agent-scope fence would write back and
// invalidate the XCD's whole L2 on every block.
template <int NV, int BS = kBlock>
__device__ __forceinline__ void block_reduce(double (&hi)[NV], double (&lo)[NV])
{
    __shared__ double sh[NV][2][BS / kWave];
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
            for (int w = 1; w < BS / kWave; ++w)
                dd_add(hi[v], lo[v], sh[v][0][w], sh[v][1][w]);
        }
    }
    __syncthreads();
}

template <int NV, int BS = kBlock, class Fin>
__device__ __forceinline__ void grid_reduce(double (&hi)[NV], double (&lo)[NV],
                                            RedBuf rb, Fin&& fin)
{
    block_reduce<NV, BS>(hi, lo);
    __shared__ int last;
    if (threadIdx.x == 0) {
        double* part = rb.partials + (size_t)blockIdx.x * NV * 2;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            __hip_atomic_store(&part[2 * v], hi[v], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&part[2 * v + 1], lo[v], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        // partials complete before arrival: the waitcnt orders the hardware;
        // the signal fence (a compiler-only barrier, no cache maintenance)
        // keeps the compiler from moving the relaxed stores past the RMW
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        // Arrivals are counted per group of blocks (blockIdx & 7), each group
        // on its own counter line: same-address atomics serialise at ~11 ns
        // each, so one counter for a 1024-2048-block grid costs 12-24 us.
        // The last block of a group then counts the group's arrival; the
        // block that completes the last group reduces.
        const unsigned groups = gridDim.x < 8 ? gridDim.x : 8;
        const unsigned grp = blockIdx.x & 7;
        const unsigned in_grp = (gridDim.x - 1 - grp) / 8 + 1;
        const unsigned t = __hip_atomic_fetch_add(
            rb.counter + grp * kCounterStride, 1u, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
        int fin_block = 0;
        if (t == in_grp - 1) {
            const unsigned u = __hip_atomic_fetch_add(
                rb.counter + 8 * kCounterStride, 1u, __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
            fin_block = (u == groups - 1);
        }
        last = fin_block;
    }
    // the partial loads below stay after the arrival count (compiler-only
    // barrier, as above)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (!last)
        return;
    double h2[NV], l2[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        h2[v] = 0.0;
        l2[v] = 0.0;
    }
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) {
        double* part = rb.partials + (size_t)b * NV * 2;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double ph = __hip_atomic_load(&part[2 * v], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            const double pl = __hip_atomic_load(
                &part[2 * v + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            dd_add(h2[v], l2[v], ph, pl);
        }
    }
    block_reduce<NV, BS>(h2, l2);
    if (threadIdx.x == 0) {
        double sum[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sum[v] = h2[v] + l2[v];
        fin(sum);
    }
    if (threadIdx.x < kCounterSlots)
        __hip_atomic_store(rb.counter + threadIdx.x * kCounterStride, 0u,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- row iteration --------------------------------------------------------------
// Wall count contribution of axis d at global index gi: the SOLID walls that
// reduce the diagonal of cell gi (BoundaryConditions.hpp:57-97).
__device__ __forceinline__ int wall_count(const Geo& g, int d, int gi)
{
    if (g.periodic[d])
        return 0;
    return (gi == 0 && g.solid[d][0] ? 1 : 0) +
           (gi == g.n[d] - 1 && g.solid[d][1] ? 1 : 0);
}

// Per-row (wave-uniform) stencil data: for the y and z axes, the linear
// offset of the -/+ neighbour row, whether the coupling is zeroed by a wall,
// and the y/z wall count of the neighbour row's cell; cnt is this row's own
// y/z wall count.
struct RowStencil
{
    int64_t off[3][2];
    bool wall[3][2];
    int ncnt[3][2];
    int cnt;
};

__device__ __forceinline__ RowStencil row_stencil(const Geo& g, int j, int k)
{
    RowStencil rs;
    const int gi[3] = {0, j, k};
    const int64_t stride[3] = {1, g.px, g.px * g.py};
    rs.cnt = wall_count(g, 1, j) + wall_count(g, 2, k);
    for (int d = 1; d < 3; ++d) {
        if (d >= g.dim) {
            rs.off[d][0] = rs.off[d][1] = 0;
            rs.wall[d][0] = rs.wall[d][1] = true;
            rs.ncnt[d][0] = rs.ncnt[d][1] = rs.cnt;
            continue;
        }
        const int n = g.n[d];
        const int other = d == 1 ? wall_count(g, 2, k) : wall_count(g, 1, j);
        int gm, gp;
        if (g.periodic[d]) {
            gm = gi[d] == 0 ? n - 1 : gi[d] - 1;
            gp = gi[d] == n - 1 ? 0 : gi[d] + 1;
            rs.wall[d][0] = rs.wall[d][1] = false;
        } else {
            gm = gi[d] - 1;
            gp = gi[d] + 1;
            rs.wall[d][0] = gi[d] <= 0;
            rs.wall[d][1] = gi[d] > n - 2;
        }
        rs.off[d][0] = (int64_t)(gm - gi[d]) * stride[d];
        rs.off[d][1] = (int64_t)(gp - gi[d]) * stride[d];
        rs.ncnt[d][0] = other + wall_count(g, d, gm);
        rs.ncnt[d][1] = other + wall_count(g, d, gp);
    }
    return rs;
}

// Per-lane x data of owned cell i: wall counts along x of the cell and of its
// -/+ neighbours, their offsets, and whether the couplings are walls.
struct LaneX
{
    int cnt, ncnt[2];
    int64_t off[2];
    bool wall[2];
};

__device__ __forceinline__ LaneX lane_x(const Geo& g, int i)
{
    LaneX lx;
    const int n = g.n[0];
    int gm, gp;
    if (g.periodic[0]) {
        gm = i == 0 ? n - 1 : i - 1;
        gp = i == n - 1 ? 0 : i + 1;
        lx.wall[0] = lx.wall[1] = false;
    } else {
        gm = i - 1;
        gp = i + 1;
        lx.wall[0] = i <= 0;
        lx.wall[1] = i > n - 2;
    }
    lx.off[0] = gm - i;
    lx.off[1] = gp - i;
    lx.cnt = wall_count(g, 0, i);
    lx.ncnt[0] = wall_count(g, 0, gm);
    lx.ncnt[1] = wall_count(g, 0, gp);
    return lx;
}

// Calls f(row_base, j, k, RowStencil) for each owned row of this block.
template <class F>
__device__ __forceinline__ void for_owned_rows(const Geo& g, F&& f)
{
    const int n1 = g.n[1];
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
        f(g.idx(g.h, g.h + j, k0 + k), j, k);
    }
}

// Grid of the row-streaming kernels: one block per owned row, rounded up to
// a multiple of 8, and at most kRedBlocks (further rows are strided).
inline int stream_blocks(const Geo& g)
{
    const int64_t rows = (int64_t)g.n[1] * g.n[2];
    const int64_t b = (rows + 7) / 8 * 8;
    return (int)(b < kRedBlocks ? b : kRedBlocks);
}

// Calls f(c, i, j, k) for every owned cell of this block (global indices).
template <class F>
__device__ __forceinline__ void for_owned_cells(const Geo& g, F&& f)
{
    const int n0 = g.n[0];
    for_owned_rows(g, [&](int64_t base, int j, int k) {
        for (int i = threadIdx.x; i < n0; i += kBlock)
            f(base + i, i, j, k);
    });
}

// ---- typed constants of the pressure operator -----------------------------------
template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c)
{
    if constexpr (std::is_same<T, float>::value)
        return fmaf(a, b, c);
    else
        return fma(a, b, c);
}

template <class T>
__device__ __forceinline__ T diag_of(const Geo& g, int cnt)
{
    if constexpr (std::is_same<T, float>::value)
        return g.diag32[cnt];
    else
        return g.diag[cnt];
}

template <class T>
__device__ __forceinline__ T dinv_of(const Geo& g, int cnt)
{
    if constexpr (std::is_same<T, float>::value)
        return g.dinv32[cnt];
    else
        return g.dinv[cnt];
}

template <class T>
__device__ __forceinline__ T offdiag(const Geo& g)
{
    if constexpr (std::is_same<T, float>::value)
        return -g.s32;
    else
        return -g.s;
}

// Value of table[ubase + wx] for a wave-uniform ubase and a per-lane
// wx in {0, 1, 2}: three uniform loads and two selects.
template <class T, class Tab>
__device__ __forceinline__ T pick3(Tab tab, int ubase, int wx)
{
    const T a = tab(ubase), b = tab(ubase + 1), c = tab(ubase + 2);
    return wx == 0 ? a : (wx == 1 ? b : c);
}

template <class T>
__device__ __forceinline__ T dinv_pick(const Geo& g, int ubase, int wx)
{
    return pick3<T>([&](int q) { return dinv_of<T>(g, q); }, ubase, wx);
}

template <class T>
__device__ __forceinline__ T diag_pick(const Geo& g, int ubase, int wx)
{
    return pick3<T>([&](int q) { return diag_of<T>(g, q); }, ubase, wx);
}

// y = A x at owned cell c (x read from memory; wall ghosts of x are 0), in
// the reference's stencil order {0}, -x, +x, -y, +y, -z, +z: y = diag x0, then
// y = fma(c_m, x_m, y) per neighbour, c_m = -scale or 0 across a wall.
template <int DIM, class T>
__device__ __forceinline__ T apply_A(const Geo& g, const RowStencil& rs,
                                     const LaneX& lx, const T* __restrict__ x,
                                     int64_t c)
{
    const T s = offdiag<T>(g);
    T y = diag_pick<T>(g, rs.cnt, lx.cnt) * x[c];
    y = fma_t<T>(lx.wall[0] ? (T)0 : s, x[c + lx.off[0]], y);
    y = fma_t<T>(lx.wall[1] ? (T)0 : s, x[c + lx.off[1]], y);
#pragma unroll
    for (int d = 1; d < DIM; ++d) {
        y = fma_t<T>(rs.wall[d][0] ? (T)0 : s, x[c + rs.off[d][0]], y);
        y = fma_t<T>(rs.wall[d][1] ? (T)0 : s, x[c + rs.off[d][1]], y);
    }
    return y;
}

} // namespace cfmi
