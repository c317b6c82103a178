// cf_init.hip -- seeded initial-state generators (the device counterpart of
// the reference's init functor, src/ProblemManager.hpp:186-263).
//
// Random draws come from a counter-based SplitMix64 (Steele, Lea & Flood,
// 2014; finaliser constants as published by Vigna): a draw depends only on
// (seed, stream, counter), so the same field comes out on any launch shape.
#include <hip/hip_runtime.h>

#include "cf_device.hpp"
#include "cf_init.hpp"

namespace cfmi
{

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t rand_u64(uint64_t seed, uint64_t stream,
                                             uint64_t idx)
{
    uint64_t key = mix64(seed + 0x632BE59BD9B4E019ULL * (stream + 1));
    return mix64(key + idx);
}

__host__ __device__ inline double rand_uniform(uint64_t seed, uint64_t stream,
                                               uint64_t idx)
{
    return (double)(rand_u64(seed, stream, idx) >> 11) * 0x1.0p-53;
}

double host_rand_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return rand_uniform(seed, stream, idx);
}

// Separation along axis d, wrapped to the nearest image on periodic axes.
__device__ inline double sep(const InitParams& P, int d, double x, double c)
{
    double r = x - c;
    if (P.periodic[d]) {
        double L = P.box_hi[d] - P.box_lo[d];
        r = r - L * nearbyint(r / L);
    }
    return r;
}

// Component m of the vector potential at point x (depends on the two axes
// other than m): sum over the tubes aligned with m of
// Gamma exp(-d^2 / (2 sigma^2)).
__device__ inline double psi(const InitParams& P, int m, const double (&x)[3])
{
    const int o1 = m == 0 ? 1 : 0;
    const int o2 = m == 2 ? 1 : 2;
    double sum = 0.0;
    for (int t = 0; t < P.ntube; ++t) {
        if (P.tube_axis[t] != m)
            continue;
        double r1 = sep(P, o1, x[o1], P.tube_a[t]);
        double r2 = sep(P, o2, x[o2], P.tube_b[t]);
        double d2 = r1 * r1 + r2 * r2;
        sum = sum + P.tube_gamma[t] * exp(-d2 / (2.0 * P.tube_sigma[t] * P.tube_sigma[t]));
    }
    return sum;
}

__device__ inline double node_x(const Geo& g, int d, int l)
{
    return g.low[d] + (double)l * g.dx;
}
__device__ inline double centre_x(const Geo& g, int d, int l)
{
    return g.low[d] + ((double)l + 0.5) * g.dx;
}

// Face velocity = discrete curl of the edge-centred potential (2-D: of the
// nodal stream function psi_z), so the discrete divergence vanishes.
template <int DIM>
__global__ __launch_bounds__(256) void k_init_tubes(Geo g, InitParams P,
                                                     Fields cur,
                                                     unsigned long long* vmax)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    double local_max = 0.0;
    for (int e = 1; e <= DIM; ++e) {
        bool own = true;
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        const int d = e - 1;
        double val;
        if (DIM == 2) {
            const int t = d == 0 ? 1 : 0;
            double xa[3] = {node_x(g, 0, i), node_x(g, 1, j), 0.0};
            double xb[3] = {xa[0], xa[1], 0.0};
            xb[t] = node_x(g, t, l[t] + 1);
            double diff = psi(P, 2, xb) - psi(P, 2, xa);
            val = d == 0 ? diff / g.dx : -diff / g.dx;
        } else {
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            double xa[3], xb[3];
            xa[d] = xb[d] = node_x(g, d, l[d]);
            xa[d2] = xb[d2] = centre_x(g, d2, l[d2]);
            xa[d1] = node_x(g, d1, l[d1]);
            xb[d1] = node_x(g, d1, l[d1] + 1);
            double t1 = (psi(P, d2, xb) - psi(P, d2, xa)) / g.dx;
            xa[d1] = xb[d1] = centre_x(g, d1, l[d1]);
            xa[d2] = node_x(g, d2, l[d2]);
            xb[d2] = node_x(g, d2, l[d2] + 1);
            double t2 = (psi(P, d1, xb) - psi(P, d1, xa)) / g.dx;
            val = t1 - t2;
        }
        cur.e[e][g.idx(i, j, k)] = val;
        local_max = fmax(local_max, fabs(val));
    }
    // |val| >= 0, so the IEEE bit pattern orders like the value.
    if (local_max > 0.0)
        atomicMax(vmax, (unsigned long long)__double_as_longlong(local_max));
}

template <int DIM>
__global__ __launch_bounds__(256) void k_scale_faces(
    Geo g, Fields cur, const unsigned long long* vmax, double peak)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    const double m = __longlong_as_double((long long)*vmax);
    if (!(m > 0.0))
        return;
    const double sc = peak / m;
    for (int e = 1; e <= DIM; ++e) {
        bool own = true;
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        const int64_t c = g.idx(i, j, k);
        cur.e[e][c] = cur.e[e][c] * sc;
    }
}

// Sum of Gaussian blobs at owned cell centres.
__global__ __launch_bounds__(256) void k_init_blobs(Geo g, InitParams P,
                                                     double* q)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    for (int d = 0; d < g.dim; ++d)
        if (l[d] >= g.own_hi[0][d])
            return;
    double x[3] = {0, 0, 0};
    for (int d = 0; d < g.dim; ++d)
        x[d] = centre_x(g, d, l[d]);
    double sum = 0.0;
    for (int b = 0; b < P.nblob; ++b) {
        double d2 = 0.0;
        for (int d = 0; d < g.dim; ++d) {
            double r = sep(P, d, x[d], P.blob_c[b][d]);
            d2 = d2 + r * r;
        }
        sum = sum + P.blob_amp[b] *
                        exp(-d2 / (2.0 * P.blob_sigma[b] * P.blob_sigma[b]));
    }
    q[g.idx(i, j, k)] = sum;
}

// i.i.d. N(0,1) (Box-Muller) on owned faces; stream = entity, counter = the
// face's index in the dense ghosted array of its entity.
template <int DIM>
__global__ __launch_bounds__(256) void k_init_random(Geo g, uint64_t seed,
                                                      Fields cur)
{
    const int i = g.h + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.h + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = g.own_lo(2) + blockIdx.z;
    const int l[3] = {i, j, k};
    for (int e = 1; e <= DIM; ++e) {
        bool own = true;
        for (int d = 0; d < DIM; ++d)
            own = own && l[d] < g.own_hi[e][d];
        if (!own)
            continue;
        const uint64_t id =
            (uint64_t)i + (uint64_t)g.ext[e][0] *
                              ((uint64_t)j + (uint64_t)g.ext[e][1] * (uint64_t)k);
        double u1 = 1.0 - rand_uniform(seed, (uint64_t)e, 2 * id);
        double u2 = rand_uniform(seed, (uint64_t)e, 2 * id + 1);
        cur.e[e][g.idx(i, j, k)] =
            sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

static dim3 owned_grid(const Geo& g, dim3 blk)
{
    const int ex = g.uni_hi[0] - g.h, ey = g.uni_hi[1] - g.h;
    const int ez = g.dim == 3 ? g.uni_hi[2] - g.h : 1;
    return dim3((ex + blk.x - 1) / blk.x, (ey + blk.y - 1) / blk.y, ez);
}

void launch_init_tubes(const Geo& g, const InitParams& P, Fields cur,
                       unsigned long long* vmax, double peak, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    dim3 grd = owned_grid(g, blk);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, vmax);
    if (g.dim == 2) {
        hipLaunchKernelGGL(k_init_tubes<2>, grd, blk, 0, s, g, P, cur, vmax);
        hipLaunchKernelGGL(k_scale_faces<2>, grd, blk, 0, s, g, cur, vmax, peak);
    } else {
        hipLaunchKernelGGL(k_init_tubes<3>, grd, blk, 0, s, g, P, cur, vmax);
        hipLaunchKernelGGL(k_scale_faces<3>, grd, blk, 0, s, g, cur, vmax, peak);
    }
}

void launch_init_blobs(const Geo& g, const InitParams& P, double* q,
                       hipStream_t s)
{
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(k_init_blobs, owned_grid(g, blk), blk, 0, s, g, P, q);
}

void launch_init_random(const Geo& g, uint64_t seed, Fields cur,
                        hipStream_t s)
{
    dim3 blk(64, 4, 1);
    if (g.dim == 2)
        hipLaunchKernelGGL(k_init_random<2>, owned_grid(g, blk), blk, 0, s, g,
                           seed, cur);
    else
        hipLaunchKernelGGL(k_init_random<3>, owned_grid(g, blk), blk, 0, s, g,
                           seed, cur);
}

// Tube / vortex parameters (stream 0, four draws per tube: position in the
// two other axes in [0.3, 0.7] of the box, sigma in [0.05, 0.1], sign of
// Gamma) and blob parameters (stream 1, eight counters per blob: centre in
// [0.1, 0.9] of the box per axis, sigma in [0.03, 0.08], amplitude in
// [0.5, 1]).
void make_tube_params(const Geo& g, const cf_config& c, uint64_t seed,
                      InitParams& P)
{
    P.ntube = 3;
    for (int t = 0; t < 3; ++t) {
        int axis = g.dim == 3 ? t : 2;
        int o1 = axis == 0 ? 1 : 0;
        int o2 = axis == 2 ? 1 : 2;
        double a = 0.3 + 0.4 * rand_uniform(seed, 0, 4 * t + 0);
        double b = 0.3 + 0.4 * rand_uniform(seed, 0, 4 * t + 1);
        double sig = 0.05 + 0.05 * rand_uniform(seed, 0, 4 * t + 2);
        double gam = rand_uniform(seed, 0, 4 * t + 3) < 0.5 ? -1.0 : 1.0;
        P.tube_axis[t] = axis;
        P.tube_a[t] = c.lo[o1] + a * (c.hi[o1] - c.lo[o1]);
        P.tube_b[t] = c.lo[o2] + b * (c.hi[o2] - c.lo[o2]);
        P.tube_sigma[t] = sig;
        P.tube_gamma[t] = gam;
    }
}

void make_blob_params(const Geo& g, const cf_config& c, uint64_t seed,
                      int nblob, InitParams& P)
{
    P.nblob = nblob;
    for (int b = 0; b < nblob; ++b) {
        uint64_t base = (uint64_t)b * 8;
        for (int d = 0; d < 3; ++d) {
            double cu = d < g.dim ? 0.1 + 0.8 * rand_uniform(seed, 1, base + d) : 0.0;
            P.blob_c[b][d] = c.lo[d] + cu * (c.hi[d] - c.lo[d]);
        }
        P.blob_sigma[b] = 0.03 + 0.05 * rand_uniform(seed, 1, base + 3);
        P.blob_amp[b] = 0.5 + 0.5 * rand_uniform(seed, 1, base + 4);
    }
}

} // namespace cfmi
