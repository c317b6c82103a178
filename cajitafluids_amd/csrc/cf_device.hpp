// cf_device.hpp -- device-side arithmetic of the hot path.
//
// Every expression is written with an explicit evaluation order and explicit
// fma(); the library is compiled with -ffp-contract=off so nothing else
// fuses.  Dot products accumulate in double-double, whose rounded value does
// not depend on the order of the terms, so the reductions' results do not
// depend on the launch shape either.
#pragma once

#include <hip/hip_runtime.h>

#include "cf_geometry.hpp"

namespace cfmi
{

constexpr int kWave = 64;

// ---- double-double accumulation ------------------------------------------
__device__ __forceinline__ void dd_add(double& hi, double& lo, double bhi,
                                       double blo)
{
    double s = hi + bhi;
    double bb = s - hi;
    double e = (hi - (s - bb)) + (bhi - bb);
    e = e + (lo + blo);
    double h2 = s + e;
    lo = e - (h2 - s);
    hi = h2;
}

__device__ __forceinline__ void dd_add_prod(double& hi, double& lo, double a,
                                            double b)
{
    double p = a * b;
    double pe = fma(a, b, -p);
    dd_add(hi, lo, p, pe);
}

// Per-lane accumulation of a product into an unnormalised pair: hi takes the
// rounded running sum, and the exact error of every addition (TwoSum) and of
// the product (TwoProd) is summed into lo.  lo is only ever a small multiple
// of ulp(hi), so its own rounding is ~2^-100 of the total; dd_norm
// renormalises before the pair enters dd_add.  EXACT: the product of the two
// operands is known to be exact in double (both are floats), so the TwoProd
// error term is skipped.
template <bool EXACT = false>
__device__ __forceinline__ void dd_acc_prod(double& hi, double& lo, double a,
                                            double b)
{
    const double p = a * b;
    const double s = hi + p;
    const double t = s - hi;
    double e = (hi - (s - t)) + (p - t);
    if constexpr (!EXACT)
        e = e + fma(a, b, -p);
    lo = lo + e;
    hi = s;
}

__device__ __forceinline__ void dd_norm(double& hi, double& lo)
{
    const double s = hi + lo;
    const double t = s - hi;
    lo = (hi - (s - t)) + (lo - t);
    hi = s;
}

// ---- B-spline sampling (Cajita SplineData / evaluateSpline / G2P::value,
// called at src/Interpolation.hpp:30-41) -------------------------------------
template <int ORDER>
__device__ __forceinline__ void spline_weights(double xi, int& s0,
                                               double (&w)[ORDER + 1])
{
    int fl = (int)xi;
    double t = xi - (double)fl;
    if constexpr (ORDER == 1) {
        s0 = fl;
        w[0] = 1.0 - t;
        w[1] = t;
    } else {
        const double c16 = 1.0 / 6.0;
        double omt = 1.0 - t;
        double t2 = t * t;
        double t3 = t2 * t;
        s0 = fl - 1;
        w[0] = ((omt * omt) * omt) * c16;
        w[1] = ((3.0 * t3 - 6.0 * t2) + 4.0) * c16;
        w[2] = (((-3.0 * t3 + 3.0 * t2) + 3.0 * t) + 1.0) * c16;
        w[3] = t3 * c16;
    }
}

// Logical coordinate along axis d, clamped so the stencil stays inside the
// ghosted array (identity for every in-range trace).
template <int ORDER>
__device__ __forceinline__ double logical(const Geo& g, int e, int d, double p)
{
    double xi = (p - g.x0[e][d]) * g.rdx[e][d];
    double lo = ORDER == 1 ? 0.0 : 1.0;
    double hi = (double)(g.ext[e][d] - (ORDER == 1 ? 2 : 3));
    return fmin(fmax(xi, lo), hi);
}

template <int DIM, int ORDER>
__device__ __forceinline__ double interpolate(const Geo& g, int e,
                                              const double* __restrict__ f,
                                              const double (&p)[3])
{
    int s0[3] = {0, 0, 0};
    double w[3][ORDER + 1];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        spline_weights<ORDER>(logical<ORDER>(g, e, d, p[d]), s0[d], w[d]);
    double val = 0.0;
    if constexpr (DIM == 2) {
#pragma unroll
        for (int a = 0; a < ORDER + 1; ++a) {
            const double* row = f + g.idx(s0[0] + a, s0[1], 0);
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b) {
                double wab = w[0][a] * w[1][b];
                val = fma(wab, row[b * g.px], val);
            }
        }
    } else {
        const int64_t pz = g.px * g.py;
#pragma unroll
        for (int a = 0; a < ORDER + 1; ++a) {
#pragma unroll
            for (int b = 0; b < ORDER + 1; ++b) {
                double wab = w[0][a] * w[1][b];
                const double* col = f + g.idx(s0[0] + a, s0[1] + b, s0[2]);
#pragma unroll
                for (int c = 0; c < ORDER + 1; ++c) {
                    double wabc = wab * w[2][c];
                    val = fma(wabc, col[c * pz], val);
                }
            }
        }
    }
    return val;
}

__device__ __forceinline__ double coord(const Geo& g, int e, int d, int l)
{
    if (e == d + 1)
        return g.low[d] + (double)l * g.dx;
    return g.low[d] + ((double)l + 0.5) * g.dx;
}

template <int DIM, int VORDER>
__device__ __forceinline__ void velocity(const Geo& g,
                                         const double* __restrict__ u,
                                         const double* __restrict__ v,
                                         const double* __restrict__ w,
                                         const double (&x)[3], double (&vel)[3])
{
    vel[0] = interpolate<DIM, VORDER>(g, 1, u, x);
    vel[1] = interpolate<DIM, VORDER>(g, 2, v, x);
    if constexpr (DIM == 3)
        vel[2] = interpolate<DIM, VORDER>(g, 3, w, x);
    else
        vel[2] = 0.0;
}

// RK3 back-trace, src/TimeIntegrator.hpp:36-77 (3/4-step probe with v0 under
// CF_QUIRK_RK3_V0, with v1 otherwise).
template <int DIM, int VORDER>
__device__ __forceinline__ void rk3(const Geo& g, const double* __restrict__ u,
                                    const double* __restrict__ v,
                                    const double* __restrict__ w,
                                    const double (&x0)[3], double (&tr)[3])
{
    const double dt = g.dt;
    double v0[3], v1[3], v2[3], x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0};
    velocity<DIM, VORDER>(g, u, v, w, x0, v0);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x1[d] = x0[d] - 0.5 * dt * v0[d];
    velocity<DIM, VORDER>(g, u, v, w, x1, v1);
    const bool q = (g.quirks & CF_QUIRK_RK3_V0) != 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        x2[d] = x0[d] - 0.75 * dt * (q ? v0[d] : v1[d]);
    velocity<DIM, VORDER>(g, u, v, w, x2, v2);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        tr[d] = x0[d] - dt * ((2.0 / 9.0) * v0[d] + (3.0 / 9.0) * v1[d] +
                              (4.0 / 9.0) * v2[d]);
    if constexpr (DIM == 2)
        tr[2] = 0.0;
}

// InflowSource membership (src/InflowSource.hpp:40-41: half-open box;
// sphere: strict).
__device__ __forceinline__ bool in_source(const Geo& g, const double (&x)[3])
{
    if (g.inflow_shape == CF_INFLOW_BOX) {
        for (int d = 0; d < g.dim; ++d)
            if (!(x[d] >= g.in_lo[d] && x[d] < g.in_hi[d]))
                return false;
        return true;
    }
    if (g.inflow_shape == CF_INFLOW_SPHERE) {
        double d2 = 0.0;
        for (int d = 0; d < g.dim; ++d) {
            double r = x[d] - g.in_c[d];
            d2 = d2 + r * r;
        }
        return d2 < g.in_r2;
    }
    return false;
}

// Number of SOLID walls adjacent to owned cell with global index gidx (the
// diagonal reductions of BoundaryConditions.hpp:57-97).
__device__ __forceinline__ int solid_count(const Geo& g, const int (&gi)[3])
{
    int k = 0;
    for (int d = 0; d < g.dim; ++d) {
        if (g.periodic[d])
            continue;
        if (gi[d] <= 0 && g.solid[d][0])
            ++k;
        if (gi[d] > g.n[d] - 2 && g.solid[d][1])
            ++k;
    }
    return k;
}

} // namespace cfmi
