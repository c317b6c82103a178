// cf_geometry.hpp -- grid geometry and device layout shared by host and
// device code.
//
// Every state array of a handle (q, u, v, w in two versions, the pressure
// lhs/rhs and the PCG vectors) uses ONE padded layout: element (i, j, k) of
// an entity's ghosted local index space lives at
//     sx + i + px * (j + py * k)
// with px, py, pz sized for the largest entity extent (n + 2h + 1 per axis).
// sx shifts the first owned column (i = h) onto a 128-byte boundary so that
// each owned x-row streams in whole cache lines.  Sharing one layout lets a
// kernel address q, u, v, w and p with the same linear index.
#pragma once

#include "cf_api.h"

#include <cstdint>

namespace cfmi
{

constexpr int kMaxEnt = 4;

struct Geo
{
    int dim, h;
    int n[3];
    int periodic[3];
    int solid[3][2];   // wall is SOLID (only meaningful on non-periodic axes)
    int ext[4][3];     // ghosted extents per entity
    int own_hi[4][3];  // owned [h, own_hi) per entity (z is [0,1) in 2-D)
    int uni_hi[3];     // union of own_hi over entities
    int64_t sx, px, py, pz;
    int64_t total;     // elements per array
    double low[3];
    double dx;
    double x0[4][3];   // coordinate of local index 0 of each entity
    double rdx[4][3];  // 1 / (coordinate(1) - coordinate(0))
    double dt;
    int quirks, field_order, vel_order;
    // inputs
    int inflow_shape;
    double in_lo[3], in_hi[3], in_c[3], in_r2;
    double U[3], Q, force[3];
    // pressure operator
    double s;          // dt / (rho dx^2): off-diagonal magnitude
    double diag[7];    // diagonal with k SOLID walls adjacent
    double dinv[7];    // 1 / diag[k]
    float s32;
    float diag32[7], dinv32[7];
    double proj_scale; // dt / (rho dx)
    double div_scale;  // 1 / dx

    __host__ __device__ inline int64_t idx(int i, int j, int k) const
    {
        return sx + i + px * (j + py * (int64_t)k);
    }
    __host__ __device__ inline int own_lo(int d) const
    {
        return d < dim ? h : 0;
    }
};

// Whether A has the constant null space: no FREE wall on a non-periodic axis.
inline bool operator_singular(const Geo& g)
{
    for (int d = 0; d < g.dim; ++d)
        if (!g.periodic[d] && (!g.solid[d][0] || !g.solid[d][1]))
            return false;
    return true;
}

// Fills every field of g from the config; returns CF_OK or CF_EINVAL and a
// message.  Host only.
cf_status build_geometry(const cf_config& c, Geo& g, const char** why);

} // namespace cfmi
