// cf_geometry.cpp -- build the device geometry from a cf_config.
#include "cf_geometry.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace cfmi
{

// Mesh (src/Mesh.hpp:41-103), the Cajita local grid index spaces for one
// rank with halo h, the dt clamp (src/Solver.hpp:93-106) and the pressure
// operator constants (src/VelocityCorrector.hpp:116-180,
// src/BoundaryConditions.hpp:57-97).
cf_status build_geometry(const cf_config& c, Geo& g, const char** why)
{
    std::memset(&g, 0, sizeof(g));
    *why = "";
    if (c.dim != 2 && c.dim != 3) {
        *why = "dim must be 2 or 3";
        return CF_EINVAL;
    }
    if (c.halo < 1 || c.halo > 8) {
        *why = "halo must be in [1, 8]";
        return CF_EINVAL;
    }
    if (c.field_order != 1 && c.field_order != 3) {
        *why = "field_order must be 1 or 3";
        return CF_EINVAL;
    }
    if (c.velocity_order != 1 && c.velocity_order != 3) {
        *why = "velocity_order must be 1 or 3";
        return CF_EINVAL;
    }
    if (c.field_order == 3 && c.halo < 2) {
        *why = "a cubic field stencil needs halo >= 2";
        return CF_EINVAL;
    }
    if (!(c.density > 0) || !(c.dt > 0)) {
        *why = "density and dt must be positive";
        return CF_EINVAL;
    }
    if (c.precision != CF_FP64 && c.precision != CF_FP32_REFINED) {
        *why = "precision must be CF_FP64 or CF_FP32_REFINED";
        return CF_EINVAL;
    }
    if (c.refine_mode != CF_REFINE_RESTART && c.refine_mode != CF_REFINE_CARRY) {
        *why = "refine_mode must be CF_REFINE_RESTART or CF_REFINE_CARRY";
        return CF_EINVAL;
    }
    if (!(c.cg_tol > 0) || !(c.inner_tol > 0) || c.max_outer < 0) {
        *why = "cg_tol and inner_tol must be positive, max_outer >= 0";
        return CF_EINVAL;
    }
    if (c.preconditioner != CF_PRECOND_JACOBI && c.preconditioner != CF_PRECOND_MG &&
        c.preconditioner != CF_PRECOND_NONE) {
        *why = "preconditioner must be CF_PRECOND_JACOBI, CF_PRECOND_MG or CF_PRECOND_NONE";
        return CF_EINVAL;
    }
    if (c.preconditioner == CF_PRECOND_MG) {
        // FP32_REFINED: the V-cycle's level 0 in FP32, on 3-D grids with
        // more than 2 cells along every axis (level 0 marched by the tiles)
        if (c.precision == CF_FP32_REFINED &&
            !(c.dim == 3 && c.num_cell[0] > 2 && c.num_cell[1] > 2 && c.num_cell[2] > 2)) {
            *why = "the FP32 multigrid V-cycle needs a 3-D grid with more than 2 cells "
                   "along every axis";
            return CF_ENOTIMPL;
        }
        if (c.mg_nu < 1 || c.mg_coarse_sweeps < 1 || !(c.mg_omega > 0) ||
            !(c.mg_omega <= 1) || !(c.mg_alpha > 0) || !(c.mg_alpha <= 2)) {
            *why = "mg_nu and mg_coarse_sweeps must be >= 1, mg_omega in (0, 1], "
                   "mg_alpha in (0, 2]";
            return CF_EINVAL;
        }
        for (int d = 0; d < c.dim; ++d)
            if (c.periodic[d] && c.num_cell[d] < 2) {
                *why = "multigrid needs at least 2 cells along a periodic axis";
                return CF_EINVAL;
            }
    }
    g.dim = c.dim;
    g.h = c.halo;
    for (int d = 0; d < 3; ++d) {
        if (d < c.dim) {
            if (c.num_cell[d] <= 0 || c.num_cell[d] > (1 << 14)) {
                *why = "num_cell must be in [1, 16384] per axis";
                return CF_EINVAL;
            }
            if (!(c.hi[d] > c.lo[d])) {
                *why = "bounding box must have hi > lo";
                return CF_EINVAL;
            }
            if (c.periodic[d] && c.num_cell[d] < c.halo) {
                *why = "a periodic axis needs at least halo cells";
                return CF_EINVAL;
            }
        }
        g.n[d] = d < c.dim ? c.num_cell[d] : 1;
        g.periodic[d] = d < c.dim ? (c.periodic[d] != 0) : 0;
        for (int s = 0; s < 2; ++s)
            g.solid[d][s] = d < c.dim && c.wall_type[d][s] == CF_SOLID;
    }
    // Mesh.hpp:49-64: dx from the x extent; every axis must divide.
    g.dx = (c.hi[0] - c.lo[0]) / c.num_cell[0];
    for (int d = 0; d < c.dim; ++d) {
        double extent = c.num_cell[d] * g.dx;
        if (std::fabs(extent - (c.hi[d] - c.lo[d])) > 10.0 * DBL_EPSILON) {
            *why = "Extent not evenly divisible by uniform cell size";
            return CF_EINVAL;
        }
    }
    for (int d = 0; d < 3; ++d)
        g.low[d] = d < c.dim ? c.lo[d] - g.dx * (double)g.h : 0.0;

    for (int e = 0; e < kMaxEnt; ++e)
        for (int d = 0; d < 3; ++d) {
            if (d >= c.dim) {
                g.ext[e][d] = 1;
                g.own_hi[e][d] = 1;
                continue;
            }
            g.ext[e][d] = g.n[d] + 2 * g.h + (e == d + 1 ? 1 : 0);
            g.own_hi[e][d] =
                g.h + g.n[d] + ((e == d + 1 && !g.periodic[d]) ? 1 : 0);
        }
    for (int d = 0; d < 3; ++d) {
        g.uni_hi[d] = 0;
        for (int e = 0; e <= c.dim; ++e)
            if (g.own_hi[e][d] > g.uni_hi[d])
                g.uni_hi[d] = g.own_hi[e][d];
    }
    // Storage: the first owned column lands on a 16-element boundary.
    g.sx = (16 - (g.h % 16)) % 16;
    int64_t row = g.sx + g.n[0] + 2 * g.h + 1;
    g.px = (row + 15) / 16 * 16;
    g.py = g.n[1] + 2 * g.h + 1;
    g.pz = c.dim == 3 ? g.n[2] + 2 * g.h + 1 : 1;
    g.total = g.px * g.py * g.pz + 16;

    // Cajita evaluateSpline: logical spacing = coordinate(1) - coordinate(0).
    for (int e = 0; e < kMaxEnt; ++e)
        for (int d = 0; d < c.dim; ++d) {
            double a = (e == d + 1) ? g.low[d] + 0.0 * g.dx
                                    : g.low[d] + (0.0 + 0.5) * g.dx;
            double b = (e == d + 1) ? g.low[d] + 1.0 * g.dx
                                    : g.low[d] + (1.0 + 0.5) * g.dx;
            g.x0[e][d] = a;
            g.rdx[e][d] = 1.0 / (b - a);
        }

    // Solver.hpp:93-106: clamp dt to dx / umax.
    double dt = c.dt;
    if (c.clamp_dt) {
        const double* f = c.body_force;
        double f2 = f[0] * f[0] + f[1] * f[1];
        if (c.dim == 3)
            f2 = f2 + f[2] * f[2];
        double forcemax = std::sqrt(f2);
        const double* U = c.inflow_velocity;
        double um = std::fmax(std::fabs(U[0]), std::fabs(U[1]));
        if (c.dim == 3)
            um = std::fmax(um, std::fabs(U[2]));
        double umax = um + std::sqrt(forcemax * g.dx);
        if (umax > 0 && dt > g.dx / umax)
            dt = g.dx / umax;
    }
    g.dt = dt;
    g.quirks = c.quirks;
    g.field_order = c.field_order;
    g.vel_order = c.velocity_order;

    g.inflow_shape = c.inflow_shape;
    for (int d = 0; d < 3; ++d) {
        g.in_lo[d] = c.inflow_location[d];
        g.in_hi[d] = c.inflow_location[d] + c.inflow_size[d];
        g.in_c[d] = c.inflow_center[d];
        g.U[d] = d < c.dim ? c.inflow_velocity[d] : 0.0;
        g.force[d] = d < c.dim ? c.body_force[d] : 0.0;
    }
    g.in_r2 = c.inflow_radius * c.inflow_radius;
    g.Q = c.inflow_quantity;

    // Pressure operator.  The reference's diagonal starts at 2*dim*scale and
    // loses one scale per adjacent SOLID wall; the value depends only on the
    // count of such walls, so it is tabulated.
    double scale = dt / (c.density * g.dx * g.dx);
    g.s = scale;
    g.s32 = (float)scale;
    for (int k = 0; k <= 6; ++k) {
        double a0 = (2.0 * c.dim) * scale;
        for (int m = 0; m < k; ++m)
            a0 -= scale;
        g.diag[k] = a0;
        g.dinv[k] = 1.0 / a0;
        g.diag32[k] = (float)a0;
        g.dinv32[k] = (float)(1.0 / a0);
    }
    g.proj_scale = dt / (c.density * g.dx);
    g.div_scale = 1.0 / g.dx;
    return CF_OK;
}

} // namespace cfmi
