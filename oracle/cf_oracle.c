/*
 * cf_oracle.c -- CPU restatement of the CajitaFluids hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  See cf_oracle.h for the reference files restated.
 * Built with -ffp-contract=off: every fused multiply-add is an explicit
 * fma() call, and every other expression rounds in source order.
 */
#include "cf_oracle_internal.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Geometry (src/Mesh.hpp:41-103; Cajita local grid / local mesh).          */
/* ------------------------------------------------------------------------ */

void cfo_owned_range(const struct cfo_state* s, int e, int lo[3], int hi[3])
{
    for (int d = 0; d < 3; ++d) {
        if (d >= s->dim) {
            lo[d] = 0;
            hi[d] = 1;
            continue;
        }
        lo[d] = s->h;
        /* A face array owns both walls of a non-periodic axis (N+1 faces);
         * a periodic axis owns the duplicated face once. */
        hi[d] = s->h + s->n[d] + ((e == d + 1 && !s->c.periodic[d]) ? 1 : 0);
    }
}

void cfo_owned(const cfo_state* s, int entity, int lo[3], int hi[3])
{
    cfo_owned_range(s, entity, lo, hi);
}

void cfo_extents(const cfo_state* s, int entity, int ext[3])
{
    for (int d = 0; d < 3; ++d)
        ext[d] = s->ext[entity][d];
}

double cfo_dt(const cfo_state* s) { return s->dt; }
double cfo_dx(const cfo_state* s) { return s->dx; }
double cfo_time(const cfo_state* s) { return s->time; }
double* cfo_field(cfo_state* s, int entity, int version)
{
    return s->f[entity][version];
}
double* cfo_lhs(cfo_state* s) { return s->lhs; }
double* cfo_rhs(cfo_state* s) { return s->rhs; }

static void* cfo_zalloc(long n, size_t sz)
{
    if (n <= 0)
        n = 1;
    return calloc((size_t)n, sz);
}

int cfo_create(const cfo_config* cfg, cfo_state** out)
{
    *out = NULL;
    if (!cfg || (cfg->dim != 2 && cfg->dim != 3) || cfg->halo < 1)
        return CFO_EINVAL;
    for (int d = 0; d < cfg->dim; ++d)
        if (cfg->num_cell[d] <= 0 || !(cfg->hi[d] > cfg->lo[d]))
            return CFO_EINVAL;
    if (cfg->field_order != 1 && cfg->field_order != 3)
        return CFO_EINVAL;
    if (cfg->velocity_order != 1 && cfg->velocity_order != 3)
        return CFO_EINVAL;
    if (!(cfg->density > 0) || !(cfg->dt > 0))
        return CFO_EINVAL;

    struct cfo_state* s = (struct cfo_state*)calloc(1, sizeof(*s));
    if (!s)
        return CFO_ENOMEM;
    s->c = *cfg;
    s->dim = cfg->dim;
    s->h = cfg->halo;
    s->nent = cfg->dim + 1;
    for (int d = 0; d < 3; ++d)
        s->n[d] = d < s->dim ? cfg->num_cell[d] : 1;

    /* Mesh.hpp:49-64: dx from the x extent; every axis must divide. */
    s->dx = (cfg->hi[0] - cfg->lo[0]) / cfg->num_cell[0];
    for (int d = 0; d < s->dim; ++d) {
        double extent = cfg->num_cell[d] * s->dx;
        if (fabs(extent - (cfg->hi[d] - cfg->lo[d])) > 10.0 * DBL_EPSILON) {
            free(s);
            return CFO_EINVAL;
        }
    }
    for (int d = 0; d < 3; ++d)
        s->low[d] = d < s->dim ? cfg->lo[d] - s->dx * (double)s->h : 0.0;

    for (int e = 0; e < 4; ++e) {
        for (int d = 0; d < 3; ++d) {
            if (d >= s->dim)
                s->ext[e][d] = 1;
            else
                s->ext[e][d] = s->n[d] + 2 * s->h + (e == d + 1 ? 1 : 0);
        }
        s->size[e] = (long)s->ext[e][0] * s->ext[e][1] * s->ext[e][2];
    }
    /* Cajita evaluateSpline: the logical spacing is the difference of the
     * coordinates of local indices 1 and 0 of the entity. */
    for (int e = 0; e < 4; ++e)
        for (int d = 0; d < s->dim; ++d) {
            double a = cfo_coord(s, e, d, 0);
            double b = cfo_coord(s, e, d, 1);
            s->x0[e][d] = a;
            s->rdx[e][d] = 1.0 / (b - a);
        }

    /* Solver.hpp:93-106: clamp dt to dx / umax. */
    double dt = cfg->dt;
    if (cfg->clamp_dt) {
        const double* f = cfg->body_force;
        double f2 = f[0] * f[0] + f[1] * f[1];
        if (s->dim == 3)
            f2 = f2 + f[2] * f[2];
        double forcemax = sqrt(f2);
        const double* U = cfg->inflow_velocity;
        double um = fmax(fabs(U[0]), fabs(U[1]));
        if (s->dim == 3)
            um = fmax(um, fabs(U[2]));
        double umax = um + sqrt(forcemax * s->dx);
        if (umax > 0 && dt > s->dx / umax)
            dt = s->dx / umax;
    }
    s->dt = dt;
    s->time = 0.0;

    int ok = 1;
    for (int e = 0; e < s->nent; ++e)
        for (int v = 0; v < 2; ++v) {
            s->f[e][v] = (double*)cfo_zalloc(s->size[e], sizeof(double));
            ok = ok && s->f[e][v];
        }
    long nc = s->size[0];
    s->lhs = (double*)cfo_zalloc(nc, sizeof(double));
    s->rhs = (double*)cfo_zalloc(nc, sizeof(double));
    s->ncoef = 2 * s->dim + 1;
    s->A = (double*)cfo_zalloc(nc * s->ncoef, sizeof(double));
    s->M = (double*)cfo_zalloc(nc, sizeof(double));
    s->r = (double*)cfo_zalloc(nc, sizeof(double));
    s->z = (double*)cfo_zalloc(nc, sizeof(double));
    s->p = (double*)cfo_zalloc(nc, sizeof(double));
    s->q = (double*)cfo_zalloc(nc, sizeof(double));
    ok = ok && s->lhs && s->rhs && s->A && s->M && s->r && s->z && s->p &&
         s->q;
    if (!ok) {
        cfo_destroy(s);
        return CFO_ENOMEM;
    }

    /* VelocityCorrector.hpp:116-144 + BoundaryConditions.hpp:57-97: the
     * stored 5-point (7-point) stencil, stencil order {0}, -x, +x, -y, +y,
     * -z, +z; couplings through a wall are zeroed and, for a SOLID wall,
     * the diagonal is reduced by one scale. */
    double scale = s->dt / (cfg->density * s->dx * s->dx);
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                long c = cfo_idx(s, 0, i, j, k);
                double* a = s->A + c * s->ncoef;
                a[0] = (2.0 * s->dim) * scale;
                for (int m = 1; m < s->ncoef; ++m)
                    a[m] = -1.0 * scale;
                int g[3] = {i - s->h, j - s->h, k - s->h};
                for (int d = 0; d < s->dim; ++d) {
                    if (cfg->periodic[d])
                        continue;
                    if (g[d] <= 0) {
                        a[1 + 2 * d] = 0;
                        if (cfg->wall_type[d][0] == CFO_SOLID)
                            a[0] -= scale;
                    }
                    if (g[d] > (s->n[d] - 1) - 1) {
                        a[2 + 2 * d] = 0;
                        if (cfg->wall_type[d][1] == CFO_SOLID)
                            a[0] -= scale;
                    }
                }
                /* VelocityCorrector.hpp:174-179: Jacobi M = 1 / A_diag;
                 * the identity for precon "None" (VelocityCorrector.hpp:327) */
                s->M[c] = cfg->preconditioner == CFO_PRECOND_NONE ? 1.0 : 1.0 / a[0];
            }

    *out = s;
    return CFO_OK;
}

void cfo_destroy(cfo_state* s)
{
    if (!s)
        return;
    for (int e = 0; e < 4; ++e)
        for (int v = 0; v < 2; ++v)
            free(s->f[e][v]);
    free(s->lhs);
    free(s->rhs);
    free(s->A);
    free(s->M);
    free(s->r);
    free(s->z);
    free(s->p);
    free(s->q);
    free(s->x32);
    free(s->r32);
    free(s->z32);
    free(s->p32);
    free(s->q32);
    cfo_mg_free(s);
    free(s);
}

/* ------------------------------------------------------------------------ */
/* Ghost fill: the single-rank form of Cajita's halo gather                 */
/* (ProblemManager.hpp:394-403, VelocityCorrector.hpp:236).  A wall axis    */
/* has no neighbour, so its ghosts are left untouched (always zero); a      */
/* periodic axis copies the opposite owned layers.  Axes are done in order  */
/* over the full extent of the others so that corners are filled.           */
/* ------------------------------------------------------------------------ */

void cfo_fill_ghosts_array(struct cfo_state* s, int e, double* a)
{
    const int* E = s->ext[e];
    for (int d = 0; d < s->dim; ++d) {
        if (!s->c.periodic[d])
            continue;
        int n = s->n[d], h = s->h;
        for (int k = 0; k < E[2]; ++k)
            for (int j = 0; j < E[1]; ++j)
                for (int i = 0; i < E[0]; ++i) {
                    int l[3] = {i, j, k};
                    int src;
                    if (l[d] < h)
                        src = l[d] + n;
                    else if (l[d] >= h + n)
                        src = l[d] - n;
                    else
                        continue;
                    int m[3] = {i, j, k};
                    m[d] = src;
                    a[cfo_idx(s, e, i, j, k)] =
                        a[cfo_idx(s, e, m[0], m[1], m[2])];
                }
    }
}

void cfo_fill_ghosts(cfo_state* s, int version)
{
    for (int e = 0; e < s->nent; ++e)
        cfo_fill_ghosts_array(s, e, s->f[e][version]);
}

void cfo_advance(cfo_state* s, int entity)
{
    /* ProblemManager.hpp:357-380: pointer swap of Current and Next. */
    double* t = s->f[entity][0];
    s->f[entity][0] = s->f[entity][1];
    s->f[entity][1] = t;
}

/* ------------------------------------------------------------------------ */
/* B-spline sampling (Interpolation.hpp:30-54 -> Cajita SplineData,          */
/* evaluateSpline, G2P::value).  Uniform B-spline of order 1 (linear) or 3  */
/* (cubic); logical coordinate xi = (p - x_E(0)) / (x_E(1) - x_E(0)).       */
/* ------------------------------------------------------------------------ */

void cfo_spline_weights(int order, double xi, int* s0, double* w)
{
    int fl = (int)xi;
    double t = xi - (double)fl;
    if (order == 1) {
        *s0 = fl;
        w[0] = 1.0 - t;
        w[1] = t;
        return;
    }
    const double c16 = 1.0 / 6.0;
    double omt = 1.0 - t;
    double t2 = t * t;
    double t3 = t2 * t;
    *s0 = fl - 1;
    w[0] = ((omt * omt) * omt) * c16;
    w[1] = ((3.0 * t3 - 6.0 * t2) + 4.0) * c16;
    w[2] = (((-3.0 * t3 + 3.0 * t2) + 3.0 * t) + 1.0) * c16;
    w[3] = t3 * c16;
}

/* Logical coordinate of p for entity e along axis d, clamped so that the
 * stencil stays inside the ghosted array (the reference reads out of bounds
 * there; in range the clamp is the identity). */
static inline double cfo_logical(const struct cfo_state* s, int e, int d,
                                 int order, double p)
{
    double xi = (p - s->x0[e][d]) * s->rdx[e][d];
    double lo = order == 1 ? 0.0 : 1.0;
    double hi = (double)(s->ext[e][d] - (order == 1 ? 2 : 3));
    return fmin(fmax(xi, lo), hi);
}

double cfo_interpolate(const cfo_state* s, int e, int order,
                       const double* field, const double p[3])
{
    int s0[3] = {0, 0, 0};
    double w[3][4];
    for (int d = 0; d < s->dim; ++d)
        cfo_spline_weights(order, cfo_logical(s, e, d, order, p[d]), &s0[d],
                           w[d]);
    int nk = order + 1;
    double val = 0.0;
    if (s->dim == 2) {
        for (int a = 0; a < nk; ++a)
            for (int b = 0; b < nk; ++b) {
                double wab = w[0][a] * w[1][b];
                val = fma(wab, field[cfo_idx(s, e, s0[0] + a, s0[1] + b, 0)],
                          val);
            }
    } else {
        for (int a = 0; a < nk; ++a)
            for (int b = 0; b < nk; ++b) {
                double wab = w[0][a] * w[1][b];
                for (int c = 0; c < nk; ++c) {
                    double wabc = wab * w[2][c];
                    val = fma(wabc,
                              field[cfo_idx(s, e, s0[0] + a, s0[1] + b,
                                            s0[2] + c)],
                              val);
                }
            }
    }
    return val;
}

static inline void cfo_velocity(const struct cfo_state* s, const double* u,
                                const double* v, const double* w,
                                const double x[3], double vel[3])
{
    int o = s->c.velocity_order;
    vel[0] = cfo_interpolate(s, CFO_FACE_X, o, u, x);
    vel[1] = cfo_interpolate(s, CFO_FACE_Y, o, v, x);
    vel[2] = s->dim == 3 ? cfo_interpolate(s, CFO_FACE_Z, o, w, x) : 0.0;
}

/* TimeIntegrator.hpp:36-77.  With CFO_QUIRK_RK3_V0 the 3/4-step probe is
 * x0 - 3/4 dt v0 (as written); otherwise x0 - 3/4 dt v1 (Ralston RK3). */
void cfo_rk3(const cfo_state* s, const double x0[3], const double* u,
             const double* v, const double* w, double dt, double trace[3])
{
    double v0[3], v1[3], v2[3], x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0};
    int D = s->dim;
    cfo_velocity(s, u, v, w, x0, v0);
    for (int d = 0; d < D; ++d)
        x1[d] = x0[d] - 0.5 * dt * v0[d];
    cfo_velocity(s, u, v, w, x1, v1);
    const double* vq = (s->c.quirks & CFO_QUIRK_RK3_V0) ? v0 : v1;
    for (int d = 0; d < D; ++d)
        x2[d] = x0[d] - 0.75 * dt * vq[d];
    cfo_velocity(s, u, v, w, x2, v2);
    for (int d = 0; d < D; ++d)
        trace[d] = x0[d] - dt * ((2.0 / 9.0) * v0[d] + (3.0 / 9.0) * v1[d] +
                                 (4.0 / 9.0) * v2[d]);
    if (D == 2)
        trace[2] = 0.0;
}

/* TimeIntegrator.hpp:81-116: per owned entity, trace back and sample the
 * current field with the field order; Next ghosts are never written. */
void cfo_advect_entity(struct cfo_state* s, int e, const double* const* cur,
                       double* next)
{
    int lo[3], hi[3];
    cfo_owned_range(s, e, lo, hi);
    const double* u = cur[1];
    const double* v = cur[2];
    const double* w = s->dim == 3 ? cur[3] : NULL;
    const double* field = cur[e];
    double dt = s->dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                int l[3] = {i, j, k};
                double x0[3] = {0, 0, 0}, tr[3];
                for (int d = 0; d < s->dim; ++d)
                    x0[d] = cfo_coord(s, e, d, l[d]);
                cfo_rk3(s, x0, u, v, w, dt, tr);
                next[cfo_idx(s, e, i, j, k)] =
                    cfo_interpolate(s, e, s->c.field_order, field, tr);
            }
}

/* TimeIntegrator.hpp:120-177: gather, advect every entity into Next, then
 * swap Current and Next. */
void cfo_time_integrate(cfo_state* s)
{
    cfo_fill_ghosts(s, CFO_CURRENT);
    const double* cur[4] = {s->f[0][0], s->f[1][0], s->f[2][0],
                            s->dim == 3 ? s->f[3][0] : NULL};
    for (int e = 0; e < s->nent; ++e)
        cfo_advect_entity(s, e, cur, s->f[e][1]);
    for (int e = 0; e < s->nent; ++e)
        cfo_advance(s, e);
}

/* ------------------------------------------------------------------------ */
/* Inputs: Solver.hpp:181-263 with InflowSource.hpp:33-78 (half-open box),  */
/* BodyForce.hpp:44-60 and BoundaryConditions.hpp:102-129.                  */
/* ------------------------------------------------------------------------ */

int cfo_in_source(const struct cfo_state* s, const double x[3])
{
    const cfo_config* c = &s->c;
    if (c->inflow_shape == CFO_INFLOW_BOX) {
        for (int d = 0; d < s->dim; ++d) {
            double lo = c->inflow_location[d];
            double hi = c->inflow_location[d] + c->inflow_size[d];
            if (!(x[d] >= lo && x[d] < hi))
                return 0;
        }
        return 1;
    }
    if (c->inflow_shape == CFO_INFLOW_SPHERE) {
        double d2 = 0.0;
        for (int d = 0; d < s->dim; ++d) {
            double r = x[d] - c->inflow_center[d];
            d2 = d2 + r * r;
        }
        return d2 < c->inflow_radius * c->inflow_radius;
    }
    return 0;
}

void cfo_add_inputs(cfo_state* s)
{
    const cfo_config* c = &s->c;
    double dt = s->dt;
    for (int e = 0; e < s->nent; ++e) {
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        double* a = s->f[e][0];
        int d = e - 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int l[3] = {i, j, k};
                    double x[3] = {0, 0, 0};
                    for (int m = 0; m < s->dim; ++m)
                        x[m] = cfo_coord(s, e, m, l[m]);
                    long id = cfo_idx(s, e, i, j, k);
                    double val = a[id];
                    int in = cfo_in_source(s, x);
                    if (e == CFO_CELL) {
                        if (in && val < c->inflow_quantity)
                            val = c->inflow_quantity;
                    } else {
                        double U = c->inflow_velocity[d];
                        if (in && fabs(val) < fabs(U))
                            val = U;
                        val = val + c->body_force[d] * dt;
                        if (!c->periodic[d]) {
                            int g = l[d] - s->h;
                            if (g <= 0 && c->wall_type[d][0] == CFO_SOLID)
                                val = 0;
                            if (g > s->n[d] - 1 &&
                                c->wall_type[d][1] == CFO_SOLID)
                                val = 0;
                        }
                    }
                    a[id] = val;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Projection pieces: VelocityCorrector.hpp:182-264.                         */
/* ------------------------------------------------------------------------ */

void cfo_divergence(struct cfo_state* s, const double* u, const double* v,
                    const double* w, double* rhs)
{
    double scale = 1.0 / s->dx;
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                double t = u[cfo_idx(s, 1, i + 1, j, k)] -
                           u[cfo_idx(s, 1, i, j, k)];
                t = t + v[cfo_idx(s, 2, i, j + 1, k)];
                t = t - v[cfo_idx(s, 2, i, j, k)];
                if (s->dim == 3) {
                    t = t + w[cfo_idx(s, 3, i, j, k + 1)];
                    t = t - w[cfo_idx(s, 3, i, j, k)];
                }
                rhs[cfo_idx(s, 0, i, j, k)] = -scale * t;
            }
}

void cfo_build_rhs(cfo_state* s)
{
    /* VelocityCorrector.hpp:182-212: gather, rhs = 0 (owned), divergence. */
    cfo_fill_ghosts(s, CFO_CURRENT);
    cfo_divergence(s, s->f[1][0], s->f[2][0], s->dim == 3 ? s->f[3][0] : NULL,
                   s->rhs);
}

void cfo_remove_mean(cfo_state* s, double* v)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    cfo_dd acc = {0, 0};
    long cnt = 0;
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                cfo_dd_add(&acc, v[cfo_idx(s, 0, i, j, k)], 0.0);
                ++cnt;
            }
    double mean = cfo_dd_value(acc) / (double)cnt;
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                long c = cfo_idx(s, 0, i, j, k);
                v[c] = v[c] - mean;
            }
}

void cfo_rhs_remove_mean(cfo_state* s) { cfo_remove_mean(s, s->rhs); }

int cfo_operator_singular(const struct cfo_state* s)
{
    for (int d = 0; d < s->dim; ++d)
        if (!s->c.periodic[d] && (s->c.wall_type[d][0] == CFO_FREE ||
                                  s->c.wall_type[d][1] == CFO_FREE))
            return 0;
    return 1;
}

void cfo_apply_pressure_arrays(struct cfo_state* s, double* u, double* v,
                               double* w, const double* p)
{
    const cfo_config* c = &s->c;
    double scale = s->dt / (c->density * s->dx);
    double* vel[3] = {u, v, w};
    for (int d = 0; d < s->dim; ++d) {
        int e = d + 1;
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        double* a = vel[d];
        /* Which array receives this loop's wall rule: the loop's own
         * component, or u for every loop after the first when the
         * as-written behaviour (VelocityCorrector.hpp:260) is requested. */
        double* bc_arr = (d > 0 && (c->quirks & CFO_QUIRK_PROJ_BC_U)) ? u : a;
        int bc_e = (bc_arr == u) ? 1 : e;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int l[3] = {i, j, k};
                    int m[3] = {i, j, k};
                    m[d] -= 1;
                    long id = cfo_idx(s, e, i, j, k);
                    double dp = p[cfo_idx(s, 0, i, j, k)] -
                                p[cfo_idx(s, 0, m[0], m[1], m[2])];
                    a[id] = a[id] - scale * dp;
                    if (!c->periodic[d]) {
                        int g = l[d] - s->h;
                        long ib = cfo_idx(s, bc_e, i, j, k);
                        if (g <= 0 && c->wall_type[d][0] == CFO_SOLID)
                            bc_arr[ib] = 0;
                        if (g > s->n[d] - 1 && c->wall_type[d][1] == CFO_SOLID)
                            bc_arr[ib] = 0;
                    }
                }
    }
}

void cfo_apply_pressure(cfo_state* s)
{
    cfo_fill_ghosts_array(s, 0, s->lhs);
    cfo_apply_pressure_arrays(s, s->f[1][0], s->f[2][0],
                              s->dim == 3 ? s->f[3][0] : NULL, s->lhs);
}

/* ------------------------------------------------------------------------ */
/* Stored-stencil SpMV and Jacobi-PCG (Cajita ReferenceConjugateGradient,   */
/* called at VelocityCorrector.hpp:276; tolerance/max-iter from the config; */
/* control flow as restated in the header of cfo_pcg64).                    */
/* ------------------------------------------------------------------------ */

void cfo_spmv(struct cfo_state* s, const double* x, double* y)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    long sy = s->ext[0][0];
    long sz = (long)s->ext[0][0] * s->ext[0][1];
    int nc = s->ncoef;
    int D = s->dim;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                long c = cfo_idx(s, 0, i, j, k);
                const double* a = s->A + c * nc;
                double acc = a[0] * x[c];
                acc = fma(a[1], x[c - 1], acc);
                acc = fma(a[2], x[c + 1], acc);
                acc = fma(a[3], x[c - sy], acc);
                acc = fma(a[4], x[c + sy], acc);
                if (D == 3) {
                    acc = fma(a[5], x[c - sz], acc);
                    acc = fma(a[6], x[c + sz], acc);
                }
                y[c] = acc;
            }
}

void cfo_apply_A(cfo_state* s, const double* x, double* y)
{
    cfo_spmv(s, x, y);
}

double cfo_dot_dd(const struct cfo_state* s, const double* a, const double* b)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    cfo_dd total = {0, 0};
#pragma omp parallel
    {
        cfo_dd acc = {0, 0};
#pragma omp for collapse(2) schedule(static) nowait
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    long c = cfo_idx(s, 0, i, j, k);
                    cfo_dd_add_prod(&acc, a[c], b[c]);
                }
#pragma omp critical
        cfo_dd_add(&total, acc.hi, acc.lo);
    }
    return cfo_dd_value(total);
}

double cfo_dot(const cfo_state* s, const double* a, const double* b)
{
    return cfo_dot_dd(s, a, b);
}

static double cfo_dot32(const struct cfo_state* s, const float* a,
                        const float* b)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    cfo_dd total = {0, 0};
#pragma omp parallel
    {
        cfo_dd acc = {0, 0};
#pragma omp for collapse(2) schedule(static) nowait
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    long c = cfo_idx(s, 0, i, j, k);
                    cfo_dd_add_prod(&acc, (double)a[c], (double)b[c]);
                }
#pragma omp critical
        cfo_dd_add(&total, acc.hi, acc.lo);
    }
    return cfo_dd_value(total);
}


/*
 * Jacobi-PCG, as the oracle defines Cajita's ReferenceConjugateGradient:
 *   r = b - A x; z = M r; p = z; rho = r.z; bnorm = ||b||
 *   (bnorm == 0: x = 0, converged in 0 iterations)
 *   for k = 1..maxit:
 *     q = A p; alpha = rho / (p.q)
 *     x = fma(alpha, p, x); r = fma(-alpha, q, r)
 *     if sqrt(r.r) / bnorm < tol: stop
 *     z = M r; rho' = r.z; beta = rho' / rho; p = fma(beta, p, z); rho = rho'
 * The stored matrix and the separate z pass keep the reference's traffic.
 */
int cfo_pcg64(struct cfo_state* s, const double* b, double* x, double tol,
              int maxit, cfo_stats* st)
{
    double *r = s->r, *z = s->z, *p = s->p, *q = s->q;
    const double* M = s->M;
    int periodic = s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2];
    double bnorm = sqrt(cfo_dot_dd(s, b, b));
    if (st) {
        st->b_norm = bnorm;
        st->outer_iterations = 0;
    }
    if (bnorm == 0.0) {
        CFO_FOR_OWNED_CELLS(s, x[c] = 0.0);
        if (st) {
            st->converged = 1;
            st->iterations = 0;
            st->final_residual = 0.0;
        }
        return CFO_OK;
    }
    if (periodic)
        cfo_fill_ghosts_array(s, 0, x);
    cfo_spmv(s, x, q);
    CFO_FOR_OWNED_CELLS(s, r[c] = b[c] - q[c]);
    CFO_FOR_OWNED_CELLS(s, z[c] = M[c] * r[c]; p[c] = z[c]);
    double rho = cfo_dot_dd(s, r, z);
    double res = 1.0;
    int k = 0, conv = 0;
    while (k < maxit) {
        ++k;
        if (periodic)
            cfo_fill_ghosts_array(s, 0, p);
        cfo_spmv(s, p, q);
        double pq = cfo_dot_dd(s, p, q);
        double alpha = rho / pq;
        CFO_FOR_OWNED_CELLS(s, x[c] = fma(alpha, p[c], x[c]);
                            r[c] = fma(-alpha, q[c], r[c]));
        double rr = cfo_dot_dd(s, r, r);
        res = sqrt(rr) / bnorm;
        if (res < tol) {
            conv = 1;
            break;
        }
        CFO_FOR_OWNED_CELLS(s, z[c] = M[c] * r[c]);
        double rho_new = cfo_dot_dd(s, r, z);
        double beta = rho_new / rho;
        CFO_FOR_OWNED_CELLS(s, p[c] = fma(beta, p[c], z[c]));
        rho = rho_new;
    }
    if (st) {
        st->converged = conv;
        st->iterations = k;
        st->final_residual = res;
    }
    return conv ? CFO_OK : CFO_ENOCONV;
}

/*
 * FP32-inner / FP64-outer defect correction (build extension).
 *   x = 0; r64 = b - A x; bnorm = ||b||
 *   loop: res = ||r64|| / bnorm; stop if res < tol
 *         r32 = (float) r64
 *         inner Jacobi-PCG in FP32 vectors and FP32 stencil from d = 0 to
 *           ||r_inner|| / ||r32|| < inner_tol; dots, alpha, beta in FP64
 *           (alpha32 = (float)alpha, beta32 = (float)beta)
 *         x = x + (double) d; r64 = b - A x
 */
void cfo_spmv32(struct cfo_state* s, const float* x, float* y)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    long sy = s->ext[0][0];
    long sz = (long)s->ext[0][0] * s->ext[0][1];
    int nc = s->ncoef;
    int D = s->dim;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                long c = cfo_idx(s, 0, i, j, k);
                const double* a = s->A + c * nc;
                float acc = (float)a[0] * x[c];
                acc = fmaf((float)a[1], x[c - 1], acc);
                acc = fmaf((float)a[2], x[c + 1], acc);
                acc = fmaf((float)a[3], x[c - sy], acc);
                acc = fmaf((float)a[4], x[c + sy], acc);
                if (D == 3) {
                    acc = fmaf((float)a[5], x[c - sz], acc);
                    acc = fmaf((float)a[6], x[c + sz], acc);
                }
                y[c] = acc;
            }
}

void cfo_fill_ghosts_f32(struct cfo_state* s, float* a)
{
    const int* E = s->ext[0];
    for (int d = 0; d < s->dim; ++d) {
        if (!s->c.periodic[d])
            continue;
        int n = s->n[d], h = s->h;
        for (int k = 0; k < E[2]; ++k)
            for (int j = 0; j < E[1]; ++j)
                for (int i = 0; i < E[0]; ++i) {
                    int l[3] = {i, j, k};
                    int src;
                    if (l[d] < h)
                        src = l[d] + n;
                    else if (l[d] >= h + n)
                        src = l[d] - n;
                    else
                        continue;
                    int m[3] = {i, j, k};
                    m[d] = src;
                    a[cfo_idx(s, 0, i, j, k)] =
                        a[cfo_idx(s, 0, m[0], m[1], m[2])];
                }
    }
}

/* Relative FP32 reduction that ends a defect-correction pass started from a
 * residual of norm rref: inner_tol, or less when half the final target
 * (0.5 tol ||b||) is already closer, so the last pass stops near the target
 * instead of a full inner_tol past it. */
static double cfo_pass_tol(const cfo_config* c0, double bnorm, double rref)
{
    return fmax(c0->inner_tol, (0.5 * c0->cg_tol * bnorm) / rref);
}

static int cfo_pcg32_inner(struct cfo_state* s, const cfo_config* c0,
                           double bnorm, int maxit)
{
    float *d = s->x32, *r = s->r32, *z = s->z32, *p = s->p32, *q = s->q32;
    const double* M = s->M;
    int periodic = s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2];
    CFO_FOR_OWNED_CELLS(s, d[c] = 0.0f; z[c] = (float)M[c] * r[c];
                        p[c] = z[c]);
    double rho = cfo_dot32(s, r, z);
    double r0 = sqrt(cfo_dot32(s, r, r));
    if (r0 == 0.0)
        return 0;
    double tol = cfo_pass_tol(c0, bnorm, r0);
    int k = 0;
    while (k < maxit) {
        ++k;
        if (periodic)
            cfo_fill_ghosts_f32(s, p);
        cfo_spmv32(s, p, q);
        double pq = cfo_dot32(s, p, q);
        double alpha = rho / pq;
        float a32 = (float)alpha;
        CFO_FOR_OWNED_CELLS(s, d[c] = fmaf(a32, p[c], d[c]);
                            r[c] = fmaf(-a32, q[c], r[c]));
        double rr = cfo_dot32(s, r, r);
        if (sqrt(rr) / r0 < tol)
            break;
        CFO_FOR_OWNED_CELLS(s, z[c] = (float)M[c] * r[c]);
        double rho_new = cfo_dot32(s, r, z);
        double beta = rho_new / rho;
        float b32 = (float)beta;
        CFO_FOR_OWNED_CELLS(s, p[c] = fmaf(b32, p[c], z[c]));
        rho = rho_new;
    }
    return k;
}

static int cfo_pcg_refined(struct cfo_state* s, const double* b, double* x,
                           cfo_stats* st)
{
    const cfo_config* c0 = &s->c;
    int periodic = c0->periodic[0] || c0->periodic[1] || c0->periodic[2];
    double* r64 = s->r;
    double* y = s->q;
    double bnorm = sqrt(cfo_dot_dd(s, b, b));
    st->b_norm = bnorm;
    st->iterations = 0;
    st->outer_iterations = 0;
    if (bnorm == 0.0) {
        CFO_FOR_OWNED_CELLS(s, x[c] = 0.0);
        st->converged = 1;
        st->final_residual = 0.0;
        return CFO_OK;
    }
    if (periodic)
        cfo_fill_ghosts_array(s, 0, x);
    cfo_spmv(s, x, y);
    CFO_FOR_OWNED_CELLS(s, r64[c] = b[c] - y[c]);
    int inner_max = c0->inner_max_iter > 0 ? c0->inner_max_iter : c0->cg_max_iter;
    for (;;) {
        double res = sqrt(cfo_dot_dd(s, r64, r64)) / bnorm;
        st->final_residual = res;
        if (res < c0->cg_tol) {
            st->converged = 1;
            return CFO_OK;
        }
        if (st->outer_iterations >= c0->max_outer ||
            st->iterations >= c0->cg_max_iter) {
            st->converged = 0;
            return CFO_ENOCONV;
        }
        float* r32 = s->r32;
        CFO_FOR_OWNED_CELLS(s, r32[c] = (float)r64[c]);
        st->iterations += cfo_pcg32_inner(s, c0, bnorm, inner_max);
        float* d = s->x32;
        CFO_FOR_OWNED_CELLS(s, x[c] = x[c] + (double)d[c]);
        if (periodic)
            cfo_fill_ghosts_array(s, 0, x);
        cfo_spmv(s, x, y);
        CFO_FOR_OWNED_CELLS(s, r64[c] = b[c] - y[c]);
        st->outer_iterations += 1;
    }
}

/*
 * Defect correction with the search direction carried across passes (a
 * reliable-update form of the same FP32-inner / FP64-outer scheme):
 *   x = 0; bnorm = ||b||; r32 = (float) b; d = 0; rref = ||r32||
 *   rho = r32.(M32 r32); one FP32 Jacobi-PCG run, iteration k:
 *     p = M32 r (k = 1) or fmaf(beta32, p, M32 r); q = A32 p
 *     alpha = rho / p.q; d = fmaf(a32, p, d); r = fmaf(-a32, q, r)
 *     if ||r|| / rref < inner_tol:                   (pass boundary)
 *        x = x + (double) d; d = 0; r64 = b - A x; res = ||r64|| / bnorm
 *        stop if res < tol (or max_outer passes)
 *        r = (float) r64; rref = ||r||; rz = r.(M32 r)
 *     else rz = r.(M32 r)
 *     beta = rz / rho; rho = rz
 */
static int cfo_pcg_carry(struct cfo_state* s, const double* b, double* x,
                         cfo_stats* st)
{
    const cfo_config* c0 = &s->c;
    int periodic = c0->periodic[0] || c0->periodic[1] || c0->periodic[2];
    float *d = s->x32, *r = s->r32, *z = s->z32, *p = s->p32, *q = s->q32;
    const double* M = s->M;
    double* r64 = s->r;
    double* y = s->q;
    double bnorm = sqrt(cfo_dot_dd(s, b, b));
    st->b_norm = bnorm;
    st->iterations = 0;
    st->outer_iterations = 0;
    CFO_FOR_OWNED_CELLS(s, x[c] = 0.0);
    if (bnorm == 0.0) {
        st->converged = 1;
        st->final_residual = 0.0;
        return CFO_OK;
    }
    st->final_residual = 1.0;
    CFO_FOR_OWNED_CELLS(s, r[c] = (float)b[c]; d[c] = 0.0f;
                        z[c] = (float)M[c] * r[c]);
    double rref = sqrt(cfo_dot32(s, r, r));
    double rho = cfo_dot32(s, r, z);
    double beta = 0.0;
    int k = 0;
    for (;;) {
        if (k >= c0->cg_max_iter) {
            st->converged = 0;
            st->iterations = k;
            return CFO_ENOCONV;
        }
        ++k;
        if (k == 1) {
            CFO_FOR_OWNED_CELLS(s, p[c] = z[c]);
        } else {
            float b32 = (float)beta;
            CFO_FOR_OWNED_CELLS(s, p[c] = fmaf(b32, p[c], z[c]));
        }
        if (periodic)
            cfo_fill_ghosts_f32(s, p);
        cfo_spmv32(s, p, q);
        double alpha = rho / cfo_dot32(s, p, q);
        float a32 = (float)alpha;
        CFO_FOR_OWNED_CELLS(s, d[c] = fmaf(a32, p[c], d[c]);
                            r[c] = fmaf(-a32, q[c], r[c]));
        double rr = cfo_dot32(s, r, r);
        if (sqrt(rr) / rref < cfo_pass_tol(c0, bnorm, rref)) {
            CFO_FOR_OWNED_CELLS(s, x[c] = x[c] + (double)d[c]; d[c] = 0.0f);
            if (periodic)
                cfo_fill_ghosts_array(s, 0, x);
            cfo_spmv(s, x, y);
            CFO_FOR_OWNED_CELLS(s, r64[c] = b[c] - y[c]);
            double res = sqrt(cfo_dot_dd(s, r64, r64)) / bnorm;
            st->outer_iterations += 1;
            st->final_residual = res;
            st->iterations = k;
            if (res < c0->cg_tol) {
                st->converged = 1;
                return CFO_OK;
            }
            if (st->outer_iterations >= c0->max_outer) {
                st->converged = 0;
                return CFO_ENOCONV;
            }
            CFO_FOR_OWNED_CELLS(s, r[c] = (float)r64[c]);
            rref = sqrt(cfo_dot32(s, r, r));
        }
        CFO_FOR_OWNED_CELLS(s, z[c] = (float)M[c] * r[c]);
        double rz = cfo_dot32(s, r, z);
        beta = rz / rho;
        rho = rz;
    }
}

int cfo_pressure_solve(cfo_state* s, int precision, cfo_stats* st)
{
    cfo_stats local;
    if (!st)
        st = &local;
    memset(st, 0, sizeof(*st));
    /* VelocityCorrector.hpp:272: lhs = 0 before every solve (no warm start). */
    CFO_FOR_OWNED_CELLS(s, s->lhs[c] = 0.0);
    if (precision == CFO_FP32_REFINED && s->c.preconditioner == CFO_PRECOND_MG)
        return cfo_pcg_mg(s, s->rhs, s->lhs, s->c.cg_tol, s->c.cg_max_iter, st, 1);
    if (precision == CFO_FP32_REFINED) {
        if (!s->x32) {
            long nc = s->size[0];
            s->x32 = (float*)cfo_zalloc(nc, sizeof(float));
            s->r32 = (float*)cfo_zalloc(nc, sizeof(float));
            s->z32 = (float*)cfo_zalloc(nc, sizeof(float));
            s->p32 = (float*)cfo_zalloc(nc, sizeof(float));
            s->q32 = (float*)cfo_zalloc(nc, sizeof(float));
            if (!s->x32 || !s->r32 || !s->z32 || !s->p32 || !s->q32)
                return CFO_ENOMEM;
        }
        int rc = s->c.refine_mode == CFO_REFINE_CARRY
                     ? cfo_pcg_carry(s, s->rhs, s->lhs, st)
                     : cfo_pcg_refined(s, s->rhs, s->lhs, st);
        if (s->c.preconditioner == CFO_PRECOND_NONE && cfo_operator_singular(s))
            cfo_remove_mean(s, s->lhs);
        return rc;
    }
    return cfo_solve64(s, s->rhs, s->lhs, s->c.cg_tol, s->c.cg_max_iter, st);
}

/* Without a preconditioner, CG from x = 0 stays in the range of A, so its
 * map is A^+ and the solution has zero mean up to rounding; on a singular
 * operator the mean is removed explicitly, as after the multigrid solve,
 * so that the map's transpose is exactly a mean removal and the same solve. */
int cfo_solve64(struct cfo_state* s, const double* b, double* x, double tol,
                int maxit, cfo_stats* st)
{
    if (s->c.preconditioner == CFO_PRECOND_MG)
        return cfo_pcg_mg(s, b, x, tol, maxit, st, 0);
    int rc = cfo_pcg64(s, b, x, tol, maxit, st);
    if (s->c.preconditioner == CFO_PRECOND_NONE && cfo_operator_singular(s))
        cfo_remove_mean(s, x);
    return rc;
}

/* VelocityCorrector.hpp:266-282.  A solve that does not converge aborts the
 * correction before the projection (the reference's solver throws). */
int cfo_correct_velocity(cfo_state* s, cfo_stats* st)
{
    cfo_build_rhs(s);
    int rc = cfo_pressure_solve(s, s->c.precision, st);
    if (rc != CFO_OK)
        return rc;
    cfo_apply_pressure(s);
    return CFO_OK;
}

/* Solver.hpp:125-133. */
int cfo_setup(cfo_state* s, cfo_stats* st)
{
    cfo_add_inputs(s);
    return cfo_correct_velocity(s, st);
}

/* Solver.hpp:135-147. */
int cfo_step(cfo_state* s, cfo_stats* st)
{
    cfo_time_integrate(s);
    cfo_add_inputs(s);
    int rc = cfo_correct_velocity(s, st);
    s->time += s->dt;
    return rc;
}

int cfo_run(cfo_state* s, int nsteps, cfo_stats* last)
{
    for (int n = 0; n < nsteps; ++n) {
        int rc = cfo_step(s, last);
        if (rc != CFO_OK)
            return rc;
    }
    return CFO_OK;
}

/* ------------------------------------------------------------------------ */
/* Seeded inputs.  Counter-based SplitMix64 (Steele, Lea, Flood 2014;        */
/* finaliser constants as published by Vigna), uniform = top 53 bits,      */
/* normal = Box-Muller.                                                     */
/* ------------------------------------------------------------------------ */

static uint64_t cfo_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t cfo_rand_u64(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t key = cfo_mix64(seed + 0x632BE59BD9B4E019ULL * (stream + 1));
    return cfo_mix64(key + idx);
}

double cfo_rand_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return (double)(cfo_rand_u64(seed, stream, idx) >> 11) * 0x1.0p-53;
}

double cfo_rand_normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = 1.0 - cfo_rand_uniform(seed, stream, 2 * idx);
    double u2 = cfo_rand_uniform(seed, stream, 2 * idx + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Separation along axis d, wrapped to the nearest image on periodic axes. */
static double cfo_sep(const struct cfo_state* s, int d, double x, double c)
{
    double r = x - c;
    if (s->c.periodic[d]) {
        double L = s->c.hi[d] - s->c.lo[d];
        r = r - L * nearbyint(r / L);
    }
    return r;
}

/*
 * Divergence-free initial velocity: discrete curl of an edge-centred vector
 * potential psi = sum_k Gamma_k exp(-d_k^2 / (2 sigma_k^2)) e_k, one Gaussian
 * vortex tube per axis (3-D) or three point vortices of a nodal stream
 * function (2-D), seeded; rescaled so that max |u| over owned faces = peak.
 * Parameters (stream 0, draws in order per tube): a, b in [0.3, 0.7] of the
 * box (the tube's position in the two other axes, ascending), sigma in
 * [0.05, 0.1], sign of Gamma from a draw < 0.5.
 */
static void cfo_tube_params(uint64_t seed, int t, double* a, double* b,
                            double* sig, double* gam)
{
    *a = 0.3 + 0.4 * cfo_rand_uniform(seed, 0, 4 * t + 0);
    *b = 0.3 + 0.4 * cfo_rand_uniform(seed, 0, 4 * t + 1);
    *sig = 0.05 + 0.05 * cfo_rand_uniform(seed, 0, 4 * t + 2);
    *gam = cfo_rand_uniform(seed, 0, 4 * t + 3) < 0.5 ? -1.0 : 1.0;
}

/* psi_k at the point (x[0], x[1], x[2]) (only the two axes != k matter). */
static double cfo_psi(const struct cfo_state* s, uint64_t seed, int k,
                      const double x[3])
{
    int o1 = k == 0 ? 1 : 0;
    int o2 = k == 2 ? 1 : 2;
    double sum = 0.0;
    int ntube = s->dim == 3 ? 1 : 3;
    for (int t = 0; t < ntube; ++t) {
        int tube = s->dim == 3 ? k : t;
        double a, b, sig, gam;
        cfo_tube_params(seed, tube, &a, &b, &sig, &gam);
        double ca = s->c.lo[o1] + a * (s->c.hi[o1] - s->c.lo[o1]);
        double cb = s->c.lo[o2] + b * (s->c.hi[o2] - s->c.lo[o2]);
        double r1 = cfo_sep(s, o1, x[o1], ca);
        double r2 = cfo_sep(s, o2, x[o2], cb);
        double d2 = r1 * r1 + r2 * r2;
        sum = sum + gam * exp(-d2 / (2.0 * sig * sig));
    }
    return sum;
}

static double cfo_node(const struct cfo_state* s, int d, int l)
{
    return s->low[d] + (double)l * s->dx;
}

static double cfo_centre(const struct cfo_state* s, int d, int l)
{
    return s->low[d] + ((double)l + 0.5) * s->dx;
}

void cfo_init_vortex_tubes(cfo_state* s, uint64_t seed, double peak)
{
    double dx = s->dx;
    double vmax = 0.0;
    for (int e = 1; e < s->nent; ++e) {
        int d = e - 1;
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        double* a = s->f[e][0];
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int l[3] = {i, j, k};
                    double val;
                    if (s->dim == 2) {
                        /* nodal stream function: u = dpsi/dy, v = -dpsi/dx */
                        int t = d == 0 ? 1 : 0;
                        double xa[3] = {cfo_node(s, 0, i), cfo_node(s, 1, j), 0};
                        double xb[3] = {xa[0], xa[1], 0};
                        xb[t] = cfo_node(s, t, l[t] + 1);
                        double diff = cfo_psi(s, seed, 2, xb) -
                                      cfo_psi(s, seed, 2, xa);
                        val = d == 0 ? diff / dx : -diff / dx;
                    } else {
                        /* (curl psi)_d = dpsi_{d2}/dx_{d1} - dpsi_{d1}/dx_{d2},
                         * (d, d1, d2) cyclic; psi_m lives on edges parallel
                         * to m: cell centre along m, nodes across it. */
                        int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
                        double xa[3], xb[3];
                        xa[d] = xb[d] = cfo_node(s, d, l[d]);
                        xa[d2] = xb[d2] = cfo_centre(s, d2, l[d2]);
                        xa[d1] = cfo_node(s, d1, l[d1]);
                        xb[d1] = cfo_node(s, d1, l[d1] + 1);
                        double t1 = (cfo_psi(s, seed, d2, xb) -
                                     cfo_psi(s, seed, d2, xa)) /
                                    dx;
                        xa[d1] = xb[d1] = cfo_centre(s, d1, l[d1]);
                        xa[d2] = cfo_node(s, d2, l[d2]);
                        xb[d2] = cfo_node(s, d2, l[d2] + 1);
                        double t2 = (cfo_psi(s, seed, d1, xb) -
                                     cfo_psi(s, seed, d1, xa)) /
                                    dx;
                        val = t1 - t2;
                    }
                    a[cfo_idx(s, e, i, j, k)] = val;
                    if (fabs(val) > vmax)
                        vmax = fabs(val);
                }
    }
    if (vmax > 0) {
        double sc = peak / vmax;
        for (int e = 1; e < s->nent; ++e) {
            int lo[3], hi[3];
            cfo_owned_range(s, e, lo, hi);
            double* a = s->f[e][0];
            for (int k = lo[2]; k < hi[2]; ++k)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int i = lo[0]; i < hi[0]; ++i) {
                        long id = cfo_idx(s, e, i, j, k);
                        a[id] = a[id] * sc;
                    }
        }
    }
}

/* Sum of nblobs seeded Gaussian blobs at owned cell centres (stream 1):
 * centre in [0.1, 0.9] of the box per axis, sigma in [0.03, 0.08],
 * amplitude in [0.5, 1]. */
void cfo_init_blobs(cfo_state* s, double* out, uint64_t seed, int nblobs)
{
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                int l[3] = {i, j, k};
                double x[3] = {0, 0, 0};
                for (int m = 0; m < s->dim; ++m)
                    x[m] = cfo_coord(s, 0, m, l[m]);
                double sum = 0.0;
                for (int b = 0; b < nblobs; ++b) {
                    uint64_t base = (uint64_t)b * 8;
                    double d2 = 0.0;
                    for (int m = 0; m < s->dim; ++m) {
                        double cu = 0.1 + 0.8 * cfo_rand_uniform(seed, 1, base + m);
                        double cm = s->c.lo[m] + cu * (s->c.hi[m] - s->c.lo[m]);
                        double r = cfo_sep(s, m, x[m], cm);
                        d2 = d2 + r * r;
                    }
                    double sig = 0.03 + 0.05 * cfo_rand_uniform(seed, 1, base + 3);
                    double amp = 0.5 + 0.5 * cfo_rand_uniform(seed, 1, base + 4);
                    sum = sum + amp * exp(-d2 / (2.0 * sig * sig));
                }
                out[cfo_idx(s, 0, i, j, k)] = sum;
            }
}

/* i.i.d. N(0,1) on every owned face; stream = entity, counter = the face's
 * linear index in its ghosted array. */
void cfo_init_random_faces(cfo_state* s, uint64_t seed)
{
    for (int e = 1; e < s->nent; ++e) {
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        double* a = s->f[e][0];
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    long id = cfo_idx(s, e, i, j, k);
                    a[id] = cfo_rand_normal(seed, (uint64_t)e, (uint64_t)id);
                }
    }
}

/* SiloWriter::writeFile's data (SiloWriter.hpp:141-197): the owned cells of
 * q (Current) and, at each owned cell centre, the velocity from linear
 * (order 1) interpolation of each face component (interpolateVelocity<2, 1>,
 * Interpolation.hpp:45-54), as compact arrays with x fastest (the writer's
 * LayoutLeft order).  vel_out holds dim arrays back to back. */
void cfo_cell_output(const cfo_state* s, double* q_out, double* vel_out)
{
    int lo[3], hi[3];
    cfo_owned_range(s, CFO_CELL, lo, hi);
    const int n0 = hi[0] - lo[0], n1 = hi[1] - lo[1], n2 = hi[2] - lo[2];
    const long N = (long)n0 * n1 * n2;
    const double* q = s->f[CFO_CELL][0];
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                long o = (long)(i - lo[0]) + (long)n0 * ((j - lo[1]) + (long)n1 * (k - lo[2]));
                int l[3] = {i, j, k};
                double x[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < s->dim; ++d)
                    x[d] = cfo_coord(s, CFO_CELL, d, l[d]);
                q_out[o] = q[cfo_idx(s, CFO_CELL, i, j, k)];
                for (int d = 0; d < s->dim; ++d)
                    vel_out[d * N + o] = cfo_interpolate(s, d + 1, 1, s->f[d + 1][0], x);
            }
}
