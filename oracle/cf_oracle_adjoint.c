/*
 * cf_oracle_adjoint.c -- reverse mode of the oracle's setup + steps.
 *
 * TEST INFRASTRUCTURE ONLY: the FP64 reference gradient the HIP adjoint is
 * checked against; itself checked against central finite differences of the
 * oracle's forward map.  The reference has no adjoint: this is the transpose
 * of the forward operators of cf_oracle.c.
 *
 *   setup:   S_in --inputs--> . --project--> S_0
 *   step n:  S_n --gather--> . --advect--> S_a --inputs--> . --project--> S_{n+1}
 *   loss:    dL/dq(S_T) supplied (owned cells)
 *
 * Backward of a step:
 *   project^T  wall/quirk mask; cell adjoint -s G^T (pressure gather folded);
 *              z = the forward PCG applied to it (see below); faces get
 *              -(1/dx) D^T z (the divergence gather folded)
 *   inputs^T   zero where the inflow overwrote a value or a wall zeroed it;
 *              the overwritten adjoint goes to dL/d(inflow velocity)
 *   advect^T   the field sample at the RK3 trace scattered with the spline
 *              weights, and the chain back through the three velocity
 *              samples using the weights' derivatives (clamped coordinates
 *              have zero derivative); the cell-state gather folded
 *
 * Pressure solve.  Jacobi-PCG from x = 0 on a consistent system converges
 * to x = L b with L = M^1/2 (M^1/2 A M^1/2)^+ M^1/2, M = diag(A)^-1: a
 * symmetric linear map, so its transpose is the same PCG.  When A is
 * singular (no FREE wall: constant null space) the adjoint right-hand side
 * y is first made consistent by y -= diag(A) * sum(y) / sum(diag(A)), which
 * is what L does to the null-space component; without the projection quirk
 * sum(y) is zero up to rounding and the shift changes nothing.  The MG-PCG
 * solve returns the zero-mean pressure on a singular operator, (I - 1 1^T /
 * n) L_B b, whose transpose is L_B (I - 1 1^T / n): there the shift is the
 * plain mean removal, as it is without a preconditioner (map A^+, also
 * returned at zero mean).
 */
#include "cf_oracle_internal.h"

#include <stdlib.h>
#include <string.h>

/* ---- spline weights with derivatives ---------------------------------------- */

/* w[k] as cfo_spline_weights and dw[k] = d w[k] / d xi_raw (zero where the
 * coordinate was clamped). */
static void spline_wd(int order, double xi_raw, double lo, double hi, int* s0,
                      double* w, double* dw)
{
    double xi = fmin(fmax(xi_raw, lo), hi);
    int clamped = !(xi_raw >= lo && xi_raw <= hi);
    cfo_spline_weights(order, xi, s0, w);
    double t = xi - (double)(int)xi;
    if (order == 1) {
        dw[0] = -1.0;
        dw[1] = 1.0;
    } else {
        double omt = 1.0 - t;
        dw[0] = -0.5 * omt * omt;
        dw[1] = 1.5 * t * t - 2.0 * t;
        dw[2] = -1.5 * t * t + t + 0.5;
        dw[3] = 0.5 * t * t;
    }
    if (clamped)
        for (int k = 0; k <= order; ++k)
            dw[k] = 0.0;
}

typedef struct {
    int s0[3];
    double w[3][4], dw[3][4]; /* dw already in units of 1/length */
    int n;                    /* taps per axis */
} stencil_t;

static void make_stencil(const cfo_state* s, int e, int order, const double p[3],
                         stencil_t* st)
{
    st->n = order + 1;
    for (int d = 0; d < 3; ++d) {
        st->s0[d] = 0;
        for (int k = 0; k < 4; ++k) {
            st->w[d][k] = (k == 0);
            st->dw[d][k] = 0.0;
        }
    }
    for (int d = 0; d < s->dim; ++d) {
        double xi = (p[d] - s->x0[e][d]) * s->rdx[e][d];
        double lo = order == 1 ? 0.0 : 1.0;
        double hi = (double)(s->ext[e][d] - (order == 1 ? 2 : 3));
        spline_wd(order, xi, lo, hi, &st->s0[d], st->w[d], st->dw[d]);
        for (int k = 0; k < st->n; ++k)
            st->dw[d][k] *= s->rdx[e][d];
    }
}

/* d/dp of the spline sample of f at the stencil's point */
static void sample_grad(const cfo_state* s, int e, const stencil_t* st,
                        const double* f, double grad[3])
{
    int D = s->dim;
    int nz = D == 3 ? st->n : 1;
    grad[0] = grad[1] = grad[2] = 0.0;
    for (int a = 0; a < st->n; ++a)
        for (int b = 0; b < st->n; ++b)
            for (int c = 0; c < nz; ++c) {
                double fv = f[cfo_idx(s, e, st->s0[0] + a, st->s0[1] + b,
                                      D == 3 ? st->s0[2] + c : 0)];
                grad[0] += st->dw[0][a] * st->w[1][b] * st->w[2][c] * fv;
                grad[1] += st->w[0][a] * st->dw[1][b] * st->w[2][c] * fv;
                if (D == 3)
                    grad[2] += st->w[0][a] * st->w[1][b] * st->dw[2][c] * fv;
            }
}

/* fbar[taps] += weight * val */
static void sample_scatter(const cfo_state* s, int e, const stencil_t* st,
                           double* fbar, double val)
{
    int D = s->dim;
    int nz = D == 3 ? st->n : 1;
    for (int a = 0; a < st->n; ++a)
        for (int b = 0; b < st->n; ++b)
            for (int c = 0; c < nz; ++c)
                fbar[cfo_idx(s, e, st->s0[0] + a, st->s0[1] + b,
                             D == 3 ? st->s0[2] + c : 0)] +=
                    st->w[0][a] * st->w[1][b] * st->w[2][c] * val;
}

/* velocity at x (forward values) and its Jacobian J[d][m] = d vel_d / d x_m */
static void velocity_jac(const cfo_state* s, double* const* uvw, const double x[3],
                         double vel[3], double J[3][3], stencil_t vst[3])
{
    int vo = s->c.velocity_order;
    for (int d = 0; d < 3; ++d) {
        vel[d] = 0.0;
        J[d][0] = J[d][1] = J[d][2] = 0.0;
    }
    for (int d = 0; d < s->dim; ++d) {
        make_stencil(s, d + 1, vo, x, &vst[d]);
        vel[d] = cfo_interpolate(s, d + 1, vo, uvw[d], x);
        sample_grad(s, d + 1, &vst[d], uvw[d], J[d]);
    }
}

/* Reverse mode of one advected point (TimeIntegrator.hpp:36-116): obar is
 * the adjoint of the Next value of entity e at x0. */
static void advect_point_adj(const cfo_state* s, int e, double* const* cur,
                             const double x0[3], double obar, double** bar)
{
    int D = s->dim;
    double dt = s->dt;
    double* uvw[3] = {cur[1], cur[2], D == 3 ? cur[3] : NULL};
    double v0[3], v1[3], v2[3], J0[3][3], J1[3][3], J2[3][3];
    double x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
    stencil_t s0v[3], s1v[3], s2v[3];
    int quirk = (s->c.quirks & CFO_QUIRK_RK3_V0) != 0;
    velocity_jac(s, uvw, x0, v0, J0, s0v);
    for (int d = 0; d < D; ++d)
        x1[d] = x0[d] - 0.5 * dt * v0[d];
    velocity_jac(s, uvw, x1, v1, J1, s1v);
    for (int d = 0; d < D; ++d)
        x2[d] = x0[d] - 0.75 * dt * (quirk ? v0[d] : v1[d]);
    velocity_jac(s, uvw, x2, v2, J2, s2v);
    for (int d = 0; d < D; ++d)
        tr[d] = x0[d] - dt * ((2.0 / 9.0) * v0[d] + (3.0 / 9.0) * v1[d] +
                              (4.0 / 9.0) * v2[d]);

    stencil_t fst;
    make_stencil(s, e, s->c.field_order, tr, &fst);
    sample_scatter(s, e, &fst, bar[e], obar);
    double g[3];
    sample_grad(s, e, &fst, cur[e], g);

    double v0b[3] = {0, 0, 0}, v1b[3] = {0, 0, 0}, v2b[3] = {0, 0, 0};
    double x1b[3] = {0, 0, 0}, x2b[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        double trb = obar * g[d];
        v0b[d] = -dt * (2.0 / 9.0) * trb;
        v1b[d] = -dt * (3.0 / 9.0) * trb;
        v2b[d] = -dt * (4.0 / 9.0) * trb;
    }
    for (int d = 0; d < D; ++d) {
        for (int m = 0; m < D; ++m)
            x2b[m] += J2[d][m] * v2b[d];
        sample_scatter(s, d + 1, &s2v[d], bar[d + 1], v2b[d]);
    }
    for (int d = 0; d < D; ++d) {
        if (quirk)
            v0b[d] += -0.75 * dt * x2b[d];
        else
            v1b[d] += -0.75 * dt * x2b[d];
    }
    for (int d = 0; d < D; ++d) {
        for (int m = 0; m < D; ++m)
            x1b[m] += J1[d][m] * v1b[d];
        sample_scatter(s, d + 1, &s1v[d], bar[d + 1], v1b[d]);
    }
    for (int d = 0; d < D; ++d) {
        v0b[d] += -0.5 * dt * x1b[d];
        sample_scatter(s, d + 1, &s0v[d], bar[d + 1], v0b[d]);
    }
    (void)J0;
}

static int is_owned(const int lo[3], const int hi[3], int i, int j, int k)
{
    return i >= lo[0] && i < hi[0] && j >= lo[1] && j < hi[1] && k >= lo[2] &&
           k < hi[2];
}

static void clear_unowned(const struct cfo_state* s, int e, double* a)
{
    int lo[3], hi[3];
    cfo_owned_range(s, e, lo, hi);
    const int* E = s->ext[e];
    for (int k = 0; k < E[2]; ++k)
        for (int j = 0; j < E[1]; ++j)
            for (int i = 0; i < E[0]; ++i)
                if (!is_owned(lo, hi, i, j, k))
                    a[cfo_idx(s, e, i, j, k)] = 0.0;
}

/* Transpose of cfo_fill_ghosts_array: periodic ghost adjoints fold onto
 * their sources (axes in reverse order), then every entry outside the owned
 * range is cleared (wall ghosts are constants). */
static void ghost_adj(const struct cfo_state* s, int e, double* a)
{
    const int* E = s->ext[e];
    for (int d = s->dim - 1; d >= 0; --d) {
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
                    long gi = cfo_idx(s, e, i, j, k);
                    a[cfo_idx(s, e, m[0], m[1], m[2])] += a[gi];
                    a[gi] = 0.0;
                }
    }
    clear_unowned(s, e, a);
}

/* advect^T: obar (adjoint of Next, owned) into bar (adjoint of Current). */
static void advect_adj(const struct cfo_state* s, double* const* cur,
                       double* const* obar, double** bar)
{
    for (int e = 0; e < s->nent; ++e) {
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    double ob = obar[e][cfo_idx(s, e, i, j, k)];
                    if (ob == 0.0)
                        continue;
                    int l[3] = {i, j, k};
                    double x0[3] = {0, 0, 0};
                    for (int d = 0; d < s->dim; ++d)
                        x0[d] = cfo_coord(s, e, d, l[d]);
                    advect_point_adj(s, e, cur, x0, ob, bar);
                }
    }
    for (int e = 0; e < s->nent; ++e)
        ghost_adj(s, e, bar[e]);
}

/* inputs^T (Solver.hpp:181-263); pre: the values before the inputs. */
static void inputs_adj(const struct cfo_state* s, double* const* pre,
                       double** bar, double dvel[3])
{
    const cfo_config* c = &s->c;
    for (int e = 0; e < s->nent; ++e) {
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        int d = e - 1;
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int l[3] = {i, j, k};
                    long id = cfo_idx(s, e, i, j, k);
                    double x[3] = {0, 0, 0};
                    for (int m = 0; m < s->dim; ++m)
                        x[m] = cfo_coord(s, e, m, l[m]);
                    int in = cfo_in_source(s, x);
                    if (e == CFO_CELL) {
                        /* an overwritten quantity has no dependence left */
                        if (in && pre[0][id] < c->inflow_quantity)
                            bar[0][id] = 0.0;
                        continue;
                    }
                    double b = bar[e][id];
                    if (!c->periodic[d]) {
                        int g = l[d] - s->h;
                        if ((g <= 0 && c->wall_type[d][0] == CFO_SOLID) ||
                            (g > s->n[d] - 1 && c->wall_type[d][1] == CFO_SOLID))
                            b = 0.0;
                    }
                    if (in && fabs(pre[e][id]) < fabs(c->inflow_velocity[d])) {
                        dvel[d] += b;
                        b = 0.0;
                    }
                    bar[e][id] = b;
                }
    }
}

/* project^T (VelocityCorrector.hpp:182-264 with the gathers of the forward) */
static int project_adj(struct cfo_state* s, double** bar, double* y, double* z)
{
    const cfo_config* c = &s->c;
    double sp = s->dt / (c->density * s->dx);
    /* 1. the wall rule applied after the face updates */
    for (int d = 0; d < s->dim; ++d) {
        if (c->periodic[d])
            continue;
        int e = d + 1;
        int bc_e = (d > 0 && (c->quirks & CFO_QUIRK_PROJ_BC_U)) ? 1 : e;
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int l[3] = {i, j, k};
                    int g = l[d] - s->h;
                    if ((g <= 0 && c->wall_type[d][0] == CFO_SOLID) ||
                        (g > s->n[d] - 1 && c->wall_type[d][1] == CFO_SOLID))
                        bar[bc_e][cfo_idx(s, bc_e, i, j, k)] = 0.0;
                }
    }
    /* 2. y = -sp G^T bar, then the pressure gather's transpose */
    memset(y, 0, (size_t)s->size[0] * sizeof(double));
    for (int d = 0; d < s->dim; ++d) {
        int e = d + 1;
        int lo[3], hi[3];
        cfo_owned_range(s, e, lo, hi);
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    int m[3] = {i, j, k};
                    m[d] -= 1;
                    double t = -sp * bar[e][cfo_idx(s, e, i, j, k)];
                    y[cfo_idx(s, 0, i, j, k)] += t;
                    y[cfo_idx(s, 0, m[0], m[1], m[2])] -= t;
                }
    }
    ghost_adj(s, 0, y);
    /* 3. consistency shift on a singular operator, then z = L y */
    int lo[3], hi[3];
    cfo_owned_range(s, 0, lo, hi);
    if (cfo_operator_singular(s) && s->c.preconditioner != CFO_PRECOND_JACOBI) {
        cfo_remove_mean(s, y);
    } else if (cfo_operator_singular(s)) {
        cfo_dd sy = {0, 0}, sd = {0, 0};
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    long cc = cfo_idx(s, 0, i, j, k);
                    cfo_dd_add(&sy, y[cc], 0.0);
                    cfo_dd_add(&sd, s->A[cc * s->ncoef], 0.0);
                }
        double mu = cfo_dd_value(sy) / cfo_dd_value(sd);
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    long cc = cfo_idx(s, 0, i, j, k);
                    y[cc] = y[cc] - s->A[cc * s->ncoef] * mu;
                }
    }
    memset(z, 0, (size_t)s->size[0] * sizeof(double));
    cfo_stats st;
    int rc = cfo_solve64(s, y, z, s->c.cg_tol, s->c.cg_max_iter, &st);
    if (rc != CFO_OK)
        return rc;
    /* 4. bar += -(1/dx) D^T z, then the divergence gather's transpose */
    double sc = -1.0 / s->dx;
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                double zc = sc * z[cfo_idx(s, 0, i, j, k)];
                for (int d = 0; d < s->dim; ++d) {
                    int e = d + 1;
                    int m[3] = {i, j, k};
                    m[d] += 1;
                    bar[e][cfo_idx(s, e, m[0], m[1], m[2])] += zc;
                    bar[e][cfo_idx(s, e, i, j, k)] -= zc;
                }
            }
    for (int e = 1; e < s->nent; ++e)
        ghost_adj(s, e, bar[e]);
    return CFO_OK;
}

/* ---- driver -------------------------------------------------------------------- */

static int save_state(const struct cfo_state* s, double** dst)
{
    for (int e = 0; e < s->nent; ++e) {
        if (!dst[e])
            dst[e] = (double*)malloc((size_t)s->size[e] * sizeof(double));
        if (!dst[e])
            return CFO_ENOMEM;
        memcpy(dst[e], s->f[e][0], (size_t)s->size[e] * sizeof(double));
    }
    return CFO_OK;
}

static void load_state(struct cfo_state* s, double* const* src)
{
    for (int e = 0; e < s->nent; ++e)
        memcpy(s->f[e][0], src[e], (size_t)s->size[e] * sizeof(double));
}

int cfo_adjoint_run(cfo_state* s, int nsteps, int ckpt, const double* dq_T,
                    double* du0, double* dv0, double* dw0, double dvel[3],
                    double* dq0)
{
    if (nsteps < 0 || ckpt <= 0 || !dq_T || !dvel)
        return CFO_EINVAL;
    int ne = s->nent;
    int nck = nsteps / ckpt + 1;
    int seglen = ckpt < nsteps ? ckpt : (nsteps > 0 ? nsteps : 1);
    double time0 = s->time, time_end = s->time;
    int rc = CFO_ENOMEM;
    cfo_stats st;
    double* init[4] = {0, 0, 0, 0};
    double* fin[4] = {0, 0, 0, 0};
    double* bar[4] = {0, 0, 0, 0};
    double* obar[4] = {0, 0, 0, 0};
    double** ck = (double**)calloc((size_t)nck * 4, sizeof(double*));
    double** seg = (double**)calloc((size_t)seglen * 4, sizeof(double*));
    double* y = (double*)calloc((size_t)s->size[0], sizeof(double));
    double* z = (double*)calloc((size_t)s->size[0], sizeof(double));
    if (!ck || !seg || !y || !z)
        goto out;
    for (int e = 0; e < ne; ++e) {
        bar[e] = (double*)calloc((size_t)s->size[e], sizeof(double));
        obar[e] = (double*)calloc((size_t)s->size[e], sizeof(double));
        if (!bar[e] || !obar[e])
            goto out;
    }

    /* forward: setup, then the steps with a checkpoint every ckpt steps */
    if ((rc = save_state(s, init)) != CFO_OK)
        goto out;
    if ((rc = cfo_setup(s, &st)) != CFO_OK)
        goto out;
    for (int n = 0; n < nsteps; ++n) {
        if (n % ckpt == 0 && (rc = save_state(s, ck + (n / ckpt) * 4)) != CFO_OK)
            goto out;
        if ((rc = cfo_step(s, &st)) != CFO_OK)
            goto out;
    }
    if ((rc = save_state(s, fin)) != CFO_OK)
        goto out;
    time_end = s->time;

    memcpy(bar[0], dq_T, (size_t)s->size[0] * sizeof(double));
    clear_unowned(s, 0, bar[0]);
    dvel[0] = dvel[1] = dvel[2] = 0.0;

    for (int c0 = nsteps > 0 ? ((nsteps - 1) / ckpt) * ckpt : -1; c0 >= 0;
         c0 -= ckpt) {
        int c1 = c0 + ckpt < nsteps ? c0 + ckpt : nsteps;
        /* recompute S_c0 .. S_{c1-1} from the checkpoint */
        load_state(s, ck + (c0 / ckpt) * 4);
        for (int n = c0; n < c1; ++n) {
            if ((rc = save_state(s, seg + (n - c0) * 4)) != CFO_OK)
                goto out;
            if (n + 1 < c1 && (rc = cfo_step(s, &st)) != CFO_OK)
                goto out;
        }
        for (int n = c1 - 1; n >= c0; --n) {
            /* S_n with its ghosts gathered, and the advected values */
            load_state(s, seg + (n - c0) * 4);
            cfo_fill_ghosts(s, CFO_CURRENT);
            double* cur[4] = {s->f[0][0], s->f[1][0], s->f[2][0],
                              ne > 3 ? s->f[3][0] : NULL};
            for (int e = 0; e < ne; ++e) {
                memset(s->f[e][1], 0, (size_t)s->size[e] * sizeof(double));
                cfo_advect_entity(s, e, (const double* const*)cur, s->f[e][1]);
            }
            double* pre[4] = {s->f[0][1], s->f[1][1], s->f[2][1],
                              ne > 3 ? s->f[3][1] : NULL};
            if ((rc = project_adj(s, bar, y, z)) != CFO_OK)
                goto out;
            inputs_adj(s, pre, bar, dvel);
            for (int e = 0; e < ne; ++e) {
                double* t = obar[e];
                obar[e] = bar[e];
                bar[e] = t;
                memset(bar[e], 0, (size_t)s->size[e] * sizeof(double));
            }
            advect_adj(s, cur, obar, bar);
        }
    }
    /* setup^T */
    if ((rc = project_adj(s, bar, y, z)) != CFO_OK)
        goto out;
    inputs_adj(s, init, bar, dvel);

    if (dq0)
        memcpy(dq0, bar[0], (size_t)s->size[0] * sizeof(double));
    if (du0)
        memcpy(du0, bar[1], (size_t)s->size[1] * sizeof(double));
    if (dv0)
        memcpy(dv0, bar[2], (size_t)s->size[2] * sizeof(double));
    if (dw0 && ne > 3)
        memcpy(dw0, bar[3], (size_t)s->size[3] * sizeof(double));
    load_state(s, fin);
    rc = CFO_OK;
out:
    s->time = rc == CFO_OK ? time_end : time0;
    for (int e = 0; e < 4; ++e) {
        free(init[e]);
        free(fin[e]);
        free(bar[e]);
        free(obar[e]);
    }
    if (ck)
        for (int i = 0; i < nck * 4; ++i)
            free(ck[i]);
    if (seg)
        for (int i = 0; i < seglen * 4; ++i)
            free(seg[i]);
    free(ck);
    free(seg);
    free(y);
    free(z);
    return rc;
}
