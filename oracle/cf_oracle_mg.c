/*
 * cf_oracle_mg.c -- geometric multigrid preconditioner for the pressure
 * solve (build extension).
 *
 * TEST INFRASTRUCTURE ONLY (see cf_oracle.h).  The reference selects HYPRE's
 * PCG with a PFMG preconditioner by default (VelocityCorrector.hpp:324-338,
 * examples/advection.cpp:186-189); HYPRE is not vendored, so this restates
 * the method class rather than PFMG itself: a symmetric V-cycle used as the
 * CG preconditioner.
 *
 *   hierarchy  level 0 is the pressure grid; an axis with more than 2 cells
 *              is coarsened by 2 (cell 2I and 2I+1 -> I, the last cell of an
 *              odd axis alone), others are kept (semi-coarsening), until
 *              every axis has at most 2 cells
 *   operators  Galerkin P^T A P with piecewise-constant P.  Every coefficient
 *              of A is an integer multiple of s = dt / (rho dx^2) (couplings
 *              -s, or 0 through a wall; diagonal (2 dim - SOLID walls) s), so
 *              the coarse operators are integer face counts times s, formed
 *              in integer arithmetic and rounded once: diag = K s, coupling
 *              to the +e_d neighbour = -C_d s (wrapped on a periodic axis, 0
 *              at a domain end)
 *   smoother   damped Jacobi x <- fma(w, b - A x, x), w = omega / diag
 *              (omega * (1 / diag) on level 0, the PCG's Jacobi table); the
 *              first sweep from x = 0 is w b
 *   V-cycle    nu sweeps, residual restricted by summing children (z, y, x
 *              order), recursion, prolongation x = fma(alpha, x_parent, x),
 *              nu sweeps; the coarsest level takes coarse_sweeps sweeps from
 *              zero.  Equal pre and post sweeps of a symmetric smoother make
 *              the cycle a symmetric preconditioner.  alpha > 1 over-corrects
 *              (Braess's remedy for the piecewise-constant P, whose coarse
 *              correction alone is too small by about half in energy)
 *   MG-PCG     as Jacobi-PCG with z = V(r) (cfo_pcg_mg)
 *   FP32 level 0  (precision CF_FP32_REFINED, 3-D, every axis > 2 cells):
 *              the V-cycle's level 0 runs on FP32 vectors with the FP32
 *              stencil (coefficients rounded from the FP64 ones, w0 rounded
 *              from omega * M): b32 = (float) r; x = w0 b32, then nu - 1
 *              sweeps fmaf(w0, b32 - A32 x, x); the residual b32 - A32 x in
 *              FP32, restricted by summing the children as doubles (z, y, x
 *              order) into level 1, which stays FP64; prolongation
 *              fmaf((float) alpha, (float) x_parent, x); nu sweeps;
 *              z = (double) x.  The CG around it (residuals, dots, alpha,
 *              beta, x) stays FP64.
 *              Level 1 runs in FP32 as well when it is a full-box level of
 *              more than 32768 cells (every level-0 axis even, every level-1
 *              axis > 2: the levels the device marches with its tiles):
 *              the level-0 sums are rounded into it, its coefficients are
 *              the FP64 ones rounded, its residuals are summed as doubles
 *              into level 2, and level 0's prolongation reads its FP32
 *              result.
 *
 * A-x order everywhere: diag x, then fma over -x, +x, -y, +y, -z, +z.
 */
#include "cf_oracle_internal.h"

#include <stdlib.h>
#include <string.h>

#define CFO_MG_MAXLEV 24

typedef struct {
    int n[3];
    int coarsen[3]; /* axis coarsened from this level to the next */
    long ncell;
    long* kdiag;    /* diagonal in units of s */
    long* kc[3];    /* coupling count to the +e_d neighbour */
    double *diag, *cp[3], *w; /* level >= 1 coefficients; w = omega / diag */
    double *x, *b, *t;        /* compact work vectors (level >= 1) */
} cfo_mglev;

struct cfo_mg {
    int nlev;
    cfo_mglev lev[CFO_MG_MAXLEV];
    double* r0;        /* level-0 residual (ghosted) */
    double* t0;        /* level-0 Jacobi buffer (ghosted) */
    double* w0;        /* level-0 omega * M (ghosted) */
    /* FP32 level 0 (ghosted, allocated on first use) */
    float *x0f, *t0f, *b0f, *r0f, *w0f;
    /* FP32 level 1 (compact), when mg_f32_level1 */
    float *x1f, *t1f, *b1f, *r1f;
};

static long lidx(const cfo_mglev* L, int i, int j, int k)
{
    return (long)i + (long)L->n[0] * ((long)j + (long)L->n[1] * k);
}

/* neighbour index along d (+1 or -1) with the periodic wrap; -1 outside */
static long lnb(const struct cfo_state* s, const cfo_mglev* L, const int c[3], int d,
                int dir)
{
    int m[3] = {c[0], c[1], c[2]};
    m[d] += dir;
    if (m[d] < 0 || m[d] >= L->n[d]) {
        if (!s->c.periodic[d])
            return -1;
        m[d] = (m[d] + L->n[d]) % L->n[d];
    }
    return lidx(L, m[0], m[1], m[2]);
}

void cfo_mg_free(struct cfo_state* s)
{
    struct cfo_mg* mg = s->mg;
    if (!mg)
        return;
    for (int l = 0; l < mg->nlev; ++l) {
        cfo_mglev* L = &mg->lev[l];
        free(L->kdiag);
        free(L->diag);
        free(L->w);
        free(L->x);
        free(L->b);
        free(L->t);
        for (int d = 0; d < 3; ++d) {
            free(L->kc[d]);
            free(L->cp[d]);
        }
    }
    free(mg->r0);
    free(mg->t0);
    free(mg->w0);
    free(mg->x0f);
    free(mg->t0f);
    free(mg->b0f);
    free(mg->r0f);
    free(mg->w0f);
    free(mg->x1f);
    free(mg->t1f);
    free(mg->b1f);
    free(mg->r1f);
    free(mg);
    s->mg = NULL;
}

static int cfo_mg_build(struct cfo_state* s)
{
    if (s->mg)
        return CFO_OK;
    const int D = s->dim;
    struct cfo_mg* mg = (struct cfo_mg*)calloc(1, sizeof(*mg));
    if (!mg)
        return CFO_ENOMEM;
    s->mg = mg;
    /* level 0 counts */
    cfo_mglev* L0 = &mg->lev[0];
    for (int d = 0; d < 3; ++d)
        L0->n[d] = d < D ? s->n[d] : 1;
    L0->ncell = (long)L0->n[0] * L0->n[1] * L0->n[2];
    L0->kdiag = (long*)calloc((size_t)L0->ncell, sizeof(long));
    for (int d = 0; d < D; ++d)
        L0->kc[d] = (long*)calloc((size_t)L0->ncell, sizeof(long));
    if (!L0->kdiag)
        return CFO_ENOMEM;
    for (int k = 0; k < L0->n[2]; ++k)
        for (int j = 0; j < L0->n[1]; ++j)
            for (int i = 0; i < L0->n[0]; ++i) {
                int c[3] = {i, j, k};
                long id = lidx(L0, i, j, k);
                long kd = 2 * D;
                for (int d = 0; d < D; ++d) {
                    if (!s->c.periodic[d]) {
                        if (c[d] == 0 && s->c.wall_type[d][0] == CFO_SOLID)
                            kd -= 1;
                        if (c[d] == L0->n[d] - 1 && s->c.wall_type[d][1] == CFO_SOLID)
                            kd -= 1;
                    }
                    L0->kc[d][id] = lnb(s, L0, c, d, +1) >= 0 ? 1 : 0;
                }
                L0->kdiag[id] = kd;
            }
    /* coarser levels */
    double sc = s->dt / (s->c.density * s->dx * s->dx);
    int l = 0;
    for (;;) {
        cfo_mglev* F = &mg->lev[l];
        int any = 0;
        for (int d = 0; d < D; ++d) {
            F->coarsen[d] = F->n[d] > 2;
            any = any || F->coarsen[d];
        }
        if (!any || l + 1 >= CFO_MG_MAXLEV)
            break;
        cfo_mglev* C = &mg->lev[l + 1];
        for (int d = 0; d < 3; ++d)
            C->n[d] = (d < D && F->coarsen[d]) ? (F->n[d] + 1) / 2 : F->n[d];
        C->ncell = (long)C->n[0] * C->n[1] * C->n[2];
        C->kdiag = (long*)calloc((size_t)C->ncell, sizeof(long));
        C->diag = (double*)calloc((size_t)C->ncell, sizeof(double));
        C->w = (double*)calloc((size_t)C->ncell, sizeof(double));
        C->x = (double*)calloc((size_t)C->ncell, sizeof(double));
        C->b = (double*)calloc((size_t)C->ncell, sizeof(double));
        C->t = (double*)calloc((size_t)C->ncell, sizeof(double));
        if (!C->kdiag || !C->diag || !C->w || !C->x || !C->b || !C->t)
            return CFO_ENOMEM;
        for (int d = 0; d < D; ++d) {
            C->kc[d] = (long*)calloc((size_t)C->ncell, sizeof(long));
            C->cp[d] = (double*)calloc((size_t)C->ncell, sizeof(double));
            if (!C->kc[d] || !C->cp[d])
                return CFO_ENOMEM;
        }
        for (int K = 0; K < C->n[2]; ++K)
            for (int J = 0; J < C->n[1]; ++J)
                for (int I = 0; I < C->n[0]; ++I) {
                    int Ic[3] = {I, J, K};
                    int lo[3], hi[3]; /* children [lo, hi] per axis */
                    for (int d = 0; d < 3; ++d) {
                        if (d < D && F->coarsen[d]) {
                            lo[d] = 2 * Ic[d];
                            hi[d] = 2 * Ic[d] + 1 < F->n[d] ? 2 * Ic[d] + 1 : 2 * Ic[d];
                        } else {
                            lo[d] = hi[d] = Ic[d];
                        }
                    }
                    long kd = 0, kcp[3] = {0, 0, 0};
                    for (int k = lo[2]; k <= hi[2]; ++k)
                        for (int j = lo[1]; j <= hi[1]; ++j)
                            for (int i = lo[0]; i <= hi[0]; ++i) {
                                int c[3] = {i, j, k};
                                long id = lidx(F, i, j, k);
                                kd += F->kdiag[id];
                                for (int d = 0; d < D; ++d) {
                                    /* a coupling to a child of the same
                                     * parent is internal: Galerkin adds it
                                     * twice to the diagonal */
                                    if (c[d] < hi[d])
                                        kd -= 2 * F->kc[d][id];
                                    else
                                        kcp[d] += F->kc[d][id];
                                }
                            }
                    long cid = lidx(C, I, J, K);
                    C->kdiag[cid] = kd;
                    C->diag[cid] = (double)kd * sc;
                    C->w[cid] = s->c.mg_omega / C->diag[cid];
                    for (int d = 0; d < D; ++d) {
                        C->kc[d][cid] = kcp[d];
                        C->cp[d][cid] = -((double)kcp[d] * sc);
                    }
                }
        ++l;
    }
    mg->nlev = l + 1;
    long nc = s->size[0];
    mg->r0 = (double*)calloc((size_t)nc, sizeof(double));
    mg->t0 = (double*)calloc((size_t)nc, sizeof(double));
    mg->w0 = (double*)calloc((size_t)nc, sizeof(double));
    if (!mg->r0 || !mg->t0 || !mg->w0)
        return CFO_ENOMEM;
    CFO_FOR_OWNED_CELLS(s, mg->w0[c] = s->c.mg_omega * s->M[c]);
    return CFO_OK;
}

/* ---- level operators ----------------------------------------------------------- */

/* y = A_l x on a compact level (l >= 1) */
static void mg_apply(const struct cfo_state* s, const cfo_mglev* L, const double* x,
                     double* y)
{
    const int D = s->dim;
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                int c[3] = {i, j, k};
                long id = lidx(L, i, j, k);
                double acc = L->diag[id] * x[id];
                for (int d = 0; d < D; ++d) {
                    long m = lnb(s, L, c, d, -1), p = lnb(s, L, c, d, +1);
                    double cm = m >= 0 ? L->cp[d][m] : 0.0;
                    double cpl = p >= 0 ? L->cp[d][id] : 0.0;
                    acc = fma(cm, m >= 0 ? x[m] : 0.0, acc);
                    acc = fma(cpl, p >= 0 ? x[p] : 0.0, acc);
                }
                y[id] = acc;
            }
}

/* one Jacobi sweep on level l >= 1: xo = fma(w, b - A xi, xi) */
static void mg_sweep(const struct cfo_state* s, const cfo_mglev* L, const double* xi,
                     const double* b, double* xo)
{
    mg_apply(s, L, xi, xo);
    for (long id = 0; id < L->ncell; ++id)
        xo[id] = fma(L->w[id], b[id] - xo[id], xi[id]);
}

/* level-0 sweep on ghosted arrays: xo = fma(w0, b - A xi, xi) */
static void mg_sweep0(struct cfo_state* s, double* xi, const double* b, double* xo)
{
    struct cfo_mg* mg = s->mg;
    if (s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2])
        cfo_fill_ghosts_array(s, 0, xi);
    cfo_spmv(s, xi, xo);
    CFO_FOR_OWNED_CELLS(s, xo[c] = fma(mg->w0[c], b[c] - xo[c], xi[c]));
}

/* owned level-0 cell (i, j, k) -> ghosted index */
static long g0(const struct cfo_state* s, int i, int j, int k)
{
    return cfo_idx(s, 0, s->h + i, s->dim > 1 ? s->h + j : 0, s->dim > 2 ? s->h + k : 0);
}

/* b_c = sum over children of (b - A x) on level f (r: level-f residual) */
static void mg_restrict(const struct cfo_state* s, int f, const double* r, double* bc)
{
    const cfo_mglev* F = &s->mg->lev[f];
    const cfo_mglev* C = &s->mg->lev[f + 1];
    const int D = s->dim;
    for (int K = 0; K < C->n[2]; ++K)
        for (int J = 0; J < C->n[1]; ++J)
            for (int I = 0; I < C->n[0]; ++I) {
                int Ic[3] = {I, J, K}, lo[3], hi[3];
                for (int d = 0; d < 3; ++d) {
                    if (d < D && F->coarsen[d]) {
                        lo[d] = 2 * Ic[d];
                        hi[d] = 2 * Ic[d] + 1 < F->n[d] ? 2 * Ic[d] + 1 : 2 * Ic[d];
                    } else {
                        lo[d] = hi[d] = Ic[d];
                    }
                }
                double sum = 0.0;
                for (int k = lo[2]; k <= hi[2]; ++k)
                    for (int j = lo[1]; j <= hi[1]; ++j)
                        for (int i = lo[0]; i <= hi[0]; ++i)
                            sum = sum + r[f == 0 ? g0(s, i, j, k) : lidx(F, i, j, k)];
                bc[lidx(C, I, J, K)] = sum;
            }
}

/* x_f = fma(alpha, x_c[parent], x_f) */
static void mg_prolong(const struct cfo_state* s, int f, const double* xc, double* xf)
{
    const cfo_mglev* F = &s->mg->lev[f];
    const cfo_mglev* C = &s->mg->lev[f + 1];
    const int D = s->dim;
    for (int k = 0; k < F->n[2]; ++k)
        for (int j = 0; j < F->n[1]; ++j)
            for (int i = 0; i < F->n[0]; ++i) {
                int c[3] = {i, j, k}, P[3];
                for (int d = 0; d < 3; ++d)
                    P[d] = (d < D && F->coarsen[d]) ? c[d] / 2 : c[d];
                long fi = f == 0 ? g0(s, i, j, k) : lidx(F, i, j, k);
                xf[fi] = fma(s->c.mg_alpha, xc[lidx(C, P[0], P[1], P[2])], xf[fi]);
            }
}

/* V-cycle on level l: x_l = B_l b_l, from x = 0 */
static void mg_vcycle_level(struct cfo_state* s, int l, double* x0, const double* b0)
{
    struct cfo_mg* mg = s->mg;
    const int last = mg->nlev - 1;
    const int nu = l == last ? s->c.mg_coarse_sweeps : s->c.mg_nu;
    if (l == 0) {
        /* first sweep from zero: w b, then nu - 1 sweeps */
        CFO_FOR_OWNED_CELLS(s, x0[c] = mg->w0[c] * b0[c]);
        for (int it = 1; it < nu; ++it) {
            mg_sweep0(s, x0, b0, mg->t0);
            CFO_FOR_OWNED_CELLS(s, x0[c] = mg->t0[c]);
        }
        if (l == last)
            return;
        if (s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2])
            cfo_fill_ghosts_array(s, 0, x0);
        cfo_spmv(s, x0, mg->r0);
        CFO_FOR_OWNED_CELLS(s, mg->r0[c] = b0[c] - mg->r0[c]);
        mg_restrict(s, 0, mg->r0, mg->lev[1].b);
        mg_vcycle_level(s, 1, NULL, NULL);
        mg_prolong(s, 0, mg->lev[1].x, x0);
        for (int it = 0; it < s->c.mg_nu; ++it) {
            mg_sweep0(s, x0, b0, mg->t0);
            CFO_FOR_OWNED_CELLS(s, x0[c] = mg->t0[c]);
        }
        return;
    }
    cfo_mglev* L = &mg->lev[l];
    for (long id = 0; id < L->ncell; ++id)
        L->x[id] = L->w[id] * L->b[id];
    for (int it = 1; it < nu; ++it) {
        mg_sweep(s, L, L->x, L->b, L->t);
        memcpy(L->x, L->t, (size_t)L->ncell * sizeof(double));
    }
    if (l == last)
        return;
    mg_apply(s, L, L->x, L->t);
    for (long id = 0; id < L->ncell; ++id)
        L->t[id] = L->b[id] - L->t[id];
    mg_restrict(s, l, L->t, mg->lev[l + 1].b);
    mg_vcycle_level(s, l + 1, NULL, NULL);
    mg_prolong(s, l, mg->lev[l + 1].x, L->x);
    for (int it = 0; it < s->c.mg_nu; ++it) {
        mg_sweep(s, L, L->x, L->b, L->t);
        memcpy(L->x, L->t, (size_t)L->ncell * sizeof(double));
    }
}

/* ---- FP32 level 0 ------------------------------------------------------------------ */

/* The FP32 level 0 runs where the device runs it: 3-D, every axis coarsened
 * at least once (more than 2 cells), so that level 0 is a marched level. */
static int mg_f32_supported(const struct cfo_state* s)
{
    return s->dim == 3 && s->n[0] > 2 && s->n[1] > 2 && s->n[2] > 2 && s->mg->nlev >= 2;
}

/* Level 1 in FP32 too: a full-box level (every level-0 axis even, every
 * level-1 axis coarsened further) of more than 32768 cells. */
static int mg_f32_level1(const struct cfo_state* s)
{
    const struct cfo_mg* mg = s->mg;
    if (mg->nlev < 3)
        return 0;
    const cfo_mglev* L1 = &mg->lev[1];
    for (int d = 0; d < 3; ++d)
        if (s->n[d] % 2 != 0 || !L1->coarsen[d])
            return 0;
    return L1->ncell > 32768;
}

static int mg_f32_alloc(struct cfo_state* s)
{
    struct cfo_mg* mg = s->mg;
    if (mg_f32_level1(s) && !mg->x1f) {
        size_t n1 = (size_t)mg->lev[1].ncell;
        mg->x1f = (float*)calloc(n1, sizeof(float));
        mg->t1f = (float*)calloc(n1, sizeof(float));
        mg->b1f = (float*)calloc(n1, sizeof(float));
        mg->r1f = (float*)calloc(n1, sizeof(float));
        if (!mg->x1f || !mg->t1f || !mg->b1f || !mg->r1f)
            return CFO_ENOMEM;
    }
    if (mg->x0f)
        return CFO_OK;
    long nc = s->size[0];
    mg->x0f = (float*)calloc((size_t)nc, sizeof(float));
    mg->t0f = (float*)calloc((size_t)nc, sizeof(float));
    mg->b0f = (float*)calloc((size_t)nc, sizeof(float));
    mg->r0f = (float*)calloc((size_t)nc, sizeof(float));
    mg->w0f = (float*)calloc((size_t)nc, sizeof(float));
    if (!mg->x0f || !mg->t0f || !mg->b0f || !mg->r0f || !mg->w0f)
        return CFO_ENOMEM;
    CFO_FOR_OWNED_CELLS(s, mg->w0f[c] = (float)mg->w0[c]);
    return CFO_OK;
}

/* level-0 FP32 sweep on ghosted arrays: xo = fmaf(w0, b - A32 xi, xi) */
static void mg_sweep0_f32(struct cfo_state* s, float* xi, const float* b, float* xo)
{
    struct cfo_mg* mg = s->mg;
    if (s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2])
        cfo_fill_ghosts_f32(s, xi);
    cfo_spmv32(s, xi, xo);
    CFO_FOR_OWNED_CELLS(s, xo[c] = fmaf(mg->w0f[c], b[c] - xo[c], xi[c]));
}

/* y = A_l x on a compact level in FP32: the FP64 coefficients rounded, the
 * order of mg_apply */
static void mg_apply_f32(const struct cfo_state* s, const cfo_mglev* L, const float* x,
                         float* y)
{
    const int D = s->dim;
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                int c[3] = {i, j, k};
                long id = lidx(L, i, j, k);
                float acc = (float)L->diag[id] * x[id];
                for (int d = 0; d < D; ++d) {
                    long m = lnb(s, L, c, d, -1), p = lnb(s, L, c, d, +1);
                    float cm = m >= 0 ? (float)L->cp[d][m] : 0.0f;
                    float cpl = p >= 0 ? (float)L->cp[d][id] : 0.0f;
                    acc = fmaf(cm, m >= 0 ? x[m] : 0.0f, acc);
                    acc = fmaf(cpl, p >= 0 ? x[p] : 0.0f, acc);
                }
                y[id] = acc;
            }
}

/* one FP32 Jacobi sweep on a compact level: xo = fmaf(w, b - A xi, xi) */
static void mg_sweep_f32(const struct cfo_state* s, const cfo_mglev* L, const float* xi,
                         const float* b, float* xo)
{
    mg_apply_f32(s, L, xi, xo);
    for (long id = 0; id < L->ncell; ++id)
        xo[id] = fmaf((float)L->w[id], b[id] - xo[id], xi[id]);
}

/* the FP32 level 1 (mg_f32_level1): b1f is its right-hand side; on return
 * x1f holds its cycle result */
static void mg_vcycle1_f32(struct cfo_state* s)
{
    struct cfo_mg* mg = s->mg;
    const cfo_mglev* L = &mg->lev[1];
    const cfo_mglev* C = &mg->lev[2];
    const int nu = s->c.mg_nu;
    float *x = mg->x1f, *t = mg->t1f, *b = mg->b1f, *rf = mg->r1f;
    for (long id = 0; id < L->ncell; ++id)
        x[id] = (float)L->w[id] * b[id];
    for (int it = 1; it < nu; ++it) {
        mg_sweep_f32(s, L, x, b, t);
        memcpy(x, t, (size_t)L->ncell * sizeof(float));
    }
    mg_apply_f32(s, L, x, rf);
    for (long id = 0; id < L->ncell; ++id)
        rf[id] = b[id] - rf[id];
    for (int K = 0; K < C->n[2]; ++K)
        for (int J = 0; J < C->n[1]; ++J)
            for (int I = 0; I < C->n[0]; ++I) {
                int Ic[3] = {I, J, K}, lo[3], hi[3];
                for (int d = 0; d < 3; ++d) {
                    lo[d] = 2 * Ic[d];
                    hi[d] = 2 * Ic[d] + 1 < L->n[d] ? 2 * Ic[d] + 1 : 2 * Ic[d];
                }
                double sum = 0.0;
                for (int k = lo[2]; k <= hi[2]; ++k)
                    for (int j = lo[1]; j <= hi[1]; ++j)
                        for (int i = lo[0]; i <= hi[0]; ++i)
                            sum = sum + (double)rf[lidx(L, i, j, k)];
                mg->lev[2].b[lidx(C, I, J, K)] = sum;
            }
    mg_vcycle_level(s, 2, NULL, NULL);
    const float a = (float)s->c.mg_alpha;
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                long fi = lidx(L, i, j, k);
                x[fi] = fmaf(a, (float)mg->lev[2].x[lidx(C, i / 2, j / 2, k / 2)], x[fi]);
            }
    for (int it = 0; it < nu; ++it) {
        mg_sweep_f32(s, L, x, b, t);
        memcpy(x, t, (size_t)L->ncell * sizeof(float));
    }
}

/* z = B r with level 0 in FP32 (levels >= 1 as mg_vcycle_level, or level 1
 * in FP32 too) */
static void mg_vcycle0_f32(struct cfo_state* s, double* z, const double* r)
{
    struct cfo_mg* mg = s->mg;
    float *x = mg->x0f, *t = mg->t0f, *b = mg->b0f, *rf = mg->r0f;
    const float* w = mg->w0f;
    const int nu = s->c.mg_nu;
    const int periodic = s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2];
    CFO_FOR_OWNED_CELLS(s, b[c] = (float)r[c]; x[c] = w[c] * b[c]);
    for (int it = 1; it < nu; ++it) {
        mg_sweep0_f32(s, x, b, t);
        CFO_FOR_OWNED_CELLS(s, x[c] = t[c]);
    }
    if (periodic)
        cfo_fill_ghosts_f32(s, x);
    cfo_spmv32(s, x, rf);
    CFO_FOR_OWNED_CELLS(s, rf[c] = b[c] - rf[c]);
    const int f1 = mg_f32_level1(s);
    /* restriction: the children's FP32 residuals summed as doubles (and
     * rounded into an FP32 level 1) */
    {
        const cfo_mglev* F = &mg->lev[0];
        const cfo_mglev* C = &mg->lev[1];
        for (int K = 0; K < C->n[2]; ++K)
            for (int J = 0; J < C->n[1]; ++J)
                for (int I = 0; I < C->n[0]; ++I) {
                    int Ic[3] = {I, J, K}, lo[3], hi[3];
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = 2 * Ic[d];
                        hi[d] = 2 * Ic[d] + 1 < F->n[d] ? 2 * Ic[d] + 1 : 2 * Ic[d];
                    }
                    double sum = 0.0;
                    for (int k = lo[2]; k <= hi[2]; ++k)
                        for (int j = lo[1]; j <= hi[1]; ++j)
                            for (int i = lo[0]; i <= hi[0]; ++i)
                                sum = sum + (double)rf[g0(s, i, j, k)];
                    if (f1)
                        mg->b1f[lidx(C, I, J, K)] = (float)sum;
                    else
                        mg->lev[1].b[lidx(C, I, J, K)] = sum;
                }
    }
    if (f1)
        mg_vcycle1_f32(s);
    else
        mg_vcycle_level(s, 1, NULL, NULL);
    /* prolongation in FP32 from the level-1 result (FP64, or FP32) */
    {
        const cfo_mglev* F = &mg->lev[0];
        const cfo_mglev* C = &mg->lev[1];
        const float a = (float)s->c.mg_alpha;
        for (int k = 0; k < F->n[2]; ++k)
            for (int j = 0; j < F->n[1]; ++j)
                for (int i = 0; i < F->n[0]; ++i) {
                    long fi = g0(s, i, j, k);
                    long ci = lidx(C, i / 2, j / 2, k / 2);
                    float xc = f1 ? mg->x1f[ci] : (float)mg->lev[1].x[ci];
                    x[fi] = fmaf(a, xc, x[fi]);
                }
    }
    for (int it = 0; it < nu; ++it) {
        mg_sweep0_f32(s, x, b, t);
        CFO_FOR_OWNED_CELLS(s, x[c] = t[c]);
    }
    CFO_FOR_OWNED_CELLS(s, z[c] = (double)x[c]);
}

/* the preconditioner of MG-PCG: z = B r */
static void mg_precondition(struct cfo_state* s, int f32, double* z, const double* r)
{
    if (f32)
        mg_vcycle0_f32(s, z, r);
    else
        mg_vcycle_level(s, 0, z, r);
}

/* ---- MG-PCG ---------------------------------------------------------------------- */

int cfo_pcg_mg(struct cfo_state* s, const double* b, double* x, double tol, int maxit,
               cfo_stats* st, int f32)
{
    int rc = cfo_mg_build(s);
    if (rc != CFO_OK)
        return rc;
    if (f32) {
        if (!mg_f32_supported(s))
            return CFO_ENOTIMPL;
        rc = mg_f32_alloc(s);
        if (rc != CFO_OK)
            return rc;
    }
    double *r = s->r, *z = s->z, *p = s->p, *q = s->q;
    int periodic = s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2];
    double bnorm = sqrt(cfo_dot_dd(s, b, b));
    if (st) {
        st->b_norm = bnorm;
        st->outer_iterations = 0;
    }
    CFO_FOR_OWNED_CELLS(s, x[c] = 0.0);
    if (bnorm == 0.0) {
        if (st) {
            st->converged = 1;
            st->iterations = 0;
            st->final_residual = 0.0;
        }
        return CFO_OK;
    }
    CFO_FOR_OWNED_CELLS(s, r[c] = b[c]);
    mg_precondition(s, f32, z, r);
    CFO_FOR_OWNED_CELLS(s, p[c] = z[c]);
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
        mg_precondition(s, f32, z, r);
        double rho_new = cfo_dot_dd(s, r, z);
        double beta = rho_new / rho;
        CFO_FOR_OWNED_CELLS(s, p[c] = fma(beta, p[c], z[c]));
        rho = rho_new;
    }
    /* On a singular operator the pressure is fixed up to a constant, and
     * MG-PCG leaves it at a level of its own, B^1/2 (B^1/2 A B^1/2)^+ B^1/2 b
     * for the V-cycle B.  Only the as-written FaceJ rule (T3) lets that
     * level reach the velocities.  The solve returns the zero-mean pressure
     * instead, so that the reverse mode can transpose the map exactly:
     * (I - 1 1^T / n) L_B has the transpose L_B (I - 1 1^T / n), a mean
     * removal followed by the same MG-PCG, whereas L_B's own null vector,
     * B^-1 1, is not at hand. */
    if (cfo_operator_singular(s))
        cfo_remove_mean(s, x);
    if (st) {
        st->converged = conv;
        st->iterations = k;
        st->final_residual = res;
    }
    return conv ? CFO_OK : CFO_ENOCONV;
}

/* ---- test access ------------------------------------------------------------------- */

int cfo_mg_nlevels(cfo_state* s)
{
    if (cfo_mg_build(s) != CFO_OK)
        return -1;
    return s->mg->nlev;
}

int cfo_mg_level_dims(cfo_state* s, int level, int n[3])
{
    if (cfo_mg_build(s) != CFO_OK || level < 0 || level >= s->mg->nlev)
        return CFO_EINVAL;
    for (int d = 0; d < 3; ++d)
        n[d] = s->mg->lev[level].n[d];
    return CFO_OK;
}

/* compact level-0 vector <-> ghosted */
static void to_ghosted(struct cfo_state* s, const double* in, double* out)
{
    const cfo_mglev* L = &s->mg->lev[0];
    memset(out, 0, (size_t)s->size[0] * sizeof(double));
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i)
                out[g0(s, i, j, k)] = in[lidx(L, i, j, k)];
}

static void from_ghosted(struct cfo_state* s, const double* in, double* out)
{
    const cfo_mglev* L = &s->mg->lev[0];
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i)
                out[lidx(L, i, j, k)] = in[g0(s, i, j, k)];
}

int cfo_mg_apply_level(cfo_state* s, int level, const double* x, double* y)
{
    if (cfo_mg_build(s) != CFO_OK || level < 0 || level >= s->mg->nlev)
        return CFO_EINVAL;
    if (level > 0) {
        mg_apply(s, &s->mg->lev[level], x, y);
        return CFO_OK;
    }
    double* gx = (double*)calloc((size_t)s->size[0], sizeof(double));
    double* gy = (double*)calloc((size_t)s->size[0], sizeof(double));
    if (!gx || !gy) {
        free(gx);
        free(gy);
        return CFO_ENOMEM;
    }
    to_ghosted(s, x, gx);
    if (s->c.periodic[0] || s->c.periodic[1] || s->c.periodic[2])
        cfo_fill_ghosts_array(s, 0, gx);
    cfo_spmv(s, gx, gy);
    from_ghosted(s, gy, y);
    free(gx);
    free(gy);
    return CFO_OK;
}

int cfo_mg_vcycle(cfo_state* s, const double* r, double* z)
{
    return cfo_mg_vcycle_prec(s, 0, r, z);
}

int cfo_mg_vcycle_prec(cfo_state* s, int f32, const double* r, double* z)
{
    if (cfo_mg_build(s) != CFO_OK)
        return CFO_ENOMEM;
    if (f32) {
        if (!mg_f32_supported(s))
            return CFO_ENOTIMPL;
        if (mg_f32_alloc(s) != CFO_OK)
            return CFO_ENOMEM;
    }
    double* gr = (double*)calloc((size_t)s->size[0], sizeof(double));
    double* gz = (double*)calloc((size_t)s->size[0], sizeof(double));
    if (!gr || !gz) {
        free(gr);
        free(gz);
        return CFO_ENOMEM;
    }
    to_ghosted(s, r, gr);
    mg_precondition(s, f32, gz, gr);
    from_ghosted(s, gz, z);
    free(gr);
    free(gz);
    return CFO_OK;
}
