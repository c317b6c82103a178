/*
 * Internal definitions shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see cf_oracle.h).
 */
#ifndef CF_ORACLE_INTERNAL_H
#define CF_ORACLE_INTERNAL_H

#include "cf_oracle.h"

#include <math.h>
#include <stddef.h>

struct cfo_state {
    cfo_config c;
    int dim, h;
    int n[3];
    int ext[4][3];  /* ghosted extents per entity */
    long size[4];   /* elements per entity array */
    double dx, dt, time;
    double low[3];  /* coordinate of the ghosted low corner, lo - h*dx */
    double x0[4][3]; /* coordinate of local index 0 of each entity, per dim */
    double rdx[4][3];
    double* f[4][2]; /* [entity][Current/Next] */
    double* lhs;
    double* rhs;
    int ncoef;
    double* A; /* stored stencil, ncoef per cell (reference layout) */
    double* M; /* stored Jacobi preconditioner */
    double *r, *z, *p, *q;
    float *x32, *r32, *z32, *p32, *q32;
    double *A_scratch;
    int nent; /* dim + 1 entities */
    struct cfo_mg* mg; /* multigrid hierarchy, built on first use */
};

/* Double-double accumulator: result rounding is independent of the order in
 * which terms are added (to within ~2^-100 relative of the exact sum). */
typedef struct {
    double hi, lo;
} cfo_dd;

static inline void cfo_dd_add(cfo_dd* acc, double bhi, double blo)
{
    double s = acc->hi + bhi;
    double bb = s - acc->hi;
    double e = (acc->hi - (s - bb)) + (bhi - bb);
    e = e + (acc->lo + blo);
    double hi = s + e;
    acc->lo = e - (hi - s);
    acc->hi = hi;
}

static inline void cfo_dd_add_prod(cfo_dd* acc, double a, double b)
{
    double p = a * b;
    double pe = fma(a, b, -p);
    cfo_dd_add(acc, p, pe);
}

static inline double cfo_dd_value(cfo_dd a) { return a.hi + a.lo; }

static inline long cfo_idx(const struct cfo_state* s, int e, int i, int j,
                           int k)
{
    return (long)i + (long)s->ext[e][0] * ((long)j + (long)s->ext[e][1] * k);
}

static inline double cfo_coord(const struct cfo_state* s, int e, int d, int l)
{
    if (e == d + 1)
        return s->low[d] + (double)l * s->dx;
    return s->low[d] + ((double)l + 0.5) * s->dx;
}

void cfo_owned_range(const struct cfo_state* s, int e, int lo[3], int hi[3]);
/* 1 when A has the constant null space (no FREE wall on a non-periodic axis) */
int cfo_operator_singular(const struct cfo_state* s);
/* v -= mean of v over the owned cells (double-double sum) */
void cfo_remove_mean(struct cfo_state* s, double* v);
void cfo_fill_ghosts_array(struct cfo_state* s, int e, double* a);
void cfo_advect_entity(struct cfo_state* s, int e, const double* const* cur,
                       double* next);
int cfo_in_source(const struct cfo_state* s, const double x[3]);
void cfo_spmv(struct cfo_state* s, const double* x, double* y);
int cfo_pcg64(struct cfo_state* s, const double* b, double* x, double tol,
              int maxit, cfo_stats* st);
/* MG-PCG; f32: the V-cycle's level 0 in FP32 (cf_oracle_mg.c) */
int cfo_pcg_mg(struct cfo_state* s, const double* b, double* x, double tol,
               int maxit, cfo_stats* st, int f32);
/* FP32 stored-stencil SpMV and periodic ghost fill on level 0 */
void cfo_spmv32(struct cfo_state* s, const float* x, float* y);
void cfo_fill_ghosts_f32(struct cfo_state* s, float* a);
/* the FP64 solve of the configured preconditioner */
int cfo_solve64(struct cfo_state* s, const double* b, double* x, double tol,
                int maxit, cfo_stats* st);
void cfo_mg_free(struct cfo_state* s);
double cfo_dot_dd(const struct cfo_state* s, const double* a, const double* b);

#define CFO_FOR_OWNED_CELLS(s, body)                                         \
    do {                                                                     \
        int lo_[3], hi_[3];                                                  \
        cfo_owned_range((s), 0, lo_, hi_);                                   \
        _Pragma("omp parallel for collapse(2) schedule(static)")             \
        for (int k_ = lo_[2]; k_ < hi_[2]; ++k_)                             \
            for (int j_ = lo_[1]; j_ < hi_[1]; ++j_)                         \
                for (int i_ = lo_[0]; i_ < hi_[0]; ++i_) {                   \
                    long c = cfo_idx((s), 0, i_, j_, k_);                    \
                    body;                                                    \
                }                                                            \
    } while (0)
void cfo_apply_pressure_arrays(struct cfo_state* s, double* u, double* v,
                               double* w, const double* p);
void cfo_divergence(struct cfo_state* s, const double* u, const double* v,
                    const double* w, double* rhs);

#endif
