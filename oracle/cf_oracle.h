/*
 * cf_oracle.h -- CPU oracle for the CajitaFluids per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithm (CUP-ECS/cajitafluids, the headers under src/) used as the checker
 * for the HIP product path and as the timed CPU baseline.  Nothing in the
 * product library links, loads or calls it; only tests/, the smoke check and
 * the CPU-baseline leg of bench.py do.
 *
 * The reference arithmetic it restates (file:line in the reference tree):
 *   geometry / dx                 src/Mesh.hpp:41-103
 *   dt clamp                      src/Solver.hpp:93-106
 *   spline sample (Cajita G2P)    src/Interpolation.hpp:30-54 (restated from
 *                                 the B-spline definition; Cajita is not
 *                                 vendored)
 *   RK3 back-trace                src/TimeIntegrator.hpp:36-77
 *   advect / step                 src/TimeIntegrator.hpp:81-177
 *   inputs (inflow/body/BC)       src/Solver.hpp:181-263, InflowSource.hpp,
 *                                 BodyForce.hpp, BoundaryConditions.hpp:102-129
 *   matrix + Jacobi               src/VelocityCorrector.hpp:116-180,
 *                                 BoundaryConditions.hpp:57-97
 *   divergence RHS                src/VelocityCorrector.hpp:182-212
 *   PCG                           Cajita ReferenceConjugateGradient, called at
 *                                 src/VelocityCorrector.hpp:266-282
 *   projection                    src/VelocityCorrector.hpp:214-264
 *   setup / step                  src/Solver.hpp:125-147
 *
 * Everything is generic over dim = 2 or 3 (the reference only has 2-D; the
 * 3-D operators generalise it per axis).  All floating-point expressions are
 * written with an explicit evaluation order and explicit fma() so that the
 * HIP path, which uses the same order, produces identical bits.  Dot products
 * accumulate in double-double, which makes their rounded result independent
 * of the summation order.
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CFO_CELL = 0, CFO_FACE_X = 1, CFO_FACE_Y = 2, CFO_FACE_Z = 3 };
enum { CFO_CURRENT = 0, CFO_NEXT = 1 };
enum { CFO_SOLID = 0, CFO_FREE = 1 };
enum { CFO_INFLOW_NONE = 0, CFO_INFLOW_BOX = 1, CFO_INFLOW_SPHERE = 2 };
enum { CFO_FP64 = 0, CFO_FP32_REFINED = 1 };

#define CFO_QUIRK_RK3_V0 1    /* 3/4-step probe uses v0 (TimeIntegrator.hpp:57-58) */
#define CFO_QUIRK_PROJ_BC_U 2 /* post-projection FaceJ/K BC hits u (VelocityCorrector.hpp:260) */

#define CFO_OK 0
#define CFO_EINVAL 1
#define CFO_ENOMEM 2
#define CFO_ENOCONV 3
#define CFO_ENOTIMPL 5 /* combination not restated (the library returns CF_ENOTIMPL too) */

/* Same field order and meaning as the product's cf_config. */
typedef struct cfo_config {
    int dim;
    int num_cell[3];
    int halo;
    int periodic[3];
    double lo[3];
    double hi[3];
    int wall_type[3][2];
    double density;
    double dt;
    int clamp_dt;
    double body_force[3];
    int inflow_shape;
    double inflow_location[3];
    double inflow_size[3];
    double inflow_center[3];
    double inflow_radius;
    double inflow_velocity[3];
    double inflow_quantity;
    int field_order;
    int velocity_order;
    double cg_tol;
    int cg_max_iter;
    int precision;
    double inner_tol;
    int inner_max_iter;
    int max_outer;
    int quirks;
    int refine_mode; /* CFO_REFINE_RESTART or CFO_REFINE_CARRY */
    int preconditioner;   /* CFO_PRECOND_JACOBI, CFO_PRECOND_MG or CFO_PRECOND_NONE */
    int mg_nu;            /* Jacobi sweeps before and after the coarse correction */
    double mg_omega;      /* Jacobi damping */
    int mg_coarse_sweeps; /* Jacobi sweeps on the coarsest level */
    double mg_alpha;      /* scale of the coarse-grid correction */
} cfo_config;

enum { CFO_REFINE_RESTART = 0, CFO_REFINE_CARRY = 1 };
enum { CFO_PRECOND_JACOBI = 0, CFO_PRECOND_MG = 1, CFO_PRECOND_NONE = 2 };

typedef struct cfo_stats {
    int converged;
    int iterations;       /* FP64 PCG iterations, or total inner iterations */
    int outer_iterations; /* defect-correction passes (FP32_REFINED) */
    double b_norm;
    double final_residual; /* ||b - A x|| / ||b|| as the solver last saw it */
} cfo_stats;

typedef struct cfo_state cfo_state;

int cfo_create(const cfo_config* cfg, cfo_state** out);
void cfo_destroy(cfo_state* s);

double cfo_dt(const cfo_state* s);
double cfo_dx(const cfo_state* s);
double cfo_time(const cfo_state* s);
void cfo_extents(const cfo_state* s, int entity, int ext[3]);
void cfo_owned(const cfo_state* s, int entity, int lo[3], int hi[3]);
double* cfo_field(cfo_state* s, int entity, int version);
double* cfo_lhs(cfo_state* s);
double* cfo_rhs(cfo_state* s);

/* Hot-path operators. */
int cfo_setup(cfo_state* s, cfo_stats* st);
int cfo_step(cfo_state* s, cfo_stats* st);
int cfo_run(cfo_state* s, int nsteps, cfo_stats* last);
void cfo_fill_ghosts(cfo_state* s, int version);
void cfo_time_integrate(cfo_state* s);
void cfo_add_inputs(cfo_state* s);
void cfo_build_rhs(cfo_state* s);
void cfo_rhs_remove_mean(cfo_state* s);
int cfo_pressure_solve(cfo_state* s, int precision, cfo_stats* st);
void cfo_apply_pressure(cfo_state* s);
int cfo_correct_velocity(cfo_state* s, cfo_stats* st);
void cfo_advance(cfo_state* s, int entity);

/* Building blocks exposed for pinning tests. */
void cfo_spline_weights(int order, double xi, int* s0, double* w);
double cfo_interpolate(const cfo_state* s, int entity, int order,
                       const double* field, const double p[3]);
void cfo_rk3(const cfo_state* s, const double x0[3], const double* u,
             const double* v, const double* w, double dt, double trace[3]);
void cfo_apply_A(cfo_state* s, const double* x, double* y); /* y = A x (owned) */
double cfo_dot(const cfo_state* s, const double* a, const double* b);

/* The writer's cell data: owned q and the linearly interpolated cell-centre
 * velocity, compact with x fastest (vel_out: dim arrays back to back). */
void cfo_cell_output(const cfo_state* s, double* q_out, double* vel_out);

/* Seeded generators (counter-based SplitMix64). */
uint64_t cfo_rand_u64(uint64_t seed, uint64_t stream, uint64_t idx);
double cfo_rand_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
double cfo_rand_normal(uint64_t seed, uint64_t stream, uint64_t idx);
void cfo_init_vortex_tubes(cfo_state* s, uint64_t seed, double peak);
void cfo_init_blobs(cfo_state* s, double* out_cells, uint64_t seed, int nblobs);
void cfo_init_random_faces(cfo_state* s, uint64_t seed);

/* Multigrid preconditioner (cf_oracle_mg.c): level count, level cell
 * counts, the level operator applied to a compact (x fastest, owned cells
 * only) vector, and one V-cycle z = B r on compact level-0 vectors. */
int cfo_mg_nlevels(cfo_state* s);
int cfo_mg_level_dims(cfo_state* s, int level, int n[3]);
int cfo_mg_apply_level(cfo_state* s, int level, const double* x, double* y);
int cfo_mg_vcycle(cfo_state* s, const double* r, double* z);
/* the V-cycle with level 0 in FP32 (f32 = 1) or FP64 (0) */
int cfo_mg_vcycle_prec(cfo_state* s, int f32, const double* r, double* z);

/* Adjoint (reverse mode) of setup + nsteps, checkpointing every ckpt steps.
 * Inputs: the initial state held in Current, the loss weights dL/dq_T in
 * dq_T (cell array).  Outputs: dL/d(initial u,v,w) into du0/dv0/dw0 (face
 * arrays, may be NULL for w in 2-D) and dL/d(inflow velocity component d)
 * into dvel[3].  The forward state after the call is the final state. */
int cfo_adjoint_run(cfo_state* s, int nsteps, int ckpt, const double* dq_T,
                    double* du0, double* dv0, double* dw0, double dvel[3],
                    double* dq0);

#ifdef __cplusplus
}
#endif
#endif
