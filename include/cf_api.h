/*
 * cf_api.h -- C ABI of the MI355X-native CajitaFluids hot path (libcfmi.so).
 *
 * One handle (cf_solver) is one rank's subdomain on one GPU: it owns every
 * device array, runs every kernel on its own HIP stream, and fills ghost
 * layers on the device in place of the reference's MPI halo exchange.
 * Plain C types only; no exceptions cross the ABI; every call returns a
 * cf_status.  Calls on one handle are stream-ordered and not thread-safe;
 * separate handles are independent.
 *
 * Each entry point names the reference interface it replaces
 * (CUP-ECS/cajitafluids, paths relative to the reference tree).
 */
#ifndef CF_API_H
#define CF_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
typedef int cf_status;
#define CF_OK 0
#define CF_EINVAL 1   /* invalid argument or configuration */
#define CF_ENOMEM 2   /* device or host allocation failed */
#define CF_ENOCONV 3  /* pressure solve hit its iteration limit */
#define CF_EDEVICE 4  /* HIP runtime error, or no usable gfx950 device */
#define CF_ENOTIMPL 5 /* combination not provided by this library */
#define CF_ENONFINITE 6 /* the pressure solve met a NaN or Inf in its right-hand
                           side or residual and stopped at once */

/* ---- tags (Cajita entity / ProblemManager version tags) ---------------- */
enum { CF_CELL = 0, CF_FACE_X = 1, CF_FACE_Y = 2, CF_FACE_Z = 3 };
enum { CF_CURRENT = 0, CF_NEXT = 1 };
enum { CF_SOLID = 0, CF_FREE = 1 }; /* BoundaryType (BoundaryConditions.hpp:22-29) */
enum { CF_INFLOW_NONE = 0, CF_INFLOW_BOX = 1, CF_INFLOW_SPHERE = 2 };
enum { CF_FP64 = 0, CF_FP32_REFINED = 1 };
enum { CF_REFINE_RESTART = 0, CF_REFINE_CARRY = 1 };
enum { CF_PRECOND_JACOBI = 0, CF_PRECOND_MG = 1, CF_PRECOND_NONE = 2 };

/* As-written reference behaviours, reproduced when set. */
#define CF_QUIRK_RK3_V0 1    /* 3/4-step probe uses v0: TimeIntegrator.hpp:57-58 */
#define CF_QUIRK_PROJ_BC_U 2 /* FaceJ(/K) wall rule applied to u after the
                                projection: VelocityCorrector.hpp:260 */

/*
 * Problem description.  Replaces the arguments of createSolver
 * (src/Solver.hpp:283-293): bounding box and cell counts (Mesh,
 * src/Mesh.hpp:41-103), BoundaryCondition::boundary_type
 * (src/BoundaryConditions.hpp:131-134; here wall_type[axis][lo/hi]),
 * InflowSource (src/InflowSource.hpp:80-90; box, or a sphere), BodyForce
 * (src/BodyForce.hpp:62-66), density, dt and the pressure-solver choice
 * (src/VelocityCorrector.hpp:297-339: Jacobi-PCG for "Reference", a
 * multigrid V-cycle standing in for HYPRE's PCG + PFMG, or plain CG; see
 * `preconditioner` below).  The initial state is set with cf_copy_in,
 * cf_copy_in_device or one of the cf_init_* generators in place of the
 * reference's init functor (src/ProblemManager.hpp:186-263); cf_entity_origin
 * gives the coordinate that functor is evaluated from.
 */
typedef struct cf_config {
    int dim;              /* 2 or 3 */
    int num_cell[3];      /* owned cells per axis (num_cell[2] = 1 in 2-D) */
    int halo;             /* ghost width; the reference uses 3 (Solver.hpp:78) */
    int periodic[3];      /* per axis; the reference is never periodic */
    double lo[3], hi[3];  /* bounding box */
    int wall_type[3][2];  /* CF_SOLID / CF_FREE per axis, [lo, hi] */
    double density;
    double dt;
    int clamp_dt;         /* apply Solver.hpp:93-106 */
    double body_force[3];
    int inflow_shape;     /* CF_INFLOW_* */
    double inflow_location[3], inflow_size[3]; /* box: half-open [loc, loc+size) */
    double inflow_center[3], inflow_radius;    /* sphere: |x - c|^2 < r^2 */
    double inflow_velocity[3];
    double inflow_quantity;
    int field_order;      /* 1 (linear) or 3 (cubic) B-spline for advected fields */
    int velocity_order;   /* B-spline order of the RK3 velocity samples (1) */
    double cg_tol;        /* relative residual ||r|| / ||b|| */
    int cg_max_iter;
    int precision;        /* CF_FP64 or CF_FP32_REFINED */
    double inner_tol;     /* FP32 inner relative tolerance */
    int inner_max_iter;
    int max_outer;        /* defect-correction passes */
    int quirks;           /* CF_QUIRK_* */
    int refine_mode;      /* CF_REFINE_RESTART: each pass restarts the FP32
                             PCG; CF_REFINE_CARRY: passes keep the search
                             direction (one FP32 PCG run whose residual is
                             replaced by the FP64 one at each pass) */
    /* Pressure-solve preconditioner (VelocityCorrector.hpp:300-339 names
     * the choice; the reference's default is HYPRE's PCG + PFMG): Jacobi,
     * a geometric multigrid V-cycle, or none ("None",
     * VelocityCorrector.hpp:327: plain CG).  With MG, CF_FP32_REFINED runs
     * the V-cycle's finest level on FP32 vectors inside the FP64 CG (3-D
     * grids with more than 2 cells along every axis; CF_ENOTIMPL otherwise).
     * With MG or none, the solve returns the zero-mean pressure on a
     * singular operator. */
    int preconditioner;   /* CF_PRECOND_JACOBI, CF_PRECOND_MG or CF_PRECOND_NONE */
    int mg_nu;            /* Jacobi sweeps before and after each coarse
                             correction (>= 1) */
    double mg_omega;      /* Jacobi damping in (0, 1] */
    int mg_coarse_sweeps; /* Jacobi sweeps on the coarsest level (>= 1) */
    double mg_alpha;      /* scale of the coarse-grid correction, x += alpha
                             P x_c, in (0, 2]: over-correction (about 1.9)
                             offsets the piecewise-constant prolongation */
} cf_config;

typedef struct cf_solve_stats {
    int converged;
    int iterations;       /* PCG iterations (FP32_REFINED: total inner) */
    int outer_iterations; /* defect-correction passes */
    double b_norm;
    double final_residual; /* ||b - A x|| / ||b|| at exit */
} cf_solve_stats;

/* Borrowed device view of one state array; valid until the next
 * cf_advance / cf_destroy on the handle.  Element (i, j, k) of the entity's
 * ghosted local index space is at data[offset + i + pitch_x*(j + pitch_y*k)]. */
typedef struct cf_view {
    double* data;
    int64_t offset, pitch_x, pitch_y;
    int ghost_ext[3];         /* ghosted extent of the entity */
    int own_lo[3], own_hi[3]; /* owned local index box [lo, hi) */
} cf_view;

typedef struct cf_solver cf_solver;

/* ---- geometry queries (host only, no device needed) -------------------- */

/* Ghosted extents and owned index box of an entity, as Cajita's local grid
 * would report them (tests/tstMesh.cpp:32-69 pins the 2-D form). */
cf_status cf_layout_query(const cf_config* cfg, int entity, int ghost_ext[3],
                          int own_lo[3], int own_hi[3]);
/* Mesh::cellSize (src/Mesh.hpp:119-122) and the effective dt after the
 * clamp of Solver.hpp:93-106. */
cf_status cf_geometry_query(const cf_config* cfg, double* dx, double* dt);
/* The padded device layout every array of a handle shares (the layout of
 * cf_view): element (i, j, k) of any entity's ghosted local index space is at
 * offset + i + pitch_x * (j + pitch_y * k), and an array holds `elems`
 * doubles.  Callers of the *_device entry points allocate arrays of this
 * size.  Stands in for the Kokkos view layout of ProblemManager::get
 * (src/ProblemManager.hpp:281-350). */
cf_status cf_padded_layout(const cf_config* cfg, int64_t* offset, int64_t* pitch_x,
                           int64_t* pitch_y, int64_t* elems);
/* Coordinate of local index 0 of an entity: LocalMesh::coordinates(entity,
 * {0, 0}, loc) as ProblemManager::initialize calls it
 * (src/ProblemManager.hpp:202-204, 222-223, 243-244).  The init functor is
 * evaluated at x = origin + cellSize * i for local index i (:210-211). */
cf_status cf_entity_origin(const cf_config* cfg, int entity, double origin[3]);

/* ---- lifecycle: Solver ctor / createSolver (src/Solver.hpp:69-123, 283-350) */
cf_status cf_create(const cf_config* cfg, cf_solver** out);
void cf_destroy(cf_solver* h);
/* Last error message of the handle, or of the last failed cf_create on this
 * thread when h is NULL. */
const char* cf_last_error(const cf_solver* h);

/* ---- the step: SolverBase (src/Solver.hpp:41-48) ------------------------ */
cf_status cf_setup(cf_solver* h, cf_solve_stats* st);  /* Solver::setup, :125-133 */
cf_status cf_step(cf_solver* h, cf_solve_stats* st);   /* Solver::step, :135-147 */
/* Solver::solve (:149-177) without I/O: exactly nsteps calls of step. */
cf_status cf_run(cf_solver* h, int nsteps, cf_solve_stats* last);
double cf_time(const cf_solver* h);
double cf_dt(const cf_solver* h);
double cf_dx(const cf_solver* h);

/* ---- the phases --------------------------------------------------------- */
/* TimeIntegrator::step (src/TimeIntegrator.hpp:120-177). */
cf_status cf_time_integrate(cf_solver* h);
/* Solver::_addInputs (src/Solver.hpp:181-263). */
cf_status cf_add_inputs(cf_solver* h);
/* VelocityCorrectorBase::correctVelocity (src/VelocityCorrector.hpp:37-42,
 * 266-282). */
cf_status cf_correct_velocity(cf_solver* h, cf_solve_stats* st);
/* VelocityCorrector::_buildRHS (src/VelocityCorrector.hpp:182-212). */
cf_status cf_build_rhs(cf_solver* h);
/* Subtract the mean of the owned RHS (makes a random pure-Neumann RHS
 * consistent; no reference counterpart). */
cf_status cf_rhs_remove_mean(cf_solver* h);
/* The pressure solve on the handle's own rhs/lhs arrays, with lhs = 0 first
 * (src/VelocityCorrector.hpp:272-276 -> Cajita ReferenceConjugateGradient::
 * solve).  precision: CF_FP64 or CF_FP32_REFINED. */
cf_status cf_pressure_solve(cf_solver* h, int precision, cf_solve_stats* st);
/* Solve A x = b for a caller-provided device rhs/lhs in the handle's cell
 * layout (cf_view of CF_CELL); lhs is zeroed first. */
cf_status cf_pressure_solve_ext(cf_solver* h, const double* rhs_dev,
                                double* lhs_dev, int precision,
                                cf_solve_stats* st);
/* VelocityCorrector::_applyPressure (src/VelocityCorrector.hpp:214-264). */
cf_status cf_apply_pressure(cf_solver* h);

/* ---- state: ProblemManager (src/ProblemManager.hpp:281-403) ------------- */
cf_status cf_field_view(cf_solver* h, int entity, int version, cf_view* out);
/* Views of the pressure lhs and rhs arrays (VelocityCorrector.hpp:83-90). */
cf_status cf_pressure_view(cf_solver* h, int which /*0 lhs, 1 rhs*/, cf_view* out);
/* Copy the full ghosted extent of an entity array, dense, x fastest. */
cf_status cf_copy_in(cf_solver* h, int entity, int version, const double* host);
cf_status cf_copy_out(cf_solver* h, int entity, int version, double* host);
/* which: 0 lhs (pressure), 1 rhs */
cf_status cf_copy_pressure_in(cf_solver* h, int which, const double* host);
cf_status cf_copy_pressure_out(cf_solver* h, int which, double* host);
/* The same copies from / into a caller's DEVICE array in the padded layout
 * (cf_padded_layout; every element, ghosts included), by a kernel on the
 * handle's stream: stream-ordered like every call on the handle.  The
 * caller's own writes must be complete before cf_copy_in_device (another
 * stream: synchronise it first), and its reads of the array written by
 * cf_copy_out_device follow cf_sync.  ProblemManager::get's device views
 * (src/ProblemManager.hpp:281-350) as copies into caller-owned memory. */
cf_status cf_copy_in_device(cf_solver* h, int entity, int version, const double* dev);
cf_status cf_copy_out_device(cf_solver* h, int entity, int version, double* dev);
cf_status cf_advance(cf_solver* h, int entity);          /* ProblemManager::advance */
cf_status cf_fill_ghosts(cf_solver* h, int version);     /* ProblemManager::gather */
cf_status cf_sync(cf_solver* h);

/* ---- seeded initial state (replaces the init functor) ------------------- */
/* Divergence-free vortex tubes (3-D) / point vortices (2-D), peak |u| = peak. */
cf_status cf_init_vortex_tubes(cf_solver* h, uint64_t seed, double peak);
/* Dye = sum of nblobs seeded Gaussian blobs. */
cf_status cf_init_blobs(cf_solver* h, uint64_t seed, int nblobs);
/* Every owned face velocity i.i.d. N(0, 1). */
cf_status cf_init_random_faces(cf_solver* h, uint64_t seed);

/* ---- measurement -------------------------------------------------------- */
/* enable > 0: bracket kernel launches with HIP events on the handle's stream
 * and accumulate per-kernel device time; the PCG kernels are timed in one
 * iteration out of every `enable` (1: every iteration).  0: off. */
cf_status cf_set_kernel_timing(cf_solver* h, int enable);
/* kernel: 0 FP64 direction+SpMV, 1 FP64 update, 2 FP64 finalize kernels
 * (unused), 3-5 the same for the FP32 inner solve, 6 advection, 7 FP64
 * defect-correction residual, 8 one multigrid V-cycle (all its kernels),
 * 9 one launch of the cooperative FP64 PCG that 2-D grids (up to 2048
 * cells along x and 256 strips of 2048 cells) run in place of 0 and 1 */
cf_status cf_kernel_time(cf_solver* h, int kernel, double* total_ms,
                         int64_t* launches);
cf_status cf_reset_kernel_time(cf_solver* h);

/* ---- adjoint (no reference counterpart; build extension) --------------- */
/* Reverse mode of setup() followed by nsteps step() calls from the state in
 * Current, recomputing from checkpoints taken every ckpt steps.  dq_T_host:
 * dL/dq at the final step (dense ghosted cell array).  Outputs (dense
 * ghosted, may be NULL): dL/d(initial u, v, w, q); dvel[3]: dL/d(inflow
 * velocity component).  dt is held fixed.  On return Current holds the final
 * forward state. */
cf_status cf_adjoint_run(cf_solver* h, int nsteps, int ckpt,
                         const double* dq_T_host, double* du0_host,
                         double* dv0_host, double* dw0_host, double* dq0_host,
                         double dvel[3]);
/* cf_adjoint_run on device arrays in the padded layout (cf_padded_layout),
 * stream-ordered on the handle's stream, with no host transfer of any array:
 * dq_T_dev's non-owned entries are ignored; each gradient array that is not
 * NULL receives dL/d(initial entity) on the owned entries and 0 elsewhere.
 * The caller's writes of dq_T_dev must be complete before the call; the
 * gradients are complete when it returns (dvel is read back to the host).
 * On any error the handle's Current state and time are those it started
 * from. */
cf_status cf_adjoint_run_device(cf_solver* h, int nsteps, int ckpt,
                                const double* dq_T_dev, double* du0_dev,
                                double* dv0_dev, double* dw0_dev, double* dq0_dev,
                                double dvel[3]);

/* ---- no-allocation pressure solve (no counterpart in the reference) ----- */
/* For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc):
 * cf_pressure_solve_ext's FP64 solve (Jacobi-PCG, or plain CG with
 * preconditioner CF_PRECOND_NONE) of A x = b for cfg, with no handle.  rhs_dev
 * and lhs_dev are the caller's device arrays in the padded layout
 * (cf_padded_layout); lhs is zeroed first.  Every work array (r, two
 * directions, the reductions' partials and counters, the PCG scalars) lives
 * in one static device array of CF_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; nothing is allocated.  Grids whose padded layout holds
 * more than the limit of cf_noalloc_limits (4,183,360 elements: with 3 ghost
 * layers, 3-D grids up to 147^3 and 2-D grids up to 2028^2) are rejected with CF_EINVAL before any launch;
 * FP32_REFINED and the multigrid preconditioner give CF_ENOTIMPL.  stream: a
 * hipStream_t, or NULL for the null stream; the call returns after the solve
 * (it polls the device for convergence).  Every call on a GPU shares that
 * GPU's scratch: calls that overlap on one GPU race.  Errors are reported
 * through cf_last_error(NULL). */
#define CF_NOALLOC_SCRATCH_BYTES (96u << 20)
cf_status cf_pressure_solve_noalloc(const cf_config* cfg, const double* rhs_dev,
                                    double* lhs_dev, void* stream, cf_solve_stats* st);
/* The scratch size and the largest padded-layout element count it serves. */
cf_status cf_noalloc_limits(size_t* scratch_bytes, int64_t* max_elems);

/* ---- output (replaces SiloWriter, src/SiloWriter.hpp:32-422) ------------ */
/* SiloWriter::writeFile's data (SiloWriter.hpp:141-197): the owned cells of
 * q (Current) and the cell-centre velocity from linear interpolation of each
 * face component, as compact host arrays with x fastest: q_host n0*n1*n2,
 * vel_host dim arrays of n0*n1*n2 back to back. */
cf_status cf_cell_output(cf_solver* h, double* q_host, double* vel_host);
/* SiloWriter::siloWrite (SiloWriter.hpp:352-418) in VTK XML instead of Silo:
 * <dir>/raw/CajitaFluidsOutput00000<step:05d>.vti (ImageData: CellData
 * "quantity" and "velocity" (3 components), FieldData CYCLE / TIME / DTIME)
 * and the master <dir>/CajitaFluids<step:05d>.pvti naming it; creates dir
 * and dir/raw. */
cf_status cf_write_output(cf_solver* h, const char* dir, int step);
/* The piece writer alone, from host arrays laid out as cf_cell_output's
 * (no handle, no device use); origin: the domain's low corner. */
cf_status cf_write_vti(const char* path, int dim, const int n[3], const double origin[3],
                       double dx, int cycle, double time, double dt, const double* q,
                       const double* vel);

#ifdef __cplusplus
}
#endif
#endif
