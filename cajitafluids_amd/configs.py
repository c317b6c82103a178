"""The five workloads of this hot path, as plain config dicts.

Field names and meaning are those of ``cf_config`` (include/cf_api.h).  The
defaults that are not workload-specific follow the reference's example
driver: density 0.1, gravity 0, four SOLID walls, zero initial state
(examples/advection.cpp:174-189, 445-454).
"""
import copy

CELL, FACE_X, FACE_Y, FACE_Z = 0, 1, 2, 3
SOLID, FREE = 0, 1
INFLOW_NONE, INFLOW_BOX, INFLOW_SPHERE = 0, 1, 2
FP64, FP32_REFINED = 0, 1
REFINE_RESTART, REFINE_CARRY = 0, 1
PRECOND_JACOBI, PRECOND_MG, PRECOND_NONE = 0, 1, 2
QUIRK_RK3_V0 = 1      # TimeIntegrator.hpp:57-58 (3/4 probe uses v0)
QUIRK_PROJ_BC_U = 2   # VelocityCorrector.hpp:260 (FaceJ BC applied to u)


def base_config(dim=2, n=128):
    return dict(
        dim=dim,
        num_cell=[n, n, n if dim == 3 else 1],
        halo=3,
        periodic=[0, 0, 0],
        lo=[0.0, 0.0, 0.0],
        hi=[1.0, 1.0, 1.0],
        wall_type=[[SOLID, SOLID], [SOLID, SOLID], [SOLID, SOLID]],
        density=0.1,
        dt=0.005,
        clamp_dt=1,
        body_force=[0.0, 0.0, 0.0],
        inflow_shape=INFLOW_NONE,
        inflow_location=[0.0, 0.0, 0.0],
        inflow_size=[0.0, 0.0, 0.0],
        inflow_center=[0.0, 0.0, 0.0],
        inflow_radius=0.0,
        inflow_velocity=[0.0, 0.0, 0.0],
        inflow_quantity=0.0,
        field_order=3,
        velocity_order=1,
        cg_tol=1e-8,
        cg_max_iter=20000,
        precision=FP64,
        inner_tol=1e-2,
        inner_max_iter=20000,
        max_outer=50,
        quirks=QUIRK_RK3_V0 | (QUIRK_PROJ_BC_U if dim == 2 else 0),
        refine_mode=REFINE_CARRY,
        preconditioner=PRECOND_JACOBI,
        mg_nu=3,
        mg_omega=0.9,
        mg_coarse_sweeps=16,
        mg_alpha=1.9,
    )


def dye_tank_2d(n=128, field_order=1):
    """The 2-D dye tank: n x n MAC grid on the unit square, FP64, dt = 0.005
    (clamped by the reference rule to dx / 3), linear field sampling, zero
    initial state, inflow box [0.45, 0.55] x [0.20, 0.22] with q = 1 and
    v = 3, SOLID walls, Jacobi-PCG to 1e-8; 50 RK3 steps."""
    c = base_config(2, n)
    c.update(
        field_order=field_order,
        inflow_shape=INFLOW_BOX,
        inflow_location=[0.45, 0.20, 0.0],
        inflow_size=[0.10, 0.02, 0.0],
        inflow_velocity=[0.0, 3.0, 0.0],
        inflow_quantity=1.0,
    )
    return c


DYE_TANK_STEPS = 50


def vortex_tank_3d(n=256, precision=FP64):
    """The 3-D vortex tank: n^3 MAC grid, FP64, cubic field sampling, dt = 0.002
    (clamped to dx / 2 = 1/512 at n = 256), initial velocity = seeded
    divergence-free vortex tubes (seed 1, peak |u| = 1), inflow sphere
    r = 0.05 at (0.5, 0.5, 0.1) with q = 1 and w = 2, SOLID walls,
    Jacobi-PCG to 1e-8; 20 RK3 steps."""
    c = base_config(3, n)
    c.update(
        dt=0.002,
        field_order=3,
        inflow_shape=INFLOW_SPHERE,
        inflow_center=[0.5, 0.5, 0.1],
        inflow_radius=0.05,
        inflow_velocity=[0.0, 0.0, 2.0],
        inflow_quantity=1.0,
        precision=precision,
    )
    return c


VORTEX_TANK_STEPS = 20
VORTEX_TANK_SEED = 1


def random_solve_3d(n=512):
    """The random-velocity pressure solve on an n^3 grid (no step).  Face velocities i.i.d.
    N(0,1) (seed 3), RHS = -div / dx with its mean removed; FP32 inner CG to
    1e-4 inside FP64 defect correction to 1e-8 (each pass restarts the inner
    CG).  `refine_mode=REFINE_CARRY, inner_tol=1e-2` is the faster variant
    measured beside it."""
    c = base_config(3, n)
    c.update(dt=0.002, clamp_dt=0, precision=FP32_REFINED, inner_tol=1e-4,
             refine_mode=REFINE_RESTART)
    return c


RANDOM_SOLVE_SEED = 3


def adjoint_tank_2d(n=512):
    """The 2-D dye-tank adjoint: n^2, FP64, linear fields, the dye tank's inflow and walls,
    dt held at the clamp value for speed 3 (dx / 3); 100 steps; loss
    0.5 ||q_T - target||^2 with target = 8 seeded Gaussian blobs (seed 4);
    checkpoint every 10 steps."""
    c = dye_tank_2d(n, field_order=1)
    c.update(dt=(1.0 / n) / 3.0, clamp_dt=0)
    return c


DYE_ADJOINT_STEPS = 100
DYE_ADJOINT_CKPT = 10
DYE_ADJOINT_TARGET_SEED = 4
DYE_ADJOINT_FD_SEED = 44


def periodic_box_3d(n=384):
    """The 3-D periodic box: n^3 single-rank subdomain with 2-cell ghosts, periodic
    in x and y, SOLID walls in z, cubic fields, FP32-refined PCG to 1e-8,
    vortex tubes (seed 5), dye = seeded Gaussian blobs (seed 55); 10-step
    adjoint checkpointed every 2 steps."""
    c = base_config(3, n)
    c.update(
        halo=2,
        periodic=[1, 1, 0],
        dt=0.002,
        field_order=3,
        precision=FP32_REFINED,
        quirks=QUIRK_RK3_V0,
    )
    return c


PERIODIC_BOX_STEPS = 10
PERIODIC_BOX_CKPT = 2
PERIODIC_BOX_VEL_SEED = 5
PERIODIC_BOX_DYE_SEED = 55


def copy_config(c):
    return copy.deepcopy(c)
