"""Python mirror of the reference's solver interface over the C ABI.

``Solver`` follows ``SolverBase`` (src/Solver.hpp:41-48: setup, step,
solve) and exposes the phases and the ProblemManager state of one handle;
``create_solver`` follows ``createSolver`` (src/Solver.hpp:283-350) with the
reference's BoundaryCondition / InflowSource / BodyForce value objects.
Every call goes to libcfmi.so; nothing here computes physics.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .configs import (CELL, FACE_X, FACE_Y, FACE_Z, FP32_REFINED, FP64,
                      INFLOW_BOX, INFLOW_NONE, PRECOND_JACOBI, PRECOND_MG, PRECOND_NONE,
                      QUIRK_PROJ_BC_U, QUIRK_RK3_V0, SOLID, base_config)

CURRENT, NEXT = 0, 1


def _check(L, h, rc, what):
    if rc != _lib.CF_OK:
        msg = L.cf_last_error(h)
        raise _lib.CfError(rc, f"{what}: {msg.decode() if msg else ''}")


def layout(cfg: dict, entity: int):
    """(ghosted extents, owned lo, owned hi) of an entity, host-only."""
    L = _lib.load()
    st = _lib.config_struct(cfg)
    ext = (C.c_int * 3)()
    lo = (C.c_int * 3)()
    hi = (C.c_int * 3)()
    rc = L.cf_layout_query(C.byref(st), entity, ext, lo, hi)
    if rc != 0:
        raise _lib.CfError(rc, "cf_layout_query")
    return tuple(ext), tuple(lo), tuple(hi)


def padded_layout(cfg: dict):
    """(offset, pitch_x, pitch_y, elems) of the padded device layout every
    array of a handle uses: element (i, j, k) at offset + i + pitch_x (j +
    pitch_y k) of an array of elems doubles.  Host-only."""
    L = _lib.load()
    st = _lib.config_struct(cfg)
    vals = [C.c_int64() for _ in range(4)]
    rc = L.cf_padded_layout(C.byref(st), *(C.byref(v) for v in vals))
    if rc != 0:
        raise _lib.CfError(rc, "cf_padded_layout")
    return tuple(v.value for v in vals)


def entity_origin(cfg: dict, entity: int):
    """Coordinate of local index 0 of an entity (LocalMesh::coordinates at
    index 0, src/ProblemManager.hpp:202-204).  Host-only."""
    L = _lib.load()
    st = _lib.config_struct(cfg)
    out = (C.c_double * 3)()
    rc = L.cf_entity_origin(C.byref(st), entity, out)
    if rc != 0:
        raise _lib.CfError(rc, "cf_entity_origin")
    return tuple(out)


def geometry(cfg: dict):
    """(dx, effective dt) of a config, host-only."""
    L = _lib.load()
    st = _lib.config_struct(cfg)
    dx = C.c_double()
    dt = C.c_double()
    rc = L.cf_geometry_query(C.byref(st), C.byref(dx), C.byref(dt))
    if rc != 0:
        raise _lib.CfError(rc, "cf_geometry_query")
    return dx.value, dt.value


class Solver:
    """One subdomain on one GPU (a cf_solver handle)."""

    def __init__(self, cfg: dict):
        self.L = _lib.load()
        self.cfg = dict(cfg)
        self.dim = cfg["dim"]
        st = _lib.config_struct(cfg)
        h = C.c_void_p()
        rc = self.L.cf_create(C.byref(st), C.byref(h))
        if rc != 0:
            msg = self.L.cf_last_error(None)
            raise _lib.CfError(rc, f"cf_create: {msg.decode() if msg else ''}")
        self.h = h
        self.last_stats = None

    def close(self):
        if getattr(self, "h", None):
            self.L.cf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- SolverBase ----------------------------------------------------------
    def _stats_call(self, fn, *args, what=""):
        st = _lib.CfSolveStats()
        rc = fn(self.h, *args, C.byref(st))
        self.last_stats = st.as_dict()
        _check(self.L, self.h, rc, what)
        return self.last_stats

    def setup(self):
        return self._stats_call(self.L.cf_setup, what="setup")

    def step(self):
        return self._stats_call(self.L.cf_step, what="step")

    def run(self, nsteps):
        return self._stats_call(self.L.cf_run, nsteps, what="run")

    def solve_steps(self, nsteps):
        """setup, then exactly nsteps steps (no output)."""
        self.setup()
        return self.run(nsteps)

    def solve(self, t_final, write_freq, output_dir="data", log=print):
        """Solver::solve (src/Solver.hpp:148-176): write step 0, setup, then
        step until time >= t_final; every write_freq-th loop counter t logs
        a line and, after that step, writes output labelled t (the state
        after t + 1 steps, as there).  output_dir None (or write_freq <= 0)
        writes nothing.  Returns the number of steps taken."""
        write = output_dir is not None and write_freq > 0
        if write:
            self.write_output(output_dir, 0)
        num_step = int(t_final / self.dt)
        self.setup()
        t = 0
        while True:
            due = write_freq > 0 and t % write_freq == 0
            if due and log is not None:
                log("Step %d / %d at time = %f" % (t, num_step, self.time))
            self.step()
            if due and write:
                self.write_output(output_dir, t)
            t += 1
            if not self.time < t_final:
                return t

    @property
    def dt(self):
        return self.L.cf_dt(self.h)

    @property
    def dx(self):
        return self.L.cf_dx(self.h)

    @property
    def time(self):
        return self.L.cf_time(self.h)

    # ---- phases ----------------------------------------------------------------
    def _call(self, fn, *args, what=""):
        rc = fn(self.h, *args)
        _check(self.L, self.h, rc, what)

    def time_integrate(self):
        self._call(self.L.cf_time_integrate, what="time_integrate")

    def add_inputs(self):
        self._call(self.L.cf_add_inputs, what="add_inputs")

    def correct_velocity(self):
        return self._stats_call(self.L.cf_correct_velocity, what="correct_velocity")

    def build_rhs(self):
        self._call(self.L.cf_build_rhs, what="build_rhs")

    def rhs_remove_mean(self):
        self._call(self.L.cf_rhs_remove_mean, what="rhs_remove_mean")

    def pressure_solve(self, precision=FP64):
        return self._stats_call(self.L.cf_pressure_solve, precision, what="pressure_solve")

    def pressure_solve_ext(self, rhs_dev: int, lhs_dev: int, precision=FP64):
        return self._stats_call(self.L.cf_pressure_solve_ext, C.c_void_p(rhs_dev),
                                C.c_void_p(lhs_dev), precision, what="pressure_solve_ext")

    def apply_pressure(self):
        self._call(self.L.cf_apply_pressure, what="apply_pressure")

    # ---- state -----------------------------------------------------------------
    def extents(self, entity):
        return layout(self.cfg, entity)[0]

    def owned(self, entity):
        _, lo, hi = layout(self.cfg, entity)
        return lo, hi

    def get(self, entity, version=CURRENT):
        """Dense copy of the ghosted array, shaped (z, y, x)."""
        ex = self.extents(entity)
        out = np.zeros((ex[2], ex[1], ex[0]))
        self._call(self.L.cf_copy_out, entity, version,
                   out.ctypes.data_as(C.POINTER(C.c_double)), what="copy_out")
        return out

    def set(self, entity, arr, version=CURRENT):
        ex = self.extents(entity)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (ex[2], ex[1], ex[0]):
            raise ValueError(f"expected shape {(ex[2], ex[1], ex[0])}, got {a.shape}")
        self._call(self.L.cf_copy_in, entity, version,
                   a.ctypes.data_as(C.POINTER(C.c_double)), what="copy_in")

    def get_pressure(self, which=0):
        ex = self.extents(CELL)
        out = np.zeros((ex[2], ex[1], ex[0]))
        self._call(self.L.cf_copy_pressure_out, which,
                   out.ctypes.data_as(C.POINTER(C.c_double)), what="copy_pressure_out")
        return out

    def set_pressure(self, arr, which=1):
        ex = self.extents(CELL)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (ex[2], ex[1], ex[0]):
            raise ValueError(f"expected shape {(ex[2], ex[1], ex[0])}, got {a.shape}")
        self._call(self.L.cf_copy_pressure_in, which,
                   a.ctypes.data_as(C.POINTER(C.c_double)), what="copy_pressure_in")

    def owned_view(self, arr, entity):
        lo, hi = self.owned(entity)
        return arr[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]

    def view(self, entity, version=CURRENT):
        v = _lib.CfView()
        self._call(self.L.cf_field_view, entity, version, C.byref(v), what="field_view")
        return v

    def pressure_view(self, which=0):
        v = _lib.CfView()
        self._call(self.L.cf_pressure_view, which, C.byref(v), what="pressure_view")
        return v

    # ---- device-resident state (caller device memory, padded layout) ---------
    def copy_in_device(self, entity, ptr, version=CURRENT):
        """State array from a device array of padded_layout(cfg)[3] doubles
        (every element, ghosts included).  The caller's writes must be
        complete (synchronise their stream first)."""
        self._call(self.L.cf_copy_in_device, entity, version, C.c_void_p(int(ptr)),
                   what="copy_in_device")

    def copy_out_device(self, entity, ptr, version=CURRENT):
        """State array into a device array in the padded layout; complete
        after sync()."""
        self._call(self.L.cf_copy_out_device, entity, version, C.c_void_p(int(ptr)),
                   what="copy_out_device")

    def advance(self, entity):
        self._call(self.L.cf_advance, entity, what="advance")

    def gather(self, version=CURRENT):
        self._call(self.L.cf_fill_ghosts, version, what="fill_ghosts")

    def sync(self):
        self._call(self.L.cf_sync, what="sync")

    # ---- seeded initial state -------------------------------------------------
    def init_vortex_tubes(self, seed, peak=1.0):
        self._call(self.L.cf_init_vortex_tubes, seed, peak, what="init_vortex_tubes")

    def init_blobs(self, seed, nblobs):
        self._call(self.L.cf_init_blobs, seed, nblobs, what="init_blobs")

    def init_random_faces(self, seed):
        self._call(self.L.cf_init_random_faces, seed, what="init_random_faces")

    # ---- measurement -----------------------------------------------------------
    def set_kernel_timing(self, on=True):
        """on: False/0 off; True/1 every PCG iteration; N > 1 one in N."""
        n = int(on) if not isinstance(on, bool) else (1 if on else 0)
        self._call(self.L.cf_set_kernel_timing, n, what="set_kernel_timing")

    def kernel_time(self, kernel):
        ms = C.c_double()
        n = C.c_int64()
        self._call(self.L.cf_kernel_time, kernel, C.byref(ms), C.byref(n), what="kernel_time")
        return ms.value, n.value

    def reset_kernel_time(self):
        self._call(self.L.cf_reset_kernel_time, what="reset_kernel_time")

    # ---- output (SiloWriter) -------------------------------------------------------
    def cell_output(self):
        """SiloWriter::writeFile's data: (q, [vel_x, vel_y(, vel_z)]), the
        owned q and the cell-centre velocity from linear interpolation of the
        faces, each an (n2, n1, n0) array."""
        g = self._owned_cells()
        N = g[0] * g[1] * g[2]
        q = np.zeros(N)
        vel = np.zeros(self.dim * N)
        self._call(self.L.cf_cell_output, q.ctypes.data_as(_lib._dp),
                   vel.ctypes.data_as(_lib._dp), what="cell_output")
        shape = (g[2], g[1], g[0])
        return q.reshape(shape), [vel[d * N:(d + 1) * N].reshape(shape) for d in range(self.dim)]

    def write_output(self, directory, step):
        """SiloWriter::siloWrite: <directory>/CajitaFluids<step>.pvti naming
        <directory>/raw/CajitaFluidsOutput00000<step>.vti (VTK XML)."""
        self._call(self.L.cf_write_output, str(directory).encode(), int(step),
                   what="write_output")

    def _owned_cells(self):
        lo, hi = self.owned(CELL)
        return [hi[d] - lo[d] for d in range(3)]

    # ---- adjoint ---------------------------------------------------------------
    def adjoint_run(self, nsteps, ckpt, dq_T):
        """Reverse mode of setup + nsteps steps; returns dL/d(initial state)
        arrays (dense ghosted) and dL/d(inflow velocity)."""
        dqT = np.ascontiguousarray(dq_T, dtype=np.float64)
        ex = {e: self.extents(e) for e in range(self.dim + 1)}
        if dqT.shape != tuple(ex[CELL][::-1]):
            raise ValueError(f"dq_T must be the dense ghosted cell array of shape "
                             f"{tuple(ex[CELL][::-1])}, got {dqT.shape}")
        du = np.zeros(ex[FACE_X][::-1])
        dv = np.zeros(ex[FACE_Y][::-1])
        dw = np.zeros(ex[FACE_Z][::-1]) if self.dim == 3 else None
        dq0 = np.zeros(ex[CELL][::-1])
        dvel = np.zeros(3)
        P = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
        self._call(self.L.cf_adjoint_run, nsteps, ckpt, P(dqT), P(du), P(dv), P(dw),
                   P(dq0), P(dvel), what="adjoint_run")
        return dict(du=du, dv=dv, dw=dw, dq0=dq0, dvel=dvel)

    def adjoint_run_device(self, nsteps, ckpt, dq_T_ptr, du_ptr=None, dv_ptr=None,
                           dw_ptr=None, dq0_ptr=None):
        """cf_adjoint_run_device: every array a device pointer in the padded
        layout (None: not wanted).  Returns dL/d(inflow velocity); the
        gradient arrays are complete on return."""
        dvel = np.zeros(3)
        P = lambda p: None if p is None else C.c_void_p(int(p))
        self._call(self.L.cf_adjoint_run_device, nsteps, ckpt, P(dq_T_ptr), P(du_ptr),
                   P(dv_ptr), P(dw_ptr), P(dq0_ptr), dvel.ctypes.data_as(_lib._dp),
                   what="adjoint_run_device")
        return dvel


# ---- the reference's value objects and factory -----------------------------------
@dataclass
class BoundaryCondition:
    """BoundaryCondition::boundary_type in the reference order: left, bottom,
    right, top (src/BoundaryConditions.hpp:131-134); 3-D appends back, front
    after the 2-D four as lo-z, hi-z."""
    boundary_type: List[int] = field(default_factory=lambda: [SOLID] * 4)


@dataclass
class InflowSource:
    """InflowSource(location, size, velocity, quantity)
    (src/InflowSource.hpp:80-90); half-open box [location, location+size)."""
    location: List[float]
    size: List[float]
    velocity: List[float]
    quantity: float


@dataclass
class BodyForce:
    """BodyForce(fx, fy[, fz]) (src/BodyForce.hpp:62-66)."""
    force: List[float]


def create_solver(device: str, global_bounding_box, global_num_cell, density: float,
                  bc: BoundaryCondition, source: Optional[InflowSource],
                  body: Optional[BodyForce], delta_t: float,
                  matrix_solver: str = "Reference", preconditioner: str = "Jacobi",
                  **options) -> Solver:
    """createSolver (src/Solver.hpp:283-350): a Solver for solver_config's
    configuration of the same arguments."""
    return Solver(solver_config(device, global_bounding_box, global_num_cell, density, bc,
                                source, body, delta_t, matrix_solver, preconditioner,
                                **options))


# HYPRE preconditioner names accepted with matrix solver "PCG"
PCG_PRECONDITIONERS = {"PFMG": PRECOND_MG, "SMG": PRECOND_MG, "Jacobi": PRECOND_JACOBI,
                       "Diagonal": PRECOND_JACOBI, "None": PRECOND_NONE, "none": PRECOND_NONE}


def solver_config(device: str, global_bounding_box, global_num_cell, density: float,
                  bc: BoundaryCondition, source: Optional[InflowSource],
                  body: Optional[BodyForce], delta_t: float,
                  matrix_solver: str = "Reference", preconditioner: str = "Jacobi",
                  **options) -> dict:
    """createSolver (src/Solver.hpp:283-350).  device must be "hip" (the only
    backend).  Solver choice as createVelocityCorrector
    (src/VelocityCorrector.hpp:297-336): matrix_solver "Reference" is the
    Jacobi-PCG (preconditioner ignored, as there); "PCG" takes the
    preconditioner "PFMG" (the reference example's default,
    examples/advection.cpp:186-189) or "SMG", both served by the library's
    geometric multigrid V-cycle, "Jacobi" or "Diagonal", or "None" / "none"
    (plain CG, VelocityCorrector.hpp:327).  Other HYPRE solvers (GMRES,
    BiCGSTAB, PFMG or SMG as the solver) are not provided.  options: any
    cf_config field (field_order, cg_tol, precision, quirks, periodic,
    mg_nu, ...)."""
    if device not in ("hip", "gfx950"):
        raise RuntimeError("invalid backend: only the HIP (gfx950) backend exists")
    if matrix_solver == "Reference":
        precond = PRECOND_JACOBI
    elif matrix_solver == "PCG" and preconditioner in PCG_PRECONDITIONERS:
        precond = PCG_PRECONDITIONERS[preconditioner]
    else:
        raise RuntimeError(f"matrix solver {matrix_solver!r} with preconditioner "
                           f"{preconditioner!r} is not provided (Reference, or PCG with "
                           "PFMG, SMG, Jacobi, Diagonal or None)")
    dim = len(global_num_cell)
    cfg = base_config(dim, global_num_cell[0])
    cfg["num_cell"] = list(global_num_cell) + [1] * (3 - dim)
    lo = list(global_bounding_box[:dim]) + [0.0] * (3 - dim)
    hi = list(global_bounding_box[dim:2 * dim]) + [1.0] * (3 - dim)
    cfg["lo"], cfg["hi"] = lo, hi
    bt = bc.boundary_type
    # reference order: lo faces of each axis, then hi faces
    cfg["wall_type"] = [[bt[d], bt[d + dim]] if d < dim else [SOLID, SOLID] for d in range(3)]
    cfg["density"] = density
    cfg["dt"] = delta_t
    cfg["preconditioner"] = precond
    # the pressure solver's settings (VelocityCorrector.hpp:103-104)
    cfg["cg_tol"] = 1.0e-6
    cfg["cg_max_iter"] = 2000
    if source is not None:
        cfg["inflow_shape"] = INFLOW_BOX
        cfg["inflow_location"] = list(source.location) + [0.0] * (3 - dim)
        cfg["inflow_size"] = list(source.size) + [0.0] * (3 - dim)
        cfg["inflow_velocity"] = list(source.velocity) + [0.0] * (3 - dim)
        cfg["inflow_quantity"] = source.quantity
    else:
        cfg["inflow_shape"] = INFLOW_NONE
    if body is not None:
        cfg["body_force"] = list(body.force) + [0.0] * (3 - len(body.force))
    for k, v in options.items():
        if k not in cfg:
            raise KeyError(f"unknown option {k!r}")
        cfg[k] = v
    return cfg
