"""ctypes binding of the CPU oracle (oracle/libcf_oracle.so).

Test infrastructure: only tests, the smoke check and the CPU-baseline leg of
bench.py use it, always as the checker, never as the thing measured for the
product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
# CF_ORACLE_LIBRARY names another build of the same oracle (the sanitizer
# build of oracle/Makefile's `asan` target)
ORACLE_SO = os.environ.get("CF_ORACLE_LIBRARY") or os.path.join(ORACLE_DIR, "libcf_oracle.so")

CELL, FACE_X, FACE_Y, FACE_Z = 0, 1, 2, 3
CURRENT, NEXT = 0, 1
SOLID, FREE = 0, 1
INFLOW_NONE, INFLOW_BOX, INFLOW_SPHERE = 0, 1, 2
FP64, FP32_REFINED = 0, 1
REFINE_RESTART, REFINE_CARRY = 0, 1
QUIRK_RK3_V0, QUIRK_PROJ_BC_U = 1, 2


class CfoConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("num_cell", C.c_int * 3),
        ("halo", C.c_int),
        ("periodic", C.c_int * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("wall_type", (C.c_int * 2) * 3),
        ("density", C.c_double),
        ("dt", C.c_double),
        ("clamp_dt", C.c_int),
        ("body_force", C.c_double * 3),
        ("inflow_shape", C.c_int),
        ("inflow_location", C.c_double * 3),
        ("inflow_size", C.c_double * 3),
        ("inflow_center", C.c_double * 3),
        ("inflow_radius", C.c_double),
        ("inflow_velocity", C.c_double * 3),
        ("inflow_quantity", C.c_double),
        ("field_order", C.c_int),
        ("velocity_order", C.c_int),
        ("cg_tol", C.c_double),
        ("cg_max_iter", C.c_int),
        ("precision", C.c_int),
        ("inner_tol", C.c_double),
        ("inner_max_iter", C.c_int),
        ("max_outer", C.c_int),
        ("quirks", C.c_int),
        ("refine_mode", C.c_int),
        ("preconditioner", C.c_int),
        ("mg_nu", C.c_int),
        ("mg_omega", C.c_double),
        ("mg_coarse_sweeps", C.c_int),
        ("mg_alpha", C.c_double),
    ]


class CfoStats(C.Structure):
    _fields_ = [
        ("converged", C.c_int),
        ("iterations", C.c_int),
        ("outer_iterations", C.c_int),
        ("b_norm", C.c_double),
        ("final_residual", C.c_double),
    ]


def fill_struct(st, cfg: dict):
    """Copy a plain config dict (the product's field names) into a ctypes
    structure with the same layout."""
    for name, _ in st._fields_:
        if name not in cfg:
            continue
        val = cfg[name]
        field = getattr(st, name)
        if isinstance(field, C.Array):
            if name == "wall_type":
                for d in range(3):
                    for s in range(2):
                        field[d][s] = int(val[d][s])
            else:
                for i, x in enumerate(val):
                    field[i] = x
        else:
            setattr(st, name, val)
    return st


_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    L = C.CDLL(ORACLE_SO)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    sig = {
        "cfo_create": (C.c_int, [C.POINTER(CfoConfig), C.POINTER(P)]),
        "cfo_destroy": (None, [P]),
        "cfo_dt": (C.c_double, [P]),
        "cfo_dx": (C.c_double, [P]),
        "cfo_time": (C.c_double, [P]),
        "cfo_extents": (None, [P, C.c_int, ip]),
        "cfo_owned": (None, [P, C.c_int, ip, ip]),
        "cfo_field": (dp, [P, C.c_int, C.c_int]),
        "cfo_lhs": (dp, [P]),
        "cfo_rhs": (dp, [P]),
        "cfo_setup": (C.c_int, [P, C.POINTER(CfoStats)]),
        "cfo_step": (C.c_int, [P, C.POINTER(CfoStats)]),
        "cfo_run": (C.c_int, [P, C.c_int, C.POINTER(CfoStats)]),
        "cfo_fill_ghosts": (None, [P, C.c_int]),
        "cfo_time_integrate": (None, [P]),
        "cfo_add_inputs": (None, [P]),
        "cfo_build_rhs": (None, [P]),
        "cfo_rhs_remove_mean": (None, [P]),
        "cfo_pressure_solve": (C.c_int, [P, C.c_int, C.POINTER(CfoStats)]),
        "cfo_apply_pressure": (None, [P]),
        "cfo_correct_velocity": (C.c_int, [P, C.POINTER(CfoStats)]),
        "cfo_advance": (None, [P, C.c_int]),
        "cfo_spline_weights": (None, [C.c_int, C.c_double, ip, dp]),
        "cfo_interpolate": (C.c_double, [P, C.c_int, C.c_int, dp, dp]),
        "cfo_cell_output": (None, [P, dp, dp]),
        "cfo_rk3": (None, [P, dp, dp, dp, dp, C.c_double, dp]),
        "cfo_apply_A": (None, [P, dp, dp]),
        "cfo_dot": (C.c_double, [P, dp, dp]),
        "cfo_rand_u64": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "cfo_rand_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "cfo_rand_normal": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "cfo_init_vortex_tubes": (None, [P, C.c_uint64, C.c_double]),
        "cfo_init_blobs": (None, [P, dp, C.c_uint64, C.c_int]),
        "cfo_init_random_faces": (None, [P, C.c_uint64]),
        "cfo_adjoint_run": (C.c_int, [P, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp]),
        "cfo_mg_nlevels": (C.c_int, [P]),
        "cfo_mg_level_dims": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
        "cfo_mg_apply_level": (C.c_int, [P, C.c_int, dp, dp]),
        "cfo_mg_vcycle": (C.c_int, [P, dp, dp]),
        "cfo_mg_vcycle_prec": (C.c_int, [P, C.c_int, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Owns one oracle state; arrays are numpy views (x fastest) shaped
    (ext_z, ext_y, ext_x) over the oracle's ghosted storage."""

    def __init__(self, cfg: dict):
        self.L = lib()
        self.cfg = dict(cfg)
        st = fill_struct(CfoConfig(), cfg)
        h = C.c_void_p()
        rc = self.L.cfo_create(C.byref(st), C.byref(h))
        if rc != 0:
            raise ValueError(f"cfo_create failed rc={rc}")
        self.h = h
        self.dim = cfg["dim"]

    def close(self):
        if self.h:
            self.L.cfo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dt(self):
        return self.L.cfo_dt(self.h)

    @property
    def dx(self):
        return self.L.cfo_dx(self.h)

    def extents(self, entity):
        e = (C.c_int * 3)()
        self.L.cfo_extents(self.h, entity, e)
        return tuple(e)

    def owned(self, entity):
        lo = (C.c_int * 3)()
        hi = (C.c_int * 3)()
        self.L.cfo_owned(self.h, entity, lo, hi)
        return tuple(lo), tuple(hi)

    def _view(self, ptr, entity):
        ex = self.extents(entity)
        n = ex[0] * ex[1] * ex[2]
        arr = np.ctypeslib.as_array(ptr, shape=(n,))
        return arr.reshape(ex[2], ex[1], ex[0])

    def field(self, entity, version=CURRENT):
        return self._view(self.L.cfo_field(self.h, entity, version), entity)

    def lhs(self):
        return self._view(self.L.cfo_lhs(self.h), CELL)

    def rhs(self):
        return self._view(self.L.cfo_rhs(self.h), CELL)

    def owned_view(self, arr, entity):
        lo, hi = self.owned(entity)
        return arr[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]

    def _stats(self, st):
        return dict(converged=st.converged, iterations=st.iterations,
                    outer_iterations=st.outer_iterations, b_norm=st.b_norm,
                    final_residual=st.final_residual)

    def setup(self):
        st = CfoStats()
        rc = self.L.cfo_setup(self.h, C.byref(st))
        return rc, self._stats(st)

    def step(self):
        st = CfoStats()
        rc = self.L.cfo_step(self.h, C.byref(st))
        return rc, self._stats(st)

    def run(self, n):
        st = CfoStats()
        rc = self.L.cfo_run(self.h, n, C.byref(st))
        return rc, self._stats(st)

    def time_integrate(self):
        self.L.cfo_time_integrate(self.h)

    def add_inputs(self):
        self.L.cfo_add_inputs(self.h)

    def build_rhs(self):
        self.L.cfo_build_rhs(self.h)

    def rhs_remove_mean(self):
        self.L.cfo_rhs_remove_mean(self.h)

    def pressure_solve(self, precision=FP64):
        st = CfoStats()
        rc = self.L.cfo_pressure_solve(self.h, precision, C.byref(st))
        return rc, self._stats(st)

    def apply_pressure(self):
        self.L.cfo_apply_pressure(self.h)

    def correct_velocity(self):
        st = CfoStats()
        rc = self.L.cfo_correct_velocity(self.h, C.byref(st))
        return rc, self._stats(st)

    def fill_ghosts(self, version=CURRENT):
        self.L.cfo_fill_ghosts(self.h, version)

    def apply_A(self, x):
        y = np.zeros_like(x)
        self.L.cfo_apply_A(self.h, _dp(x), _dp(y))
        return y

    def dot(self, a, b):
        return self.L.cfo_dot(self.h, _dp(a), _dp(b))

    def cell_output(self):
        """(q, vel): owned q and the cell-centre velocity, each as
        (n2, n1, n0) arrays (vel: dim of them)."""
        lo, hi = self.owned(CELL)
        n = [hi[d] - lo[d] for d in range(3)]
        N = n[0] * n[1] * n[2]
        q = np.zeros(N)
        vel = np.zeros(self.dim * N)
        self.L.cfo_cell_output(self.h, _dp(q), _dp(vel))
        shape = (n[2], n[1], n[0])
        return q.reshape(shape), [vel[d * N:(d + 1) * N].reshape(shape) for d in range(self.dim)]

    def interpolate(self, entity, order, field, p):
        p = np.ascontiguousarray(np.asarray(list(p) + [0.0] * (3 - len(p)), dtype=np.float64))
        return self.L.cfo_interpolate(self.h, entity, order, _dp(field), _dp(p))

    def rk3(self, x0, dt):
        x0 = np.ascontiguousarray(np.asarray(list(x0) + [0.0] * (3 - len(x0)), dtype=np.float64))
        tr = np.zeros(3)
        u = self.field(FACE_X)
        v = self.field(FACE_Y)
        w = self.field(FACE_Z) if self.dim == 3 else None
        self.L.cfo_rk3(self.h, _dp(x0), _dp(u), _dp(v), _dp(w), dt, _dp(tr))
        return tr[:self.dim]

    def init_vortex_tubes(self, seed, peak=1.0):
        self.L.cfo_init_vortex_tubes(self.h, seed, peak)

    def init_blobs(self, seed, nblobs):
        out = np.zeros(self.extents(CELL)[::-1])
        self.L.cfo_init_blobs(self.h, _dp(out), seed, nblobs)
        return out

    def init_random_faces(self, seed):
        self.L.cfo_init_random_faces(self.h, seed)

    # ---- multigrid (compact vectors: owned cells, x fastest) -----------------
    def mg_nlevels(self):
        return self.L.cfo_mg_nlevels(self.h)

    def mg_level_dims(self, level):
        n = (C.c_int * 3)()
        rc = self.L.cfo_mg_level_dims(self.h, level, n)
        if rc != 0:
            raise ValueError(f"cfo_mg_level_dims rc={rc}")
        return tuple(n)

    def mg_apply(self, level, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros_like(x)
        rc = self.L.cfo_mg_apply_level(self.h, level, _dp(x), _dp(y))
        if rc != 0:
            raise ValueError(f"cfo_mg_apply_level rc={rc}")
        return y

    def mg_vcycle(self, r, f32=False):
        """z = B r on compact level-0 vectors; f32: level 0 in FP32."""
        r = np.ascontiguousarray(r, dtype=np.float64).ravel()
        z = np.zeros_like(r)
        rc = self.L.cfo_mg_vcycle_prec(self.h, 1 if f32 else 0, _dp(r), _dp(z))
        if rc != 0:
            raise ValueError(f"cfo_mg_vcycle rc={rc}")
        return z

    def adjoint_run(self, nsteps, ckpt, dq_T):
        dq_T = np.ascontiguousarray(dq_T, dtype=np.float64)
        du = np.zeros(self.extents(FACE_X)[::-1])
        dv = np.zeros(self.extents(FACE_Y)[::-1])
        dw = np.zeros(self.extents(FACE_Z)[::-1]) if self.dim == 3 else None
        dq0 = np.zeros(self.extents(CELL)[::-1])
        dvel = np.zeros(3)
        rc = self.L.cfo_adjoint_run(self.h, nsteps, ckpt, _dp(dq_T), _dp(du), _dp(dv),
                                    _dp(dw), _dp(dvel), _dp(dq0))
        return rc, dict(du=du, dv=dv, dw=dw, dvel=dvel, dq0=dq0)


def spline_weights(order, xi):
    s0 = C.c_int()
    w = (C.c_double * 4)()
    lib().cfo_spline_weights(order, xi, C.byref(s0), w)
    return s0.value, list(w)[: order + 1]


def rand_u64(seed, stream, idx):
    return lib().cfo_rand_u64(seed, stream, idx)


def rand_uniform(seed, stream, idx):
    return lib().cfo_rand_uniform(seed, stream, idx)


def rand_normal(seed, stream, idx):
    return lib().cfo_rand_normal(seed, stream, idx)
