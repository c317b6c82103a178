"""ctypes binding of libcfmi.so (include/cf_api.h).

The library is built in-tree (``cajitafluids_amd/libcfmi.so``).  Loading
fails loudly when it is missing: there is no CPU fallback for the product
path.
"""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# CFMI_LIBRARY names another build of the same library (A/B timing of
# kernel variants); the default is the in-tree build.
LIB_PATH = os.environ.get("CFMI_LIBRARY") or os.path.join(HERE, "libcfmi.so")

CF_OK, CF_EINVAL, CF_ENOMEM, CF_ENOCONV, CF_EDEVICE, CF_ENOTIMPL, CF_ENONFINITE = range(7)
STATUS_NAMES = {0: "CF_OK", 1: "CF_EINVAL", 2: "CF_ENOMEM", 3: "CF_ENOCONV",
                4: "CF_EDEVICE", 5: "CF_ENOTIMPL", 6: "CF_ENONFINITE"}

# kernel ids of cf_kernel_time
(TK_DIR_SPMV, TK_UPDATE, TK_FINALIZE, TK_DIR_SPMV32, TK_UPDATE32, TK_FINALIZE32,
 TK_ADVECT, TK_RESIDUAL, TK_VCYCLE, TK_PCG_COOP) = range(10)


class CfConfig(C.Structure):
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


class CfSolveStats(C.Structure):
    _fields_ = [
        ("converged", C.c_int),
        ("iterations", C.c_int),
        ("outer_iterations", C.c_int),
        ("b_norm", C.c_double),
        ("final_residual", C.c_double),
    ]

    def as_dict(self):
        return dict(converged=self.converged, iterations=self.iterations,
                    outer_iterations=self.outer_iterations, b_norm=self.b_norm,
                    final_residual=self.final_residual)


class CfView(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("offset", C.c_int64),
        ("pitch_x", C.c_int64),
        ("pitch_y", C.c_int64),
        ("ghost_ext", C.c_int * 3),
        ("own_lo", C.c_int * 3),
        ("own_hi", C.c_int * 3),
    ]


def config_struct(cfg: dict) -> CfConfig:
    st = CfConfig()
    for name, _ in CfConfig._fields_:
        if name not in cfg:
            raise KeyError(f"config is missing {name!r}")
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


# name -> (restype, argtypes): every symbol include/cf_api.h declares.
_P = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_SST = C.POINTER(CfSolveStats)
_i64p = C.POINTER(C.c_int64)
SIGNATURES = {
    "cf_layout_query": (C.c_int, [C.POINTER(CfConfig), C.c_int, _ip, _ip, _ip]),
    "cf_geometry_query": (C.c_int, [C.POINTER(CfConfig), _dp, _dp]),
    "cf_padded_layout": (C.c_int, [C.POINTER(CfConfig), _i64p, _i64p, _i64p, _i64p]),
    "cf_entity_origin": (C.c_int, [C.POINTER(CfConfig), C.c_int, _dp]),
    "cf_create": (C.c_int, [C.POINTER(CfConfig), C.POINTER(_P)]),
    "cf_destroy": (None, [_P]),
    "cf_last_error": (C.c_char_p, [_P]),
    "cf_setup": (C.c_int, [_P, _SST]),
    "cf_step": (C.c_int, [_P, _SST]),
    "cf_run": (C.c_int, [_P, C.c_int, _SST]),
    "cf_time": (C.c_double, [_P]),
    "cf_dt": (C.c_double, [_P]),
    "cf_dx": (C.c_double, [_P]),
    "cf_time_integrate": (C.c_int, [_P]),
    "cf_add_inputs": (C.c_int, [_P]),
    "cf_correct_velocity": (C.c_int, [_P, _SST]),
    "cf_build_rhs": (C.c_int, [_P]),
    "cf_rhs_remove_mean": (C.c_int, [_P]),
    "cf_pressure_solve": (C.c_int, [_P, C.c_int, _SST]),
    "cf_pressure_solve_ext": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int, _SST]),
    "cf_apply_pressure": (C.c_int, [_P]),
    "cf_field_view": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(CfView)]),
    "cf_pressure_view": (C.c_int, [_P, C.c_int, C.POINTER(CfView)]),
    "cf_copy_in": (C.c_int, [_P, C.c_int, C.c_int, _dp]),
    "cf_copy_out": (C.c_int, [_P, C.c_int, C.c_int, _dp]),
    "cf_copy_pressure_in": (C.c_int, [_P, C.c_int, _dp]),
    "cf_copy_pressure_out": (C.c_int, [_P, C.c_int, _dp]),
    "cf_copy_in_device": (C.c_int, [_P, C.c_int, C.c_int, C.c_void_p]),
    "cf_copy_out_device": (C.c_int, [_P, C.c_int, C.c_int, C.c_void_p]),
    "cf_advance": (C.c_int, [_P, C.c_int]),
    "cf_fill_ghosts": (C.c_int, [_P, C.c_int]),
    "cf_sync": (C.c_int, [_P]),
    "cf_init_vortex_tubes": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "cf_init_blobs": (C.c_int, [_P, C.c_uint64, C.c_int]),
    "cf_init_random_faces": (C.c_int, [_P, C.c_uint64]),
    "cf_set_kernel_timing": (C.c_int, [_P, C.c_int]),
    "cf_kernel_time": (C.c_int, [_P, C.c_int, _dp, C.POINTER(C.c_int64)]),
    "cf_reset_kernel_time": (C.c_int, [_P]),
    "cf_adjoint_run": (C.c_int, [_P, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
    "cf_adjoint_run_device": (C.c_int, [_P, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    "cf_pressure_solve_noalloc": (C.c_int, [C.POINTER(CfConfig), C.c_void_p, C.c_void_p,
                                            C.c_void_p, _SST]),
    "cf_noalloc_limits": (C.c_int, [C.POINTER(C.c_size_t), _i64p]),
    "cf_cell_output": (C.c_int, [_P, _dp, _dp]),
    "cf_write_output": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "cf_write_vti": (C.c_int, [C.c_char_p, C.c_int, _ip, _dp, C.c_double, C.c_int,
                               C.c_double, C.c_double, _dp, _dp]),
}

_lib = None


class CfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def _torch_runtime_first():
    """The torch wheel carries its own HIP/HSA runtime next to the system one
    that libcfmi.so links.  Both can serve one process only if torch's
    initialises first: in the other order torch finds no device.  So when
    the process has imported torch, its runtime is brought up before
    libcfmi.so loads.  A process without torch keeps a single runtime (which
    is what profilers that intercept the runtime need)."""
    torch = sys.modules.get("torch")
    if torch is None:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def load():
    """Load libcfmi.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import "
            "__graft_entry__ as g; g.build()'` (make -C cajitafluids_amd/csrc)")
    _torch_runtime_first()
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
