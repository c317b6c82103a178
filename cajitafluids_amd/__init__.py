"""MI355X-native CajitaFluids hot path: HIP kernels for gfx950 behind a C ABI
(include/cf_api.h), with a thin Python mirror of the reference's solver
interface."""
from . import configs  # noqa: F401
from .configs import (CELL, FACE_X, FACE_Y, FACE_Z, FP32_REFINED, FP64,  # noqa: F401
                      QUIRK_PROJ_BC_U, QUIRK_RK3_V0)


def __getattr__(name):
    # Loading the solver module is deferred so that `import cajitafluids_amd`
    # works for configs alone; the library itself is loaded on first use.
    if name in ("Solver", "create_solver", "solver_config", "BoundaryCondition", "InflowSource",
                "BodyForce", "layout", "geometry", "CURRENT", "NEXT"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
