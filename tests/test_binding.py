"""The reference-side adapter (include/cf_hip_solver.hpp): a C++ caller of
libcfmi.so shaped like the reference's examples/advection.cpp, built by
tests/binding/Makefile against tests/binding/reference_types.hpp (the
reference's value objects by their member names).

CPU: the C ABI header is valid C, the adapter compiles with -Wall -Werror,
and the demo fails cleanly without a device.  GPU: the demo's run (setup + steps through SolverBase)
is bit-identical to the same run through the Python mirror, with the init
functor's values computed from the same coordinates.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from cajitafluids_amd import configs as K

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "binding", "_build", "binding_demo")


def _gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_header_is_plain_c(tmp_path):
    src = tmp_path / "c_caller.c"
    src.write_text('#include "cf_api.h"\n'
                   "int main(void) { cf_config c = {0}; c.dim = 2; "
                   "return cf_layout_query(&c, CF_CELL, 0, 0, 0) == CF_OK; }\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(REPO, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adapter_builds_and_fails_cleanly_without_device():
    r = subprocess.run(["make", "-B", "-C", os.path.join(REPO, "tests", "binding")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    if _gpu():
        pytest.skip("a GPU is present")
    r = subprocess.run([BIN, "16", "1", os.devnull], capture_output=True, text=True)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr


def _exact_extent(lo, hi):
    s = hi - lo
    while lo + s < hi:
        s = math.nextafter(s, math.inf)
    while lo + s > hi:
        s = math.nextafter(s, -math.inf)
    return s


def demo_config(n):
    """The configuration HipSolver2 builds for binding_demo's arguments."""
    c = K.base_config(2, n)
    lo, size = [0.45, 0.20], [0.10, 0.02]
    bb = [lo[0], lo[1], lo[0] + size[0], lo[1] + size[1]]
    c.update(dt=0.005, clamp_dt=1, body_force=[0.0, -0.5, 0.0], inflow_shape=K.INFLOW_BOX,
             inflow_location=[bb[0], bb[1], 0.0],
             inflow_size=[_exact_extent(bb[0], bb[2]), _exact_extent(bb[1], bb[3]), 0.0],
             inflow_velocity=[0.0, 3.0, 0.0], inflow_quantity=1.0, field_order=3,
             velocity_order=1, cg_tol=1e-6, cg_max_iter=2000, inner_max_iter=2000,
             quirks=K.QUIRK_RK3_V0 | K.QUIRK_PROJ_BC_U, preconditioner=K.PRECOND_JACOBI)
    return c


def test_exact_extent_reproduces_the_box_corner():
    for lo, size in ((0.45, 0.10), (0.20, 0.02), (0.01, 0.49), (0.3, 1e-7)):
        hi = lo + size
        assert lo + _exact_extent(lo, hi) == hi


@pytest.mark.gpu
def test_adapter_run_matches_python_mirror(tmp_path):
    from cajitafluids_amd.solver import Solver, entity_origin, geometry, layout
    n, nsteps = 128, 3  # dx / (3 + sqrt(0.5 dx)) < 0.005: the dt clamp applies
    out = tmp_path / "demo.bin"
    print(f"binding_demo {n} {nsteps}", flush=True)
    r = subprocess.run([BIN, str(n), str(nsteps), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Reducting timestep to" in r.stderr  # the reference's clamp message
    cfg = demo_config(n)
    dx, _ = geometry(cfg)
    raw = np.fromfile(out, dtype=np.float64)
    with Solver(cfg) as sol:
        for e in range(3):
            ext, lo, hi = layout(cfg, e)
            o = entity_origin(cfg, e)
            a = np.zeros((ext[2], ext[1], ext[0]))
            i = np.arange(lo[0], hi[0], dtype=np.float64)
            j = np.arange(lo[1], hi[1], dtype=np.float64)
            x = o[0] + dx * i[None, :]
            y = o[1] + dx * j[:, None]
            if e == 0:
                v = (x * (1.0 - x)) * y
            elif e == 1:
                v = x * y - 0.25
            else:
                v = 0.5 - x * y
            a[0, lo[1]:hi[1], lo[0]:hi[0]] = v
            sol.set(e, a)
        sol.setup()
        for _ in range(nsteps):
            sol.step()
        off = 0
        for e in range(3):
            ext = layout(cfg, e)[0]
            cnt = ext[0] * ext[1] * ext[2]
            got = raw[off:off + cnt].reshape(ext[2], ext[1], ext[0])
            off += cnt
            assert np.array_equal(got, sol.get(e)), e
    assert off == raw.size
