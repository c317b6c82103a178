"""The no-allocation pressure solve (cf_pressure_solve_noalloc) through the
C ABI against the handle's cf_pressure_solve_ext on the same device
arrays: the same kernels, so the same iteration count and the same bits, on
two calls in a row (the second on a smaller grid, over the scratch the
first left behind), plus the rejections (a grid past the scratch limit,
FP32-refined, multigrid)."""
import ctypes as C

import numpy as np
import pytest

from cajitafluids_amd import _lib
from cajitafluids_amd import configs as K

pytestmark = pytest.mark.gpu


def _padded(cf, cfg, dense):
    import torch
    off, px, py, elems = cf.padded_layout(cfg)
    p = np.zeros(elems)
    ext = dense.shape[::-1]
    for k in range(ext[2]):
        for j in range(ext[1]):
            s = off + px * (j + py * k)
            p[s:s + ext[0]] = dense[k, j]
    return torch.tensor(p, dtype=torch.float64, device="cuda")


def _rhs(cf, cfg, seed):
    with cf.Solver(cfg) as sol:
        sol.init_random_faces(seed)
        sol.build_rhs()
        sol.rhs_remove_mean()
        return sol.get_pressure(1)


def _noalloc(cfg, rhs_ptr, lhs_ptr):
    L = _lib.load()
    st = _lib.CfSolveStats()
    cs = _lib.config_struct(cfg)
    rc = L.cf_pressure_solve_noalloc(C.byref(cs), C.c_void_p(rhs_ptr), C.c_void_p(lhs_ptr),
                                     None, C.byref(st))
    return rc, st.as_dict()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_pressure_solve_ext(order):
    import torch
    from cajitafluids_amd import solver as cf

    big = K.vortex_tank_3d(64)
    big.update(clamp_dt=0, dt=0.002)
    small = K.dye_tank_2d(96)
    small.update(preconditioner=K.PRECOND_NONE)
    for n, cfg in enumerate((big, small)):
        print(f"no-alloc call {n}: dim {cfg['dim']} n {cfg['num_cell'][0]}", flush=True)
        rhs = _padded(cf, cfg, _rhs(cf, cfg, 3 + n))
        elems = cf.padded_layout(cfg)[3]
        ref = torch.full((elems,), 5.0, dtype=torch.float64, device="cuda")
        got = torch.full((elems,), -5.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with cf.Solver(cfg) as sol:
            st_ref = sol.pressure_solve_ext(rhs.data_ptr(), ref.data_ptr(), K.FP64)
            sol.sync()
        rc, st = _noalloc(cfg, rhs.data_ptr(), got.data_ptr())
        assert rc == _lib.CF_OK, _lib.load().cf_last_error(None)
        assert st["converged"] and st["iterations"] == st_ref["iterations"], (st, st_ref)
        assert st["final_residual"] == st_ref["final_residual"]
        # the owned cells carry the solution; both leave the rest as the
        # solve wrote it (x = 0 on every cell of the padded layout first)
        _, lo, hi = cf.layout(cfg, 0)
        off, px, py, _ = cf.padded_layout(cfg)
        a, b = ref.cpu().numpy(), got.cpu().numpy()
        for k in range(lo[2], hi[2]):
            for j in range(lo[1], hi[1]):
                s = off + lo[0] + px * (j + py * k)
                assert np.array_equal(a[s:s + hi[0] - lo[0]], b[s:s + hi[0] - lo[0]]), (n, j, k)


def test_noalloc_rejections():
    L = _lib.load()
    size = C.c_size_t()
    lim = C.c_int64()
    assert L.cf_noalloc_limits(C.byref(size), C.byref(lim)) == _lib.CF_OK
    assert size.value == 96 << 20 and lim.value == 4183360
    for cfg, code in ((K.vortex_tank_3d(148), _lib.CF_EINVAL),
                      (K.vortex_tank_3d(32, precision=K.FP32_REFINED), _lib.CF_ENOTIMPL),
                      (dict(K.vortex_tank_3d(32), preconditioner=K.PRECOND_MG),
                       _lib.CF_ENOTIMPL)):
        rc, _ = _noalloc(cfg, 16, 32)  # rejected before any pointer is used
        assert rc == code, (rc, L.cf_last_error(None))
