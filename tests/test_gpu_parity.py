"""Parity of the HIP path against the CPU oracle, through the C ABI.

The oracle and the kernels evaluate every expression in the same order with
the same explicit fma() calls, and their dot products are double-double, so
the expected difference is zero; the asserted bar is the hot path's 1e-12
relative L2 (FP64) with identical PCG iteration counts.
"""
import numpy as np
import pytest

from cajitafluids_amd import configs as K
from helpers import compare_state, gpu_to_oracle, rel_l2, small_2d, small_3d

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def cf():
    from cajitafluids_amd import solver
    return solver


def _pair(cf, oracle_mod, cfg):
    sol = cf.Solver(cfg)
    orc = oracle_mod.Oracle(cfg)
    return sol, orc


def _assert_state(sol, orc, tol=TOL, what=""):
    res = compare_state(sol, orc)
    for e, (rl2, mx) in res.items():
        assert rl2 <= tol, f"{what} entity {e}: rel L2 {rl2:.3e} (max abs {mx:.3e})"


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 12)])
def test_pcg_fp64_random_rhs(cf, oracle_mod, dim, n):
    cfg = K.base_config(dim, n)
    cfg.update(clamp_dt=0, dt=0.002)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_random_faces(3)
    gpu_to_oracle(sol, orc)
    sol.build_rhs()
    sol.rhs_remove_mean()
    orc.build_rhs()
    orc.rhs_remove_mean()
    b_g = sol.owned_view(sol.get_pressure(1), 0)
    b_o = orc.owned_view(orc.rhs(), 0)
    assert rel_l2(b_g, b_o) <= 1e-15
    st_g = sol.pressure_solve(K.FP64)
    rc, st_o = orc.pressure_solve(K.FP64)
    assert rc == 0 and st_g["converged"]
    assert st_g["iterations"] == st_o["iterations"]
    x_g = sol.owned_view(sol.get_pressure(0), 0)
    x_o = orc.owned_view(orc.lhs(), 0)
    assert rel_l2(x_g, x_o) <= TOL
    assert st_g["final_residual"] < cfg["cg_tol"]


@pytest.mark.parametrize("mode,itol", [(K.REFINE_CARRY, 1e-2), (K.REFINE_RESTART, 1e-4),
                                       (K.REFINE_CARRY, 1e-4)])
@pytest.mark.parametrize("dim,n", [(2, 32), (3, 12), (3, 40)])
def test_pcg_fp32_refined_random_rhs(cf, oracle_mod, dim, n, mode, itol):
    cfg = K.base_config(dim, n)
    cfg.update(clamp_dt=0, dt=0.002, precision=K.FP32_REFINED, refine_mode=mode,
               inner_tol=itol)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_random_faces(3)
    gpu_to_oracle(sol, orc)
    sol.build_rhs()
    sol.rhs_remove_mean()
    orc.build_rhs()
    orc.rhs_remove_mean()
    st_g = sol.pressure_solve(K.FP32_REFINED)
    rc, st_o = orc.pressure_solve(K.FP32_REFINED)
    assert rc == 0 and st_g["converged"]
    assert st_g["iterations"] == st_o["iterations"]
    assert st_g["outer_iterations"] == st_o["outer_iterations"]
    assert st_g["final_residual"] < 1e-8
    x_g = sol.owned_view(sol.get_pressure(0), 0)
    x_o = orc.owned_view(orc.lhs(), 0)
    assert rel_l2(x_g, x_o) <= TOL


@pytest.mark.parametrize("periodic", [[0, 0, 0], [1, 1, 0], [0, 1, 1]])
@pytest.mark.parametrize("precision", [K.FP64, K.FP32_REFINED])
@pytest.mark.parametrize("nz", [40, 33])
def test_pcg_3d_ragged_tiles(cf, oracle_mod, periodic, precision, nz):
    """Non-cubic 70 x 36 x nz grid: partial 64 x 16 tiles in x and y and
    several z chunks, with walls or periodic axes.  nz = 33 ends in a
    one-plane chunk (16 + 16 + 1 planes for FP64, 4 x 8 + 1 for FP32), the
    odd tail of both the upward and the downward march."""
    n = [70, 36, nz]
    cfg = K.base_config(3, 8)
    dx = 1.0 / 70
    cfg.update(num_cell=n, hi=[1.0, 36 * dx, nz * dx], clamp_dt=0, dt=0.002,
               periodic=periodic, halo=2, precision=precision)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_random_faces(7)
    gpu_to_oracle(sol, orc)
    sol.build_rhs()
    sol.rhs_remove_mean()
    orc.build_rhs()
    orc.rhs_remove_mean()
    st_g = sol.pressure_solve(precision)
    rc, st_o = orc.pressure_solve(precision)
    assert rc == 0 and st_g["converged"]
    assert st_g["iterations"] == st_o["iterations"]
    x_g = sol.owned_view(sol.get_pressure(0), 0)
    x_o = orc.owned_view(orc.lhs(), 0)
    assert rel_l2(x_g, x_o) <= TOL


def _cfg2(n0, n1, periodic=(0, 0), walls=None, **kw):
    c = K.base_config(2, n0)
    c["num_cell"] = [n0, n1, 1]
    c["hi"] = [1.0, n1 / n0, 1.0]
    c.update(clamp_dt=0, dt=0.002, periodic=[periodic[0], periodic[1], 0])
    if walls is not None:
        c["wall_type"] = walls
    c.update(kw)
    return c


# 2-D solves run as one cooperative launch of one workgroup per strip of
# rows (2048 cells a strip): strip counts 1 to 34, a last strip shorter than
# the others, periodic axes wrapping inside a strip and across strips, FREE
# walls, a 1000-cell row, and no preconditioner
COOP_CASES = {
    "solid_128": _cfg2(128, 128),
    "periodic_x_300x200": _cfg2(300, 200, periodic=(1, 0)),
    "periodic_y_100x90": _cfg2(100, 90, periodic=(0, 1)),
    "periodic_xy_64x48": _cfg2(64, 48, periodic=(1, 1)),
    "periodic_y_one_strip_40x20": _cfg2(40, 20, periodic=(0, 1)),
    "free_walls_77x61": _cfg2(77, 61, walls=[[K.FREE, K.SOLID], [K.SOLID, K.FREE],
                                             [K.SOLID, K.SOLID]]),
    "wide_rows_1000x20": _cfg2(1000, 20),
    "unpreconditioned_96x80": _cfg2(96, 80, periodic=(1, 0), preconditioner=K.PRECOND_NONE),
    # workgroup partials (75 and 138 strips)
    "periodic_y_512x300": _cfg2(512, 300, periodic=(0, 1)),
    "tall_256x1100": _cfg2(256, 1100),
}


def _coop_solve(cf, oracle_mod, cfg, seed):
    from cajitafluids_amd import _lib
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.set_kernel_timing(True)
    sol.init_random_faces(seed)
    gpu_to_oracle(sol, orc)
    sol.build_rhs()
    sol.rhs_remove_mean()
    orc.build_rhs()
    orc.rhs_remove_mean()
    try:
        st_g = sol.pressure_solve(K.FP64)
        code = 0
    except _lib.CfError as e:
        st_g, code = sol.last_stats, e.code
    rc, st_o = orc.pressure_solve(K.FP64)
    assert code == rc
    assert st_g["iterations"] == st_o["iterations"]
    # the cooperative kernel ran (and nothing else of the PCG)
    assert sol.kernel_time(_lib.TK_PCG_COOP)[1] >= 1
    assert sol.kernel_time(_lib.TK_DIR_SPMV)[1] == 0
    x_g = sol.owned_view(sol.get_pressure(0), 0)
    x_o = orc.owned_view(orc.lhs(), 0)
    assert np.any(x_o != 0.0)
    assert np.array_equal(x_g, x_o)
    return st_g


@pytest.mark.parametrize("name", list(COOP_CASES))
def test_pcg_2d_cooperative_strips(cf, oracle_mod, name):
    """Converged 2-D solves: same iteration count and x bit for bit."""
    print(f"cooperative pcg: {name}", flush=True)
    st = _coop_solve(cf, oracle_mod, COOP_CASES[name], 7)
    assert st["converged"] and st["final_residual"] < 1e-8


@pytest.mark.parametrize("name,maxit", [("periodic_x_300x200", 1), ("periodic_x_300x200", 2),
                                        ("periodic_x_300x200", 5), ("periodic_x_300x200", 300),
                                        ("periodic_xy_64x48", 1), ("periodic_xy_64x48", 6),
                                        ("solid_128", 300)])
def test_pcg_2d_cooperative_truncated(cf, oracle_mod, name, maxit):
    """Solves stopped by the iteration cap, inside the first launch and (300)
    in a second one that resumes from the stored state (the first launch of
    a handle's first solve runs 256 iterations), with strips of 8 cells per
    lane (300x200) and of 4 (64x48, 128^2)."""
    print(f"cooperative pcg truncated: {name} maxit {maxit}", flush=True)
    cfg = dict(COOP_CASES[name], cg_max_iter=maxit, cg_tol=1e-30)
    st = _coop_solve(cf, oracle_mod, cfg, 9)
    assert st["iterations"] == maxit and not st["converged"]


@pytest.mark.parametrize("precision", [K.FP64, K.FP32_REFINED])
@pytest.mark.parametrize("maxit", [1, 2, 5, 6])
def test_pcg_3d_truncated_runs(cf, oracle_mod, precision, maxit):
    """Solves stopped by the iteration cap after odd and even counts.  The
    3-D update leaves x alone on every second iteration and applies both
    steps on the next, so a run that stops after a skipped x step has to
    complete it; x must come out with the same bits as the oracle's, which
    takes the x step every iteration.  FP32: one restart pass of `maxit`
    inner iterations, added to the FP64 x."""
    from cajitafluids_amd import _lib
    cfg = K.base_config(3, 12)
    cfg.update(clamp_dt=0, dt=0.002, cg_max_iter=maxit, precision=precision,
               refine_mode=K.REFINE_RESTART, inner_max_iter=0)
    print(f"truncated pcg: precision {precision} maxit {maxit}", flush=True)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_random_faces(5)
    gpu_to_oracle(sol, orc)
    sol.build_rhs()
    sol.rhs_remove_mean()
    orc.build_rhs()
    orc.rhs_remove_mean()
    with pytest.raises(_lib.CfError) as ei:
        sol.pressure_solve(precision)
    assert ei.value.code == _lib.CF_ENOCONV
    rc, st_o = orc.pressure_solve(precision)
    assert rc == _lib.CF_ENOCONV
    assert sol.last_stats["iterations"] == st_o["iterations"] == maxit
    x_g = sol.owned_view(sol.get_pressure(0), 0)
    x_o = orc.owned_view(orc.lhs(), 0)
    assert np.any(x_o != 0.0)
    assert np.array_equal(x_g, x_o)


def test_dye_tank_small_steps(cf, oracle_mod):
    cfg = small_2d(32)
    sol, orc = _pair(cf, oracle_mod, cfg)
    assert sol.dt == orc.dt
    st = sol.setup()
    rc, so = orc.setup()
    assert rc == 0 and st["iterations"] == so["iterations"]
    _assert_state(sol, orc, what="setup")
    for n in range(5):
        st = sol.step()
        rc, so = orc.step()
        assert rc == 0 and st["iterations"] == so["iterations"], n
        _assert_state(sol, orc, what=f"step {n}")


def test_dye_tank_128_full(cf, oracle_mod):
    """The 2-D dye tank exactly: 128^2, linear fields, 50 RK3 steps."""
    cfg = K.dye_tank_2d(128)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.setup()
    orc.setup()
    its_g, its_o = [], []
    for _ in range(K.DYE_TANK_STEPS):
        its_g.append(sol.step()["iterations"])
        rc, so = orc.step()
        assert rc == 0
        its_o.append(so["iterations"])
    assert its_g == its_o
    _assert_state(sol, orc, what="dye tank 128^2 final")


def test_dye_tank_cubic_as_written(cf, oracle_mod):
    """The as-written field order (cubic) on the dye tank's setup at 64^2."""
    cfg = small_2d(64, field_order=3)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.setup()
    orc.setup()
    for _ in range(5):
        sol.step()
        orc.step()
    _assert_state(sol, orc, what="cubic")


@pytest.mark.parametrize("quirks", [K.QUIRK_RK3_V0, 0, K.QUIRK_RK3_V0 | K.QUIRK_PROJ_BC_U])
def test_3d_vortex_steps(cf, oracle_mod, quirks):
    cfg = small_3d(16, quirks=quirks)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_vortex_tubes(K.VORTEX_TANK_SEED, 1.0)
    gpu_to_oracle(sol, orc)
    sol.setup()
    orc.setup()
    _assert_state(sol, orc, what="3d setup")
    for n in range(3):
        st = sol.step()
        rc, so = orc.step()
        assert rc == 0 and st["iterations"] == so["iterations"]
        _assert_state(sol, orc, what=f"3d step {n}")


@pytest.mark.parametrize("precision", [K.FP64, K.FP32_REFINED])
@pytest.mark.parametrize("dim", [2, 3])
def test_unpreconditioned_steps(cf, oracle_mod, dim, precision):
    """Preconditioner "None" (plain CG; the zero-mean pressure on a singular
    operator): setup + 3 steps against the oracle, with the FaceJ quirk on,
    so that the pressure's level reaches the velocities."""
    quirks = K.QUIRK_RK3_V0 | K.QUIRK_PROJ_BC_U
    cfg = small_2d(32, quirks=quirks) if dim == 2 else small_3d(16, quirks=quirks)
    cfg.update(preconditioner=K.PRECOND_NONE, precision=precision)
    sol, orc = _pair(cf, oracle_mod, cfg)
    if dim == 3:
        sol.init_vortex_tubes(K.VORTEX_TANK_SEED, 1.0)
        gpu_to_oracle(sol, orc)
    sol.setup()
    orc.setup()
    for n in range(3):
        st = sol.step()
        rc, so = orc.step()
        assert rc == 0 and st["iterations"] == so["iterations"]
        _assert_state(sol, orc, what=f"none {dim}d step {n}")
    p_g = sol.owned_view(sol.get_pressure(0), 0)
    assert abs(float(p_g.mean())) <= 1e-14 * float(np.abs(p_g).max())


def test_3d_periodic_fp32_refined(cf, oracle_mod):
    """The periodic box's shape at 16^3: h = 2, periodic x/y, walls in z, cubic,
    FP32-refined pressure, vortex tubes + blob dye."""
    cfg = K.periodic_box_3d(16)
    sol, orc = _pair(cf, oracle_mod, cfg)
    sol.init_vortex_tubes(K.PERIODIC_BOX_VEL_SEED, 1.0)
    sol.init_blobs(K.PERIODIC_BOX_DYE_SEED, 8)
    gpu_to_oracle(sol, orc)
    sol.setup()
    orc.setup()
    for n in range(3):
        st = sol.step()
        rc, so = orc.step()
        assert rc == 0
        assert st["iterations"] == so["iterations"]
        assert st["outer_iterations"] == so["outer_iterations"]
        _assert_state(sol, orc, what=f"periodic step {n}")


def test_init_generators_match_oracle(cf, oracle_mod):
    for cfg in (small_3d(16), K.periodic_box_3d(16), small_2d(32)):
        sol, orc = _pair(cf, oracle_mod, cfg)
        sol.init_vortex_tubes(5, 1.0)
        orc.init_vortex_tubes(5, 1.0)
        for e in range(1, cfg["dim"] + 1):
            g = sol.owned_view(sol.get(e), e)
            o = orc.owned_view(orc.field(e), e)
            assert rel_l2(g, o) <= 1e-13
        sol.init_blobs(55, 8)
        qo = orc.init_blobs(55, 8)
        assert rel_l2(sol.owned_view(sol.get(0), 0), orc.owned_view(qo, 0)) <= 1e-13


def test_ghost_fill_periodic(cf, oracle_mod):
    cfg = K.periodic_box_3d(12)
    sol, orc = _pair(cf, oracle_mod, cfg)
    rng = np.random.default_rng(0)
    for e in range(4):
        a = np.zeros(sol.extents(e)[::-1])
        lo, hi = sol.owned(e)
        a[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = rng.standard_normal(
            (hi[2] - lo[2], hi[1] - lo[1], hi[0] - lo[0]))
        sol.set(e, a)
        orc.field(e)[...] = a
    sol.gather()
    orc.fill_ghosts()
    for e in range(4):
        assert np.array_equal(sol.get(e), np.array(orc.field(e)))


def test_reference_halo_semantics_single_rank(cf):
    """tests/tstProblemManager.cpp:61-97 on one rank with walls: no neighbour,
    so gather leaves every ghost untouched (zero)."""
    cfg = K.base_config(2, 16)
    sol = cf.Solver(cfg)
    ex = sol.extents(K.FACE_X)
    lo, hi = sol.owned(K.FACE_X)
    a = np.zeros(ex[::-1])
    for j in range(lo[1], hi[1]):
        for i in range(lo[0], hi[0]):
            a[0, j, i] = i * 100 + j * 10 + 1
    sol.set(K.FACE_X, a)
    sol.gather()
    assert np.array_equal(sol.get(K.FACE_X), a)


def test_advance_swaps(cf):
    """tests/tstProblemManager.cpp:23-59: advance makes Next current."""
    cfg = K.base_config(2, 16)
    sol = cf.Solver(cfg)
    ex = sol.extents(K.CELL)
    lo, hi = sol.owned(K.CELL)
    cur = np.zeros(ex[::-1])
    nxt = np.zeros(ex[::-1])
    for j in range(lo[1], hi[1]):
        for i in range(lo[0], hi[0]):
            cur[0, j, i] = i * 100 + j * 10
            nxt[0, j, i] = i * 100 + j * 10 + 5
    sol.set(K.CELL, cur, 0)
    sol.set(K.CELL, nxt, 1)
    sol.advance(K.CELL)
    assert np.array_equal(sol.get(K.CELL, 0), nxt)


def test_nonconvergence_reports_enoconv(cf):
    from cajitafluids_amd import _lib
    cfg = small_2d(32, cg_max_iter=3)
    sol = cf.Solver(cfg)
    with pytest.raises(_lib.CfError) as ei:
        sol.setup()
    assert ei.value.code == _lib.CF_ENOCONV
    assert sol.last_stats["iterations"] == 3
