"""FP32-inner / FP64-outer pressure solves against the FP64 PCG on the same
right-hand side, at the full 512^3 random-velocity workload: both reach the
1e-8 relative residual, and the pressure difference (means removed) is
reported and bounded.

The FP32-refined solves themselves are checked bit for bit against the
oracle's restatement at small sizes (test_gpu_parity.py); this file states
their error against the FP64 solution at full size.
"""
import numpy as np
import pytest

from cajitafluids_amd import configs as K

pytestmark = pytest.mark.gpu


def _solve(cf, cfg, precision):
    with cf.Solver(cfg) as sol:
        sol.init_random_faces(K.RANDOM_SOLVE_SEED)
        sol.build_rhs()
        sol.rhs_remove_mean()
        st = sol.pressure_solve(precision)
        p = sol.owned_view(sol.get_pressure(0), 0).copy()
    p -= p.mean()
    return st, p


@pytest.fixture(scope="module")
def cf():
    from cajitafluids_amd import solver
    return solver


@pytest.fixture(scope="module")
def fp64_512(cf):
    cfg = K.random_solve_3d(512)
    print("512^3 FP64 PCG", flush=True)
    st, p = _solve(cf, cfg, K.FP64)
    assert st["converged"] and st["final_residual"] < 1e-8
    return st, p


@pytest.mark.parametrize("mode,itol", [(K.REFINE_RESTART, 1e-4), (K.REFINE_CARRY, 1e-2)])
def test_refined_vs_fp64_512(cf, fp64_512, mode, itol):
    st64, p64 = fp64_512
    cfg = K.random_solve_3d(512)
    cfg.update(refine_mode=mode, inner_tol=itol)
    print(f"512^3 FP32-refined mode={mode} inner_tol={itol}", flush=True)
    st, p = _solve(cf, cfg, K.FP32_REFINED)
    assert st["converged"] and st["final_residual"] < 1e-8
    err = float(np.linalg.norm(p - p64) / np.linalg.norm(p64))
    print(f"  outer passes {st['outer_iterations']}, inner iterations {st['iterations']} "
          f"(FP64 {st64['iterations']}), |p - p64| / |p64| = {err:.3e}", flush=True)
    # Both solutions have residuals below tol ||b||, so their difference is
    # A^+ of a vector of norm < 2 tol ||b||: relative to ||p|| >= ||b|| /
    # lambda_max it is at most 2 tol kappa, with lambda_max < 12 s
    # (Gershgorin) and the smallest nonzero Neumann eigenvalue
    # 4 s sin^2(pi / 2n).
    n = cfg["num_cell"][0]
    kappa = 12.0 / (4.0 * np.sin(np.pi / (2 * n)) ** 2)
    assert err < 2 * cfg["cg_tol"] * kappa
    # That bound (6.4e-3 here) is loose: the solves are deterministic, and
    # the errors measured on MI355X are 1.45e-5 (restart, inner 1e-4) and
    # 9.4e-6 (carry, inner 1e-2).  The test holds them to 2x those values.
    measured = {K.REFINE_RESTART: 1.45e-5, K.REFINE_CARRY: 9.4e-6}[mode]
    assert err <= 2.0 * measured, (err, measured)
