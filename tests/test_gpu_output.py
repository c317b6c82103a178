"""The output writer on the GPU (cf_cell_output, cf_write_output) against the
oracle's cell data (oracle/cf_oracle.c cfo_cell_output, SiloWriter.hpp:
141-197), and the reference's write schedule (Solver.hpp:148-176)."""
import os

import numpy as np
import pytest

from cajitafluids_amd import configs as K
from cajitafluids_amd import vtk
from helpers import gpu_to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cf():
    from cajitafluids_amd import solver
    return solver


def _cases():
    c2 = K.dye_tank_2d(48)
    c3 = K.vortex_tank_3d(16)
    c3["periodic"] = [1, 0, 1]
    c3["clamp_dt"] = 0
    c3["dt"] = 0.002
    return {"2d_dye_tank_48": c2, "3d_periodic_16": c3}


@pytest.mark.parametrize("name", sorted(_cases()))
def test_cell_output_matches_oracle(cf, oracle_mod, name):
    cfg = _cases()[name]
    print(f"cell output vs oracle: {name}", flush=True)
    sol = cf.Solver(cfg)
    orc = oracle_mod.Oracle(cfg)
    sol.init_random_faces(9)
    gpu_to_oracle(sol, orc)
    orc.fill_ghosts()
    sol.gather()
    q_g, v_g = sol.cell_output()
    q_o, v_o = orc.cell_output()
    np.testing.assert_array_equal(q_g, q_o)
    for d in range(cfg["dim"]):
        np.testing.assert_array_equal(v_g[d], v_o[d])
    sol.close()
    orc.close()


def test_write_output_files(cf, tmp_path):
    cfg = K.dye_tank_2d(32)
    with cf.Solver(cfg) as sol:
        sol.setup()
        for _ in range(3):
            sol.step()
        q, vel = sol.cell_output()
        sol.write_output(tmp_path / "data", 3)
        master, piece = vtk.step_files(tmp_path / "data", 3)
        m = vtk.read_pvti(master)
        assert m["pieces"] == [piece] and m["arrays"] == ["quantity", "velocity"]
        assert m["extent"] == [0, 32, 0, 32, 0, 0]
        r = vtk.read_vti(piece)
        assert r["cycle"] == 3 and r["time"] == sol.time and r["dtime"] == sol.dt
        assert r["origin"][:2] == cfg["lo"][:2] and r["spacing"][0] == sol.dx
        np.testing.assert_array_equal(r["quantity"].reshape(q.shape), q)
        v = r["velocity"].reshape(q.shape + (3,))
        for d in range(2):
            np.testing.assert_array_equal(v[..., d], vel[d])


def test_solve_writes_on_the_reference_schedule(cf, tmp_path):
    """Solver::solve writes label 0 before setup, then after the step of
    every loop counter t with t % write_freq == 0 (label t, the state after
    t + 1 steps), until time >= t_final."""
    cfg = K.dye_tank_2d(32)
    out = tmp_path / "data"
    logs = []
    with cf.Solver(cfg) as sol:
        nsteps = sol.solve(5.5 * sol.dt, 2, output_dir=out, log=logs.append)
        assert nsteps == 6
    assert [s.split(" at ")[0] for s in logs] == ["Step %d / 5" % t for t in (0, 2, 4)]
    for t in (0, 2, 4):
        master, piece = vtk.step_files(out, t)
        assert os.path.exists(master)
        assert vtk.read_vti(piece)["cycle"] == t
    for t in (1, 3, 5):
        assert not os.path.exists(vtk.step_files(out, t)[1])
    # label 2 holds the state after 3 steps
    with cf.Solver(cfg) as ref:
        ref.setup()
        for _ in range(3):
            ref.step()
        q, _ = ref.cell_output()
    r = vtk.read_vti(vtk.step_files(out, 2)[1])
    np.testing.assert_array_equal(r["quantity"].reshape(q.shape), q)


def test_cli_runs_the_example(tmp_path, capsys):
    """advection.cpp's main at 32^2 for t = 0.2 (40 steps, PCG + PFMG):
    the banner, the step lines and the step-0, 20 and 40 outputs."""
    from cajitafluids_amd import cli
    out = tmp_path / "data"
    rc = cli.main(["-n", "32", "-t", "0.2", "--output-dir", str(out)])
    assert rc == 0
    text = capsys.readouterr().out
    assert text.startswith("CajitaFluids\n")
    assert "Step 0 / 40 at time = 0.000000" in text and "Step 20 / 40" in text
    for t in (0, 20):
        master, piece = vtk.step_files(out, t)
        assert os.path.exists(master) and os.path.exists(piece)
        r = vtk.read_vti(piece)
        assert r["quantity"].shape == (1, 32, 32)
        assert r["quantity"].max() > 0  # the inflow has added dye
