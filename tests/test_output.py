"""The output writer's data and files without a GPU: the oracle's cell data
(SiloWriter.hpp:141-197) against its closed form, and the VTK piece writer
(cf_write_vti) against a numpy reader."""
import numpy as np
import pytest

from cajitafluids_amd import configs as K
from cajitafluids_amd import vtk


@pytest.mark.parametrize("dim,periodic", [(2, [0, 0, 0]), (3, [1, 0, 1])])
def test_oracle_cell_output_is_the_face_average(oracle_mod, dim, periodic):
    """At a cell centre the linear interpolation of a face component has
    weights 1/2, 1/2 along its axis and 1, 0 along the others, so it is
    (f_i + f_{i+1}) / 2 up to the rounding of the computed logical
    coordinates (weights off by an ulp), i.e. to about 1e-14 of the field."""
    cfg = K.base_config(dim, 12)
    cfg["num_cell"] = [12, 10, 8][:dim] + [1] * (3 - dim)
    cfg["hi"] = [1.0, 10 / 12, 8 / 12 if dim == 3 else 1.0]
    cfg["periodic"] = periodic
    o = oracle_mod.Oracle(cfg)
    o.init_random_faces(5)
    o.fill_ghosts()
    q, vel = o.cell_output()
    np.testing.assert_array_equal(q, o.owned_view(o.field(0), 0))
    lo, hi = o.owned(0)
    for d in range(dim):
        f = o.field(d + 1)
        sl0 = [slice(lo[a], hi[a]) for a in range(3)]
        sl1 = list(sl0)
        sl1[d] = slice(lo[d] + 1, hi[d] + 1)
        a = f[tuple(sl0[::-1])]
        b = f[tuple(sl1[::-1])]
        np.testing.assert_allclose(vel[d], 0.5 * a + 0.5 * b, rtol=0,
                                   atol=1e-13 * np.abs(f).max())
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_write_vti_round_trip(tmp_path, dim):
    rng = np.random.default_rng(dim)
    shape = (5, 4, 3)[3 - dim:]
    q = rng.normal(size=shape)
    vel = [rng.normal(size=shape) for _ in range(dim)]
    path = tmp_path / "piece.vti"
    vtk.write_vti(path, q, vel, origin=[0.25, -1.0, 2.0][:dim], dx=0.125, cycle=7, time=1.5,
                  dt=0.01)
    r = vtk.read_vti(path)
    n = list(shape[::-1])
    assert r["extent"] == [0, n[0], 0, n[1], 0, n[2] if dim == 3 else 0]
    assert r["origin"][:dim] == [0.25, -1.0, 2.0][:dim]
    assert r["spacing"] == [0.125] * 3
    assert (r["cycle"], r["time"], r["dtime"]) == (7, 1.5, 0.01)
    np.testing.assert_array_equal(r["quantity"].reshape(shape), q)
    v = r["velocity"].reshape(shape + (3,))
    for d in range(dim):
        np.testing.assert_array_equal(v[..., d], vel[d])
    if dim == 2:
        assert not v[..., 2].any()


def test_write_vti_rejects_bad_arguments(tmp_path):
    from cajitafluids_amd import _lib
    with pytest.raises(_lib.CfError):
        vtk.write_vti(tmp_path / "no_such_dir" / "x.vti", np.zeros((2, 2)),
                      [np.zeros((2, 2))] * 2, origin=[0, 0], dx=1.0)
