"""Shared helpers for the parity tests: state transfer between the HIP
solver and the oracle, and error norms."""
import numpy as np

from cajitafluids_amd import configs as K


def entities(dim):
    return list(range(dim + 1))


def gpu_to_oracle(sol, orc, version=0):
    """Copy every state array of the GPU solver into the oracle."""
    for e in entities(sol.dim):
        orc.field(e, version)[...] = sol.get(e, version)


def oracle_to_gpu(orc, sol, version=0):
    for e in entities(orc.dim):
        sol.set(e, np.array(orc.field(e, version)), version)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    if nb == 0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / nb)


def compare_state(sol, orc, version=0):
    """Max relative L2 over q, u, v (w) owned regions, and max abs diff."""
    out = {}
    for e in entities(sol.dim):
        g = sol.owned_view(sol.get(e, version), e)
        o = orc.owned_view(orc.field(e, version), e)
        out[e] = (rel_l2(g, o), float(np.max(np.abs(g - o))) if g.size else 0.0)
    return out


def small_2d(n=32, **kw):
    c = K.dye_tank_2d(n)
    c.update(kw)
    return c


def small_3d(n=16, **kw):
    c = K.vortex_tank_3d(n)
    c.update(kw)
    return c
