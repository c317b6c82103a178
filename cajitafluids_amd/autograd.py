"""torch.autograd bridge for a rollout of the hot path.

``rollout(cfg, nsteps, q0, u0, v0, w0, inflow_velocity, ckpt)`` runs
Solver::setup and nsteps Solver::step (src/Solver.hpp:125-147) on the GPU
from the given initial state and returns the final quantity q_T.  Its
backward is the hand-written reverse mode of the library (cf_adjoint_run):
gradients with respect to the owned initial q, u, v, (w) and the three
inflow velocity components, recomputed from checkpoints every ``ckpt``
steps.  The time step is held fixed: with ``clamp_dt`` on, the clamp value
is taken at the given inflow velocity and not differentiated.

Tensors are float64, shaped like the owned region of each entity,
(z, y, x) with z = 1 in 2-D (``owned_shape``); they may live on any device
(the library reads them through host copies).  The pressure solves are
treated as exact, so gradients match finite differences to the PCG
tolerance.
"""
import numpy as np
import torch

from .configs import CELL
from .solver import Solver, layout

__all__ = ["Rollout", "rollout", "owned_shape"]


def owned_shape(cfg, entity):
    """(z, y, x) shape of the owned region of an entity."""
    _, lo, hi = layout(cfg, entity)
    return tuple(hi[d] - lo[d] for d in (2, 1, 0))


def _dense(cfg, entity, owned):
    ext, lo, hi = layout(cfg, entity)
    a = np.zeros((ext[2], ext[1], ext[0]))
    a[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = owned
    return a


def _owned(cfg, entity, dense):
    _, lo, hi = layout(cfg, entity)
    return dense[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]


def _host(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _with_inflow(cfg, inflow_velocity):
    c = dict(cfg)
    c["inflow_velocity"] = [float(x) for x in _host(inflow_velocity).reshape(-1)[:3]]
    return c


def _load_state(sol, cfg, state):
    for e, t in enumerate(state):
        if t is None:
            continue
        sol.set(e, _dense(cfg, e, _host(t)))


class Rollout(torch.autograd.Function):
    """q_T = rollout(q0, u0, v0, w0, U); backward by the library's adjoint."""

    @staticmethod
    def forward(ctx, cfg, nsteps, ckpt, q0, u0, v0, w0, inflow_velocity):
        run_cfg = _with_inflow(cfg, inflow_velocity)
        state = (q0, u0, v0, w0 if run_cfg["dim"] == 3 else None)
        with Solver(run_cfg) as sol:
            _load_state(sol, run_cfg, state)
            sol.setup()
            for _ in range(nsteps):
                sol.step()
            q_T = _owned(run_cfg, CELL, sol.get(CELL))
        ctx.run_cfg, ctx.nsteps, ctx.ckpt = run_cfg, nsteps, ckpt
        ctx.save_for_backward(*(t for t in state if t is not None))
        ctx.has_w = state[3] is not None
        ctx.u_device = inflow_velocity.device
        return torch.as_tensor(np.ascontiguousarray(q_T), dtype=torch.float64,
                               device=q0.device)

    @staticmethod
    def backward(ctx, g_qT):
        saved = list(ctx.saved_tensors)
        state = saved + ([None] if not ctx.has_w else [])
        cfg = ctx.run_cfg
        with Solver(cfg) as sol:
            _load_state(sol, cfg, state)
            g = sol.adjoint_run(ctx.nsteps, ctx.ckpt, _dense(cfg, CELL, _host(g_qT)))
        dev = state[0].device

        def out(key, entity):
            if g[key] is None:
                return None
            return torch.as_tensor(np.ascontiguousarray(_owned(cfg, entity, g[key])),
                                   dtype=torch.float64, device=dev)

        grads = [out("dq0", 0), out("du", 1), out("dv", 2),
                 out("dw", 3) if ctx.has_w else None]
        d_u = torch.as_tensor(g["dvel"], dtype=torch.float64, device=ctx.u_device)
        return (None, None, None, *grads, d_u)


def rollout(cfg, nsteps, q0, u0, v0, w0, inflow_velocity, ckpt=10):
    """Final quantity after setup + nsteps steps, differentiable in the
    initial state and the inflow velocity (w0 may be None in 2-D)."""
    return Rollout.apply(cfg, nsteps, ckpt, q0, u0, v0, w0, inflow_velocity)
