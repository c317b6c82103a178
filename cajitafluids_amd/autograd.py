"""torch.autograd bridge for a rollout of the hot path.

``rollout(cfg, nsteps, q0, u0, v0, w0, inflow_velocity, ckpt)`` runs
Solver::setup and nsteps Solver::step (src/Solver.hpp:125-147) on the GPU
from the given initial state and returns the final quantity q_T.  Its
backward is the hand-written reverse mode of the library: gradients with
respect to the owned initial q, u, v, (w) and the inflow velocity
components, recomputed from checkpoints every ``ckpt`` steps.  The time
step is held fixed: with ``clamp_dt`` on, the clamp value is taken at the
given inflow velocity and not differentiated.

Tensors are float64, shaped like the owned region of each entity,
(z, y, x) with z = 1 in 2-D (``owned_shape``).  State tensors on the GPU
stay there: they are handed to the library as device pointers in its padded
layout (cf_copy_in_device / cf_copy_out_device / cf_adjoint_run_device), so
no state array crosses PCIe.  State tensors on the CPU take the host entry
points (cf_copy_in / cf_copy_out / cf_adjoint_run).  The inflow velocity is
a parameter of the handle's configuration and is read on the host either
way.  The pressure solves are treated as exact, so gradients match finite
differences to the PCG tolerance.
"""
import numpy as np
import torch

from .configs import CELL
from .solver import Solver, layout, padded_layout

__all__ = ["Rollout", "rollout", "owned_shape"]


def owned_shape(cfg, entity):
    """(z, y, x) shape of the owned region of an entity."""
    _, lo, hi = layout(cfg, entity)
    return tuple(hi[d] - lo[d] for d in (2, 1, 0))


# ---- host path: dense ghosted numpy arrays ------------------------------------
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


# ---- device path: padded device tensors -----------------------------------------
def _owned_view(cfg, entity, padded):
    """The owned box of an entity inside a padded 1-D device tensor, as a
    (z, y, x) strided view."""
    off, px, py, _ = padded_layout(cfg)
    _, lo, hi = layout(cfg, entity)
    shape = tuple(hi[d] - lo[d] for d in (2, 1, 0))
    start = off + lo[0] + px * (lo[1] + py * lo[2])
    return padded.as_strided(shape, (px * py, px, 1), start)


def _padded_from(cfg, entity, owned):
    """A zero padded tensor on owned's device holding owned in its box."""
    elems = padded_layout(cfg)[3]
    p = torch.zeros(elems, dtype=torch.float64, device=owned.device)
    _owned_view(cfg, entity, p).copy_(owned.detach().to(torch.float64))
    return p


def _with_inflow(cfg, inflow_velocity):
    c = dict(cfg)
    u = _host(inflow_velocity).reshape(-1)
    if u.size > 3:
        raise ValueError("inflow_velocity has at most 3 components")
    vel = [0.0, 0.0, 0.0]
    vel[:u.size] = [float(x) for x in u]
    c["inflow_velocity"] = vel
    return c


def _load_state_host(sol, cfg, state):
    for e, t in enumerate(state):
        if t is not None:
            sol.set(e, _dense(cfg, e, _host(t)))


def _load_state_device(sol, cfg, state):
    """Padded device copies of the initial state, handed to the library after
    torch's stream has finished writing them.  Returns the tensors, which
    must outlive the library's stream-ordered copies (until sol.sync())."""
    pads = [None if t is None else _padded_from(cfg, e, t) for e, t in enumerate(state)]
    torch.cuda.synchronize()
    for e, p in enumerate(pads):
        if p is not None:
            sol.copy_in_device(e, p.data_ptr())
    return pads


class Rollout(torch.autograd.Function):
    """q_T = rollout(q0, u0, v0, w0, U); backward by the library's adjoint."""

    @staticmethod
    def forward(ctx, cfg, nsteps, ckpt, q0, u0, v0, w0, inflow_velocity):
        run_cfg = _with_inflow(cfg, inflow_velocity)
        state = (q0, u0, v0, w0 if run_cfg["dim"] == 3 else None)
        on_gpu = q0.is_cuda
        with Solver(run_cfg) as sol:
            if on_gpu:
                pads = _load_state_device(sol, run_cfg, state)
                sol.setup()
                for _ in range(nsteps):
                    sol.step()
                out = torch.empty(padded_layout(run_cfg)[3], dtype=torch.float64,
                                  device=q0.device)
                sol.copy_out_device(CELL, out.data_ptr())
                sol.sync()
                del pads
                q_T = _owned_view(run_cfg, CELL, out).clone()
            else:
                _load_state_host(sol, run_cfg, state)
                sol.setup()
                for _ in range(nsteps):
                    sol.step()
                q_T = torch.as_tensor(
                    np.ascontiguousarray(_owned(run_cfg, CELL, sol.get(CELL))),
                    dtype=torch.float64, device=q0.device)
        ctx.run_cfg, ctx.nsteps, ctx.ckpt = run_cfg, nsteps, ckpt
        ctx.save_for_backward(*(t for t in state if t is not None))
        ctx.has_w = state[3] is not None
        ctx.u_device = inflow_velocity.device
        ctx.u_shape = inflow_velocity.shape
        return q_T

    @staticmethod
    def backward(ctx, g_qT):
        saved = list(ctx.saved_tensors)
        state = saved + ([None] if not ctx.has_w else [])
        cfg = ctx.run_cfg
        dev = state[0].device
        ne = 4 if ctx.has_w else 3
        with Solver(cfg) as sol:
            if state[0].is_cuda:
                pads = _load_state_device(sol, cfg, state)
                g_pad = _padded_from(cfg, CELL, g_qT.to(dev))
                elems = padded_layout(cfg)[3]
                outs = [torch.empty(elems, dtype=torch.float64, device=dev)
                        for _ in range(ne)]
                torch.cuda.synchronize()
                ptr = [o.data_ptr() for o in outs] + [None]
                dvel = sol.adjoint_run_device(ctx.nsteps, ctx.ckpt, g_pad.data_ptr(),
                                              du_ptr=ptr[1], dv_ptr=ptr[2],
                                              dw_ptr=ptr[3] if ctx.has_w else None,
                                              dq0_ptr=ptr[0])
                del pads, g_pad
                grads = [_owned_view(cfg, e, outs[e]).clone() for e in range(ne)]
            else:
                _load_state_host(sol, cfg, state)
                g = sol.adjoint_run(ctx.nsteps, ctx.ckpt, _dense(cfg, CELL, _host(g_qT)))
                keys = ["dq0", "du", "dv", "dw"]
                grads = [torch.as_tensor(np.ascontiguousarray(_owned(cfg, e, g[keys[e]])),
                                         dtype=torch.float64, device=dev)
                         for e in range(ne)]
                dvel = g["dvel"]
        grads += [None] * (4 - len(grads))
        n_u = int(np.prod(ctx.u_shape)) if len(ctx.u_shape) else 1
        d_u = torch.as_tensor(np.asarray(dvel[:n_u]), dtype=torch.float64,
                              device=ctx.u_device).reshape(ctx.u_shape)
        return (None, None, None, *grads, d_u)


def rollout(cfg, nsteps, q0, u0, v0, w0, inflow_velocity, ckpt=10):
    """Final quantity after setup + nsteps steps, differentiable in the
    initial state and the inflow velocity (w0 may be None in 2-D)."""
    return Rollout.apply(cfg, nsteps, ckpt, q0, u0, v0, w0, inflow_velocity)
