"""A whole setup + steps on an 8 x 8 grid restated independently in numpy,
with a dense pseudo-inverse pressure solve, against the oracle.

The numpy step follows the reference directly (Solver.hpp:125-147 setup /
step order; TimeIntegrator.hpp:36-116 RK3 and advection with the B-spline
weights of Cajita's order-1 / order-3 splines; Solver.hpp:181-263 inputs
with InflowSource.hpp:33-78 and BoundaryConditions.hpp:102-129;
VelocityCorrector.hpp:182-282 divergence, solve and projection including
the FaceJ rule applied to u) with its own index arithmetic, point by point.
The only shared choices are the documented ones: traces leaving the
ghosted array are clamped, and the solve is the minimum-norm solution plus
the oracle's constant.  With only SOLID walls the pressure is fixed up to a
constant, which the Jacobi-preconditioned CG leaves at a level of its own
(its iterates start from D^-1 b, not from the range of A); the projection
does not see it, except through the as-written FaceJ rule, which leaves the
wall-normal v at -scale (p - 0) on the walls.  So each numpy solve takes
the oracle's mean pressure of the same solve, one scalar.  A wrong phase
order, sign, ghost rule or quirk in the oracle fails this test; agreement
is to the pressure solver's tolerance (1e-14 here).
"""
import numpy as np
import pytest

from cajitafluids_amd import configs as K

N, H = 8, 3


def _cfg(field_order, quirks, top=K.SOLID):
    c = K.base_config(2, N)
    c.update(field_order=field_order, velocity_order=1, quirks=quirks, clamp_dt=0, dt=0.02,
             cg_tol=1e-14, cg_max_iter=5000, inflow_shape=K.INFLOW_BOX,
             inflow_location=[0.375, 0.125, 0.0], inflow_size=[0.25, 0.25, 0.0],
             inflow_velocity=[0.5, 3.0, 0.0], inflow_quantity=1.0,
             body_force=[0.3, -0.7, 0.0])
    c["wall_type"] = [[K.SOLID, K.SOLID], [K.SOLID, top], [K.SOLID, K.SOLID]]
    return c


class NumpyFluid:
    """2-D state on ghosted arrays indexed [j, i]: q (cells), u (x faces),
    v (y faces); ghosts stay zero (walls, one rank)."""

    def __init__(self, cfg):
        self.c = cfg
        self.dx = 1.0 / N
        self.dt = cfg["dt"]
        self.ext = {0: (N + 2 * H, N + 2 * H), 1: (N + 2 * H + 1, N + 2 * H),
                    2: (N + 2 * H, N + 2 * H + 1)}  # (x, y) extents
        self.f = {e: np.zeros(self.ext[e][::-1]) for e in range(3)}

    # Cajita local mesh coordinates of entity e at local index l along axis d
    def coord(self, e, d, l):
        face = e == d + 1
        return (l - H + (0.0 if face else 0.5)) * self.dx

    def owned(self, e):
        nx = N + 1 if e == 1 else N
        ny = N + 1 if e == 2 else N
        return range(H, H + nx), range(H, H + ny)

    @staticmethod
    def weights(order, xi):
        s = int(xi)
        t = xi - s
        if order == 1:
            return s, [1.0 - t, t]
        return s - 1, [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6,
                       (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]

    def sample(self, e, order, field, p):
        idx, ws = [], []
        for d in range(2):
            c0, c1 = self.coord(e, d, 0), self.coord(e, d, 1)
            xi = (p[d] - c0) / (c1 - c0)
            lo, hi = (0.0, self.ext[e][d] - 2.0) if order == 1 else (1.0, self.ext[e][d] - 3.0)
            s, w = self.weights(order, min(max(xi, lo), hi))
            idx.append(s)
            ws.append(w)
        val = 0.0
        for a, wa in enumerate(ws[0]):
            for b, wb in enumerate(ws[1]):
                val += wa * wb * field[idx[1] + b, idx[0] + a]
        return val

    def velocity(self, p):
        return np.array([self.sample(1, 1, self.f[1], p), self.sample(2, 1, self.f[2], p)])

    def advect(self):
        dt, quirk = self.dt, self.c["quirks"] & K.QUIRK_RK3_V0
        nxt = {e: np.zeros_like(self.f[e]) for e in range(3)}
        for e in range(3):
            ri, rj = self.owned(e)
            for j in rj:
                for i in ri:
                    x0 = np.array([self.coord(e, 0, i), self.coord(e, 1, j)])
                    v0 = self.velocity(x0)
                    v1 = self.velocity(x0 - 0.5 * dt * v0)
                    v2 = self.velocity(x0 - 0.75 * dt * (v0 if quirk else v1))
                    tr = x0 - dt * (2 / 9 * v0 + 3 / 9 * v1 + 4 / 9 * v2)
                    nxt[e][j, i] = self.sample(e, self.c["field_order"], self.f[e], tr)
        self.f = nxt

    def inputs(self):
        lo, sz = self.c["inflow_location"], self.c["inflow_size"]
        U, Q, fb = self.c["inflow_velocity"], self.c["inflow_quantity"], self.c["body_force"]
        solid = [[w == K.SOLID for w in self.c["wall_type"][d]] for d in range(2)]

        def inside(x, y):
            return lo[0] <= x < lo[0] + sz[0] and lo[1] <= y < lo[1] + sz[1]
        for e in range(3):
            ri, rj = self.owned(e)
            for j in rj:
                for i in ri:
                    x, y = self.coord(e, 0, i), self.coord(e, 1, j)
                    a = self.f[e]
                    if e == 0:
                        if inside(x, y):
                            a[j, i] = max(a[j, i], Q)
                        continue
                    d = e - 1
                    if inside(x, y) and abs(a[j, i]) < abs(U[d]):
                        a[j, i] = U[d]
                    a[j, i] += fb[d] * self.dt
                    g = (i if d == 0 else j) - H
                    if (g <= 0 and solid[d][0]) or (g > N - 1 and solid[d][1]):
                        a[j, i] = 0.0

    def correct_velocity(self, p_mean=None):
        u, v = self.f[1], self.f[2]
        rhs = np.zeros((N, N))
        for j in range(N):
            for i in range(N):
                I, J = i + H, j + H
                rhs[j, i] = -(1.0 / self.dx) * (u[J, I + 1] - u[J, I] + v[J + 1, I] - v[J, I])
        s = self.dt / (self.c["density"] * self.dx ** 2)
        free = [[w == K.FREE for w in self.c["wall_type"][d]] for d in range(2)]
        A = np.zeros((N * N, N * N))
        for j in range(N):
            for i in range(N):
                r = j * N + i
                for d, (di, dj) in enumerate(((1, 0), (0, 1))):
                    for sgn in (-1, 1):
                        ii, jj = i + sgn * di, j + sgn * dj
                        if 0 <= ii < N and 0 <= jj < N:
                            A[r, r] += s
                            A[r, jj * N + ii] -= s
                        elif free[d][0 if sgn < 0 else 1]:
                            A[r, r] += s  # Dirichlet p = 0 in the ghost
        p = np.zeros((N + 2 * H, N + 2 * H))
        sol = (np.linalg.pinv(A) @ rhs.ravel()).reshape(N, N)
        if p_mean is not None:
            sol = sol - sol.mean() + p_mean
        p[H:H + N, H:H + N] = sol
        scale = self.dt / (self.c["density"] * self.dx)
        solid = [[w == K.SOLID for w in self.c["wall_type"][d]] for d in range(2)]
        ri, rj = self.owned(1)
        for j in rj:
            for i in ri:
                u[j, i] -= scale * (p[j, i] - p[j, i - 1])
                gi = i - H
                if (gi <= 0 and solid[0][0]) or (gi > N - 1 and solid[0][1]):
                    u[j, i] = 0.0
        target = u if self.c["quirks"] & K.QUIRK_PROJ_BC_U else v
        ri, rj = self.owned(2)
        for j in rj:
            for i in ri:
                v[j, i] -= scale * (p[j, i] - p[j - 1, i])
                gj = j - H
                if (gj <= 0 and solid[1][0]) or (gj > N - 1 and solid[1][1]):
                    target[j, i] = 0.0

    def setup(self, p_mean=None):
        self.inputs()
        self.correct_velocity(p_mean)

    def step(self, p_mean=None):
        self.advect()
        self.inputs()
        self.correct_velocity(p_mean)


@pytest.mark.parametrize("field_order,quirks,top", [
    (1, K.QUIRK_RK3_V0 | K.QUIRK_PROJ_BC_U, K.SOLID),
    (3, K.QUIRK_RK3_V0 | K.QUIRK_PROJ_BC_U, K.SOLID),
    (3, 0, K.SOLID),
    (1, 0, K.FREE),
])
def test_numpy_step_matches_oracle(oracle_mod, field_order, quirks, top):
    cfg = _cfg(field_order, quirks, top)
    npf = NumpyFluid(cfg)
    orc = oracle_mod.Oracle(cfg)
    rng = np.random.default_rng(21)
    for e in range(3):
        ri, rj = npf.owned(e)
        sl = (slice(rj[0], rj[-1] + 1), slice(ri[0], ri[-1] + 1))
        vals = rng.normal(size=(len(rj), len(ri))) * (0.5 if e else 1.0)
        npf.f[e][sl] = vals
        orc.field(e)[0][sl] = vals
    singular = top == K.SOLID

    def p_mean():
        return float(orc.owned_view(orc.lhs(), 0).mean()) if singular else None
    rc, _ = orc.setup()
    assert rc == 0
    npf.setup(p_mean())
    for n in range(3):
        rc, _ = orc.step()
        assert rc == 0
        npf.step(p_mean())
    for e in range(3):
        ri, rj = npf.owned(e)
        sl = (slice(rj[0], rj[-1] + 1), slice(ri[0], ri[-1] + 1))
        a = orc.field(e)[0][sl]
        b = npf.f[e][sl]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err <= 1e-11, (e, err)
    orc.close()
