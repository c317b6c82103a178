// cf_adjoint_run.cpp -- Solver::adjointRun: reverse mode of setup() followed by
// nsteps step() calls, with checkpoints every ckpt steps.
//
// Memory: the states at steps 0, ckpt, 2 ckpt, ... stay on the device; the
// backward sweep takes the segments last to first, recomputes each
// segment's states from its checkpoint (ckpt - 1 forward steps), and walks
// them in reverse.  Peak device memory is (nsteps / ckpt + ckpt + 6) state
// arrays per entity beyond the solver's own.
//
// Per reversed step n (state S_n): gather S_n's ghosts and re-advect it into
// Next (the values the inputs saw), then project^T, inputs^T, and advect^T
// linearised at S_n.  The pressure solves of project^T run the forward's
// PCG (same precision and tolerance) on the adjoint right-hand side.
#include <cstring>
#include <memory>
#include <vector>

#include "cf_adjoint.hpp"
#include "cf_solver.hpp"

namespace cfmi
{

namespace
{

using Arr = std::unique_ptr<DevArray<double>>;

struct StateSet
{
    Arr a[4];
};

Arr make_array(const Geo& g, hipStream_t s)
{
    Arr a(new DevArray<double>());
    a->alloc(g.total, s);
    return a;
}

void alloc_set(const Geo& g, StateSet& set, hipStream_t s)
{
    for (int e = 0; e <= g.dim; ++e)
        if (!set.a[e])
            set.a[e] = make_array(g, s);
}

Fields fields_of(const Geo& g, const StateSet& set)
{
    Fields f;
    for (int e = 0; e < 4; ++e)
        f.e[e] = e <= g.dim ? set.a[e]->ptr : nullptr;
    return f;
}

void copy_arrays(const Geo& g, Fields dst, Fields src, hipStream_t s)
{
    for (int e = 0; e <= g.dim; ++e)
        CF_HIP(hipMemcpyAsync(dst.e[e], src.e[e], (size_t)g.total * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
}

// sum of diag(A) over the owned cells: the wall count of a cell is the sum of
// independent per-axis counts, so its distribution is a convolution.
double diag_sum(const Geo& g)
{
    double dist[7] = {1, 0, 0, 0, 0, 0, 0};
    for (int d = 0; d < g.dim; ++d) {
        double axis[3] = {(double)g.n[d], 0, 0};
        if (!g.periodic[d]) {
            axis[0] = 0;
            for (int i = 0; i < g.n[d]; ++i) {
                int w = (i == 0 && g.solid[d][0]) + (i == g.n[d] - 1 && g.solid[d][1]);
                axis[w] += 1;
            }
        }
        double nd[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 7; ++k)
            for (int w = 0; w < 3 && k + w < 7; ++w)
                nd[k + w] += dist[k] * axis[w];
        std::memcpy(dist, nd, sizeof(dist));
    }
    double s = 0.0;
    for (int k = 0; k < 7; ++k)
        s += dist[k] * g.diag[k];
    return s;
}

} // namespace

cf_status Solver::adjointRun(int nsteps, int ckpt, const double* dq_T,
                             double* const* grad, double dvel[3])
{
    if (nsteps < 0 || ckpt <= 0 || !dq_T || !grad || !dvel)
        throw CfError(CF_EINVAL, "adjoint: nsteps >= 0, ckpt > 0 and non-null "
                                 "arrays required");
    const Geo& g = _mesh->geo();
    const int ne = g.dim + 1;
    const int nck = nsteps / ckpt + 1;
    const int seglen = nsteps < ckpt ? (nsteps > 0 ? nsteps : 1) : ckpt;
    const double time0 = _time;
    const bool singular = operator_singular(g);
    const double dsum = diag_sum(g);
    const RedBuf rb = _vc->redBuf();
    double* ysum = _vc->scratch() + 4;

    StateSet init, fin, bar, obar;
    std::vector<StateSet> ck(nck), seg(seglen);
    alloc_set(g, init, _s);
    alloc_set(g, fin, _s);
    alloc_set(g, bar, _s);
    alloc_set(g, obar, _s);
    Arr y = make_array(g, _s), z = make_array(g, _s);
    DevArray<double> dvel_dev;
    dvel_dev.alloc(4, _s);

    cf_solve_stats st;
    // A failed run leaves the handle as the caller had it: Current holds the
    // initial state at time0, or the final forward state at time_end once
    // the forward sweep has finished (the state cf_adjoint_run returns with).
    bool fwd_done = false;
    double time_end = time0;
    auto fail = [&](cf_status rc) {
        copy_arrays(g, _pm->fields(CF_CURRENT), fields_of(g, fwd_done ? fin : init), _s);
        CF_HIP(hipStreamSynchronize(_s));
        _time = fwd_done ? time_end : time0;
        return rc;
    };
    auto project_adj = [&](Fields b) -> cf_status {
        launch_adj_mask(g, b, _s);
        launch_adj_gradT(g, b, y->ptr, ysum, rb, _s);
        // Jacobi-PCG's map kills diag(A) 1, so the shift follows diag(A); the
        // MG-PCG and unpreconditioned solves return the zero-mean pressure,
        // whose transpose removes the mean of y first
        if (singular && _cfg.preconditioner != CF_PRECOND_JACOBI)
            launch_rhs_shift(g, y->ptr, ysum, _s);
        else if (singular)
            launch_adj_shift(g, y->ptr, ysum, dsum, _s);
        cf_solve_stats ss;
        cf_status rc = _vc->solve(_cfg.precision, y->ptr, z->ptr, &ss);
        if (rc != CF_OK)
            return rc;
        launch_adj_divT(g, z->ptr, b, _s);
        return CF_OK;
    };

    TraceRange tr("Adjoint");
    // forward sweep with checkpoints
    copy_arrays(g, fields_of(g, init), _pm->fields(CF_CURRENT), _s);
    cf_status rc = setup(&st);
    if (rc != CF_OK)
        return fail(rc);
    for (int n = 0; n < nsteps; ++n) {
        if (n % ckpt == 0) {
            alloc_set(g, ck[n / ckpt], _s);
            copy_arrays(g, fields_of(g, ck[n / ckpt]), _pm->fields(CF_CURRENT), _s);
        }
        if ((rc = step(&st)) != CF_OK)
            return fail(rc);
    }
    copy_arrays(g, fields_of(g, fin), _pm->fields(CF_CURRENT), _s);
    time_end = _time;
    fwd_done = true;

    // backward sweep (dq_T and the gradients may be caller memory: kernel
    // copies only)
    launch_copy_padded(g, CF_CELL, bar.a[0]->ptr, dq_T, true, _s);
    for (int c0 = nsteps > 0 ? ((nsteps - 1) / ckpt) * ckpt : -1; c0 >= 0;
         c0 -= ckpt) {
        const int c1 = c0 + ckpt < nsteps ? c0 + ckpt : nsteps;
        copy_arrays(g, _pm->fields(CF_CURRENT), fields_of(g, ck[c0 / ckpt]), _s);
        for (int n = c0; n < c1; ++n) {
            alloc_set(g, seg[n - c0], _s);
            copy_arrays(g, fields_of(g, seg[n - c0]), _pm->fields(CF_CURRENT), _s);
            if (n + 1 < c1 && (rc = step(&st)) != CF_OK)
                return fail(rc);
        }
        for (int n = c1 - 1; n >= c0; --n) {
            copy_arrays(g, _pm->fields(CF_CURRENT), fields_of(g, seg[n - c0]), _s);
            _pm->gather(CF_CURRENT);
            launch_advect(g, _pm->fields(CF_CURRENT), _pm->fields(CF_NEXT), _s);
            if ((rc = project_adj(fields_of(g, bar))) != CF_OK)
                return fail(rc);
            launch_adj_inputs(g, _pm->fields(CF_NEXT), fields_of(g, bar),
                              dvel_dev.ptr, _s);
            for (int e = 0; e < ne; ++e) {
                std::swap(bar.a[e], obar.a[e]);
                CF_HIP(hipMemsetAsync(bar.a[e]->ptr, 0,
                                      (size_t)g.total * sizeof(double), _s));
            }
            launch_adj_advect(g, _pm->fields(CF_CURRENT), fields_of(g, obar),
                              fields_of(g, bar), _s);
        }
    }
    // setup^T
    if ((rc = project_adj(fields_of(g, bar))) != CF_OK)
        return fail(rc);
    launch_adj_inputs(g, fields_of(g, init), fields_of(g, bar), dvel_dev.ptr, _s);

    for (int e = 0; e < ne; ++e)
        if (grad[e])
            launch_copy_padded(g, e, grad[e], bar.a[e]->ptr, true, _s);
    double hv[4];
    CF_HIP(hipMemcpyAsync(hv, dvel_dev.ptr, sizeof(hv), hipMemcpyDeviceToHost, _s));
    copy_arrays(g, _pm->fields(CF_CURRENT), fields_of(g, fin), _s);
    CF_HIP(hipStreamSynchronize(_s));
    for (int d = 0; d < 3; ++d)
        dvel[d] = hv[d];
    _time = time_end;
    return CF_OK;
}

} // namespace cfmi
