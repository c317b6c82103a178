// cf_api.cpp -- the C ABI (include/cf_api.h): a thin shell over
// cfmi::Solver that turns exceptions into status codes.
#include "cf_api.h"

#include <cstring>
#include <string>
#include <vector>

#include "cf_solver.hpp"

using cfmi::CfError;

struct cf_solver
{
    cfmi::Solver* s = nullptr;
    std::string err;
};

static thread_local std::string g_create_error;

template <class F>
static cf_status guarded(cf_solver* h, F&& f)
{
    if (!h || !h->s)
        return CF_EINVAL;
    try {
        return f(*h->s);
    } catch (const CfError& e) {
        h->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        h->err = "host allocation failed";
        return CF_ENOMEM;
    } catch (const std::exception& e) {
        h->err = e.what();
        return CF_EDEVICE;
    }
}

static bool valid_entity(const cfmi::Geo& g, int e)
{
    return e >= 0 && e <= g.dim;
}

extern "C" {

cf_status cf_layout_query(const cf_config* cfg, int entity, int ghost_ext[3],
                          int own_lo[3], int own_hi[3])
{
    if (!cfg)
        return CF_EINVAL;
    cfmi::Geo g;
    const char* why;
    cf_status rc = cfmi::build_geometry(*cfg, g, &why);
    if (rc != CF_OK)
        return rc;
    if (!valid_entity(g, entity))
        return CF_EINVAL;
    for (int d = 0; d < 3; ++d) {
        if (ghost_ext)
            ghost_ext[d] = g.ext[entity][d];
        if (own_lo)
            own_lo[d] = g.own_lo(d);
        if (own_hi)
            own_hi[d] = g.own_hi[entity][d];
    }
    return CF_OK;
}

cf_status cf_geometry_query(const cf_config* cfg, double* dx, double* dt)
{
    if (!cfg)
        return CF_EINVAL;
    cfmi::Geo g;
    const char* why;
    cf_status rc = cfmi::build_geometry(*cfg, g, &why);
    if (rc != CF_OK)
        return rc;
    if (dx)
        *dx = g.dx;
    if (dt)
        *dt = g.dt;
    return CF_OK;
}

cf_status cf_padded_layout(const cf_config* cfg, int64_t* offset, int64_t* pitch_x,
                           int64_t* pitch_y, int64_t* elems)
{
    if (!cfg)
        return CF_EINVAL;
    cfmi::Geo g;
    const char* why;
    cf_status rc = cfmi::build_geometry(*cfg, g, &why);
    if (rc != CF_OK)
        return rc;
    if (offset)
        *offset = g.sx;
    if (pitch_x)
        *pitch_x = g.px;
    if (pitch_y)
        *pitch_y = g.py;
    if (elems)
        *elems = g.total;
    return CF_OK;
}

cf_status cf_entity_origin(const cf_config* cfg, int entity, double origin[3])
{
    if (!cfg || !origin)
        return CF_EINVAL;
    cfmi::Geo g;
    const char* why;
    cf_status rc = cfmi::build_geometry(*cfg, g, &why);
    if (rc != CF_OK)
        return rc;
    if (!valid_entity(g, entity))
        return CF_EINVAL;
    for (int d = 0; d < 3; ++d)
        origin[d] = d < g.dim ? g.x0[entity][d] : 0.0;
    return CF_OK;
}

cf_status cf_create(const cf_config* cfg, cf_solver** out)
{
    if (!out)
        return CF_EINVAL;
    *out = nullptr;
    if (!cfg) {
        g_create_error = "null config";
        return CF_EINVAL;
    }
    cf_solver* h = new (std::nothrow) cf_solver();
    if (!h)
        return CF_ENOMEM;
    try {
        h->s = new cfmi::Solver(*cfg);
    } catch (const CfError& e) {
        g_create_error = e.what();
        delete h;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_create_error = "host allocation failed";
        delete h;
        return CF_ENOMEM;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        delete h;
        return CF_EDEVICE;
    }
    *out = h;
    return CF_OK;
}

void cf_destroy(cf_solver* h)
{
    if (!h)
        return;
    try {
        delete h->s;
    } catch (...) {
    }
    delete h;
}

const char* cf_last_error(const cf_solver* h)
{
    if (!h)
        return g_create_error.c_str();
    return h->err.c_str();
}

cf_status cf_setup(cf_solver* h, cf_solve_stats* st)
{
    return guarded(h, [&](cfmi::Solver& s) { return s.setup(st); });
}

cf_status cf_step(cf_solver* h, cf_solve_stats* st)
{
    return guarded(h, [&](cfmi::Solver& s) { return s.step(st); });
}

cf_status cf_run(cf_solver* h, int nsteps, cf_solve_stats* last)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (nsteps < 0)
            throw CfError(CF_EINVAL, "nsteps must be >= 0");
        return s.run(nsteps, last);
    });
}

double cf_time(const cf_solver* h) { return h && h->s ? h->s->time() : 0.0; }
double cf_dt(const cf_solver* h) { return h && h->s ? h->s->geo().dt : 0.0; }
double cf_dx(const cf_solver* h) { return h && h->s ? h->s->geo().dx : 0.0; }

cf_status cf_time_integrate(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.timeIntegrate();
        return CF_OK;
    });
}

cf_status cf_add_inputs(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.addInputs();
        return CF_OK;
    });
}

cf_status cf_correct_velocity(cf_solver* h, cf_solve_stats* st)
{
    return guarded(h, [&](cfmi::Solver& s) { return s.vc().correctVelocity(st); });
}

cf_status cf_build_rhs(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.vc().buildRHS();
        return CF_OK;
    });
}

cf_status cf_rhs_remove_mean(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.vc().removeMean();
        return CF_OK;
    });
}

cf_status cf_pressure_solve(cf_solver* h, int precision, cf_solve_stats* st)
{
    return guarded(h, [&](cfmi::Solver& s) {
        return s.vc().solve(precision, s.vc().rhs(), s.vc().lhs(), st);
    });
}

cf_status cf_pressure_solve_ext(cf_solver* h, const double* rhs_dev,
                                double* lhs_dev, int precision,
                                cf_solve_stats* st)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!rhs_dev || !lhs_dev || rhs_dev == lhs_dev)
            throw CfError(CF_EINVAL, "rhs and lhs must be distinct device arrays");
        return s.vc().solve(precision, rhs_dev, lhs_dev, st);
    });
}

cf_status cf_apply_pressure(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.vc().applyPressure();
        return CF_OK;
    });
}

static void fill_view(const cfmi::Geo& g, double* data, int entity, cf_view* out)
{
    out->data = data;
    out->offset = g.sx;
    out->pitch_x = g.px;
    out->pitch_y = g.py;
    for (int d = 0; d < 3; ++d) {
        out->ghost_ext[d] = g.ext[entity][d];
        out->own_lo[d] = g.own_lo(d);
        out->own_hi[d] = g.own_hi[entity][d];
    }
}

cf_status cf_field_view(cf_solver* h, int entity, int version, cf_view* out)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!out || !valid_entity(s.geo(), entity) || (version != 0 && version != 1))
            throw CfError(CF_EINVAL, "bad entity/version/out");
        fill_view(s.geo(), s.pm().get(entity, version), entity, out);
        return CF_OK;
    });
}

cf_status cf_pressure_view(cf_solver* h, int which, cf_view* out)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!out || (which != 0 && which != 1))
            throw CfError(CF_EINVAL, "which must be 0 (lhs) or 1 (rhs)");
        fill_view(s.geo(), which == 0 ? s.vc().lhs() : s.vc().rhs(), CF_CELL, out);
        return CF_OK;
    });
}

// Dense ghosted host array <-> padded device layout: one contiguous host
// transfer through a dense device buffer, repacked on the device
// (plane-by-plane 2-D copies from pageable memory ran at about 3 ms per
// plane).
static void copy_dense(cfmi::Solver& s, double* dev, int entity, void* host,
                       bool to_device)
{
    const cfmi::Geo& g = s.geo();
    const int E0 = g.ext[entity][0], E1 = g.ext[entity][1], E2 = g.ext[entity][2];
    const size_t n = (size_t)E0 * E1 * E2;
    cfmi::DevArray<double> dense;
    dense.alloc((int64_t)n, s.stream());
    if (to_device) {
        CF_HIP(hipMemcpyAsync(dense.ptr, host, n * sizeof(double), hipMemcpyHostToDevice,
                              s.stream()));
        cfmi::launch_repack(g, entity, dev, dense.ptr, true, s.stream());
    } else {
        cfmi::launch_repack(g, entity, dev, dense.ptr, false, s.stream());
        CF_HIP(hipMemcpyAsync(host, dense.ptr, n * sizeof(double), hipMemcpyDeviceToHost,
                              s.stream()));
    }
    CF_HIP(hipStreamSynchronize(s.stream()));
}

cf_status cf_copy_in(cf_solver* h, int entity, int version, const double* host)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!host || !valid_entity(s.geo(), entity) || (version != 0 && version != 1))
            throw CfError(CF_EINVAL, "bad entity/version/host");
        copy_dense(s, s.pm().get(entity, version), entity, (void*)host, true);
        return CF_OK;
    });
}

cf_status cf_copy_out(cf_solver* h, int entity, int version, double* host)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!host || !valid_entity(s.geo(), entity) || (version != 0 && version != 1))
            throw CfError(CF_EINVAL, "bad entity/version/host");
        copy_dense(s, s.pm().get(entity, version), entity, host, false);
        return CF_OK;
    });
}

cf_status cf_copy_in_device(cf_solver* h, int entity, int version, const double* dev)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!dev || !valid_entity(s.geo(), entity) || (version != 0 && version != 1))
            throw CfError(CF_EINVAL, "bad entity/version/device pointer");
        cfmi::launch_copy_padded(s.geo(), entity, s.pm().get(entity, version), dev, false,
                                 s.stream());
        CF_HIP(hipGetLastError());
        return CF_OK;
    });
}

cf_status cf_copy_out_device(cf_solver* h, int entity, int version, double* dev)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!dev || !valid_entity(s.geo(), entity) || (version != 0 && version != 1))
            throw CfError(CF_EINVAL, "bad entity/version/device pointer");
        cfmi::launch_copy_padded(s.geo(), entity, dev, s.pm().get(entity, version), false,
                                 s.stream());
        CF_HIP(hipGetLastError());
        return CF_OK;
    });
}

cf_status cf_copy_pressure_in(cf_solver* h, int which, const double* host)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!host || (which != 0 && which != 1))
            throw CfError(CF_EINVAL, "which must be 0 (lhs) or 1 (rhs)");
        copy_dense(s, which == 0 ? s.vc().lhs() : s.vc().rhs(), CF_CELL,
                   (void*)host, true);
        return CF_OK;
    });
}

cf_status cf_copy_pressure_out(cf_solver* h, int which, double* host)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!host || (which != 0 && which != 1))
            throw CfError(CF_EINVAL, "which must be 0 (lhs) or 1 (rhs)");
        copy_dense(s, which == 0 ? s.vc().lhs() : s.vc().rhs(), CF_CELL, host,
                   false);
        return CF_OK;
    });
}

cf_status cf_advance(cf_solver* h, int entity)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!valid_entity(s.geo(), entity))
            throw CfError(CF_EINVAL, "bad entity");
        s.pm().advance(entity);
        return CF_OK;
    });
}

cf_status cf_fill_ghosts(cf_solver* h, int version)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (version != 0 && version != 1)
            throw CfError(CF_EINVAL, "bad version");
        s.pm().gather(version);
        return CF_OK;
    });
}

cf_status cf_sync(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        CF_HIP(hipStreamSynchronize(s.stream()));
        return CF_OK;
    });
}

cf_status cf_init_vortex_tubes(cf_solver* h, uint64_t seed, double peak)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.initVortexTubes(seed, peak);
        return CF_OK;
    });
}

cf_status cf_init_blobs(cf_solver* h, uint64_t seed, int nblobs)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.initBlobs(seed, nblobs);
        return CF_OK;
    });
}

cf_status cf_init_random_faces(cf_solver* h, uint64_t seed)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.initRandomFaces(seed);
        return CF_OK;
    });
}

cf_status cf_set_kernel_timing(cf_solver* h, int enable)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.timer().enabled = enable != 0;
        s.timer().sample = enable > 1 ? enable : 1;
        return CF_OK;
    });
}

cf_status cf_kernel_time(cf_solver* h, int kernel, double* total_ms,
                         int64_t* launches)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (kernel < 0 || kernel >= cfmi::TK_COUNT)
            throw CfError(CF_EINVAL, "bad kernel id");
        CF_HIP(hipStreamSynchronize(s.stream()));
        s.timer().resolve(INT32_MAX);
        if (total_ms)
            *total_ms = s.timer().total_ms[kernel];
        if (launches)
            *launches = s.timer().launches[kernel];
        return CF_OK;
    });
}

cf_status cf_reset_kernel_time(cf_solver* h)
{
    return guarded(h, [&](cfmi::Solver& s) {
        s.timer().reset();
        return CF_OK;
    });
}

cf_status cf_adjoint_run(cf_solver* h, int nsteps, int ckpt,
                         const double* dq_T_host, double* du0_host,
                         double* dv0_host, double* dw0_host, double* dq0_host,
                         double dvel[3])
{
    return guarded(h, [&](cfmi::Solver& s) -> cf_status {
        if (!dq_T_host || !dvel)
            throw CfError(CF_EINVAL, "dq_T and dvel are required");
        const cfmi::Geo& g = s.geo();
        // the loss is defined on the owned cells: adjointRun ignores the
        // other entries of dq_T
        cfmi::DevArray<double> dq_dev, gr[4];
        dq_dev.alloc(g.total, s.stream());
        copy_dense(s, dq_dev.ptr, CF_CELL, (void*)dq_T_host, true);
        double* host_out[4] = {dq0_host, du0_host, dv0_host, dw0_host};
        double* dev_out[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int e = 0; e <= g.dim; ++e)
            if (host_out[e]) {
                gr[e].alloc(g.total, s.stream());
                dev_out[e] = gr[e].ptr;
            }
        cf_status rc = s.adjointRun(nsteps, ckpt, dq_dev.ptr, dev_out, dvel);
        if (rc != CF_OK)
            return rc;
        for (int e = 0; e <= g.dim; ++e)
            if (host_out[e])
                copy_dense(s, gr[e].ptr, e, host_out[e], false);
        return CF_OK;
    });
}

cf_status cf_adjoint_run_device(cf_solver* h, int nsteps, int ckpt,
                                const double* dq_T_dev, double* du0_dev,
                                double* dv0_dev, double* dw0_dev, double* dq0_dev,
                                double dvel[3])
{
    return guarded(h, [&](cfmi::Solver& s) -> cf_status {
        if (!dq_T_dev || !dvel)
            throw CfError(CF_EINVAL, "dq_T and dvel are required");
        double* grads[4] = {dq0_dev, du0_dev, dv0_dev, s.geo().dim == 3 ? dw0_dev : nullptr};
        return s.adjointRun(nsteps, ckpt, dq_T_dev, grads, dvel);
    });
}

cf_status cf_cell_output(cf_solver* h, double* q_host, double* vel_host)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!q_host || !vel_host)
            throw CfError(CF_EINVAL, "q_host and vel_host must be non-null");
        s.cellOutput(q_host, vel_host);
        return CF_OK;
    });
}

cf_status cf_write_output(cf_solver* h, const char* dir, int step)
{
    return guarded(h, [&](cfmi::Solver& s) {
        if (!dir || step < 0)
            throw CfError(CF_EINVAL, "dir must be non-null and step >= 0");
        cfmi::write_output(s, dir, step);
        return CF_OK;
    });
}

// No handle: errors are reported through cf_last_error(NULL).
cf_status cf_pressure_solve_noalloc(const cf_config* cfg, const double* rhs_dev,
                                    double* lhs_dev, void* stream, cf_solve_stats* st)
{
    try {
        if (!cfg)
            throw CfError(CF_EINVAL, "null config");
        return cfmi::pressure_solve_noalloc(*cfg, rhs_dev, lhs_dev,
                                            static_cast<hipStream_t>(stream), st);
    } catch (const CfError& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return CF_EDEVICE;
    }
}

cf_status cf_noalloc_limits(size_t* scratch_bytes, int64_t* max_elems)
{
    if (scratch_bytes)
        *scratch_bytes = cfmi::noalloc_scratch_bytes();
    if (max_elems)
        *max_elems = cfmi::noalloc_max_elems();
    return CF_OK;
}

// No handle: errors are reported through cf_last_error(NULL).
cf_status cf_write_vti(const char* path, int dim, const int n[3], const double origin[3],
                       double dx, int cycle, double time, double dt, const double* q,
                       const double* vel)
{
    try {
        if (!path || (dim != 2 && dim != 3) || !n || !origin || !q || !vel)
            throw CfError(CF_EINVAL, "cf_write_vti: bad arguments");
        for (int d = 0; d < dim; ++d)
            if (n[d] <= 0)
                throw CfError(CF_EINVAL, "cf_write_vti: n must be positive");
        cfmi::write_vti(path, dim, n, origin, dx, cycle, time, dt, q, vel);
        return CF_OK;
    } catch (const CfError& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return CF_EINVAL;
    }
}

} // extern "C"
