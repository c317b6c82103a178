// cf_solver.cpp -- host-side classes (Mesh, ProblemManager,
// VelocityCorrector, TimeIntegrator, Solver) driving the HIP kernels.
#include "cf_solver.hpp"

#include <climits>
#include <cmath>
#include <cstring>

#include "cf_init.hpp"

namespace cfmi
{

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw CfError(e == hipErrorOutOfMemory ? CF_ENOMEM : CF_EDEVICE,
                      std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void DevArray<T>::alloc(int64_t count, hipStream_t s)
{
    release();
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw CfError(CF_ENOMEM, "hipMalloc of " +
                                     std::to_string((size_t)count * sizeof(T)) +
                                     " bytes failed");
    }
    ptr = (T*)p;
    n = count;
    CF_HIP(hipMemsetAsync(ptr, 0, (size_t)count * sizeof(T), s));
}

template struct DevArray<double>;
template struct DevArray<float>;
template struct DevArray<unsigned>;
template struct DevArray<unsigned long long>;

// ---- KernelTimer -----------------------------------------------------------
KernelTimer::~KernelTimer()
{
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_)
        (void)hipEventDestroy(e);
}

hipEvent_t KernelTimer::get()
{
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    CF_HIP(hipEventCreate(&e));
    return e;
}

int KernelTimer::begin(hipStream_t s)
{
    if (!enabled)
        return -1;
    Pending p;
    p.a = get();
    p.b = nullptr;
    p.kernel = -1;
    p.iter = 0;
    CF_HIP(hipEventRecord(p.a, s));
    pending_.push_back(p);
    return (int)pending_.size() - 1;
}

void KernelTimer::end(int slot, hipStream_t s, int kernel, int iter)
{
    if (slot < 0)
        return;
    Pending& p = pending_[slot];
    p.b = get();
    p.kernel = kernel;
    p.iter = iter;
    CF_HIP(hipEventRecord(p.b, s));
}

void KernelTimer::resolve(int limit)
{
    for (auto& p : pending_) {
        if (p.b && p.iter < limit) {
            float ms = 0.f;
            CF_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            total_ms[p.kernel] += ms;
            launches[p.kernel] += 1;
        }
        pool_.push_back(p.a);
        if (p.b)
            pool_.push_back(p.b);
    }
    pending_.clear();
}

void KernelTimer::reset()
{
    for (int k = 0; k < TK_COUNT; ++k) {
        total_ms[k] = 0.0;
        launches[k] = 0;
    }
}

// ---- Mesh --------------------------------------------------------------------
Mesh::Mesh(const cf_config& c)
{
    const char* why = "";
    cf_status rc = build_geometry(c, _geo, &why);
    if (rc != CF_OK)
        throw CfError(rc, why);
}

// ---- ProblemManager ------------------------------------------------------------
ProblemManager::ProblemManager(const Mesh& mesh, hipStream_t s)
    : _mesh(mesh), _s(s)
{
    const Geo& g = mesh.geo();
    // ProblemManager.hpp:136-165: every array zeroed over its ghosted extent.
    for (int e = 0; e <= g.dim; ++e)
        for (int v = 0; v < 2; ++v) {
            _f[e][v].reset(new DevArray<double>());
            _f[e][v]->alloc(g.total, s);
        }
}

Fields ProblemManager::fields(int version) const
{
    Fields f;
    const int D = _mesh.geo().dim;
    for (int e = 0; e < 4; ++e)
        f.e[e] = e <= D ? get(e, version) : nullptr;
    return f;
}

void ProblemManager::gather(int version)
{
    // ProblemManager.hpp:394-403 on one rank: periodic axes copy, walls keep
    // their zero ghosts.
    const Geo& g = _mesh.geo();
    if (!(g.periodic[0] || g.periodic[1] || g.periodic[2]))
        return;
    for (int e = 0; e <= g.dim; ++e)
        launch_ghost_fill(g, e, get(e, version), _s);
}

// ---- TimeIntegrator ------------------------------------------------------------
namespace TimeIntegrator
{
void step(ProblemManager& pm, hipStream_t s, KernelTimer& timer)
{
    TraceRange tr("TimeIntegrator::Step");
    const Geo& g = pm.mesh().geo();
    {
        TraceRange t1("TimeIntegrator::Step::Gather");
        pm.gather(CF_CURRENT);
    }
    {
        // one launch advects every entity (the reference's
        // ::Advect::Quantity / ::Velocity::FaceI/J/K regions)
        TraceRange t2("TimeIntegrator::Step::Advect");
        int slot = timer.begin(s);
        launch_advect(g, pm.fields(CF_CURRENT), pm.fields(CF_NEXT), s);
        timer.end(slot, s, TK_ADVECT, 0);
    }
    TraceRange t3("CajitaFluids::TimeIntegrator::Advance");
    for (int e = 0; e <= g.dim; ++e)
        pm.advance(e);
}
} // namespace TimeIntegrator

// ---- VelocityCorrector -----------------------------------------------------------
VelocityCorrector::VelocityCorrector(ProblemManager& pm, const cf_config& cfg,
                                     hipStream_t s, KernelTimer& timer)
    : _pm(pm), _cfg(cfg), _s(s), _timer(timer), _g(pm.mesh().geo())
{
    const int64_t n = _g.total;
    _g_unit = _g;
    for (int k = 0; k <= 6; ++k) {
        _g_unit.dinv[k] = 1.0;
        _g_unit.dinv32[k] = 1.0f;
    }
    // VelocityCorrector.hpp:83-90: lhs / rhs zeroed over the ghosted extent.
    _lhs.alloc(n, s);
    _rhs.alloc(n, s);
    _r.alloc(n, s);
    _p[0].alloc(n, s);
    _p[1].alloc(n, s);
    _red_partials.alloc((int64_t)reduction_blocks_max(_g) * 3 * 2, s);
    _red_counter.alloc(kCounterSlots * kCounterStride, s);
    _scratch.alloc(8, s);
    void* p = nullptr;
    CF_HIP(hipMalloc(&p, sizeof(PcgState)));
    _st = (PcgState*)p;
    CF_HIP(hipMalloc(&p, sizeof(RefineState)));
    _rs = (RefineState*)p;
    CF_HIP(hipHostMalloc(&p, sizeof(PcgState), hipHostMallocDefault));
    _st_host = (PcgState*)p;
    CF_HIP(hipHostMalloc(&p, sizeof(RefineState), hipHostMallocDefault));
    _rs_host = (RefineState*)p;
    CF_HIP(hipMemsetAsync(_st, 0, sizeof(PcgState), s));
    CF_HIP(hipMemsetAsync(_rs, 0, sizeof(RefineState), s));
    if (cfg.precision == CF_FP32_REFINED)
        ensure32();
}

VelocityCorrector::~VelocityCorrector()
{
    if (_st)
        (void)hipFree(_st);
    if (_rs)
        (void)hipFree(_rs);
    if (_st_host)
        (void)hipHostFree(_st_host);
    if (_rs_host)
        (void)hipHostFree(_rs_host);
}

void VelocityCorrector::ensure32()
{
    if (_d32.ptr)
        return;
    const int64_t n = _g.total;
    _d32.alloc(n, _s);
    _r32.alloc(n, _s);
    _p32[0].alloc(n, _s);
    _p32[1].alloc(n, _s);
}

void VelocityCorrector::readState(PcgState& host)
{
    CF_HIP(hipMemcpyAsync(_st_host, _st, sizeof(PcgState),
                          hipMemcpyDeviceToHost, _s));
    CF_HIP(hipStreamSynchronize(_s));
    host = *_st_host;
}

void VelocityCorrector::buildRHS()
{
    // VelocityCorrector.hpp:182-212.
    TraceRange tr("VelocityCorrector::BuildProblem:buildRHS");
    {
        TraceRange t1("VelocityCorrector::BuildProblem:buildRHS:Gather");
        _pm.gather(CF_CURRENT);
    }
    launch_divergence(_g, _pm.fields(CF_CURRENT), _rhs.ptr, _s);
}

void VelocityCorrector::removeMean()
{
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_rhs_sum(_g, _rhs.ptr, _scratch.ptr, rb, _s);
    launch_rhs_shift(_g, _rhs.ptr, _scratch.ptr, _s);
}

void VelocityCorrector::applyPressure()
{
    // VelocityCorrector.hpp:214-264: pressure halo, then the face updates.
    TraceRange tr("VelocityCorrector::ApplyPressure");
    if (_g.periodic[0] || _g.periodic[1] || _g.periodic[2]) {
        TraceRange t1("VelocityCorrector::ApplyPressure:Gather");
        launch_ghost_fill(_g, CF_CELL, _lhs.ptr, _s);
    }
    launch_project(_g, _pm.fields(CF_CURRENT), _lhs.ptr, _s);
}

// The update's x mode for the n-th iteration (from 0) of a run: with pairs,
// X_SKIP then X_PAIR, so that every run starts and ends on a whole x.
static int xmode(bool pairs, int n)
{
    return !pairs ? X_EACH : (n & 1) ? X_PAIR : X_SKIP;
}

// Queue batches of iterations without waiting; poll the device flag once
// per batch.  The first batch is sized from a prediction (the previous
// solve's, or pass's, count) so that few no-op launches follow the stop.
// p[cur] holds the current direction; the kernels write the next one into
// p[cur ^ 1].  Returns the device's iteration counter.
template <class T>
int VelocityCorrector::runIterations(T* x, T* r, DevArray<T> (&p)[2], int& cur,
                                     int predict, int tk0)
{
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    PcgState hs;
    readState(hs);
    if (hs.done)
        return hs.iter;
    const int it0 = hs.iter;
    const bool pairs = pcg_pairs_x<T>(pcgGeo());
    int launched = 0;
    int batch = predict > 4 ? predict - 2 : 16;
    for (;;) {
        for (int c = 0; c < batch; ++c) {
            const int it = it0 + launched + c + 1;
            int t0 = _timer.begin_iter(_s, it);
            launch_pcg_dir_spmv<T>(pcgGeo(), r, p[cur].ptr, p[cur ^ 1].ptr, _st, rb, _s);
            _timer.end(t0, _s, tk0 + 0, it);
            int t2 = _timer.begin_iter(_s, it);
            launch_pcg_update<T>(pcgGeo(), x, r, p[cur ^ 1].ptr, _st, rb, _s,
                                 xmode(pairs, launched + c), p[cur].ptr);
            _timer.end(t2, _s, tk0 + 1, it);
            cur ^= 1;
        }
        launched += batch;
        readState(hs);
        _timer.resolve(hs.done ? hs.iter + 1 : INT_MAX);
        if (hs.done)
            break;
        batch = 4;
    }
    // The no-op launches after the stop still flipped `cur`: recover the
    // buffer that holds the last real direction.
    const int real = hs.iter - it0;
    cur ^= (launched - real) & 1;
    if (pairs && (real & 1)) // the last real update was an X_SKIP
        launch_pcg_xfix<T>(pcgGeo(), x, p[cur].ptr, _st, _s);
    return hs.iter;
}

int VelocityCorrector::runPcg64(const double* b, double* x, double tol,
                                int maxit)
{
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_pcg_init64(pcgGeo(), b, x, _r.ptr, _st, tol, maxit, rb, _s);
    int cur = 0;
    int its = !_coop_off && pcg_coop_fits(pcgGeo()) ? runCoop(x) : -1;
    // the streaming kernels, or where a cooperative launch was refused, on
    // from the state it left (p[0] holds the current direction)
    if (its < 0)
        its = runIterations<double>(x, _r.ptr, _p, cur, _predict64, TK_DIR_SPMV);
    _predict64 = its;
    return its;
}

// Iterations of a launch of the cooperative PCG: the first covers the
// previous solve's count with a margin (at least kCoopItsPerLaunch); the
// host reads the state after each launch, and a solve that has not stopped
// goes on in launches of kCoopItsPerLaunch from the stored x, r, direction
// and state.  Returns -1 (and turns the path off for the handle) if the
// runtime refuses a cooperative launch.
constexpr int kCoopItsPerLaunch = 256;

int VelocityCorrector::runCoop(double* x)
{
    if (!_xch.ptr)
        _xch.alloc((int64_t)(coop_xch_bytes(pcgGeo()) + sizeof(double) - 1) / sizeof(double), _s);
    PcgState hs;
    int its = _predict64 + 64 > kCoopItsPerLaunch ? _predict64 + 64 : kCoopItsPerLaunch;
    for (;; its = kCoopItsPerLaunch) {
        const int t0 = _timer.begin(_s);
        const bool ran = launch_pcg_solve_coop(pcgGeo(), x, _r.ptr, _p[0].ptr, _st, _xch.ptr,
                                               its, _s);
        _timer.end(t0, _s, TK_PCG_COOP, 0);
        if (!ran) {
            _timer.resolve(INT_MAX);
            _coop_off = true;
            return -1;
        }
        readState(hs);
        _timer.resolve(INT_MAX);
        if (hs.stalled)
            throw CfError(CF_EDEVICE, "cooperative PCG: a workgroup did not arrive");
        if (hs.done)
            return hs.iter;
    }
}

cf_status VelocityCorrector::solve64(const double* b, double* x,
                                     cf_solve_stats* st)
{
    const int its = runPcg64(b, x, _cfg.cg_tol, _cfg.cg_max_iter);
    PcgState hs = *_st_host;
    if (st) {
        st->converged = hs.converged;
        st->iterations = its;
        st->outer_iterations = 0;
        st->b_norm = hs.bnorm;
        st->final_residual = hs.res;
    }
    if (hs.nonfinite)
        return CF_ENONFINITE;
    return hs.converged ? CF_OK : CF_ENOCONV;
}

void VelocityCorrector::readRefine()
{
    CF_HIP(hipMemcpyAsync(_rs_host, _rs, sizeof(RefineState),
                          hipMemcpyDeviceToHost, _s));
    CF_HIP(hipStreamSynchronize(_s));
}

int& VelocityCorrector::passPrediction(int pass)
{
    if ((int)_predict_pass.size() <= pass)
        _predict_pass.resize(pass + 1, 0);
    return _predict_pass[pass];
}

// Restart mode: every pass runs a fresh FP32 PCG on the FP32 residual.
cf_status VelocityCorrector::solveRestart(const double* b, double* x,
                                          cf_solve_stats& loc)
{
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_refine_init(pcgGeo(), b, x, _r32.ptr, _rs, rb, _s);
    readRefine();
    loc.b_norm = _rs_host->bnorm;
    if (!std::isfinite(_rs_host->bnorm)) {
        loc.final_residual = _rs_host->bnorm;
        return CF_ENONFINITE;
    }
    if (_rs_host->bnorm == 0.0) {
        loc.converged = 1;
        loc.final_residual = 0.0;
        return CF_OK;
    }
    const int inner_max =
        _cfg.inner_max_iter > 0 ? _cfg.inner_max_iter : _cfg.cg_max_iter;
    for (;;) {
        loc.final_residual = _rs_host->res;
        if (_rs_host->res < _cfg.cg_tol) {
            loc.converged = 1;
            return CF_OK;
        }
        if (loc.outer_iterations >= _cfg.max_outer ||
            loc.iterations >= _cfg.cg_max_iter)
            return CF_ENOCONV;
        launch_pcg_init32(pcgGeo(), _r32.ptr, _d32.ptr, _st, _rs, _cfg.inner_tol,
                          _cfg.cg_tol, inner_max, rb, _s);
        int cur = 0;
        int& pred = passPrediction(loc.outer_iterations);
        const int its =
            runIterations<float>(_d32.ptr, _r32.ptr, _p32, cur, pred, TK_DIR_SPMV32);
        pred = its;
        loc.iterations += its;
        if (_st_host->nonfinite)
            return CF_ENONFINITE;
        launch_refine_add(pcgGeo(), x, _d32.ptr, _s);
        int t = _timer.begin(_s);
        launch_refine_residual(pcgGeo(), b, x, _r32.ptr, _rs, rb, _s);
        _timer.end(t, _s, TK_RESIDUAL, 0);
        readRefine();
        _timer.resolve(INT_MAX);
        loc.outer_iterations += 1;
        if (!std::isfinite(_rs_host->res)) {
            loc.final_residual = _rs_host->res;
            return CF_ENONFINITE;
        }
    }
}

// Carry mode: one FP32 PCG run whose residual is replaced by the FP64 one at
// each pass boundary, keeping the search direction.
cf_status VelocityCorrector::solveCarry(const double* b, double* x,
                                        cf_solve_stats& loc)
{
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_refine_init_carry(pcgGeo(), b, x, _r32.ptr, _d32.ptr, _rs, _st,
                             _cfg.inner_tol, _cfg.cg_tol, _cfg.cg_max_iter, rb, _s);
    readRefine();
    loc.b_norm = _rs_host->bnorm;
    loc.final_residual = _rs_host->res;
    if (!std::isfinite(_rs_host->bnorm)) {
        loc.final_residual = _rs_host->bnorm;
        return CF_ENONFINITE;
    }
    if (_rs_host->bnorm == 0.0) {
        loc.converged = 1;
        return CF_OK;
    }
    int cur = 0;
    int done_its = 0;
    for (;;) {
        int& pred = passPrediction(loc.outer_iterations);
        const int its =
            runIterations<float>(_d32.ptr, _r32.ptr, _p32, cur, pred, TK_DIR_SPMV32);
        pred = its - done_its;
        done_its = its;
        loc.iterations = its;
        if (_st_host->nonfinite)
            return CF_ENONFINITE;
        if (!_st_host->converged) // iteration cap reached inside the pass
            return CF_ENOCONV;
        launch_refine_add(pcgGeo(), x, _d32.ptr, _s);
        int t = _timer.begin(_s);
        launch_refine_residual_carry(pcgGeo(), b, x, _r32.ptr, _rs, _st, _cfg.inner_tol,
                                     _cfg.cg_tol, rb, _s);
        _timer.end(t, _s, TK_RESIDUAL, 0);
        readRefine();
        _timer.resolve(INT_MAX);
        loc.outer_iterations += 1;
        loc.final_residual = _rs_host->res;
        if (!std::isfinite(_rs_host->res))
            return CF_ENONFINITE;
        if (_rs_host->res < _cfg.cg_tol) {
            loc.converged = 1;
            return CF_OK;
        }
        if (loc.outer_iterations >= _cfg.max_outer)
            return CF_ENOCONV;
    }
}

cf_status VelocityCorrector::solveRefined(const double* b, double* x,
                                          cf_solve_stats* st)
{
    ensure32();
    cf_solve_stats loc;
    std::memset(&loc, 0, sizeof(loc));
    cf_status rc = _cfg.refine_mode == CF_REFINE_CARRY ? solveCarry(b, x, loc)
                                                       : solveRestart(b, x, loc);
    if (st)
        *st = loc;
    return rc;
}

cf_status VelocityCorrector::solve(int precision, const double* b, double* x,
                                   cf_solve_stats* st)
{
    if (precision == CF_FP32_REFINED && _cfg.preconditioner == CF_PRECOND_MG)
        return solveMg(b, x, st, true);
    if (precision == CF_FP32_REFINED) {
        const cf_status rc = solveRefined(b, x, st);
        zeroMeanIfUnpreconditioned(x);
        return rc;
    }
    if (precision != CF_FP64)
        throw CfError(CF_EINVAL, "precision must be CF_FP64 or CF_FP32_REFINED");
    if (_cfg.preconditioner == CF_PRECOND_MG)
        return solveMg(b, x, st, false);
    const cf_status rc = solve64(b, x, st);
    zeroMeanIfUnpreconditioned(x);
    return rc;
}

// Without a preconditioner, CG from x = 0 stays in the range of A: its map
// is A^+ and the solution has zero mean up to rounding.  On a singular
// operator the mean is removed explicitly, as after the multigrid solve, so
// that the map's transpose is exactly a mean removal and the same solve.
void VelocityCorrector::zeroMeanIfUnpreconditioned(double* x)
{
    if (_cfg.preconditioner != CF_PRECOND_NONE || !operator_singular(_g))
        return;
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_rhs_sum(_g, x, _scratch.ptr + 1, rb, _s);
    launch_rhs_shift(_g, x, _scratch.ptr + 1, _s);
}

// MG-PCG: Jacobi-PCG's kernels with z = B r from one V-cycle (with f32, a
// V-cycle whose level 0 runs on FP32 vectors; every CG vector, dot and
// scalar stays FP64, so the solve reaches the FP64 tolerance).  The
// direction kernel runs on a geometry whose Jacobi table is 1, so it forms
// p = fma(beta, p_old, z) (z alone on the first iteration); the update
// kernel leaves beta and rho to the V-cycle's last sweep, which reduces r.z.
// Iterations are queued in batches like runIterations; every kernel of an
// iteration queued after convergence returns at once.
cf_status VelocityCorrector::solveMg(const double* b, double* x, cf_solve_stats* st,
                                     bool f32)
{
    if (!_mg)
        _mg.reset(new Multigrid(_g, _cfg, _s));
    if (f32 && !_mg->f32_supported())
        throw CfError(CF_ENOTIMPL, "the FP32 multigrid V-cycle needs a 3-D grid with more "
                                   "than 2 cells along every axis");
    RedBuf rb{_red_partials.ptr, _red_counter.ptr};
    launch_mgpcg_init(_g, b, x, _r.ptr, _st, _cfg.cg_tol, _cfg.cg_max_iter, rb, _s);
    // the V-cycle's z: FP64, or FP32 from the FP32 finest level
    const double* z = nullptr;
    const float* z32 = nullptr;
    if (f32)
        z32 = _mg->vcycle32(_r.ptr, _st, rb, false);
    else
        z = _mg->vcycle(_r.ptr, _st, rb);
    PcgState hs;
    readState(hs);
    int cur = 0, launched = 0;
    int batch = _predict_mg > 3 ? _predict_mg - 2 : 8;
    const int it0 = hs.iter;
    const bool pairs = pcg_pairs_x<double>(_g);
    while (!hs.done) {
        for (int c = 0; c < batch; ++c) {
            const int it = it0 + launched + c + 1;
            int t0 = _timer.begin_iter(_s, it);
            if (f32)
                launch_pcg_dir_spmv_z32(_g_unit, z32, _p[cur].ptr, _p[cur ^ 1].ptr, _st, rb,
                                        _s);
            else
                launch_pcg_dir_spmv<double>(_g_unit, z, _p[cur].ptr, _p[cur ^ 1].ptr, _st, rb,
                                            _s);
            _timer.end(t0, _s, TK_DIR_SPMV, it);
            int t1 = _timer.begin_iter(_s, it);
            // with the FP32 V-cycle the update also writes its right-hand
            // side, (float) r
            launch_pcg_update<double>(_g, x, _r.ptr, _p[cur ^ 1].ptr, _st, rb, _s,
                                      xmode(pairs, launched + c), _p[cur].ptr,
                                      f32 ? _mg->b32() : nullptr);
            _timer.end(t1, _s, TK_UPDATE, it);
            // the V-cycle prepares iteration it + 1 (a no-op after the stop)
            int t2 = _timer.begin_iter(_s, it + 1);
            if (f32)
                z32 = _mg->vcycle32(_r.ptr, _st, rb, true);
            else
                z = _mg->vcycle(_r.ptr, _st, rb);
            _timer.end(t2, _s, TK_VCYCLE, it + 1);
            cur ^= 1;
        }
        launched += batch;
        readState(hs);
        _timer.resolve(hs.done ? hs.iter + 1 : INT_MAX);
        batch = 2;
    }
    const int real = hs.iter - it0;
    if (pairs && (real & 1)) { // the last real update was an X_SKIP
        cur ^= (launched - real) & 1;
        launch_pcg_xfix<double>(_g, x, _p[cur].ptr, _st, _s);
    }
    // On a singular operator the pressure is fixed up to a constant, which
    // MG-PCG leaves at a level of the V-cycle's (only the as-written FaceJ
    // rule lets it reach the velocities).  Return the zero-mean pressure:
    // that map's transpose is a mean removal and the same solve, which the
    // reverse mode applies.
    if (operator_singular(_g)) {
        launch_rhs_sum(_g, x, _scratch.ptr + 1, rb, _s);
        launch_rhs_shift(_g, x, _scratch.ptr + 1, _s);
    }
    _predict_mg = hs.iter;
    if (st) {
        st->converged = hs.converged;
        st->iterations = hs.iter;
        st->outer_iterations = 0;
        st->b_norm = hs.bnorm;
        st->final_residual = hs.res;
    }
    if (hs.nonfinite)
        return CF_ENONFINITE;
    return hs.converged ? CF_OK : CF_ENOCONV;
}

cf_status VelocityCorrector::correctVelocity(cf_solve_stats* st)
{
    // VelocityCorrector.hpp:266-282; a solve that does not converge aborts
    // the correction before the projection (the reference's solver throws).
    TraceRange tr("VelocityCorrector");
    {
        TraceRange t1("VelocityCorrector::BuildProblem");
        buildRHS();
    }
    cf_status rc;
    {
        // the reference's region name, spelling included
        TraceRange t2("VelocityCorrector::PresureSolve");
        rc = solve(_cfg.precision, _rhs.ptr, _lhs.ptr, st);
    }
    if (rc != CF_OK)
        return rc;
    applyPressure();
    return CF_OK;
}

// ---- Solver ----------------------------------------------------------------------
Solver::Solver(const cf_config& cfg) : _cfg(cfg)
{
    _mesh.reset(new Mesh(cfg));
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        throw CfError(CF_EDEVICE, "no HIP device available");
    }
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CF_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw CfError(CF_EDEVICE, std::string("device is ") + prop.gcnArchName +
                                      "; this library is built for gfx950");
    CF_HIP(hipStreamCreateWithFlags(&_s, hipStreamNonBlocking));
    _pm.reset(new ProblemManager(*_mesh, _s));
    _vc.reset(new VelocityCorrector(*_pm, _cfg, _s, _timer));
    _vmax.alloc(1, _s);
    CF_HIP(hipStreamSynchronize(_s));
}

Solver::~Solver()
{
    if (_s)
        (void)hipStreamSynchronize(_s);
    _vc.reset();
    _pm.reset();
    _vmax.release();
    if (_s)
        (void)hipStreamDestroy(_s);
}

void Solver::timeIntegrate() { TimeIntegrator::step(*_pm, _s, _timer); }

void Solver::addInputs()
{
    // one launch for every entity (the reference's Solve::AddInputs::Cell /
    // ::FaceI / ::FaceJ regions)
    TraceRange tr("Solve::AddInputs");
    launch_inputs(_mesh->geo(), _pm->fields(CF_CURRENT), _s);
}

cf_status Solver::setup(cf_solve_stats* st)
{
    // Solver.hpp:125-133.
    addInputs();
    return _vc->correctVelocity(st);
}

cf_status Solver::step(cf_solve_stats* st)
{
    // Solver.hpp:135-147.
    timeIntegrate();
    addInputs();
    cf_status rc = _vc->correctVelocity(st);
    _time += _mesh->geo().dt;
    return rc;
}

cf_status Solver::run(int nsteps, cf_solve_stats* last)
{
    TraceRange tr("Solve");
    for (int n = 0; n < nsteps; ++n) {
        cf_status rc = step(last);
        if (rc != CF_OK)
            return rc;
    }
    return CF_OK;
}

void Solver::initVortexTubes(uint64_t seed, double peak)
{
    const Geo& g = _mesh->geo();
    InitParams P;
    std::memset(&P, 0, sizeof(P));
    for (int d = 0; d < 3; ++d) {
        P.periodic[d] = g.periodic[d];
        P.box_lo[d] = _cfg.lo[d];
        P.box_hi[d] = _cfg.hi[d];
    }
    make_tube_params(g, _cfg, seed, P);
    launch_init_tubes(g, P, _pm->fields(CF_CURRENT), _vmax.ptr, peak, _s);
}

void Solver::initBlobs(uint64_t seed, int nblobs)
{
    const Geo& g = _mesh->geo();
    if (nblobs < 0 || nblobs > kMaxBlobs)
        throw CfError(CF_EINVAL, "nblobs must be in [0, 32]");
    InitParams P;
    std::memset(&P, 0, sizeof(P));
    for (int d = 0; d < 3; ++d) {
        P.periodic[d] = g.periodic[d];
        P.box_lo[d] = _cfg.lo[d];
        P.box_hi[d] = _cfg.hi[d];
    }
    make_blob_params(g, _cfg, seed, nblobs, P);
    launch_init_blobs(g, P, _pm->get(CF_CELL, CF_CURRENT), _s);
}

void Solver::initRandomFaces(uint64_t seed)
{
    launch_init_random(_mesh->geo(), seed, _pm->fields(CF_CURRENT), _s);
}

void Solver::cellOutput(double* q_host, double* vel_host)
{
    const Geo& g = _mesh->geo();
    const int64_t N = (int64_t)g.n[0] * g.n[1] * g.n[2];
    DevArray<double> buf;
    buf.alloc((1 + g.dim) * N, _s);
    launch_cell_output(g, _pm->fields(CF_CURRENT), buf.ptr, buf.ptr + N, _s);
    CF_HIP(hipMemcpyAsync(q_host, buf.ptr, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, _s));
    CF_HIP(hipMemcpyAsync(vel_host, buf.ptr + N, (size_t)(g.dim * N) * sizeof(double),
                          hipMemcpyDeviceToHost, _s));
    CF_HIP(hipStreamSynchronize(_s));
}

} // namespace cfmi
