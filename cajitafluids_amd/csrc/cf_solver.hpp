// cf_solver.hpp -- host-side classes mirroring the reference's interfaces:
// Mesh (src/Mesh.hpp), ProblemManager (src/ProblemManager.hpp),
// VelocityCorrector (src/VelocityCorrector.hpp), TimeIntegrator
// (src/TimeIntegrator.hpp) and Solver (src/Solver.hpp).  The C ABI in
// cf_api.cpp is a thin shell over Solver.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cf_api.h"
#include "cf_geometry.hpp"
#include "cf_kernels.hpp"
#include "cf_mg.hpp"

namespace cfmi
{

struct CfError : std::runtime_error
{
    cf_status code;
    CfError(cf_status c, const std::string& m) : std::runtime_error(m), code(c)
    {
    }
};

void hip_check(hipError_t e, const char* what);
#define CF_HIP(call) ::cfmi::hip_check((call), #call)

// A named range on the host timeline (roctx; rocprofv3 --marker-trace shows
// the kernels dispatched inside it).  Ranges carry the names of the
// reference's Kokkos profiling regions around the same phases
// (Solver.hpp:154-241, TimeIntegrator.hpp:125-163,
// VelocityCorrector.hpp:184-275).
struct TraceRange
{
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

// One device allocation in the handle's padded layout.
template <class T>
struct DevArray
{
    T* ptr = nullptr;
    int64_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void alloc(int64_t count, hipStream_t s);
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
};

// Per-kernel device time from HIP events on the handle's stream.
enum TimedKernel
{
    TK_DIR_SPMV = 0,   // k_pcg_dir_spmv<double>
    TK_UPDATE = 1,     // k_pcg_update<double>
    TK_FINALIZE = 2,   // the two one-block finalize kernels of an iteration
    TK_DIR_SPMV32 = 3, // FP32 inner solve
    TK_UPDATE32 = 4,
    TK_FINALIZE32 = 5,
    TK_ADVECT = 6,
    TK_RESIDUAL = 7,   // FP64 defect-correction residual
    TK_VCYCLE = 8,     // one multigrid V-cycle
    TK_PCG_COOP = 9,   // one launch of the cooperative 2-D PCG
    TK_COUNT = 10
};

class KernelTimer
{
  public:
    ~KernelTimer();
    bool enabled = false;
    int sample = 1; // PCG kernels: time two consecutive iterations in `sample`
    // Record a start event; returns a slot for end().
    int begin(hipStream_t s);
    // begin() for a kernel of PCG iteration `it` (1-based; sampled).  The
    // samples come in pairs of iterations so that the update's X_SKIP and
    // X_PAIR variants, which alternate, are timed alike.
    int begin_iter(hipStream_t s, int it)
    {
        const int half = sample / 2 > 1 ? sample / 2 : 1;
        return (sample <= 1 || ((it - 1) / 2) % half == 0) ? begin(s) : -1;
    }
    void end(int slot, hipStream_t s, int kernel, int iter);
    // Resolve pending pairs (stream must be synchronised); pairs whose tag is
    // not below `limit` were no-op launches after convergence and are dropped.
    void resolve(int limit);
    double total_ms[TK_COUNT] = {};
    int64_t launches[TK_COUNT] = {};
    void reset();

  private:
    struct Pending
    {
        hipEvent_t a, b;
        int kernel, iter;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t get();
};

// Mesh: src/Mesh.hpp:32-145.
class Mesh
{
  public:
    explicit Mesh(const cf_config& c);
    const Geo& geo() const { return _geo; }
    double cellSize() const { return _geo.dx; }
    std::array<int, 3> minDomainGlobalCellIndex() const { return {0, 0, 0}; }
    std::array<int, 3> maxDomainGlobalCellIndex() const
    {
        return {_geo.n[0] - 1, _geo.n[1] - 1, _geo.n[2] - 1};
    }

  private:
    Geo _geo;
};

// ProblemManager: src/ProblemManager.hpp:97-418.  Current/Next versions of
// q, u, v (w) with zeroed ghosts; advance swaps; gather fills ghosts.
class ProblemManager
{
  public:
    ProblemManager(const Mesh& mesh, hipStream_t s);
    double* get(int entity, int version) const
    {
        return _f[entity][_cur[entity] ^ version]->ptr;
    }
    Fields fields(int version) const;
    void advance(int entity) { _cur[entity] ^= 1; }
    void gather(int version);
    const Mesh& mesh() const { return _mesh; }

  private:
    const Mesh& _mesh;
    hipStream_t _s;
    std::unique_ptr<DevArray<double>> _f[4][2];
    int _cur[4] = {0, 0, 0, 0};
};

// VelocityCorrector: src/VelocityCorrector.hpp:50-295 with the Jacobi-PCG of
// the "Reference" solver choice run as matrix-free HIP kernels.
class VelocityCorrector
{
  public:
    VelocityCorrector(ProblemManager& pm, const cf_config& cfg, hipStream_t s,
                      KernelTimer& timer);
    ~VelocityCorrector();
    cf_status correctVelocity(cf_solve_stats* st);
    void buildRHS();
    void removeMean();
    cf_status solve(int precision, const double* b, double* x,
                    cf_solve_stats* st);
    void applyPressure();
    double* lhs() const { return _lhs.ptr; }
    double* rhs() const { return _rhs.ptr; }
    RedBuf redBuf() const { return RedBuf{_red_partials.ptr, _red_counter.ptr}; }
    double* scratch() const { return _scratch.ptr; }

  private:
    cf_status solve64(const double* b, double* x, cf_solve_stats* st);
    cf_status solveRefined(const double* b, double* x, cf_solve_stats* st);
    cf_status solveRestart(const double* b, double* x, cf_solve_stats& loc);
    cf_status solveCarry(const double* b, double* x, cf_solve_stats& loc);
    // f32: the V-cycle's level 0 in FP32 (CF_FP32_REFINED with the multigrid
    // preconditioner); the CG around it stays FP64
    cf_status solveMg(const double* b, double* x, cf_solve_stats* st, bool f32);
    template <class T>
    int runIterations(T* x, T* r, DevArray<T> (&p)[2], int& cur, int predict,
                      int tk0);
    int runPcg64(const double* b, double* x, double tol, int maxit);
    // the whole FP64 solve by the cooperative 2-D kernel (pcg_coop_fits)
    int runCoop(double* x);
    void readState(PcgState& host);
    void readRefine();
    int& passPrediction(int pass);
    void ensure32();

    ProblemManager& _pm;
    cf_config _cfg;
    hipStream_t _s;
    KernelTimer& _timer;
    const Geo& _g;
    DevArray<double> _lhs, _rhs, _r, _p[2];
    DevArray<float> _d32, _r32, _p32[2];
    DevArray<double> _red_partials, _scratch;
    DevArray<double> _xch; // exchange buffer of the cooperative 2-D PCG
    bool _coop_off = false; // the runtime refused a cooperative launch
    DevArray<unsigned> _red_counter;
    PcgState* _st = nullptr;       // device
    RefineState* _rs = nullptr;    // device
    PcgState* _st_host = nullptr;  // pinned
    RefineState* _rs_host = nullptr;
    int _predict64 = 0;
    int _predict_mg = 0;
    std::unique_ptr<Multigrid> _mg; // built on the first MG solve
    Geo _g_unit;                    // _g with a unit Jacobi table: the
                                    // direction kernel then reads z as is
    // geometry of the Jacobi-PCG kernels: _g, or _g_unit without a
    // preconditioner (z = r)
    const Geo& pcgGeo() const
    {
        return _cfg.preconditioner == CF_PRECOND_NONE ? _g_unit : _g;
    }
    void zeroMeanIfUnpreconditioned(double* x);
    std::vector<int> _predict_pass; // FP32 iterations of each refinement pass
};

// Solver: src/Solver.hpp:54-279 (setup / step / _addInputs; the dt clamp is
// applied when the Mesh geometry is built).
class Solver
{
  public:
    explicit Solver(const cf_config& cfg);
    ~Solver();
    cf_status setup(cf_solve_stats* st);
    cf_status step(cf_solve_stats* st);
    cf_status run(int nsteps, cf_solve_stats* last);
    void timeIntegrate();
    void addInputs();

    const cf_config& config() const { return _cfg; }
    const Geo& geo() const { return _mesh->geo(); }
    hipStream_t stream() const { return _s; }
    ProblemManager& pm() { return *_pm; }
    VelocityCorrector& vc() { return *_vc; }
    KernelTimer& timer() { return _timer; }
    double time() const { return _time; }
    std::string last_error;

    // Reverse mode of setup() + nsteps step() from the state in Current
    // (cf_adjoint_run.cpp).  dq_T: dL/dq of the final state (device, padded
    // layout, ghost entries ignored); grad[e]: device arrays that receive
    // dL/d(initial entity e) (owned entries; ghosts zero); dvel: host.
    cf_status adjointRun(int nsteps, int ckpt, const double* dq_T,
                         double* const* grad, double dvel[3]);

    // The output writer's cell data into host arrays (cf_output.cpp).
    void cellOutput(double* q_host, double* vel_host);

    void initVortexTubes(uint64_t seed, double peak);
    void initBlobs(uint64_t seed, int nblobs);
    void initRandomFaces(uint64_t seed);

  private:
    cf_config _cfg;
    hipStream_t _s = nullptr;
    KernelTimer _timer;
    std::unique_ptr<Mesh> _mesh;
    std::unique_ptr<ProblemManager> _pm;
    std::unique_ptr<VelocityCorrector> _vc;
    DevArray<unsigned long long> _vmax;
    double _time = 0.0;
};

// The output writer (cf_output.cpp).
void write_vti(const std::string& path, int dim, const int n[3], const double origin[3], double dx,
               int cycle, double time, double dt, const double* q, const double* vel);
void write_output(Solver& s, const std::string& dir, int step);

// The no-allocation pressure solve (cf_noalloc.hip): FP64 Jacobi or
// unpreconditioned PCG on caller device arrays, every work array in one
// static device array per GPU.
cf_status pressure_solve_noalloc(const cf_config& cfg, const double* b, double* x,
                                 hipStream_t s, cf_solve_stats* st);
size_t noalloc_scratch_bytes();
int64_t noalloc_max_elems();

// TimeIntegrator::step (src/TimeIntegrator.hpp:120-177): gather, advect all
// entities into Next in one launch, swap.
namespace TimeIntegrator
{
void step(ProblemManager& pm, hipStream_t s, KernelTimer& timer);
}

} // namespace cfmi
