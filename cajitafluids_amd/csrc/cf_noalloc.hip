// cf_noalloc.hip -- the pressure solve for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): cf_pressure_solve_ext's FP64 PCG on the
// caller's rhs / lhs arrays with no handle, every work array in one static
// device array.
//
// `scratch` below is that array.  In host code its name is only a host-side
// stand-in: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that address is ever passed to a kernel.
//
// Layout of the scratch (bytes):
//   [0, 256)          PcgState
//   [256, 1024)       scalar sums (the zero-mean step of an unpreconditioned
//                     solve)
//   [1024, 3328)      the grid reductions' arrival counters (kCounterSlots
//                     lines of kCounterStride unsigned), cleared by a kernel
//                     at the start of every call
//   [4096, 256 KiB)   the grid reductions' per-block partials
//   [256 KiB, ...)    r, p[0], p[1]: the padded layout's `elems` doubles
//                     each, 256-byte aligned
// Every call on a GPU shares that GPU's scratch, which keeps the last call's
// contents; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "cf_solver.hpp"

namespace cfmi
{

constexpr size_t kScratchBytes = size_t(96) << 20;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace
{
constexpr size_t kStateOff = 0, kScalarOff = 256, kCounterOff = 1024, kPartialOff = 4096;
constexpr size_t kVectorOff = size_t(256) << 10;
static_assert(sizeof(PcgState) <= kScalarOff - kStateOff, "PcgState fits its slot");
static_assert(kCounterOff + sizeof(unsigned) * kCounterSlots * kCounterStride <= kPartialOff,
              "counters fit their slot");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_scratch[kMaxDevices]; // device address of each GPU's scratch
int g_predict[kMaxDevices];            // last solve's iteration count per GPU

// The current device's scratch address, taken on its first use.
unsigned char* scratch_address(int* dev_out)
{
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw CfError(CF_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *dev_out = dev;
    return g_scratch[dev];
}

__global__ __launch_bounds__(256) void k_zero_u32(unsigned* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0u;
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

} // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Largest padded-layout element count the scratch holds three vectors of.
int64_t noalloc_max_elems()
{
    return (int64_t)((kScratchBytes - kVectorOff) / 3 / sizeof(double)) / 32 * 32;
}

cf_status pressure_solve_noalloc(const cf_config& cfg, const double* b, double* x,
                                 hipStream_t s, cf_solve_stats* st)
{
    Geo g;
    const char* why = "";
    cf_status rc = build_geometry(cfg, g, &why);
    if (rc != CF_OK)
        throw CfError(rc, why);
    if (cfg.precision != CF_FP64 || cfg.preconditioner == CF_PRECOND_MG)
        throw CfError(CF_ENOTIMPL, "the no-allocation solve runs FP64 Jacobi or "
                                   "unpreconditioned PCG");
    if (!b || !x || b == x)
        throw CfError(CF_EINVAL, "rhs and lhs must be distinct device arrays");
    if (g.total > noalloc_max_elems())
        throw CfError(CF_EINVAL, "grid exceeds the no-allocation scratch: " +
                                     std::to_string(g.total) + " padded elements, at most " +
                                     std::to_string(noalloc_max_elems()));
    const int nblk = reduction_blocks_max(g);
    if ((size_t)nblk * 3 * 2 * sizeof(double) > kVectorOff - kPartialOff)
        throw CfError(CF_EINVAL, "grid exceeds the no-allocation reduction buffer");

    int dev = 0;
    unsigned char* base = scratch_address(&dev);
    PcgState* state = reinterpret_cast<PcgState*>(base + kStateOff);
    double* scal = reinterpret_cast<double*>(base + kScalarOff);
    RedBuf rb{reinterpret_cast<double*>(base + kPartialOff),
              reinterpret_cast<unsigned*>(base + kCounterOff)};
    const size_t vb = round256((size_t)g.total * sizeof(double));
    double* r = reinterpret_cast<double*>(base + kVectorOff);
    double* p[2] = {reinterpret_cast<double*>(base + kVectorOff + vb),
                    reinterpret_cast<double*>(base + kVectorOff + 2 * vb)};

    Geo gp = g; // the PCG's geometry: a unit Jacobi table without a preconditioner
    if (cfg.preconditioner == CF_PRECOND_NONE)
        for (int k = 0; k <= 6; ++k) {
            gp.dinv[k] = 1.0;
            gp.dinv32[k] = 1.0f;
        }

    // the counters start at zero (a kernel, not a memset: callers may
    // capture this stream)
    hipLaunchKernelGGL(k_zero_u32, dim3(3), dim3(256), 0, s, rb.counter,
                       kCounterSlots * kCounterStride);
    launch_pcg_init64(gp, b, x, r, state, cfg.cg_tol, cfg.cg_max_iter, rb, s);
    PcgState hs;
    auto read_state = [&]() {
        CF_HIP(hipMemcpyAsync(&hs, state, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        CF_HIP(hipStreamSynchronize(s));
    };
    read_state();
    int cur = 0;
    int batch;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        batch = g_predict[dev] > 4 ? g_predict[dev] - 2 : 16;
    }
    while (!hs.done) {
        for (int c = 0; c < batch; ++c) {
            launch_pcg_dir_spmv<double>(gp, r, p[cur], p[cur ^ 1], state, rb, s);
            launch_pcg_update<double>(gp, x, r, p[cur ^ 1], state, rb, s);
            cur ^= 1;
        }
        CF_HIP(hipGetLastError());
        read_state();
        batch = 4;
    }
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_predict[dev] = hs.iter;
    }
    // without a preconditioner the solve returns the zero-mean pressure on a
    // singular operator, as the handle's solve does
    if (cfg.preconditioner == CF_PRECOND_NONE && operator_singular(g)) {
        launch_rhs_sum(g, x, scal, rb, s);
        launch_rhs_shift(g, x, scal, s);
        CF_HIP(hipStreamSynchronize(s));
    }
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

} // namespace cfmi
