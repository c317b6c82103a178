// cf_kernels.hpp -- host-callable launchers for the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "cf_geometry.hpp"

namespace cfmi
{

// Device-resident state of one PCG solve.  The last block of each reducing
// kernel computes the scalars, so the host never waits on a reduction; it
// polls `done` once per batch of iterations.
struct PcgState
{
    double rho, alpha, beta, bnorm, res, tol;
    int iter, maxit, done, converged;
};

// Scratch of a grid reduction: per-block double-double partials and the
// arrival counter (wraps back to 0 after every launch).
struct RedBuf
{
    double* partials; // kRedBlocks * 4 * 2 doubles
    unsigned* counter;
};

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048; // grid of the streaming kernels (multiple of 8)

int stream_grid(const Geo& g);

// State arrays of one version: q, u, v, w (w null in 2-D).
struct Fields
{
    double* e[4];
};

void launch_advect(const Geo& g, Fields cur, Fields next, hipStream_t s);
void launch_inputs(const Geo& g, Fields cur, hipStream_t s);
void launch_ghost_fill(const Geo& g, int entity, double* a, hipStream_t s);
void launch_divergence(const Geo& g, Fields cur, double* rhs, hipStream_t s);
void launch_rhs_sum(const Geo& g, const double* rhs, double* out_sum,
                    RedBuf rb, hipStream_t s);
void launch_rhs_shift(const Geo& g, double* rhs, const double* sum,
                      hipStream_t s);
void launch_project(const Geo& g, Fields cur, const double* p, hipStream_t s);

// FP64 Jacobi-PCG (x = 0 start).
void launch_pcg_init(const Geo& g, const double* b, double* x, double* r,
                     double* p, PcgState* st, double tol, int maxit,
                     RedBuf rb, hipStream_t s);
void launch_pcg_spmv_dot(const Geo& g, const double* p, double* q,
                         PcgState* st, RedBuf rb, hipStream_t s);
void launch_pcg_update(const Geo& g, double* x, double* r, const double* p,
                       const double* q, PcgState* st, RedBuf rb,
                       hipStream_t s);
void launch_pcg_dir(const Geo& g, const double* r, double* p,
                    const PcgState* st, hipStream_t s);

// FP32-inner / FP64-outer defect correction.  The host runs the outer loop
// from the device-computed norms (one read-back per pass).
struct RefineState
{
    double bnorm, res;
};
void launch_refine_init(const Geo& g, const double* b, double* x64,
                        float* r32, RefineState* rs, RedBuf rb, hipStream_t s);
void launch_refine_add(const Geo& g, double* x64, const float* d32,
                       hipStream_t s);
void launch_refine_residual(const Geo& g, const double* b, const double* x64,
                            float* r32, RefineState* rs, RedBuf rb,
                            hipStream_t s);
void launch_pcg32_init(const Geo& g, const float* r, float* d, float* p,
                       PcgState* st, double tol, int maxit, RedBuf rb,
                       hipStream_t s);
void launch_pcg32_spmv_dot(const Geo& g, const float* p, float* q,
                           PcgState* st, RedBuf rb, hipStream_t s);
void launch_pcg32_update(const Geo& g, float* d, float* r, const float* p,
                         const float* q, PcgState* st, RedBuf rb,
                         hipStream_t s);
void launch_pcg32_dir(const Geo& g, const float* r, float* p,
                      const PcgState* st, hipStream_t s);

} // namespace cfmi
