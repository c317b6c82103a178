// cf_kernels.hpp -- host-callable launchers for the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cf_geometry.hpp"

namespace cfmi
{

// Device-resident state of one PCG solve.  Finalize kernels update the
// scalars on the device, so the host never waits on a reduction; it polls
// `done` once per batch of iterations.
struct PcgState
{
    double rho, alpha, beta, bnorm, res, tol;
    int iter, maxit, done, converged;
    int mg; // z comes from the multigrid V-cycle: the update kernel leaves
            // beta and rho to k_mgpcg_rz
    int nonfinite; // ||b|| or ||r|| came out NaN or Inf: the solve stopped
    double alpha_lag; // alpha of an X_SKIP update, applied by the next X_PAIR
    int stalled;      // a cooperative launch timed out waiting for a workgroup
};

// Grid-reduction scratch of the streaming kernels: per-block double-double
// partials (up to reduction_blocks_max * 3 values * 2 doubles) and the
// arrival counters -- one per group of blocks (blockIdx & 7) and one for the
// groups, each on its own 256-byte line (kCounterSlots * kCounterStride
// unsigned) -- which the last block resets to 0.
struct RedBuf
{
    double* partials;
    unsigned* counter;
};

constexpr int kCounterStride = 64; // unsigned per counter line
constexpr int kCounterSlots = 9;   // 8 group counters + the group counter

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048; // largest grid of the streaming kernels (multiple of 8)

// State arrays of one version: q, u, v, w (w null in 2-D).
struct Fields
{
    double* e[4];
};

void launch_advect(const Geo& g, Fields cur, Fields next, hipStream_t s);
// Solver::_addInputs.  With no body force only the inflow region's
// bounding box and the SOLID wall planes of the face entities are visited
// (at most kMaxInputBoxes index boxes); with one, every owned entity.
constexpr int kMaxInputBoxes = 12;
void launch_inputs(const Geo& g, Fields cur, hipStream_t s);
bool inputs_are_sparse(const Geo& g);
void launch_ghost_fill(const Geo& g, int entity, double* a, hipStream_t s);
void launch_divergence(const Geo& g, Fields cur, double* rhs, hipStream_t s);
void launch_rhs_sum(const Geo& g, const double* rhs, double* out_sum,
                    RedBuf rb, hipStream_t s);
void launch_rhs_shift(const Geo& g, double* rhs, const double* sum,
                      hipStream_t s);
void launch_project(const Geo& g, Fields cur, const double* p, hipStream_t s);
// Dense ghosted array <-> padded layout (host transfers).
void launch_repack(const Geo& g, int entity, double* padded, double* dense, bool to_padded,
                   hipStream_t s);
// A whole padded array copied by a kernel, dst = src over [0, g.total):
// the device entry points take caller memory that another allocator (another
// HIP runtime's, in a process that also runs torch) may own, which the
// runtime's own copy path does not know.  owned_only: entries outside the
// entity's owned box are written as 0.
void launch_copy_padded(const Geo& g, int entity, double* dst, const double* src,
                        bool owned_only, hipStream_t s);
// The output writer's cell data (owned q, cell-centre velocity; compact).
void launch_cell_output(const Geo& g, Fields cur, double* q_out, double* vel_out, hipStream_t s);

// Jacobi-PCG (cf_pcg.hip).  init64: r = b, x = 0; init32: d = 0 for the
// FP32 inner solve of the residual already in r.
void launch_pcg_init64(const Geo& g, const double* b, double* x, double* r,
                       PcgState* st, double tol, int maxit, RedBuf rb,
                       hipStream_t s);
struct RefineState;
// init32 ends the pass at ||r|| / ||r0|| < max(itol, 0.5 gtol ||b|| / ||r0||)
// with ||b|| = rs->bnorm.
void launch_pcg_init32(const Geo& g, const float* r, float* d, PcgState* st,
                       const RefineState* rs, double itol, double gtol,
                       int maxit, RedBuf rb, hipStream_t s);
template <class T>
void launch_pcg_dir_spmv(const Geo& g, const T* r, const T* p_old, T* p_new,
                         PcgState* st, RedBuf rb, hipStream_t s);
// The FP64 3-D direction kernel on a unit-Jacobi geometry reading an FP32 z
// (an FP32 multigrid V-cycle's output): p_new = fma(beta, p_old, (double) z).
void launch_pcg_dir_spmv_z32(const Geo& g, const float* z, const double* p_old, double* p_new,
                             PcgState* st, RedBuf rb, hipStream_t s);
// The update's treatment of x.  The 3-D tile kernels can leave x alone on
// every second iteration and apply both steps on the next, in the order of
// two single steps (same bits):
//   X_EACH: x = fma(a_k, p_k, x)                                 5 words
//   X_SKIP: x unchanged; the state keeps a_k as alpha_lag         3 words
//   X_PAIR: x = fma(a_k, p_k, fma(alpha_lag, p_{k-1}, x))         6 words
// p_{k-1} is the other direction buffer, which the direction kernel of
// iteration k read and left intact.  A run that stops after an X_SKIP
// update completes x with launch_pcg_xfix.
enum XUpdate { X_EACH = 0, X_SKIP = 1, X_PAIR = 2 };
// r32 (3-D FP64 only): also write (float) r there, the right-hand side of
// an FP32 finest multigrid level.
template <class T>
void launch_pcg_update(const Geo& g, T* x, T* r, const T* p, PcgState* st,
                       RedBuf rb, hipStream_t s, int xmode = X_EACH,
                       const T* p_lag = nullptr, float* r32 = nullptr);
// x = fma(alpha, p, x) over the owned cells with the state's alpha: the x
// step of a run's last iteration when that was an X_SKIP update.
template <class T>
void launch_pcg_xfix(const Geo& g, T* x, const T* p, const PcgState* st, hipStream_t s);
// Whether launch_pcg_update<T> takes X_SKIP / X_PAIR on this geometry: the
// 3-D tile kernels do; the 2-D kernels update x every iteration.
template <class T>
inline bool pcg_pairs_x(const Geo& g)
{
    return g.dim == 3;
}
// Largest grid any reducing kernel of this geometry launches (sizes the
// partials buffer).
int reduction_blocks_max(const Geo& g);

// 2-D grids: FP64 Jacobi-PCG iterations in one cooperative launch of one
// workgroup per strip of rows, until the state's stop test or `max_its`
// iterations of this launch (x, r and the direction p are read at the start
// and written back at the end, so launches chain).  The same x, r, direction
// and state bits as the dir_spmv / update pair.  pcg_coop_fits: the grid
// qualifies (2-D, at most 2048 cells along x and 256 strips of 2048 cells);
// xch: coop_xch_bytes(g) bytes of device memory.  A launch that waited too
// long for another workgroup sets PcgState::stalled.  Returns false when the
// runtime refuses the cooperative launch (the workgroups cannot all be
// resident, or the device has no cooperative launch); nothing ran then.
bool pcg_coop_fits(const Geo& g);
size_t coop_xch_bytes(const Geo& g);
bool launch_pcg_solve_coop(const Geo& g, double* x, double* r, double* p, PcgState* st,
                           void* xch, int max_its, hipStream_t s);

// FP32-inner / FP64-outer defect correction.  The host runs the outer loop
// from the device-computed norms (one read-back per pass).
struct RefineState
{
    double bnorm, res;
};
void launch_refine_init(const Geo& g, const double* b, double* x64,
                        float* r32, RefineState* rs, RedBuf rb, hipStream_t s);
// x64 += (double) d32; d32 = 0
void launch_refine_add(const Geo& g, double* x64, float* d32, hipStream_t s);
void launch_refine_residual(const Geo& g, const double* b, const double* x64,
                            float* r32, RefineState* rs, RedBuf rb,
                            hipStream_t s);
// Carried-direction mode (cf_pcg.hip): start, and the pass-boundary
// residual replacement that re-arms the FP32 PCG in `st`.
void launch_refine_init_carry(const Geo& g, const double* b, double* x64,
                              float* r32, float* d32, RefineState* rs,
                              PcgState* st, double itol, double gtol, int maxit,
                              RedBuf rb, hipStream_t s);
void launch_refine_residual_carry(const Geo& g, const double* b,
                                  const double* x64, float* r32,
                                  RefineState* rs, PcgState* st, double itol,
                                  double gtol, RedBuf rb, hipStream_t s);

} // namespace cfmi
