// cf_mg.hpp -- geometric multigrid preconditioner for the FP64 pressure
// solve: level descriptions and kernel launchers (cf_mg.hip), and the host
// hierarchy (cf_mg_host.cpp).
//
// Level 0 is the pressure grid in the handle's padded layout with the
// matrix-free operator.  Coarser levels are compact arrays (x fastest, no
// ghosts).  Their Galerkin operators P^T A P (P piecewise constant over the
// aggregation) are matrix-free too: a coarse cell is a box of level-0 cells,
// every level-0 coupling is s = dt / (rho dx^2), so
//   coupling to the -/+ e_d neighbour = -(area_d s)   (0 without neighbour)
//   diagonal                         = K s,
// where area_d is the number of level-0 faces between two neighbouring boxes
// along d (the product of the box's extents along the other axes) and K
// counts the box's boundary faces that are not SOLID domain walls.  These
// are the integer face counts the aggregation produces level by level,
// rounded once; the Jacobi weight is omega / (K s).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "cf_kernels.hpp"

namespace cfmi
{

constexpr int kMgMaxLevels = 24;
// The coarsest levels whose vectors (3 per level) fit this much LDS run
// inside one workgroup (k_mg_tail).
constexpr int kMgTailLdsBytes = 64 * 1024;
constexpr int kMgTailThreads = 1024;

// omega * (1 / diag) on level 0, by the cell's SOLID wall count.
struct MgTab
{
    double v[7];
};

// One compact level as the kernels see it.
struct MgLev
{
    int n[3];
    int coarsen[3];   // axis coarsened from this level to the next
    int periodic[3];
    int shift[3];     // a cell spans 2^shift[d] level-0 cells along d (fewer at the end)
    int n0[3];        // level-0 cells per axis
    int solid[3][2];  // SOLID domain ends of non-periodic axes
    int dim;
    int ncell;
    double s;         // dt / (rho dx^2)
    double omega;
    double alpha;     // coarse-correction scale: x = fma(alpha, x_parent, x)
};

// The levels that k_mg_tail runs (l0 .. last): their LDS offsets (X, T, B
// of level l from element off[l]), and the global arrays it reads (B[l0])
// and writes (res[l0]).  Each level starts a cycle in X and ends it in X or
// T by the parity of its sweep count.
struct MgTail
{
    int l0, last, nu, coarse;
    MgLev lev[kMgMaxLevels];
    int off[kMgMaxLevels];
    double* B[kMgMaxLevels];
    double* res[kMgMaxLevels];
};

// Linear index of the parent of cell (i, j, k) in the next level's array:
// base + i/2 + sy j/2 + sz k/2 (compact or padded layout).
struct MgParentMap
{
    int64_t base, sy, sz;
};

// Source of the values a smoothing sweep reads:
//   first:   x = w b (the sweep from zero, computed on the fly)
//   plain:   x = xi
//   prolong: x = fma(alpha, xc[parent], xi)
enum MgSrc
{
    MG_FIRST = 0,
    MG_PLAIN = 1,
    MG_PROLONG = 2
};

// Padded levels by row kernels (level 0 of 2-D or thin grids, and small
// tiled levels): xo = fma(w, b - A x, x) with x from `src` (the parent's
// value read at pm, scaled by alpha); with rz, also r.z = b.xo reduced into
// st (beta = r.z / rho after the first iteration, rho = r.z).
// launch_mg0_first: x = w b alone (a single sweep from zero).
void launch_mg0_smooth(const Geo& g, const MgTab& wd0, int src, bool rz, const double* xi,
                       const double* b, MgParentMap pm, double alpha, const double* xc,
                       double* xo, PcgState* st, RedBuf rb, hipStream_t s);
void launch_mg0_first(const Geo& g, const MgTab& wd0, const double* b, double* x,
                      const PcgState* st, hipStream_t s);
// Next level's b1 = sum over children (z, y, x order) of (b - A x); c1: the
// next level's dims, pm: where its cells are stored.
void launch_mg0_resrestrict(const Geo& g, const MgLev& c1, MgParentMap pm, const double* x,
                            const double* b, double* b1, const PcgState* st, hipStream_t s);

// Tiled levels (3-D, every axis coarsened further): level 0 and the coarse
// levels of full 2^l boxes in a padded layout with their own Geo.  The same
// sweeps (the parent read through pm, scaled by alpha) and
// residual-restriction as z-marching tiles; gw is the level's geometry with
// the Jacobi table omega / diag.
bool mg_tiled3(const Geo& g);
void launch_mg_smooth3(const Geo& g, const Geo& gw, int src, bool rz, const double* xi,
                       const double* b, const double* xc, MgParentMap pm, double alpha,
                       double* xo, PcgState* st, RedBuf rb, hipStream_t s);
void launch_mg_resrestrict3(const Geo& g, const double* x, const double* b, double* bc,
                            MgParentMap pm, const PcgState* st, hipStream_t s);

// FP32 level 0 (3-D, marched), and level 1 when it is a marched tiled level:
// the same sweeps and restriction on FP32 vectors.  b32 = (float) r (launch_mg_cast32, or the PCG update); a sweep
// with rz reads r64 instead of b32 (its sweep uses (float) r64) and reduces
// r.z with the FP64 r; the restriction sums the FP32 residuals as doubles
// into the FP64 level 1; the prolongation rounds the FP64 parent to FP32.
void launch_mg_cast32(const Geo& g, const double* r, float* b32, const PcgState* st,
                      hipStream_t s);
void launch_mg0_first_f32(const Geo& g, const MgTab& wd0, const float* b, float* x,
                          const PcgState* st, hipStream_t s);
// xc32 (non-null): an FP32 parent level instead of xc; cpl: cells per lane
// (2 on level 0, 1 on an FP32 level 1).
void launch_mg_smooth3_f32(const Geo& g, const Geo& gw, int src, bool rz, const float* xi,
                           const float* b32, const double* r64, const double* xc,
                           const float* xc32, MgParentMap pm, double alpha, float* xo,
                           PcgState* st, RedBuf rb, hipStream_t s, int cpl);
// bc32 (non-null): the sums rounded into an FP32 level 1 instead of bc.
void launch_mg_resrestrict3_f32(const Geo& g, const float* x, const float* b, double* bc,
                                float* bc32, MgParentMap pm, const PcgState* st, hipStream_t s,
                                int cpl);

// Compact levels.
void launch_mg_first(const MgLev& L, const double* b, double* x, const PcgState* st,
                     hipStream_t s);
void launch_mg_smooth(const MgLev& L, int src, const double* xi, const double* b,
                      const MgLev& C, const double* xc, double* xo, const PcgState* st,
                      hipStream_t s);
void launch_mg_resrestrict(const MgLev& F, const MgLev& C, const double* x, const double* b,
                           double* bc, const PcgState* st, hipStream_t s);
// One workgroup runs the V-cycle of levels t->l0 .. t->last in lds_bytes of LDS.
void launch_mg_tail(const MgTail* t_dev, size_t lds_bytes, const PcgState* st, hipStream_t s);

// MG-PCG: r = b, x = 0, ||b|| (sets st->mg); and rho / beta from r.z.
void launch_mgpcg_init(const Geo& g, const double* b, double* x, double* r,
                       PcgState* st, double tol, int maxit, RedBuf rb, hipStream_t s);
void launch_mgpcg_rz(const Geo& g, const double* r, const double* z, PcgState* st,
                     RedBuf rb, hipStream_t s);

// The device hierarchy of one handle.
class Multigrid
{
  public:
    Multigrid(const Geo& g, const cf_config& cfg, hipStream_t s);
    ~Multigrid();
    // z = B r (one V-cycle), then rho / beta from r.z; returns the level-0
    // buffer that holds z (one of the two the hierarchy owns).  Kernels
    // return at once when st->done.
    const double* vcycle(const double* r, PcgState* st, RedBuf rb);
    // The same with level 0 on FP32 vectors (3-D grids with more than 2
    // cells per axis, f32_supported()), the coarser levels in FP64; z comes
    // back in FP32.  b32_ready: the caller has written (float) r into b32()
    // (the PCG update does), so the V-cycle skips its own cast.
    const float* vcycle32(const double* r, PcgState* st, RedBuf rb, bool b32_ready);
    bool f32_supported() const { return _nt >= 1 && _lev.size() >= 2; }
    // the FP32 level 0's right-hand side (null before the first FP32 V-cycle)
    float* b32() const;
    int levels() const { return (int)_lev.size(); }
    const MgLev& level(int l) const { return _lev[l]; }

  private:
    struct Buffers;
    const void* cycle(const double* r, PcgState* st, RedBuf rb, bool f32, bool b32_ready);
    void first(int l, const double* b, double* x, PcgState* st);
    void smooth(int l, int src, bool rz, const double* xi, const double* b, double* xo,
                PcgState* st, RedBuf rb);
    void resrestrict(int l, const double* x, const double* b, PcgState* st);
    // one FP32 sweep on level l (launch_mg_smooth3_f32)
    void smooth32(int l, int src, bool rz, const float* xi, const float* b32, const double* r,
                  float* xo, const float* xc32, PcgState* st, RedBuf rb);
    // level 1 runs in FP32 too (with level 0): it is a marched tile level
    bool f32_level1() const;
    bool marched(int l) const;
    const Geo& _g;
    hipStream_t _s;
    int _nu, _coarse;
    int _tail0;              // first level run by k_mg_tail
    size_t _tail_lds;        // its LDS bytes
    int _nt;                 // levels 0 .. _nt-1 run the z-marching tile kernels
    std::vector<Geo> _lg;    // Geo of each tiled level (padded layout)
    std::vector<Geo> _lgw;   // ... with the Jacobi table omega / diag
    std::vector<MgTab> _wt;  // level 0 and tiled levels: omega / diag by wall count
    std::vector<MgParentMap> _pm; // parent index of level l's cells in level l + 1
    std::vector<MgLev> _lev; // _lev[0]: level-0 dims and coarsening only
    std::unique_ptr<Buffers> _buf;
};

// Sweep-count parity of one cycle on a level: the cycle starts in X and
// ends in T when odd.
__host__ __device__ inline int mg_level_swaps(bool last, int nu, int coarse)
{
    if (last)
        return coarse >= 2 ? coarse - 2 : 0;
    return (nu >= 2 ? nu - 2 : 0) + nu;
}

} // namespace cfmi
