// cf_mg.hpp -- geometric multigrid preconditioner for the FP64 pressure
// solve: level descriptions and kernel launchers (cf_mg.hip), and the host
// hierarchy (cf_mg_host.cpp).
//
// Level 0 is the pressure grid in the handle's padded layout with the
// matrix-free operator.  Coarser levels are compact arrays (x fastest, no
// ghosts) with stored Galerkin coefficients: diag, and the coupling cp[d]
// of each cell to its +e_d neighbour (wrapped on a periodic axis, 0 at a
// domain end); the -e_d coupling of a cell is cp[d] of that neighbour.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "cf_kernels.hpp"

namespace cfmi
{

constexpr int kMgMaxLevels = 24;

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
    int dim;
    int64_t ncell;
    const double* diag;
    const double* cp[3];
    const double* w;  // omega / diag
};

// Level 0 (padded layout): x = wd0[walls] * b; one damped Jacobi sweep
// xo = fma(wd0, b - A xi, xi); the residual r = b - A x.
void launch_mg0_first(const Geo& g, const MgTab& wd0, const double* b, double* x,
                      const PcgState* st, hipStream_t s);
void launch_mg0_sweep(const Geo& g, const MgTab& wd0, const double* xi,
                      const double* b, double* xo, const PcgState* st, hipStream_t s);
void launch_mg0_residual(const Geo& g, const double* x, const double* b, double* r,
                         const PcgState* st, hipStream_t s);
// Level 0 -> 1: b1 = sum over children of r (z, y, x order).
void launch_mg0_restrict(const Geo& g, const MgLev& c1, const double* r, double* b1,
                         const PcgState* st, hipStream_t s);
// Level 1 -> 0: x += x1[parent].
void launch_mg0_prolong(const Geo& g, const MgLev& c1, const double* x1, double* x,
                        const PcgState* st, hipStream_t s);

// Compact levels.
void launch_mg_first(const MgLev& L, const double* b, double* x, const PcgState* st,
                     hipStream_t s);
void launch_mg_sweep(const MgLev& L, const double* xi, const double* b, double* xo,
                     const PcgState* st, hipStream_t s);
// Level f -> f+1: bc = sum over children of (b - A x).
void launch_mg_residual_restrict(const MgLev& F, const MgLev& C, const double* x,
                                 const double* b, double* bc, const PcgState* st,
                                 hipStream_t s);
void launch_mg_prolong(const MgLev& F, const MgLev& C, const double* xc, double* x,
                       const PcgState* st, hipStream_t s);

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
    // z = B r (one V-cycle); returns the level-0 buffer that holds z (one of
    // the two the hierarchy owns).  Kernels return at once when st->done.
    const double* vcycle(const double* r, const PcgState* st);
    int levels() const { return (int)_lev.size(); }
    const MgLev& level(int l) const { return _lev[l]; }

  private:
    struct Buffers;
    const Geo& _g;
    hipStream_t _s;
    int _nu, _coarse;
    std::vector<MgLev> _lev; // _lev[0]: dims only
    std::unique_ptr<Buffers> _buf;
    void cycle(int l, const double* b, const PcgState* st);
};

} // namespace cfmi
