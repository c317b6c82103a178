// cf_mg_host.cpp -- the multigrid hierarchy of a handle (cf_mg.hpp).
//
// Levels: level 0 is the pressure grid; an axis with more than 2 cells is
// coarsened by 2 (cells 2I and 2I+1 form cell I; the last cell of an odd
// axis stands alone), other axes are kept, until every axis has at most 2
// cells (at most kMgMaxLevels levels).  A coarse cell is then a box of
// level-0 cells, which is all its matrix-free Galerkin operator needs.
//
// Three kinds of level:
//   tiled    (3-D) level 0, and the levels below it whose boxes are all full
//            2^l cubes while every axis still coarsens: padded layout and a
//            Geo of their own (couplings -(4^l s), diagonal 4^l (6 - SOLID
//            walls) s, the coarse-level values bit for bit), run by the
//            z-marching tile kernels
//   compact  the other large levels (and level 0 of 2-D or thin grids, in
//            its padded layout, by the row kernels)
//   tail     the coarsest levels, as many as fit kMgTailLdsBytes of LDS, in
//            one workgroup
// One V-cycle launches the tiled and compact levels from here.  Each level's
// cycle starts in its X buffer and ends in X or T by the parity of its sweep
// count, fixed per level, so the prolongation of the level above knows
// where to read.
#include <climits>
#include <cstring>
#include <utility>

#include "cf_mg.hpp"
#include "cf_solver.hpp"

namespace cfmi
{

struct Multigrid::Buffers
{
    DevArray<double> z[2];
    DevArray<float> f[3];  // FP32 level 0: x, t, b32 (allocated on first use)
    DevArray<float> f1[3]; // FP32 level 1: x, t, b (when f32_level1())
    std::vector<std::unique_ptr<DevArray<double>>> vec;
    std::vector<double*> X, T, B, res;
    MgTail* tail = nullptr; // device copy
    ~Buffers()
    {
        if (tail)
            (void)hipFree(tail);
    }
};

namespace
{

// Padded Geo of tiled coarse level l (3-D; every box 2^l cells a side).
void tiled_level_geo(const Geo& g, const MgLev& L, int l, double omega, Geo& lg, Geo& lgw)
{
    lg = Geo{};
    lg.dim = 3;
    lg.h = 1;
    for (int d = 0; d < 3; ++d) {
        lg.n[d] = L.n[d];
        lg.periodic[d] = g.periodic[d];
        lg.solid[d][0] = g.solid[d][0];
        lg.solid[d][1] = g.solid[d][1];
    }
    lg.sx = (16 - (lg.h % 16)) % 16;
    const int64_t row = lg.sx + lg.n[0] + 2 * lg.h + 1;
    lg.px = (row + 15) / 16 * 16;
    lg.py = lg.n[1] + 2 * lg.h + 1;
    lg.pz = lg.n[2] + 2 * lg.h + 1;
    lg.total = lg.px * lg.py * lg.pz + 16;
    const int64_t A = (int64_t)1 << (2 * l); // level-0 faces per box face
    lg.s = (double)A * g.s;
    lg.s32 = (float)lg.s;
    lgw = lg;
    for (int k = 0; k <= 6; ++k) {
        const int64_t K = A * (6 - k);
        lg.diag[k] = (double)K * g.s;
        lg.dinv[k] = K > 0 ? 1.0 / lg.diag[k] : 0.0;
        lgw.diag[k] = lg.diag[k];
        lgw.dinv[k] = K > 0 ? omega / lg.diag[k] : 0.0;
        // an FP32 level: the FP64 coefficients rounded
        lg.diag32[k] = lgw.diag32[k] = (float)lg.diag[k];
        lg.dinv32[k] = (float)lg.dinv[k];
        lgw.dinv32[k] = (float)lgw.dinv[k];
    }
}

MgParentMap padded_map(const Geo& lg)
{
    return MgParentMap{lg.idx(lg.h, lg.h, lg.h), lg.px, lg.px * lg.py};
}

MgParentMap compact_map(const MgLev& L)
{
    return MgParentMap{0, L.n[0], (int64_t)L.n[0] * L.n[1]};
}

} // namespace

Multigrid::Multigrid(const Geo& g, const cf_config& cfg, hipStream_t s)
    : _g(g), _s(s), _nu(cfg.mg_nu), _coarse(cfg.mg_coarse_sweeps), _tail_lds(0),
      _buf(new Buffers)
{
    const int D = g.dim;
    MgLev L0{};
    for (int d = 0; d < 3; ++d) {
        L0.n[d] = d < D ? g.n[d] : 1;
        L0.n0[d] = L0.n[d];
        L0.periodic[d] = d < D ? g.periodic[d] : 0;
        L0.coarsen[d] = d < D && L0.n[d] > 2;
        L0.shift[d] = 0;
        L0.solid[d][0] = d < D && !L0.periodic[d] && g.solid[d][0];
        L0.solid[d][1] = d < D && !L0.periodic[d] && g.solid[d][1];
    }
    L0.dim = D;
    const int64_t n0cells = (int64_t)L0.n[0] * L0.n[1] * L0.n[2];
    L0.ncell = n0cells <= INT_MAX ? (int)n0cells : 0; // level 0 is addressed by Geo
    L0.s = g.s;
    L0.omega = cfg.mg_omega;
    L0.alpha = cfg.mg_alpha;
    _lev.push_back(L0);
    while ((int)_lev.size() < kMgMaxLevels) {
        const MgLev F = _lev.back();
        if (!(F.coarsen[0] || F.coarsen[1] || F.coarsen[2]))
            break;
        MgLev C = F;
        int64_t nc = 1;
        for (int d = 0; d < 3; ++d) {
            C.n[d] = F.coarsen[d] ? (F.n[d] + 1) / 2 : F.n[d];
            C.shift[d] = F.shift[d] + (F.coarsen[d] ? 1 : 0);
            C.coarsen[d] = d < D && C.n[d] > 2;
            nc *= C.n[d];
        }
        if (nc > INT_MAX)
            throw CfError(CF_ENOTIMPL, "multigrid: a coarse level has more than 2^31 cells");
        C.ncell = (int)nc;
        _lev.push_back(C);
    }
    const int nl = (int)_lev.size();
    const int last = nl - 1;
    if (_lev[last].ncell > 64)
        throw CfError(CF_ENOTIMPL, "multigrid: the coarsest level must fit one wave");
    // the tail: the coarsest levels whose X, T, B fit the LDS budget
    _tail0 = last;
    {
        int64_t cells = 0;
        for (int l = last; l >= 1; --l) {
            cells += _lev[l].ncell;
            if (3 * cells * (int64_t)sizeof(double) > kMgTailLdsBytes)
                break;
            _tail0 = l;
        }
    }
    // tiled levels
    MgTab wd0;
    for (int k = 0; k < 7; ++k)
        wd0.v[k] = cfg.mg_omega * g.dinv[k];
    _nt = 0;
    if (last >= 1 && mg_tiled3(g)) {
        Geo gw = g;
        for (int k = 0; k < 7; ++k) {
            gw.dinv[k] = wd0.v[k];
            gw.dinv32[k] = (float)wd0.v[k]; // the FP32 level 0's w
        }
        _lg.push_back(g);
        _lgw.push_back(gw);
        _nt = 1;
        for (int l = 1; l < _tail0; ++l) {
            const MgLev& L = _lev[l];
            bool ok = L.coarsen[0] && L.coarsen[1] && L.coarsen[2];
            for (int d = 0; d < 3; ++d)
                ok = ok && L.shift[d] == l && g.n[d] % (1 << l) == 0;
            if (!ok)
                break;
            Geo lg, lgw;
            tiled_level_geo(g, L, l, cfg.mg_omega, lg, lgw);
            _lg.push_back(lg);
            _lgw.push_back(lgw);
            _nt = l + 1;
        }
    }
    _wt.assign(nl, MgTab{});
    _wt[0] = wd0;
    for (int l = 1; l < _nt; ++l)
        for (int k = 0; k < 7; ++k)
            _wt[l].v[k] = _lgw[l].dinv[k];
    _pm.assign(nl, MgParentMap{0, 0, 0});
    for (int l = 0; l + 1 < nl; ++l)
        _pm[l] = l + 1 < _nt ? padded_map(_lg[l + 1]) : compact_map(_lev[l + 1]);
    // buffers
    Buffers& B = *_buf;
    B.z[0].alloc(g.total, s);
    B.z[1].alloc(g.total, s);
    B.X.assign(nl, nullptr);
    B.T.assign(nl, nullptr);
    B.B.assign(nl, nullptr);
    B.res.assign(nl, nullptr);
    B.X[0] = B.z[0].ptr;
    B.T[0] = B.z[1].ptr;
    for (int l = 1; l < nl; ++l) {
        const int64_t len = l < _nt ? _lg[l].total : (int64_t)_lev[l].ncell;
        for (double** dst : {&B.X[l], &B.T[l], &B.B[l]}) {
            std::unique_ptr<DevArray<double>> a(new DevArray<double>());
            a->alloc(len, s);
            *dst = a->ptr;
            B.vec.push_back(std::move(a));
        }
    }
    for (int l = 0; l < nl; ++l)
        B.res[l] = (mg_level_swaps(l == last, _nu, _coarse) & 1) ? B.T[l] : B.X[l];
    if (last >= 1) {
        MgTail h{};
        h.l0 = _tail0;
        h.last = last;
        h.nu = _nu;
        h.coarse = _coarse;
        int off = 0;
        for (int l = 0; l < nl; ++l) {
            h.lev[l] = _lev[l];
            h.B[l] = B.B[l];
            h.res[l] = B.res[l];
            h.off[l] = off;
            if (l >= _tail0)
                off += 3 * _lev[l].ncell;
        }
        _tail_lds = (size_t)off * sizeof(double);
        void* p = nullptr;
        CF_HIP(hipMalloc(&p, sizeof(MgTail)));
        B.tail = (MgTail*)p;
        CF_HIP(hipMemcpyAsync(B.tail, &h, sizeof(MgTail), hipMemcpyHostToDevice, s));
    }
    CF_HIP(hipStreamSynchronize(s));
}

Multigrid::~Multigrid() = default;

// x = w b on level l
void Multigrid::first(int l, const double* b, double* x, PcgState* st)
{
    if (l == 0 || l < _nt)
        launch_mg0_first(l == 0 ? _g : _lg[l], _wt[l], b, x, st, _s);
    else
        launch_mg_first(_lev[l], b, x, st, _s);
}

// Sweeps on padded levels of at most this many cells run the row kernels:
// one pass over memory per launch, where the z march's prologue and
// plane-by-plane chain make a small grid latency-bound (measured at 32^3:
// 4.5 against 7.5 µs; at 64^3 the march is faster).  Restrictions keep the
// march, whose lane-pair sums beat the row kernel's thread per parent.
constexpr int64_t kMgMarchMinCells = 1 << 15;

bool Multigrid::marched(int l) const
{
    return l < _nt && (l == 0 || (int64_t)_lev[l].ncell > kMgMarchMinCells);
}

// one sweep on level l (l below the coarsest); the prolongation source is
// level l + 1's cycle result
void Multigrid::smooth(int l, int src, bool rz, const double* xi, const double* b, double* xo,
                       PcgState* st, RedBuf rb)
{
    const double* xc = _buf->res[l + 1];
    const double alpha = _lev[0].alpha;
    if (marched(l))
        launch_mg_smooth3(_lg[l], _lgw[l], src, rz, xi, b, xc, _pm[l], alpha, xo, st, rb, _s);
    else if (l == 0 || l < _nt)
        launch_mg0_smooth(l == 0 ? _g : _lg[l], _wt[l], src, rz, xi, b, _pm[l], alpha, xc, xo,
                          st, rb, _s);
    else
        launch_mg_smooth(_lev[l], src, xi, b, _lev[l + 1], xc, xo, st, _s);
}

// level l + 1's right-hand side from level l
void Multigrid::resrestrict(int l, const double* x, const double* b, PcgState* st)
{
    double* bc = _buf->B[l + 1];
    if (l < _nt)
        launch_mg_resrestrict3(_lg[l], x, b, bc, _pm[l], st, _s);
    else if (l == 0)
        launch_mg0_resrestrict(_g, _lev[1], _pm[0], x, b, bc, st, _s);
    else
        launch_mg_resrestrict(_lev[l], _lev[l + 1], x, b, bc, st, _s);
}

// one FP32 sweep on level l (0, or 1 when f32_level1()); the prolongation
// source is level l + 1's cycle result (xc32 when that level is FP32)
void Multigrid::smooth32(int l, int src, bool rz, const float* xi, const float* b32,
                         const double* r, float* xo, const float* xc32, PcgState* st, RedBuf rb)
{
    Buffers& B = *_buf;
    launch_mg_smooth3_f32(_lg[l], _lgw[l], src, rz, xi, b32, r, xc32 ? nullptr : B.res[l + 1],
                          xc32, _pm[l], _lev[0].alpha, xo, st, rb, _s, l == 0 ? 2 : 1);
}

bool Multigrid::f32_level1() const
{
    return f32_supported() && _nt >= 2 && marched(1) && _tail0 > 1 && _lev.size() >= 3;
}

float* Multigrid::b32() const { return _buf->f[2].ptr; }

const double* Multigrid::vcycle(const double* r, PcgState* st, RedBuf rb)
{
    return static_cast<const double*>(cycle(r, st, rb, false, false));
}

const float* Multigrid::vcycle32(const double* r, PcgState* st, RedBuf rb, bool b32_ready)
{
    return static_cast<const float*>(cycle(r, st, rb, true, b32_ready));
}

// One V-cycle; returns z (double, or float with f32).
const void* Multigrid::cycle(const double* r, PcgState* st, RedBuf rb, bool f32,
                             bool b32_ready)
{
    Buffers& B = *_buf;
    const int last = (int)_lev.size() - 1;
    if (f32 && !f32_supported())
        throw CfError(CF_ENOTIMPL, "the FP32 multigrid V-cycle needs a 3-D grid with more "
                                   "than 2 cells along every axis");
    if (last == 0) {
        // level 0 is the coarsest: coarse sweeps from zero
        double* x = B.X[0];
        double* t = B.T[0];
        const MgParentMap none{0, 0, 0};
        if (_coarse == 1)
            first(0, r, x, st);
        else
            launch_mg0_smooth(_g, _wt[0], MG_FIRST, false, nullptr, r, none, 0.0, nullptr, x, st,
                              rb, _s);
        for (int it = 2; it < _coarse; ++it) {
            launch_mg0_smooth(_g, _wt[0], MG_PLAIN, false, x, r, none, 0.0, nullptr, t, st, rb,
                              _s);
            std::swap(x, t);
        }
        launch_mgpcg_rz(_g, r, x, st, rb, _s);
        return x;
    }
    // FP32 level 0 (and level 1 when it is a marched tile level): b32 =
    // (float) r, pre-smoothing and restriction down to the first FP64 level
    // (the up phase below ends these levels in FP32 too)
    const bool f1 = f32 && f32_level1();
    const int l0 = f32 ? (f1 ? 2 : 1) : 0; // first level of the FP64 loops
    float *x32 = nullptr, *x1f = nullptr;
    if (f32) {
        for (int i = 0; i < 3; ++i) {
            if (!B.f[i].ptr)
                B.f[i].alloc(_g.total, _s);
            if (f1 && !B.f1[i].ptr)
                B.f1[i].alloc(_lg[1].total, _s);
        }
        float *x = B.f[0].ptr, *t = B.f[1].ptr;
        float* b0 = B.f[2].ptr;
        if (!b32_ready)
            launch_mg_cast32(_g, r, b0, st, _s);
        if (_nu == 1)
            launch_mg0_first_f32(_g, _wt[0], b0, x, st, _s);
        else
            smooth32(0, MG_FIRST, false, nullptr, b0, r, x, nullptr, st, rb);
        for (int it = 2; it < _nu; ++it) {
            smooth32(0, MG_PLAIN, false, x, b0, r, t, nullptr, st, rb);
            std::swap(x, t);
        }
        launch_mg_resrestrict3_f32(_lg[0], x, b0, B.B[1], f1 ? B.f1[2].ptr : nullptr, _pm[0], st,
                                   _s, 2);
        x32 = x;
        if (f1) {
            float *y = B.f1[0].ptr, *u = B.f1[1].ptr;
            const float* b1 = B.f1[2].ptr;
            if (_nu == 1)
                launch_mg0_first_f32(_lg[1], _wt[1], b1, y, st, _s);
            else
                smooth32(1, MG_FIRST, false, nullptr, b1, nullptr, y, nullptr, st, rb);
            for (int it = 2; it < _nu; ++it) {
                smooth32(1, MG_PLAIN, false, y, b1, nullptr, u, nullptr, st, rb);
                std::swap(y, u);
            }
            launch_mg_resrestrict3_f32(_lg[1], y, b1, B.B[2], nullptr, _pm[1], st, _s, 1);
            x1f = y;
        }
    }
    // down: pre-smoothing from zero and restriction
    for (int l = l0; l < _tail0; ++l) {
        const double* b = l == 0 ? r : B.B[l];
        double* x = B.X[l];
        double* t = B.T[l];
        if (_nu == 1)
            first(l, b, x, st);
        else
            smooth(l, MG_FIRST, false, nullptr, b, x, st, rb);
        for (int it = 2; it < _nu; ++it) {
            smooth(l, MG_PLAIN, false, x, b, t, st, rb);
            std::swap(x, t);
        }
        resrestrict(l, x, b, st);
    }
    launch_mg_tail(B.tail, _tail_lds, st, _s);
    // up: the first post-sweep reads the coarse correction; on level 0 the
    // last sweep also reduces r.z
    const bool pre_odd = (_nu >= 2 ? _nu - 2 : 0) & 1;
    double* x = nullptr;
    for (int l = _tail0 - 1; l >= l0; --l) {
        const double* b = l == 0 ? r : B.B[l];
        x = pre_odd ? B.T[l] : B.X[l];
        double* t = pre_odd ? B.X[l] : B.T[l];
        smooth(l, MG_PROLONG, l == 0 && _nu == 1, x, b, t, st, rb);
        std::swap(x, t);
        for (int it = 1; it < _nu; ++it) {
            smooth(l, MG_PLAIN, l == 0 && it == _nu - 1, x, b, t, st, rb);
            std::swap(x, t);
        }
    }
    if (f32) {
        const float* xc32 = nullptr; // level 0's parent when level 1 is FP32
        if (f1) {
            float* y = x1f;
            float* u = y == B.f1[0].ptr ? B.f1[1].ptr : B.f1[0].ptr;
            const float* b1 = B.f1[2].ptr;
            smooth32(1, MG_PROLONG, false, y, b1, nullptr, u, nullptr, st, rb);
            std::swap(y, u);
            for (int it = 1; it < _nu; ++it) {
                smooth32(1, MG_PLAIN, false, y, b1, nullptr, u, nullptr, st, rb);
                std::swap(y, u);
            }
            xc32 = y;
        }
        // level 0 in FP32: the last sweep writes z and reduces r.z from the
        // FP64 r; z stays in FP32 (the direction kernel widens it)
        float* xf = x32;
        float* tf = xf == B.f[0].ptr ? B.f[1].ptr : B.f[0].ptr;
        const float* b0 = B.f[2].ptr;
        smooth32(0, MG_PROLONG, _nu == 1, xf, b0, r, tf, xc32, st, rb);
        std::swap(xf, tf);
        for (int it = 1; it < _nu; ++it) {
            smooth32(0, MG_PLAIN, it == _nu - 1, xf, b0, r, tf, xc32, st, rb);
            std::swap(xf, tf);
        }
        return xf;
    }
    return x;
}

} // namespace cfmi
