// cf_mg_host.cpp -- the multigrid hierarchy of a handle (cf_mg.hpp).
//
// Levels: level 0 is the pressure grid; an axis with more than 2 cells is
// coarsened by 2 (cells 2I and 2I+1 form cell I; the last cell of an odd
// axis stands alone), other axes are kept, until every axis has at most 2
// cells.  Coarse operators are Galerkin P^T A P for the piecewise-constant
// P.  Every coefficient of the pressure operator is an integer multiple of
// s = dt / (rho dx^2), so the coarse ones are formed as integer face counts
// (K for the diagonal, C_d for the coupling to the +e_d neighbour) and
// rounded once: diag = K s, cp[d] = -(C_d s), w = omega / diag.
#include <cstring>
#include <utility>

#include "cf_mg.hpp"
#include "cf_solver.hpp"

namespace cfmi
{

struct Multigrid::Buffers
{
    DevArray<double> z[2], r0;
    std::vector<std::unique_ptr<DevArray<double>>> coef, vec;
    std::vector<const double*> xcur; // per level: the buffer holding x
    std::vector<double*> x, t, b;    // per compact level
    MgTab wd0;
};

namespace
{

int64_t flat3(const int (&n)[3], int i, int j, int k)
{
    return (int64_t)i + (int64_t)n[0] * ((int64_t)j + (int64_t)n[1] * k);
}

// index of the +e_d neighbour (wrapped when periodic), or -1
int64_t plus_nbr(const int (&n)[3], const int (&per)[3], const int (&c)[3], int d)
{
    int m[3] = {c[0], c[1], c[2]};
    m[d] += 1;
    if (m[d] >= n[d]) {
        if (!per[d])
            return -1;
        m[d] = 0;
    }
    return flat3(n, m[0], m[1], m[2]);
}

} // namespace

Multigrid::Multigrid(const Geo& g, const cf_config& cfg, hipStream_t s)
    : _g(g), _s(s), _nu(cfg.mg_nu), _coarse(cfg.mg_coarse_sweeps), _buf(new Buffers)
{
    const int D = g.dim;
    int per[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d)
        per[d] = g.periodic[d];
    // level-0 integer counts
    int n0[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d)
        n0[d] = g.n[d];
    std::vector<int64_t> K((size_t)n0[0] * n0[1] * n0[2]);
    std::vector<int64_t> Cc[3];
    for (int d = 0; d < D; ++d)
        Cc[d].assign(K.size(), 0);
    for (int k = 0; k < n0[2]; ++k)
        for (int j = 0; j < n0[1]; ++j)
            for (int i = 0; i < n0[0]; ++i) {
                const int c[3] = {i, j, k};
                const int64_t id = flat3(n0, i, j, k);
                int64_t kd = 2 * D;
                for (int d = 0; d < D; ++d) {
                    if (!per[d]) {
                        if (c[d] == 0 && g.solid[d][0])
                            --kd;
                        if (c[d] == n0[d] - 1 && g.solid[d][1])
                            --kd;
                    }
                    Cc[d][id] = plus_nbr(n0, per, c, d) >= 0 ? 1 : 0;
                }
                K[id] = kd;
            }
    MgLev L0{};
    for (int d = 0; d < 3; ++d) {
        L0.n[d] = n0[d];
        L0.periodic[d] = per[d];
        L0.coarsen[d] = d < D && n0[d] > 2;
    }
    L0.dim = D;
    L0.ncell = (int64_t)K.size();
    _lev.push_back(L0);
    // coarser levels
    const double sc = g.s;
    const double omega = cfg.mg_omega;
    while ((int)_lev.size() < kMgMaxLevels) {
        const MgLev& F = _lev.back();
        if (!(F.coarsen[0] || F.coarsen[1] || F.coarsen[2]))
            break;
        MgLev C{};
        for (int d = 0; d < 3; ++d) {
            C.n[d] = F.coarsen[d] ? (F.n[d] + 1) / 2 : F.n[d];
            C.periodic[d] = per[d];
        }
        C.dim = D;
        C.ncell = (int64_t)C.n[0] * C.n[1] * C.n[2];
        for (int d = 0; d < 3; ++d)
            C.coarsen[d] = d < D && C.n[d] > 2;
        std::vector<int64_t> Kc((size_t)C.ncell, 0), Ccc[3];
        for (int d = 0; d < D; ++d)
            Ccc[d].assign((size_t)C.ncell, 0);
        for (int Kz = 0; Kz < C.n[2]; ++Kz)
            for (int J = 0; J < C.n[1]; ++J)
                for (int I = 0; I < C.n[0]; ++I) {
                    const int P[3] = {I, J, Kz};
                    int lo[3], hi[3];
                    for (int d = 0; d < 3; ++d) {
                        if (F.coarsen[d]) {
                            lo[d] = 2 * P[d];
                            hi[d] = 2 * P[d] + 1 < F.n[d] ? 2 * P[d] + 1 : 2 * P[d];
                        } else {
                            lo[d] = hi[d] = P[d];
                        }
                    }
                    int64_t kd = 0, cp[3] = {0, 0, 0};
                    for (int k = lo[2]; k <= hi[2]; ++k)
                        for (int j = lo[1]; j <= hi[1]; ++j)
                            for (int i = lo[0]; i <= hi[0]; ++i) {
                                const int c[3] = {i, j, k};
                                const int64_t fid = flat3(F.n, i, j, k);
                                kd += K[(size_t)fid];
                                for (int d = 0; d < D; ++d) {
                                    // a coupling between two children of
                                    // one parent: twice into the diagonal
                                    if (c[d] < hi[d])
                                        kd -= 2 * Cc[d][(size_t)fid];
                                    else
                                        cp[d] += Cc[d][(size_t)fid];
                                }
                            }
                    const int64_t cid = flat3(C.n, I, J, Kz);
                    Kc[(size_t)cid] = kd;
                    for (int d = 0; d < D; ++d)
                        Ccc[d][(size_t)cid] = cp[d];
                }
        // device coefficients
        std::vector<double> hd((size_t)C.ncell), hw((size_t)C.ncell), hc((size_t)C.ncell);
        auto upload = [&](const std::vector<double>& h) -> const double* {
            std::unique_ptr<DevArray<double>> a(new DevArray<double>());
            a->alloc(C.ncell, s);
            CF_HIP(hipMemcpyAsync(a->ptr, h.data(), h.size() * sizeof(double),
                                  hipMemcpyHostToDevice, s));
            CF_HIP(hipStreamSynchronize(s));
            const double* p = a->ptr;
            _buf->coef.push_back(std::move(a));
            return p;
        };
        for (int64_t i = 0; i < C.ncell; ++i) {
            hd[(size_t)i] = (double)Kc[(size_t)i] * sc;
            hw[(size_t)i] = omega / hd[(size_t)i];
        }
        C.diag = upload(hd);
        C.w = upload(hw);
        for (int d = 0; d < D; ++d) {
            for (int64_t i = 0; i < C.ncell; ++i)
                hc[(size_t)i] = -((double)Ccc[d][(size_t)i] * sc);
            C.cp[d] = upload(hc);
        }
        K.swap(Kc);
        for (int d = 0; d < D; ++d)
            Cc[d].swap(Ccc[d]);
        _lev.push_back(C);
    }
    // work vectors
    _buf->z[0].alloc(g.total, s);
    _buf->z[1].alloc(g.total, s);
    _buf->r0.alloc(g.total, s);
    const int nl = (int)_lev.size();
    _buf->xcur.assign(nl, nullptr);
    _buf->x.assign(nl, nullptr);
    _buf->t.assign(nl, nullptr);
    _buf->b.assign(nl, nullptr);
    for (int l = 1; l < nl; ++l) {
        for (double** dst : {&_buf->x[l], &_buf->t[l], &_buf->b[l]}) {
            std::unique_ptr<DevArray<double>> a(new DevArray<double>());
            a->alloc(_lev[l].ncell, s);
            *dst = a->ptr;
            _buf->vec.push_back(std::move(a));
        }
    }
    for (int k = 0; k < 7; ++k)
        _buf->wd0.v[k] = omega * g.dinv[k];
    CF_HIP(hipStreamSynchronize(s));
}

Multigrid::~Multigrid() = default;

// V-cycle on level l >= 1 with its right-hand side in b[l]; x ends in xcur[l].
void Multigrid::cycle(int l, const double* b, const PcgState* st)
{
    Buffers& B = *_buf;
    const int last = (int)_lev.size() - 1;
    const MgLev& L = _lev[l];
    double* x = B.x[l];
    double* t = B.t[l];
    const int pre = l == last ? _coarse : _nu;
    launch_mg_first(L, b, x, st, _s);
    for (int it = 1; it < pre; ++it) {
        launch_mg_sweep(L, x, b, t, st, _s);
        std::swap(x, t);
    }
    if (l < last) {
        launch_mg_residual_restrict(L, _lev[l + 1], x, b, B.b[l + 1], st, _s);
        cycle(l + 1, B.b[l + 1], st);
        launch_mg_prolong(L, _lev[l + 1], B.xcur[l + 1], x, st, _s);
        for (int it = 0; it < _nu; ++it) {
            launch_mg_sweep(L, x, b, t, st, _s);
            std::swap(x, t);
        }
    }
    B.xcur[l] = x;
    // keep the two buffers distinct for the next cycle
    B.x[l] = x;
    B.t[l] = t;
}

const double* Multigrid::vcycle(const double* r, const PcgState* st)
{
    Buffers& B = *_buf;
    const int last = (int)_lev.size() - 1;
    double* x = B.z[0].ptr;
    double* t = B.z[1].ptr;
    const int pre = last == 0 ? _coarse : _nu;
    launch_mg0_first(_g, B.wd0, r, x, st, _s);
    for (int it = 1; it < pre; ++it) {
        launch_mg0_sweep(_g, B.wd0, x, r, t, st, _s);
        std::swap(x, t);
    }
    if (last > 0) {
        launch_mg0_residual(_g, x, r, B.r0.ptr, st, _s);
        launch_mg0_restrict(_g, _lev[1], B.r0.ptr, B.b[1], st, _s);
        cycle(1, B.b[1], st);
        launch_mg0_prolong(_g, _lev[1], B.xcur[1], x, st, _s);
        for (int it = 0; it < _nu; ++it) {
            launch_mg0_sweep(_g, B.wd0, x, r, t, st, _s);
            std::swap(x, t);
        }
    }
    return x;
}

} // namespace cfmi
