// binding_demo.cpp -- a C++ caller of libcfmi.so through the reference-side
// adapter (include/cf_hip_solver.hpp), shaped like examples/advection.cpp's
// advect(): SOLID walls, an InflowSource box, a BodyForce, an init functor,
// createSolver's arguments.  Runs setup + nsteps steps of an n x n dye tank
// and writes q, u, v (dense ghosted, binary doubles) to out_path.
//
// usage: binding_demo n nsteps out_path
// The init functor is not constant (q = x (1 - x) y, u = x y - 1/4,
// v = 1/2 - x y) so that the coordinates it receives are checked too.
#include "reference_types.hpp"

#include <cf_hip_solver.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

struct DemoInit
{
    bool operator()(Cajita::Cell, CajitaFluids::Field::Quantity, const int[2],
                    const double x[2], double& q) const
    {
        q = (x[0] * (1.0 - x[0])) * x[1];
        return true;
    }
    bool operator()(Cajita::Face<Cajita::Dim::I>, CajitaFluids::Field::Velocity,
                    const int[2], const double x[2], double& u) const
    {
        u = x[0] * x[1] - 0.25;
        return true;
    }
    bool operator()(Cajita::Face<Cajita::Dim::J>, CajitaFluids::Field::Velocity,
                    const int[2], const double x[2], double& v) const
    {
        v = 0.5 - x[0] * x[1];
        return true;
    }
};

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s n nsteps out_path\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const int nsteps = std::atoi(argv[2]);
    try {
        CajitaFluids::BoundaryCondition<2> bc;
        bc.boundary_type = {CajitaFluids::BoundaryType::SOLID, CajitaFluids::BoundaryType::SOLID,
                            CajitaFluids::BoundaryType::SOLID, CajitaFluids::BoundaryType::SOLID};
        CajitaFluids::InflowSource<2> source({0.45, 0.20}, {0.10, 0.02}, {0.0, 3.0}, 1.0);
        CajitaFluids::BodyForce<2> body(0.0, -0.5);
        Kokkos::Array<double, 4> bbox = {0.0, 0.0, 1.0, 1.0};
        CajitaFluids::HipSolver2 solver(bbox, {n, n}, 0.1, DemoInit(), bc, source, body, 0.005,
                                        "Reference", "None", "");
        solver.setup();
        for (int s = 0; s < nsteps; ++s)
            solver.step();
        FILE* f = std::fopen(argv[3], "wb");
        if (!f)
            throw std::runtime_error("cannot open output");
        for (int e = CF_CELL; e <= CF_FACE_Y; ++e) {
            int ext[3], lo[3], hi[3];
            cf_layout_query(&solver.config(), e, ext, lo, hi);
            std::vector<double> a((size_t)ext[0] * ext[1] * ext[2]);
            CajitaFluids::hipcf::check(solver.handle(),
                                       cf_copy_out(solver.handle(), e, CF_CURRENT, a.data()),
                                       "cf_copy_out");
            std::fwrite(a.data(), sizeof(double), a.size(), f);
        }
        std::fclose(f);
        std::printf("binding_demo: %d^2, %d steps, time %.6f\n", n, nsteps,
                    cf_time(solver.handle()));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "binding_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
