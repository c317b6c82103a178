/*
 * cf_hip_solver.hpp -- the reference-side adapter: a CajitaFluids::SolverBase
 * (src/Solver.hpp:41-48) that runs on libcfmi.so, for a maintainer to add to
 * the reference tree next to src/Solver.hpp and return from createSolver
 * (src/Solver.hpp:283-350) for the device name "hip".
 *
 * Include it after the reference's <Solver.hpp>: it reads the reference's own
 * value objects by their members --
 *   BoundaryCondition<2>::boundary_type (BoundaryConditions.hpp:131-134;
 *     order left, bottom, right, top: BoundaryConditions.hpp:66-97),
 *   InflowSource<2>::_bounding_box, _velocity, _quantity
 *     (InflowSource.hpp:80-97; the box is [bb[0], bb[2]) x [bb[1], bb[3])),
 *   BodyForce<2>::_force (BodyForce.hpp:62-68) --
 * and calls the init functor exactly as ProblemManager::initialize does
 * (ProblemManager.hpp:186-263): create_functor(entity tag, field tag,
 * coords, x, value) over the owned index space of each entity, with local
 * coords {i, j} and x = loc + cellSize * i, loc being the coordinate of local
 * index 0.  The library runs one rank: the communicator must have size 1.
 */
#ifndef CF_HIP_SOLVER_HPP
#define CF_HIP_SOLVER_HPP

#include <cf_api.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace CajitaFluids
{
namespace hipcf
{
inline void check(const cf_solver* h, cf_status rc, const char* what)
{
    if (rc != CF_OK)
        throw std::runtime_error(std::string("libcfmi ") + what + ": " + cf_last_error(h));
}

// A size s with lo + s == hi exactly, so that the library's half-open box
// [lo, lo + s) is InflowSource's [_bounding_box[d], _bounding_box[d + 2]) to
// the last bit (InflowSource.hpp:40-41 compares x < _bounding_box[2]).
inline double exact_extent(double lo, double hi)
{
    double s = hi - lo;
    while (lo + s < hi)
        s = std::nextafter(s, HUGE_VAL);
    while (lo + s > hi)
        s = std::nextafter(s, -HUGE_VAL);
    return s;
}

// createVelocityCorrector's choice (VelocityCorrector.hpp:297-339):
// "Reference" -> Jacobi-PCG; "PCG" + "PFMG"/"SMG" -> the multigrid V-cycle,
// + "None"/"none" -> plain CG, + "Jacobi"/"Diagonal" -> Jacobi-PCG.
inline int preconditioner_of(const std::string& solver, const std::string& precon)
{
    if (solver == "Reference")
        return CF_PRECOND_JACOBI;
    if (solver != "PCG")
        throw std::runtime_error("libcfmi: matrix solver " + solver + " is not provided");
    if (precon == "PFMG" || precon == "SMG")
        return CF_PRECOND_MG;
    if (precon == "None" || precon == "none")
        return CF_PRECOND_NONE;
    if (precon == "Jacobi" || precon == "Diagonal")
        return CF_PRECOND_JACOBI;
    throw std::runtime_error("libcfmi: preconditioner " + precon + " is not provided");
}

// ProblemManager::initialize (ProblemManager.hpp:186-263) into the handle's
// Current arrays: every owned entity of each kind gets the functor's value;
// ghosts stay 0 (ProblemManager.hpp:149-165).
template <class InitFunc>
void initialize(cf_solver* h, const cf_config& c, const InitFunc& create_functor)
{
    double cell_size = 0.0;
    check(nullptr, cf_geometry_query(&c, &cell_size, nullptr), "cf_geometry_query");
    for (int e = CF_CELL; e <= CF_FACE_Y; ++e) {
        int ext[3], lo[3], hi[3];
        double loc[3];
        check(nullptr, cf_layout_query(&c, e, ext, lo, hi), "cf_layout_query");
        check(nullptr, cf_entity_origin(&c, e, loc), "cf_entity_origin");
        std::vector<double> a((size_t)ext[0] * ext[1] * ext[2], 0.0);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j) {
                int coords[2] = {i, j};
                double x[2];
                x[0] = loc[0] + cell_size * i;
                x[1] = loc[1] + cell_size * j;
                double& value = a[(size_t)j * ext[0] + i];
                if (e == CF_CELL)
                    create_functor(Cajita::Cell(), Field::Quantity(), coords, x, value);
                else if (e == CF_FACE_X)
                    create_functor(Cajita::Face<Cajita::Dim::I>(), Field::Velocity(), coords,
                                   x, value);
                else
                    create_functor(Cajita::Face<Cajita::Dim::J>(), Field::Velocity(), coords,
                                   x, value);
            }
        check(h, cf_copy_in(h, e, CF_CURRENT, a.data()), "cf_copy_in");
    }
}
} // namespace hipcf

class HipSolver2 : public SolverBase
{
  public:
    // The arguments of Solver<2, ...>'s constructor (Solver.hpp:69-77)
    // without the communicator and partitioner (one rank).
    template <class InitFunc>
    HipSolver2(const Kokkos::Array<double, 4>& global_bounding_box,
               const std::array<int, 2>& global_num_cell, const double density,
               const InitFunc& create_functor, const BoundaryCondition<2>& bc,
               const InflowSource<2>& source, const BodyForce<2>& body,
               const double delta_t, const std::string& matrix_solver,
               const std::string& preconditioner, const std::string& output_dir = "data")
        : _output_dir(output_dir)
    {
        cf_config c = {};
        c.dim = 2;
        c.num_cell[0] = global_num_cell[0];
        c.num_cell[1] = global_num_cell[1];
        c.num_cell[2] = 1;
        c.halo = 3; // Solver.hpp:78
        for (int d = 0; d < 2; ++d) {
            c.lo[d] = global_bounding_box[d];
            c.hi[d] = global_bounding_box[d + 2];
        }
        c.lo[2] = 0.0;
        c.hi[2] = 1.0;
        auto wall = [](int t) { return t == BoundaryType::SOLID ? CF_SOLID : CF_FREE; };
        c.wall_type[0][0] = wall(bc.boundary_type[0]); // left
        c.wall_type[1][0] = wall(bc.boundary_type[1]); // bottom
        c.wall_type[0][1] = wall(bc.boundary_type[2]); // right
        c.wall_type[1][1] = wall(bc.boundary_type[3]); // top
        c.wall_type[2][0] = c.wall_type[2][1] = CF_SOLID;
        c.density = density;
        c.dt = delta_t;
        c.clamp_dt = 1; // Solver.hpp:93-106
        c.body_force[0] = body._force[0];
        c.body_force[1] = body._force[1];
        c.inflow_shape = CF_INFLOW_BOX;
        for (int d = 0; d < 2; ++d) {
            c.inflow_location[d] = source._bounding_box[d];
            c.inflow_size[d] =
                hipcf::exact_extent(source._bounding_box[d], source._bounding_box[d + 2]);
            c.inflow_velocity[d] = source._velocity[d];
        }
        c.inflow_quantity = source._quantity;
        c.field_order = 3;    // TimeIntegrator.hpp:113
        c.velocity_order = 1; // TimeIntegrator.hpp:45, 53, 61
        c.cg_tol = 1e-6;      // VelocityCorrector.hpp:103
        c.cg_max_iter = 2000; // VelocityCorrector.hpp:104
        c.precision = CF_FP64;
        c.inner_tol = 1e-2;
        c.inner_max_iter = 2000;
        c.max_outer = 50;
        c.quirks = CF_QUIRK_RK3_V0 | CF_QUIRK_PROJ_BC_U; // as written
        c.refine_mode = CF_REFINE_CARRY;
        c.preconditioner = hipcf::preconditioner_of(matrix_solver, preconditioner);
        c.mg_nu = 3;
        c.mg_omega = 0.9;
        c.mg_coarse_sweeps = 16;
        c.mg_alpha = 1.9;
        _cfg = c;
        hipcf::check(nullptr, cf_create(&_cfg, &_h), "cf_create");
        double dt = delta_t;
        hipcf::check(_h, cf_geometry_query(&_cfg, nullptr, &dt), "cf_geometry_query");
        if (dt != delta_t) // the reference's message (Solver.hpp:101-105)
            std::cerr << "Reducting timestep to " << dt
                      << " given mesh size and inflow velocity.\n";
        hipcf::initialize(_h, _cfg, create_functor);
    }
    HipSolver2(const HipSolver2&) = delete;
    HipSolver2& operator=(const HipSolver2&) = delete;
    ~HipSolver2() override { cf_destroy(_h); }

    void setup() override { hipcf::check(_h, cf_setup(_h, nullptr), "cf_setup"); }
    void step() override { hipcf::check(_h, cf_step(_h, nullptr), "cf_step"); }

    // Solver::solve (Solver.hpp:148-176) with the library's writer in place
    // of SiloWriter::siloWrite (the same file names, as VTK XML).
    void solve(const double t_final, const int write_freq) override
    {
        int t = 0;
        write(t);
        const int num_step = (int)(t_final / cf_dt(_h));
        setup();
        do {
            if (0 == t % write_freq)
                std::printf("Step %d / %d at time = %f\n", t, num_step, cf_time(_h));
            step();
            if (0 == t % write_freq)
                write(t);
            t++;
        } while (cf_time(_h) < t_final);
    }

    cf_solver* handle() const { return _h; }
    const cf_config& config() const { return _cfg; }

  private:
    void write(int t)
    {
        if (!_output_dir.empty())
            hipcf::check(_h, cf_write_output(_h, _output_dir.c_str(), t), "cf_write_output");
    }
    cf_config _cfg;
    cf_solver* _h = nullptr;
    std::string _output_dir;
};

} // namespace CajitaFluids

#endif
