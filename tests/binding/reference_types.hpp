/*
 * reference_types.hpp -- compile harness for include/cf_hip_solver.hpp.
 *
 * The adapter is meant to be included after the reference's <Solver.hpp>,
 * which needs Kokkos, Cabana/Cajita and MPI (none of which are in this
 * image).  This file declares only what the adapter itself reads, with the
 * reference's names, member names and layouts, so that the adapter compiles
 * and runs against libcfmi.so here:
 *   Kokkos::Array<T, N> (operator[]), the Cajita entity tags Cell and
 *   Face<Dim::I / J>, CajitaFluids::Field::{Quantity, Velocity},
 *   BoundaryType::Values (BoundaryConditions.hpp:29-36),
 *   BoundaryCondition<2>::boundary_type (:131-134), InflowSource<2>'s
 *   constructor and members (InflowSource.hpp:80-97), BodyForce<2>'s
 *   constructor and _force (BodyForce.hpp:62-68), and SolverBase
 *   (Solver.hpp:41-48).
 * It is not a build of the reference: no behaviour of the reference is
 * restated here.
 */
#pragma once

#include <array>
#include <cstddef>

namespace Kokkos
{
template <class T, std::size_t N>
struct Array
{
    T data_[N];
    T& operator[](std::size_t i) { return data_[i]; }
    const T& operator[](std::size_t i) const { return data_[i]; }
};
} // namespace Kokkos

namespace Cajita
{
struct Cell
{
};
struct Dim
{
    enum Values
    {
        I = 0,
        J = 1,
        K = 2
    };
};
template <int D>
struct Face
{
};
} // namespace Cajita

namespace CajitaFluids
{
namespace Field
{
struct Quantity
{
};
struct Velocity
{
};
} // namespace Field

struct BoundaryType
{
    enum Values
    {
        SOLID = 0,
        FREE = 1
    };
};

template <std::size_t NumSpaceDim>
struct BoundaryCondition;
template <>
struct BoundaryCondition<2>
{
    Kokkos::Array<int, 4> boundary_type;
    Kokkos::Array<int, 2> min;
    Kokkos::Array<int, 2> max;
};

template <std::size_t NumSpaceDim>
struct InflowSource;
template <>
struct InflowSource<2>
{
    InflowSource(std::array<double, 2> location, std::array<double, 2> size,
                 std::array<double, 2> velocity, double quantity)
        : _quantity(quantity)
    {
        _bounding_box[0] = location[0];
        _bounding_box[1] = location[1];
        _bounding_box[2] = location[0] + size[0];
        _bounding_box[3] = location[1] + size[1];
        _velocity[0] = velocity[0];
        _velocity[1] = velocity[1];
    }
    Kokkos::Array<double, 4> _bounding_box;
    double _quantity;
    Kokkos::Array<double, 2> _velocity;
};

template <std::size_t NumSpaceDim>
struct BodyForce;
template <>
struct BodyForce<2>
{
    BodyForce(double fx, double fy)
    {
        _force[0] = fx;
        _force[1] = fy;
    }
    Kokkos::Array<double, 2> _force;
};

class SolverBase
{
  public:
    virtual ~SolverBase() = default;
    virtual void setup(void) = 0;
    virtual void step(void) = 0;
    virtual void solve(const double t_final, const int write_freq) = 0;
};
} // namespace CajitaFluids
