// 2-D compressible Euler on a mesh read from a gmsh file through the C++ drop-in driver
// (setMesh(path, generate), the reference's System::setMesh, SystemControl.cpp:55-61): an existing
// .msh is read; for a path that does not exist the callback generates a triangulation of
// [0, 2] x [0, 1] with jittered interior nodes in memory. Physical group 1 is a Riemann far field
// (x = 0 and x = 2), group 2 an adiabatic slip wall (y = 0 and y = 1); a uniform stream along x
// must stay uniform (free-stream preservation on any affine mesh), so the residual monitor stays
// at rounding level.
// Usage: mesh_file_2d_ceuler <mesh.msh> [order: 2 or 3] [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "subrosadg_amd/SubrosaDG.hpp"

template <SubrosaDG::PolynomialOrderEnum Order>
using Control = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D2, Order, SubrosaDG::BoundaryTimeEnum::Steady,
                            SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Triangle, SubrosaDG::ShockCapturingEnum::None,
                                SubrosaDG::LimiterEnum::None, SubrosaDG::InitialConditionEnum::Function,
                                SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleEulerVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                        SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::ConvectiveFluxEnum::HLLC>>;
using P2 = Control<SubrosaDG::PolynomialOrderEnum::P2>;
using P3 = Control<SubrosaDG::PolynomialOrderEnum::P3>;

namespace {
constexpr std::array<double, 4> kStream{1.0, 0.5, 0.0, 1.0};  // rho, u, v, T (R = 1)
}

template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::InitialCondition<P2>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x) const {
  return kStream;
}
template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::InitialCondition<P3>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x) const {
  return kStream;
}
template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::BoundaryCondition<P2>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x, [[maybe_unused]] SubrosaDG::Real time,
    SubrosaDG::Isize physical_index) const {
  if (physical_index != 1 && physical_index != 2) throw std::runtime_error("unexpected physical group");
  return kStream;
}
template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::BoundaryCondition<P3>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x, [[maybe_unused]] SubrosaDG::Real time,
    SubrosaDG::Isize physical_index) const {
  if (physical_index != 1 && physical_index != 2) throw std::runtime_error("unexpected physical group");
  return kStream;
}

// nx x ny cells of two triangles each, interior nodes moved by up to 0.2 of a cell
void generateMesh(SubrosaDG::ArrayMesh& mesh) {
  constexpr int nx = 10, ny = 5;
  mesh.element = SDG_ELEM_TRIANGLE;
  const auto id = [](int i, int j) { return static_cast<int64_t>(j * (nx + 1) + i); };
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      double x = 2.0 * i / nx, y = 1.0 * j / ny;
      if (i > 0 && i < nx && j > 0 && j < ny) {
        x += 0.2 * (2.0 / nx) * std::sin(3.1 * i + 1.7 * j);
        y += 0.2 * (1.0 / ny) * std::cos(2.3 * i - 0.9 * j);
      }
      mesh.nodes.push_back({x, y, 0.0});
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      mesh.elements.insert(mesh.elements.end(), {id(i, j), id(i + 1, j), id(i + 1, j + 1)});
      mesh.elements.insert(mesh.elements.end(), {id(i, j), id(i + 1, j + 1), id(i, j + 1)});
    }
  for (int j = 0; j < ny; j++) {
    mesh.boundary_faces.insert(mesh.boundary_faces.end(), {id(0, j), id(0, j + 1), id(nx, j), id(nx, j + 1)});
    mesh.boundary_physical.insert(mesh.boundary_physical.end(), {1, 1});
  }
  for (int i = 0; i < nx; i++) {
    mesh.boundary_faces.insert(mesh.boundary_faces.end(), {id(i, 0), id(i + 1, 0), id(i, ny), id(i + 1, ny)});
    mesh.boundary_physical.insert(mesh.boundary_physical.end(), {2, 2});
  }
}

template <typename SC>
int run(const std::string& file, int steps) {
  SubrosaDG::System<SC> system;
  system.setMesh(std::filesystem::path(file), generateMesh);
  system.template addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::RiemannFarfield>(1);
  system.template addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::AdiabaticSlipWall>(2);
  system.template setThermodynamicModel<SC::kThermodynamicModel>(3.5, 2.5);
  system.setTimeIntegration(0.1, {0, steps});
  double worst = 0.0;
  system.setStepCallback([&worst](int, const std::array<double, 4>& r) {
    for (double v : r) worst = std::fmax(worst, std::isfinite(v) ? v : 1e300);
  });
  system.synchronize();
  system.solve();
  std::printf("mesh file: %d steps, largest residual monitor %.3e\n", steps, worst);
  return worst < 1e-10 ? EXIT_SUCCESS : EXIT_FAILURE;
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <mesh.msh> [order 2|3] [steps]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const int order = argc > 2 ? std::atoi(argv[2]) : 3;
  const int steps = argc > 3 ? std::atoi(argv[3]) : 50;
  try {
    return order == 2 ? run<P2>(argv[1], steps) : run<P3>(argv[1], steps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
