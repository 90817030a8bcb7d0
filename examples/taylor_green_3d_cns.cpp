// 3-D compressible Navier-Stokes Taylor-Green vortex (Re 1600, Ma 0.1) on the GPU engine through
// the C++ drop-in driver: the same SimulationControl / System usage as the reference's examples
// (e.g. examples/periodic_3d_ceuler.cpp), with a structured periodic box instead of a gmsh mesh.
// Usage: taylor_green_3d_cns [cells per direction] [steps] [output directory for the final raw state]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "subrosadg_amd/SubrosaDG.hpp"

using SimulationControl = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D3, SubrosaDG::PolynomialOrderEnum::P3,
                            SubrosaDG::BoundaryTimeEnum::Steady, SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Hexahedron, SubrosaDG::ShockCapturingEnum::None,
                                SubrosaDG::LimiterEnum::None, SubrosaDG::InitialConditionEnum::Function,
                                SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleNSVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                     SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::TransportModelEnum::Constant,
                                     SubrosaDG::ConvectiveFluxEnum::HLLC, SubrosaDG::ViscousFluxEnum::BR2>>;

namespace {
constexpr double kGamma = 1.4, kMach = 0.1;
constexpr double kP0 = 1.0 / (kGamma * kMach * kMach);  // R = p0, T0 = 1
}  // namespace

template <>
std::array<SubrosaDG::Real, 5> SubrosaDG::InitialCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 3>& x) const {
  const double p = kP0 + (std::cos(2 * x[0]) + std::cos(2 * x[1])) * (std::cos(2 * x[2]) + 2.0) / 16.0;
  return {p / kP0, std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]), -std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]),
          0.0, 1.0};
}

int main(int argc, char* argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 8;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 10;
  SubrosaDG::System<SimulationControl> system;
  system.setMesh({{n, n, n}, {0.0, 0.0, 0.0}, {2 * M_PI, 2 * M_PI, 2 * M_PI}});
  system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::Periodic>(1);
  const double cv = kP0 / (kGamma - 1.0);
  system.setThermodynamicModel<SimulationControl::kThermodynamicModel>(kGamma * cv, cv);
  system.setTransportModel<SimulationControl::kTransportModel>(1.0 / 1600.0);
  system.setTimeIntegration(0.1, {0, steps});
  if (argc > 3) system.setViewConfig(argv[3], "tgv", steps);  // final state: raw/tgv_<steps>.zst
  system.setStepCallback([](int i, const std::array<double, 5>& r) {
    std::printf("step %d  monitor %.6e %.6e %.6e %.6e %.6e\n", i, r[0], r[1], r[2], r[3], r[4]);
  });
  try {
    system.synchronize();
    system.solve();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
