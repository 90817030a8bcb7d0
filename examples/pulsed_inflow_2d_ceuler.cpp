// 2-D compressible Euler channel with time-varying boundary data (BoundaryTimeEnum::TimeVarying)
// through the C++ drop-in driver: a velocity inflow at x = 0 whose velocity and temperature pulse in
// time, a pressure outflow at x = 1 and a far-field stream turning in time at y = 0 and y = 1; P3
// quadrangles, Roe flux, SSP-RK3 with a fixed time step. The boundary callback is evaluated before
// every step at the time of the previous step (BoundaryCondition.cpp:29-50, TimeIntegration.cpp:334-336).
// Usage: pulsed_inflow_2d_ceuler <output directory> [cells per direction] [steps] [time step]
// Writes <output directory>/raw/pulse_<steps>.zst (the final state) and error.txt.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "subrosadg_amd/SubrosaDG.hpp"

using SimulationControl = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D2, SubrosaDG::PolynomialOrderEnum::P3,
                            SubrosaDG::BoundaryTimeEnum::TimeVarying, SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Quadrangle, SubrosaDG::ShockCapturingEnum::None,
                                SubrosaDG::LimiterEnum::None, SubrosaDG::InitialConditionEnum::Function,
                                SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleEulerVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                        SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::ConvectiveFluxEnum::Roe>>;

namespace {
constexpr int kInflow = 1, kOutflow = 2, kFarfield = 3;
constexpr double kPi = 3.14159265358979323846;
}  // namespace

template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::InitialCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x) const {
  return {1.0, 0.3, 0.2, 1.0};
}

template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::BoundaryCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 2>& x, const SubrosaDG::Real time, const SubrosaDG::Isize physical_index) const {
  if (physical_index == kInflow)
    return {1.0, 0.3 + 0.1 * std::sin(40.0 * time) * std::cos(kPi * x[1]), 0.2 * std::cos(25.0 * time),
            1.0 + 0.05 * std::sin(30.0 * time)};
  if (physical_index == kFarfield) return {1.0, 0.3, 0.2 * std::cos(25.0 * time), 1.0};
  return {1.0, 0.3, 0.2, 1.0};  // outflow: the pressure of rho = 1, T = 1
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <output directory> [cells] [steps] [time step]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const std::string dir = argv[1];
  const int n = argc > 2 ? std::atoi(argv[2]) : 8;
  const int steps = argc > 3 ? std::atoi(argv[3]) : 10;
  const double dt = argc > 4 ? std::atof(argv[4]) : 2.0e-3;
  try {
    SubrosaDG::System<SimulationControl> system;
    system.setMesh({{n, n, 1}, {0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {kInflow, kOutflow, kFarfield, kFarfield, -1, -1}});
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::VelocityInflow>(kInflow);
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::PressureOutflow>(kOutflow);
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::RiemannFarfield>(kFarfield);
    system.setThermodynamicModel<SimulationControl::kThermodynamicModel>(3.5, 2.5);
    system.setTimeIntegration(0.1, {0, steps});
    system.setDeltaTime(dt);
    system.setViewConfig(dir, "pulse", steps);
    system.synchronize();
    system.solve();
    std::printf("pulsed inflow: %d steps of %.6e\n", steps, dt);
    return EXIT_SUCCESS;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
