// 2-D compressible Navier-Stokes double shear layer through the C++ drop-in driver, with raw state
// files (setViewConfig) and the two file-based starts of the reference:
//   1. Function initial condition, N steps, a raw file every N/2 steps (raw/shear_<step>.zst);
//   2. InitialConditionEnum::LastStep from step N/2 to N: must end bit-identical to run 1;
//   3. InitialConditionEnum::SpecificFile: an order P+1 run started from run 1's final order-P file;
//   4. LastStep again with no setDeltaTime: the time step comes from run 1's error.txt (the time of
//      its first step, TimeIntegration.cpp:138-146) and the run must again end bit-identical.
// Usage: shear_layer_2d_cns_restart <output directory> [cells per direction] [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "subrosadg_amd/SubrosaDG.hpp"

namespace {
constexpr double kGamma = 1.4, kMach = 0.1;
constexpr double kP0 = 1.0 / (kGamma * kMach * kMach);  // R = p0, T0 = 1
constexpr double kPi = 3.14159265358979323846;

template <SubrosaDG::PolynomialOrderEnum Order, SubrosaDG::InitialConditionEnum Start>
using Control = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D2, Order, SubrosaDG::BoundaryTimeEnum::Steady,
                            SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Quadrangle, SubrosaDG::ShockCapturingEnum::None,
                                SubrosaDG::LimiterEnum::None, Start, SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleNSVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                     SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::TransportModelEnum::Constant,
                                     SubrosaDG::ConvectiveFluxEnum::Roe, SubrosaDG::ViscousFluxEnum::BR2>>;
using Fresh = Control<SubrosaDG::PolynomialOrderEnum::P3, SubrosaDG::InitialConditionEnum::Function>;
using Restart = Control<SubrosaDG::PolynomialOrderEnum::P3, SubrosaDG::InitialConditionEnum::LastStep>;
using Refined = Control<SubrosaDG::PolynomialOrderEnum::P4, SubrosaDG::InitialConditionEnum::SpecificFile>;

template <typename SC>
void configure(SubrosaDG::System<SC>& system, int n, int first, int last, double dt, const std::string& dir) {
  system.setMesh({{n, n, 1}, {0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}});
  system.template addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::Periodic>(1);
  const double cv = kP0 / (kGamma - 1.0);
  system.template setThermodynamicModel<SC::kThermodynamicModel>(kGamma * cv, cv);
  system.template setTransportModel<SC::kTransportModel>(1.0e-4);
  system.setTimeIntegration(0.1, {first, last});
  if (dt > 0.0) system.setDeltaTime(dt);
  system.setViewConfig(dir, "shear", (last - first) > 1 ? (last - first) / 2 : 1);
}
}  // namespace

template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::InitialCondition<Fresh>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 2>& x) const {
  const double u = x[1] <= 0.5 ? std::tanh(80.0 * (x[1] - 0.25)) : std::tanh(80.0 * (0.75 - x[1]));
  return {1.0, u, 0.05 * std::sin(2.0 * kPi * x[0]), 1.0};
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <output directory> [cells] [steps]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const std::string dir = argv[1];
  const int n = argc > 2 ? std::atoi(argv[2]) : 16;
  const int steps = argc > 3 ? std::atoi(argv[3]) : 20;
  const int half = steps / 2;
  const double dt = 2.0e-5;
  try {
    SubrosaDG::System<Fresh> first;
    configure(first, n, 0, steps, dt, dir);
    first.synchronize();
    first.solve();
    const auto reference = first.state();

    SubrosaDG::System<Restart> second;  // reads raw/shear_<half>.zst written by the first run
    configure(second, n, half, steps, dt, dir + "/restart");
    second.setViewConfig(dir, "shear", 1 << 30);  // read the first run's files, write none
    second.synchronize();
    second.solve();
    const auto restarted = second.state();
    double diff = 0.0;
    for (size_t i = 0; i < reference.size(); i++) diff = std::fmax(diff, std::fabs(restarted[i] - reference[i]));
    std::printf("last-step restart: max |difference| = %.3e over %zu coefficients\n", diff, reference.size());

    SubrosaDG::System<Refined> third;  // order 4 from the order-3 file of the last step
    configure(third, n, 0, 2, dt, dir + "/refined");
    third.addInitialCondition<SubrosaDG::InitialConditionEnum::SpecificFile>(dir + "/raw/shear_" +
                                                                           std::to_string(steps) + ".zst");
    third.synchronize();
    third.solve();
    std::printf("specific-file start at order 4: %zu coefficients, 2 steps\n", third.state().size());

    SubrosaDG::System<Restart> fourth;  // time step from the log of the first run
    configure(fourth, n, half, steps, 0.0, dir);
    fourth.setViewConfig(dir, "shear", 1 << 30);
    fourth.synchronize();
    fourth.solve();
    const auto logged = fourth.state();
    double diff4 = 0.0;
    for (size_t i = 0; i < reference.size(); i++) diff4 = std::fmax(diff4, std::fabs(logged[i] - reference[i]));
    std::printf("last-step restart with the logged time step %.6e: max |difference| = %.3e\n", fourth.deltaTime(),
                diff4);
    return diff == 0.0 && diff4 == 0.0 ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
