// 3-D compressible Navier-Stokes differentially heated cavity with the Boussinesq source term
// (SourceTermEnum::Boussinesq, setSourceTerm; the compressible counterpart of the reference's
// thermalcavity examples) through the C++ drop-in driver: a hot isothermal wall at x = 0, a cold
// one at x = 1, adiabatic no-slip walls at z = 0 and z = 1, periodic in y; buoyancy
// rho beta (T - T_ref) acts on the z momentum.
// Usage: thermal_cavity_3d_cns [cells per direction] [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "subrosadg_amd/SubrosaDG.hpp"

using SimulationControl = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D3, SubrosaDG::PolynomialOrderEnum::P3,
                            SubrosaDG::BoundaryTimeEnum::Steady, SubrosaDG::SourceTermEnum::Boussinesq>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Hexahedron, SubrosaDG::ShockCapturingEnum::None,
                                SubrosaDG::LimiterEnum::None, SubrosaDG::InitialConditionEnum::Function,
                                SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleNSVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                     SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::TransportModelEnum::Constant,
                                     SubrosaDG::ConvectiveFluxEnum::Roe, SubrosaDG::ViscousFluxEnum::BR2>>;

namespace {
constexpr double kGamma = 1.4, kMach = 0.1;
constexpr double kP0 = 1.0 / (kGamma * kMach * kMach);  // R = p0, T0 = 1
constexpr int kHot = 1, kCold = 2, kLid = 3;
}  // namespace

template <>
std::array<SubrosaDG::Real, 5> SubrosaDG::InitialCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 3>& x) const {
  return {1.0, 0.0, 0.0, 0.0, 1.0};
}

template <>
std::array<SubrosaDG::Real, 5> SubrosaDG::BoundaryCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 3>& x, [[maybe_unused]] SubrosaDG::Real time,
    SubrosaDG::Isize physical_index) const {
  const double t = physical_index == kHot ? 1.05 : (physical_index == kCold ? 0.95 : 1.0);
  return {1.0, 0.0, 0.0, 0.0, t};
}

int main(int argc, char* argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 6;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 100;
  try {
    SubrosaDG::System<SimulationControl> system;
    system.setMesh({{n, n, n}, {0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {kHot, kCold, -1, -1, kLid, kLid}});
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::IsoThermalNonSlipWall>(kHot);
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::IsoThermalNonSlipWall>(kCold);
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::AdiabaticNonSlipWall>(kLid);
    const double cv = kP0 / (kGamma - 1.0);
    system.setThermodynamicModel<SimulationControl::kThermodynamicModel>(kGamma * cv, cv);
    system.setTransportModel<SimulationControl::kTransportModel>(1.0e-2);
    system.setSourceTerm(1.0, 1.0);  // beta, T_ref
    system.setTimeIntegration(0.1, {0, steps});
    system.setStepCallback([steps](int i, const std::array<double, 5>& r) {
      if (i % 50 == 0 || i == steps)
        std::printf("step %d  monitor %.6e %.6e %.6e %.6e %.6e\n", i, r[0], r[1], r[2], r[3], r[4]);
    });
    system.synchronize();
    system.solve();
    // the z momentum, zero at the start: largest |rho w| coefficient (NV = 5, rho w = index 3)
    const auto c = system.state();
    double wmax = 0.0;
    bool finite = true;
    for (size_t i = 0; i < c.size(); i++) {
      finite = finite && std::isfinite(c[i]);
      if (i % 5 == 3) wmax = std::fmax(wmax, std::fabs(c[i]));
    }
    std::printf("max |rho w| coefficient %.6e\n", wmax);
    return finite && wmax > 0.0 ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
