// 2-D compressible Euler explosion (the reference's explosion_2d_ceuler case) through the C++
// drop-in driver with artificial-viscosity shock capturing: a circle of radius 0.4 around (1, 1)
// at rho = 1, p = 1 in a gas at rho = 0.125, p = 0.1, in the box [0, 2]^2 closed by adiabatic slip
// walls; P3 quadrangles on a structured grid, HLLC flux, SSP-RK3, setArtificialViscosity(4.0).
// The reference meshes the circle into the element edges (gmsh); on the structured grid the circle
// cuts through elements, where the least-squares projection of the jump would undershoot to
// negative pressure, so the initial jump is spread over one cell (a tanh profile of width h/2).
// Usage: explosion_2d_ceuler [cells per direction] [steps] [output directory]
// With an output directory: raw files every steps/2 steps and their VTU files (system.view()).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "subrosadg_amd/SubrosaDG.hpp"

using SimulationControl = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D2, SubrosaDG::PolynomialOrderEnum::P3,
                            SubrosaDG::BoundaryTimeEnum::Steady, SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Quadrangle,
                                SubrosaDG::ShockCapturingEnum::ArtificialViscosity, SubrosaDG::LimiterEnum::None,
                                SubrosaDG::InitialConditionEnum::Function, SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleEulerVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                        SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::ConvectiveFluxEnum::HLLC>>;

namespace {
constexpr int kWall = 1;
constexpr double kGasConstant = 2.5 - 25.0 / 14.0;  // R = cp - cv = 5/7
double g_width = 0.025;                               // half a cell
}  // namespace

// inside: rho = 1, p = 1 (T = 1.4); outside: rho = 0.125, p = 0.1 (T = 1.12)
template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::InitialCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 2>& x) const {
  const double r = std::sqrt((x[0] - 1.0) * (x[0] - 1.0) + (x[1] - 1.0) * (x[1] - 1.0));
  const double w = 0.5 * (1.0 - std::tanh((r - 0.4) / g_width));
  const double rho = 0.125 + 0.875 * w, p = 0.1 + 0.9 * w;
  return {rho, 0.0, 0.0, p / (rho * kGasConstant)};
}

template <>
std::array<SubrosaDG::Real, 4> SubrosaDG::BoundaryCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    [[maybe_unused]] const std::array<SubrosaDG::Real, 2>& x, [[maybe_unused]] SubrosaDG::Real time,
    [[maybe_unused]] SubrosaDG::Isize physical_index) const {
  return {0.0, 0.0, 0.0, 0.0};  // slip walls take no boundary data
}

int main(int argc, char* argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 40;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 200;
  const std::string dir = argc > 3 ? argv[3] : "";
  g_width = 1.0 / n;
  try {
    SubrosaDG::System<SimulationControl> system;
    system.setMesh({{n, n, 1}, {0.0, 0.0, 0.0}, {2.0, 2.0, 1.0}, {kWall, kWall, kWall, kWall, -1, -1}});
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::AdiabaticSlipWall>(kWall);
    system.setThermodynamicModel<SimulationControl::kThermodynamicModel>(2.5, 25.0 / 14.0);
    system.setArtificialViscosity(4.0);
    system.setTimeIntegration(0.1, {0, steps});
    if (!dir.empty()) system.setViewConfig(dir, "explosion", steps > 1 ? steps / 2 : 1);
    bool finite = true;
    system.setStepCallback([&finite, steps](int i, const std::array<double, 4>& r) {
      for (double v : r) finite = finite && std::isfinite(v);
      if (i % 50 == 0 || i == steps) std::printf("step %d  monitor %.6e %.6e %.6e %.6e\n", i, r[0], r[1], r[2], r[3]);
    });
    system.synchronize();
    system.solve();
    if (!dir.empty()) system.view();
    const auto av = system.artificialViscosity();
    const auto c = system.state();
    double rho_min = 1e300;  // vertex-mode density coefficients: the density at the element corners
    for (size_t e = 0; e + 16 * 4 <= c.size(); e += 16 * 4)
      for (int b = 0; b < 4; b++) rho_min = std::min(rho_min, c[e + static_cast<size_t>(b) * 4]);
    const double av_max = *std::max_element(av.begin(), av.end());
    const size_t on = static_cast<size_t>(std::count_if(av.begin(), av.end(), [](double v) { return v > 0.0; }));
    std::printf("explosion: %d steps, nodes with viscosity %zu of %zu, max %.4e, min corner density %.4f\n", steps, on,
                av.size(), av_max, rho_min);
    return finite && rho_min > 0.0 && av_max > 0.0 && on < av.size() ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
