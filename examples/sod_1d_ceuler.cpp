// 1-D compressible Euler shock tube (the reference's sod_1d_ceuler case) through the C++ drop-in
// driver: DimensionEnum::D1 on a line mesh of [0, 1] with Riemann far-field ends carrying the two
// states, HLLC flux, SSP-RK3, artificial viscosity (setArtificialViscosity(0.5)); rho = 1, p = 1
// left of x = 0.5 and rho = 0.125, p = 0.1 right of it (R = 1). Prints the density's L1 distance
// from the exact Riemann solution at t = 0.2.
// Usage: sod_1d_ceuler [cells]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "subrosadg_amd/SubrosaDG.hpp"

using SimulationControl = SubrosaDG::SimulationControl<
    SubrosaDG::SolveControl<SubrosaDG::DimensionEnum::D1, SubrosaDG::PolynomialOrderEnum::P2,
                            SubrosaDG::BoundaryTimeEnum::Steady, SubrosaDG::SourceTermEnum::None>,
    SubrosaDG::NumericalControl<SubrosaDG::MeshModelEnum::Line, SubrosaDG::ShockCapturingEnum::ArtificialViscosity,
                                SubrosaDG::LimiterEnum::None, SubrosaDG::InitialConditionEnum::Function,
                                SubrosaDG::TimeIntegrationEnum::SSPRK3>,
    SubrosaDG::CompresibleEulerVariable<SubrosaDG::ThermodynamicModelEnum::Constant,
                                        SubrosaDG::EquationOfStateEnum::IdealGas, SubrosaDG::ConvectiveFluxEnum::HLLC>>;

namespace {
constexpr int kEnd = 1;
std::array<double, 3> sod(double x) { return x < 0.5 ? std::array<double, 3>{1.0, 0.0, 1.0} : std::array<double, 3>{0.125, 0.0, 0.8}; }

// exact density of the tube at (x, t): star state by Newton iteration on the pressure functions
double exactDensity(double x, double t) {
  const double g = 1.4, rl = 1.0, pl = 1.0, rr = 0.125, pr = 0.1;
  const double cl = std::sqrt(g * pl / rl), cr = std::sqrt(g * pr / rr);
  auto f = [&](double p, double rk, double pk, double ck, double& d) {
    if (p > pk) {
      const double a = 2.0 / ((g + 1.0) * rk), b = (g - 1.0) / (g + 1.0) * pk;
      d = std::sqrt(a / (b + p)) * (1.0 - (p - pk) / (2.0 * (b + p)));
      return (p - pk) * std::sqrt(a / (p + b));
    }
    d = 1.0 / (rk * ck) * std::pow(p / pk, -(g + 1.0) / (2.0 * g));
    return 2.0 * ck / (g - 1.0) * (std::pow(p / pk, (g - 1.0) / (2.0 * g)) - 1.0);
  };
  double p = 0.3, dl = 0.0, dr = 0.0;
  for (int i = 0; i < 50; i++) p -= (f(p, rl, pl, cl, dl) + f(p, rr, pr, cr, dr)) / (dl + dr);
  const double u = 0.5 * (f(p, rr, pr, cr, dr) - f(p, rl, pl, cl, dl));
  const double q = (g - 1.0) / (g + 1.0);
  const double rls = rl * std::pow(p / pl, 1.0 / g), rrs = rr * ((p / pr + q) / (q * p / pr + 1.0));
  const double cls = cl * std::pow(p / pl, (g - 1.0) / (2.0 * g)), shock = rrs * u / (rrs - rr);
  const double s = (x - 0.5) / t;
  if (s <= -cl) return rl;
  if (s < u - cls) return rl * std::pow(2.0 / (g + 1.0) - (g - 1.0) / ((g + 1.0) * cl) * s, 2.0 / (g - 1.0));
  if (s < u) return rls;
  return s < shock ? rrs : rr;
}
}  // namespace

template <>
std::array<SubrosaDG::Real, 3> SubrosaDG::InitialCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 1>& x) const {
  return sod(x[0]);
}
template <>
std::array<SubrosaDG::Real, 3> SubrosaDG::BoundaryCondition<SimulationControl>::calculatePrimitiveFromCoordinate(
    const std::array<SubrosaDG::Real, 1>& x, [[maybe_unused]] SubrosaDG::Real time,
    [[maybe_unused]] SubrosaDG::Isize physical_index) const {
  return sod(x[0]);
}

int main(int argc, char* argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 200;
  try {
    const double t_end = 0.2, dt = 0.1 * (1.0 / n) / (1.5 * 9.0);  // below CFL 0.1 for |u| + c < 1.5
    const int steps = static_cast<int>(std::ceil(t_end / dt));
    SubrosaDG::System<SimulationControl> system;
    system.setMesh({{n, 1, 1}, {0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {kEnd, kEnd, -1, -1, -1, -1}});
    system.addBoundaryCondition<SubrosaDG::BoundaryConditionEnum::RiemannFarfield>(kEnd);
    system.setThermodynamicModel<SimulationControl::kThermodynamicModel>(3.5, 2.5);
    system.setArtificialViscosity(0.5);
    system.setTimeIntegration(0.1, {0, steps});
    system.setDeltaTime(t_end / steps);
    system.synchronize();
    system.solve();
    // density at the element midpoints (xi = 0): each vertex mode is 1/2 there, the P2 Lobatto
    // bubble (L_2 - L_0) / sqrt(6) is -1.5 / sqrt(6)
    const auto c = system.state();
    double l1 = 0.0;
    for (int e = 0; e < n; e++) {
      const double xm = (e + 0.5) / n;
      const double r0 = c[static_cast<size_t>(e) * 9 + 0], r1 = c[static_cast<size_t>(e) * 9 + 3],
                   r2 = c[static_cast<size_t>(e) * 9 + 6];
      const double bubble = -1.5 / std::sqrt(6.0);
      l1 += std::fabs(0.5 * (r0 + r1) + r2 * bubble - exactDensity(xm, t_end)) / n;
    }
    std::printf("sod 1-D: %d cells, %d steps, L1(rho at midpoints) = %.4f\n", n, steps, l1);
    return l1 < 0.02 ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
}
