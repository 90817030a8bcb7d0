// SubrosaDG.hpp -- C++ drop-in for the reference's compile-time configuration and System driver
// (src/Solver/SimulationControl.cpp:1197-1279, src/Utils/SystemControl.cpp:42-236), running the
// hot path on the GPU through the C-ABI of libsdg_hip.so (sdg.h).
//
// Kept from the reference: the enum vocabulary (src/Utils/Enum.cpp:22-233), the
// SimulationControl<SolveControl<...>, NumericalControl<...>, *Variable<...>> composition, the
// System setters and solve() loop with the all-NaN stop (SystemControl.cpp:159-195), and the
// user-specialised InitialCondition<SC>::calculatePrimitiveFromCoordinate.
// Changed: coordinates and primitive states are std::array instead of Eigen vectors; meshes are
// structured boxes (setMesh(StructuredMesh)), each box side carrying a physical index (or -1:
// periodic direction), or existing gmsh MSH 4.1 files (setMesh(path)) whose boundary groups carry
// their physical tags -- there is no mesh-generating callback, as that needs the gmsh library.
// Raw files (setViewConfig) and VTU files (view()) are written. Boundary data come from the
// user-specialised BoundaryCondition<SC>::calculatePrimitiveFromCoordinate(x, t, physical_index),
// evaluated at the boundary face points at start-up and, for BoundaryTimeEnum::TimeVarying,
// before every step (BoundaryCondition.cpp:29-50).
#ifndef SUBROSA_DG_AMD_HPP_
#define SUBROSA_DG_AMD_HPP_

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../sdg.h"

namespace SubrosaDG {

using Real = double;
using Isize = int;

enum class DimensionEnum { D1 = 1, D2, D3 };
enum class PolynomialOrderEnum { P1 = 1, P2, P3, P4, P5 };
enum class MeshModelEnum {
  Line, Triangle, Quadrangle, TriangleQuadrangle, Tetrahedron, Hexahedron, TetrahedronPyramidHexahedron
};
enum class EquationModelEnum { CompresibleEuler, CompresibleNS, IncompresibleEuler, IncompresibleNS, CompresibleRANS };
enum class SourceTermEnum { None, Boussinesq };
enum class ShockCapturingEnum { None, ArtificialViscosity };
enum class LimiterEnum { None, PositivityPreserving };
enum class InitialConditionEnum { Function, SpecificFile, LastStep };
enum class BoundaryConditionEnum {
  RiemannFarfield, VelocityInflow, PressureOutflow, IsoThermalNonSlipWall, AdiabaticSlipWall, AdiabaticNonSlipWall,
  Periodic
};
enum class BoundaryTimeEnum { Steady, TimeVarying };
enum class ConvectiveFluxEnum { Central, LaxFriedrichs, HLLC, Roe, Exact };
enum class ViscousFluxEnum { None, BR1, BR2 };
enum class ThermodynamicModelEnum { Constant };
enum class EquationOfStateEnum { IdealGas, WeakCompressibleFluid };
enum class TransportModelEnum { None, Constant, Sutherland };
enum class TimeIntegrationEnum { ForwardEuler, HeunRK2, SSPRK3 };
enum class PrecisionEnum { FP64, Mixed };  // not in the reference: the library's precision modes

template <DimensionEnum Dimension, PolynomialOrderEnum PolynomialOrder, BoundaryTimeEnum BoundaryTimeType,
          SourceTermEnum SourceTermType>
struct SolveControl {
  inline static constexpr int kDimension{static_cast<int>(Dimension)};
  inline static constexpr int kPolynomialOrder{static_cast<int>(PolynomialOrder)};
  inline static constexpr BoundaryTimeEnum kBoundaryTime{BoundaryTimeType};
  inline static constexpr SourceTermEnum kSourceTerm{SourceTermType};
};

template <MeshModelEnum MeshModelType, ShockCapturingEnum ShockCapturingType, LimiterEnum LimiterType,
          InitialConditionEnum InitialConditionType, TimeIntegrationEnum TimeIntegrationType>
struct NumericalControl {
  inline static constexpr MeshModelEnum kMeshModel{MeshModelType};
  inline static constexpr InitialConditionEnum kInitialCondition{InitialConditionType};
  inline static constexpr ShockCapturingEnum kShockCapturing{ShockCapturingType};
  inline static constexpr LimiterEnum kLimiter{LimiterType};
  inline static constexpr TimeIntegrationEnum kTimeIntegration{TimeIntegrationType};
};

template <ThermodynamicModelEnum ThermodynamicModelType, EquationOfStateEnum EquationOfStateType,
          ConvectiveFluxEnum ConvectiveFluxType>
struct CompresibleEulerVariable {
  inline static constexpr EquationModelEnum kEquationModel{EquationModelEnum::CompresibleEuler};
  inline static constexpr ThermodynamicModelEnum kThermodynamicModel{ThermodynamicModelType};
  inline static constexpr EquationOfStateEnum kEquationOfState{EquationOfStateType};
  inline static constexpr TransportModelEnum kTransportModel{TransportModelEnum::None};
  inline static constexpr ConvectiveFluxEnum kConvectiveFlux{ConvectiveFluxType};
  inline static constexpr ViscousFluxEnum kViscousFlux{ViscousFluxEnum::None};
};

template <ThermodynamicModelEnum ThermodynamicModelType, EquationOfStateEnum EquationOfStateType,
          TransportModelEnum TransportModelType, ConvectiveFluxEnum ConvectiveFluxType, ViscousFluxEnum ViscousFluxType>
struct CompresibleNSVariable {
  inline static constexpr EquationModelEnum kEquationModel{EquationModelEnum::CompresibleNS};
  inline static constexpr ThermodynamicModelEnum kThermodynamicModel{ThermodynamicModelType};
  inline static constexpr EquationOfStateEnum kEquationOfState{EquationOfStateType};
  inline static constexpr TransportModelEnum kTransportModel{TransportModelType};
  inline static constexpr ConvectiveFluxEnum kConvectiveFlux{ConvectiveFluxType};
  inline static constexpr ViscousFluxEnum kViscousFlux{ViscousFluxType};
};

template <typename SolveControlT, typename NumericalControlT, typename EquationVariable>
struct SimulationControl : SolveControlT, NumericalControlT, EquationVariable {
  inline static constexpr int kConservedVariableNumber{SolveControlT::kDimension + 2};
  inline static constexpr int kComputationalVariableNumber{SolveControlT::kDimension + 3};
  inline static constexpr int kPrimitiveVariableNumber{SolveControlT::kDimension + 2};
};

// User hook, specialised per example like the reference's InitialCondition<SC>
// (InitialCondition.cpp:85-116): primitive (rho, u, v[, w], T) at a physical point.
template <typename SC>
struct InitialCondition {
  std::array<Real, SC::kPrimitiveVariableNumber> calculatePrimitiveFromCoordinate(
      const std::array<Real, SC::kDimension>& coordinate) const;
};

// Specialise calculatePrimitiveFromCoordinate for meshes with boundary sides; the default is only
// reached when a boundary exists without one.
template <typename SC>
struct BoundaryCondition {
  std::array<Real, SC::kPrimitiveVariableNumber> calculatePrimitiveFromCoordinate(
      const std::array<Real, SC::kDimension>& /*coordinate*/, Real /*time*/, Isize physical_index) const {
    throw std::runtime_error("SubrosaDG: BoundaryCondition<SC>::calculatePrimitiveFromCoordinate is not specialised "
                             "(boundary physical index " + std::to_string(physical_index) + ")");
  }
};

// Structured box (replaces the gmsh generate-mesh callback of the examples). physical_index gives
// the boundary physical index of each side (x low, x high, y low, y high, z low, z high); a
// direction whose two sides are -1 is periodic.
// A mesh in host arrays (the analogue of the gmsh model a setMesh(path, generate_mesh_function)
// callback builds, SystemControl.cpp:55-61): node coordinates, element vertex lists (0-based, gmsh
// vertex order), boundary faces with their physical index, and translational periodic node pairs.
struct ArrayMesh {
  int element = SDG_ELEM_QUADRANGLE;                    // SDG_ELEM_TRIANGLE / QUADRANGLE / TETRAHEDRON / HEXAHEDRON
  int geometry_order = 1;                               // gmsh order of the element node lists (curved elements)
  std::vector<std::array<Real, 3>> nodes;
  std::vector<int64_t> elements;                        // element_count x geometry nodes, gmsh order
  std::vector<int64_t> boundary_faces;                  // count x face vertices
  std::vector<Isize> boundary_physical;                 // per boundary face
  std::vector<std::vector<std::pair<int64_t, int64_t>>> periodic;  // per link: (slave, master) pairs
};

struct StructuredMesh {
  std::array<int, 3> cells{2, 2, 2};
  std::array<Real, 3> lower{0.0, 0.0, 0.0};
  std::array<Real, 3> upper{1.0, 1.0, 1.0};
  std::array<Isize, 6> physical_index{-1, -1, -1, -1, -1, -1};
};

template <typename SC>
struct System {
  static_assert(SC::kEquationModel == EquationModelEnum::CompresibleEuler ||
                    SC::kEquationModel == EquationModelEnum::CompresibleNS,
                "the GPU engine implements the compressible Euler and Navier-Stokes models");

  InitialCondition<SC> initial_condition_;
  BoundaryCondition<SC> boundary_condition_;
  std::vector<std::pair<Isize, BoundaryConditionEnum>> boundary_types_;
  StructuredMesh mesh_desc_;
  std::string mesh_file_;  // gmsh MSH 4.1 file (setMesh(path)); empty for a structured box
  ArrayMesh array_mesh_;   // setMesh(path, generate) without the file: the callback's arrays
  bool use_arrays_ = false;
  bool mesh_set_ = false;
  sdg_problem_desc problem_{};
  sdg_mesh* mesh_ = nullptr;
  sdg_ctx* ctx_ = nullptr;
  int iteration_start_ = 0, iteration_end_ = 0;
  Real cfl_ = 0.0, delta_time_ = 0.0;
  int io_interval_ = 0;
  std::string output_directory_, output_prefix_, initial_condition_file_;
  std::vector<std::array<Real, SC::kConservedVariableNumber>> relative_error_history_;
  std::function<void(int, const std::array<Real, SC::kConservedVariableNumber>&)> on_step_;

  System() {
    problem_.abi_version = SDG_ABI_VERSION;
    problem_.equation = SC::kEquationModel == EquationModelEnum::CompresibleNS ? SDG_EQ_COMPRESSIBLE_NS
                                                                                : SDG_EQ_COMPRESSIBLE_EULER;
    problem_.convective_flux = static_cast<int>(SC::kConvectiveFlux);
    problem_.viscous_flux = static_cast<int>(SC::kViscousFlux);
    problem_.time_integration = static_cast<int>(SC::kTimeIntegration);
    problem_.equation_of_state = static_cast<int>(SC::kEquationOfState);
    problem_.transport = static_cast<int>(SC::kTransportModel);
    problem_.precision = SDG_PREC_FP64;
    problem_.source_term = static_cast<int>(SC::kSourceTerm);
    problem_.shock_capturing = static_cast<int>(SC::kShockCapturing);
    problem_.artificial_viscosity_factor = 1.0;
  }
  ~System() {
    if (ctx_) sdg_destroy(ctx_);
    if (mesh_) sdg_mesh_destroy(mesh_);
  }
  System(const System&) = delete;
  System& operator=(const System&) = delete;

  static void check(int rc, const sdg_ctx* ctx) {
    if (rc != SDG_OK) throw std::runtime_error(std::string("SubrosaDG: ") + sdg_last_error(ctx));
  }

  // SystemControl.cpp:55-61 (structured periodic box instead of a gmsh file)
  // SystemControl.cpp:55-61 with an existing gmsh file (MSH 4.1 ASCII, affine elements of one type):
  // the boundary faces keep the physical tags of their groups for addBoundaryCondition and the
  // BoundaryCondition callback. (The reference's mesh-generating callback needs the gmsh library.)
  void setMesh(const std::filesystem::path& mesh_file) {
    mesh_file_ = mesh_file.string();
    use_arrays_ = false;
    mesh_set_ = true;
  }
  void setMesh(const StructuredMesh& mesh) {
    mesh_file_.clear();
    use_arrays_ = false;
    mesh_desc_ = mesh;
    mesh_set_ = true;
  }
  // SystemControl.cpp:55-61 with a generating callback: an existing mesh file is read; otherwise the
  // callback fills an ArrayMesh in memory (where the reference's callback drives the gmsh library).
  void setMesh(const std::filesystem::path& mesh_file, const std::function<void(ArrayMesh&)>& generate_mesh_function) {
    if (std::filesystem::exists(mesh_file)) {
      setMesh(mesh_file);
      return;
    }
    array_mesh_ = ArrayMesh{};
    generate_mesh_function(array_mesh_);
    mesh_file_.clear();
    use_arrays_ = true;
    mesh_set_ = true;
  }
  // SystemControl.cpp:76-80: boundary type of a physical index (Periodic: nothing to record,
  // periodic directions are the ones whose sides carry index -1)
  template <BoundaryConditionEnum BoundaryConditionType>
  void addBoundaryCondition(const Isize physical_index) {
    if constexpr (BoundaryConditionType != BoundaryConditionEnum::Periodic)
      boundary_types_.emplace_back(physical_index, BoundaryConditionType);
  }
  // SystemControl.cpp:82-88
  template <ThermodynamicModelEnum ThermodynamicModelType>
  void setThermodynamicModel(const Real specific_heat_constant_pressure, const Real specific_heat_constant_volume) {
    problem_.cp = specific_heat_constant_pressure;
    problem_.cv = specific_heat_constant_volume;
  }
  // SystemControl.cpp:90-96
  template <EquationOfStateEnum EquationOfStateType>
  void setEquationOfState(const Real reference_sound_speed, const Real reference_density) {
    problem_.reference_sound_speed = reference_sound_speed;
    problem_.reference_density = reference_density;
  }
  // SystemControl.cpp:98-103 (thermal conductivity cp mu / Pr is derived inside the engine)
  template <TransportModelEnum TransportModelType>
  void setTransportModel(const Real dynamic_viscosity) {
    problem_.dynamic_viscosity = dynamic_viscosity;
  }
  // SystemControl.cpp:65-68: Boussinesq thermal expansion coefficient and reference temperature
  void setSourceTerm(const Real thermal_expansion_coefficient, const Real reference_temperature) {
    static_assert(SC::kSourceTerm == SourceTermEnum::Boussinesq, "setSourceTerm needs SourceTermEnum::Boussinesq");
    problem_.thermal_expansion_coefficient = thermal_expansion_coefficient;
    problem_.reference_temperature = reference_temperature;
  }
  // SystemControl.cpp:105-108: empirical tolerance k of the sensor ramp and the viscosity factor
  void setArtificialViscosity(const Real empirical_tolerance, const Real artificial_viscosity_factor = 1.0) {
    static_assert(SC::kShockCapturing == ShockCapturingEnum::ArtificialViscosity,
                  "setArtificialViscosity needs ShockCapturingEnum::ArtificialViscosity");
    problem_.empirical_tolerance = empirical_tolerance;
    problem_.artificial_viscosity_factor = artificial_viscosity_factor;
  }
  // SystemControl.cpp:110-121 (no interactive prompt: the range is required)
  void setTimeIntegration(const Real courant_friedrichs_lewy_number, const std::pair<int, int> iteration_range) {
    cfl_ = courant_friedrichs_lewy_number;
    iteration_start_ = iteration_range.first;
    iteration_end_ = iteration_range.second;
  }
  void setDeltaTime(const Real delta_time) { delta_time_ = delta_time; }
  Real deltaTime() const { return delta_time_; }
  // Not in the reference (its SUBROSA_DG_SINGLE_PRECISION switch is uniform FP32): FP32 matrix-core
  // contractions with FP64 state, geometry, pointwise physics and RK accumulation. Call before
  // synchronize().
  void setPrecision(const PrecisionEnum precision) {
    problem_.precision = precision == PrecisionEnum::Mixed ? SDG_PREC_MIXED : SDG_PREC_FP64;
  }
  // SystemControl.cpp:124-140: raw state files <output_directory>/raw/<prefix>_<step>.zst every
  // io_interval steps (and at the start unless restarting); -1 = only at the end. No prompt.
  void setViewConfig(const std::string& output_directory, const std::string& prefix, const int io_interval = -1) {
    output_directory_ = output_directory;
    output_prefix_ = prefix;
    io_interval_ = io_interval;
  }
  // SystemControl.cpp:69-74: the order-(P-1) raw file of InitialConditionEnum::SpecificFile
  template <InitialConditionEnum InitialConditionType>
  void addInitialCondition(const std::string& initial_condition_file) {
    static_assert(InitialConditionType == InitialConditionEnum::SpecificFile, "only SpecificFile takes a file");
    initial_condition_file_ = initial_condition_file;
  }
  std::string rawPath(const int step) const {
    return output_directory_ + "/raw/" + output_prefix_ + "_" + std::to_string(step) + ".zst";
  }
  // RawBinary.cpp:165-191 via the library; node count of the mesh (the vertex grid of the box)
  void writeRaw(const int step) const {
    if (output_directory_.empty()) return;
    std::filesystem::create_directories(output_directory_ + "/raw");
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    check(sdg_raw_binary_write(ctx_, rawPath(step).c_str(), info.nodes), ctx_);
  }
  // The monitor log <output_directory>/error.txt (CommandLine.cpp:70-127, IOControl.cpp:318-333):
  // a header line, the line of time 0, then one line per step of "|time|monitor...|" with each
  // value in %.5e centred in 13 columns.
  static std::string centred(const std::string& v) {
    const size_t pad = v.size() < 13 ? 13 - v.size() : 0;
    return std::string(pad / 2, ' ') + v + std::string(pad - pad / 2, ' ');
  }
  static std::string errorLine(const Real time, const std::array<Real, SC::kConservedVariableNumber>& err) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.5e", static_cast<double>(time));
    std::string line = "|" + centred(buf) + "|";
    for (const Real e : err) {
      std::snprintf(buf, sizeof(buf), "%.5e", static_cast<double>(e));
      line += centred(buf) + "|";
    }
    return line;
  }
  static std::string errorHeader() {
    const char* names1[] = {"rho", "rho*u", "rho*E"};
    const char* names2[] = {"rho", "rho*u", "rho*v", "rho*E"};
    const char* names3[] = {"rho", "rho*u", "rho*v", "rho*w", "rho*E"};
    const char** names = SC::kDimension == 1 ? names1 : (SC::kDimension == 2 ? names2 : names3);
    std::string line = "|" + centred("Time") + "|";
    for (int k = 0; k < SC::kConservedVariableNumber; k++) line += centred(names[k]) + "|";
    return line;
  }
  std::string errorPath() const { return output_directory_ + "/error.txt"; }
  // Open the log for a run: a fresh file, or for LastStep the existing one cut after the line of
  // iteration_start (CommandLine.cpp:108-127). Returns the lines kept.
  std::vector<std::string> openErrorLog() {
    std::vector<std::string> kept;
    if (output_directory_.empty()) return kept;
    std::filesystem::create_directories(output_directory_);
    if (SC::kInitialCondition == InitialConditionEnum::LastStep) {
      std::ifstream in(errorPath());
      std::string line;
      for (int i = 0; i < iteration_start_ + 2 && std::getline(in, line); i++) kept.push_back(line);
    } else {
      kept.push_back(errorHeader());
      kept.push_back(errorLine(0.0, std::array<Real, SC::kConservedVariableNumber>{}));
    }
    std::ofstream out(errorPath(), std::ios::trunc);
    for (const auto& l : kept) out << l << '\n';
    return kept;
  }
  // Solver::calculateDeltaTime for LastStep (TimeIntegration.cpp:138-146): the time value of the
  // third line of the log, i.e. the first step's time, is the time step of the earlier run.
  Real deltaTimeFromErrorLog() const {
    std::ifstream in(errorPath());
    std::string line;
    for (int i = 0; i < 3; i++)
      if (!std::getline(in, line)) throw std::runtime_error("SubrosaDG: LastStep needs the earlier run's error.txt");
    std::stringstream ss(line);
    Real dt = 0.0;
    ss.ignore(2) >> dt;
    if (!(dt > 0.0)) throw std::runtime_error("SubrosaDG: no time step in " + errorPath());
    return dt;
  }

  // Called after every step with the step number and the residual monitor.
  void setStepCallback(std::function<void(int, const std::array<Real, SC::kConservedVariableNumber>&)> f) {
    on_step_ = std::move(f);
  }

  // SystemControl.cpp:146-157: build the mesh and the device context
  void synchronize() {
    if (!mesh_set_) throw std::runtime_error("SubrosaDG: setMesh was not called");
    if (use_arrays_) {
      const ArrayMesh& am = array_mesh_;
      std::vector<double> xyz;
      for (const auto& p : am.nodes) xyz.insert(xyz.end(), p.begin(), p.end());
      std::vector<int> btype, bphys;
      for (const Isize ph : am.boundary_physical) {
        btype.push_back(boundaryType(ph));
        bphys.push_back(static_cast<int>(ph));
      }
      std::vector<int64_t> sizes, pairs;
      for (const auto& link : am.periodic) {
        sizes.push_back(static_cast<int64_t>(link.size()));
        for (const auto& pr : link) pairs.insert(pairs.end(), {pr.first, pr.second});
      }
      const int g = am.geometry_order, g1 = g + 1;
      const int nvert = am.element == SDG_ELEM_TRIANGLE      ? g1 * (g1 + 1) / 2
                        : am.element == SDG_ELEM_QUADRANGLE  ? g1 * g1
                        : am.element == SDG_ELEM_TETRAHEDRON ? g1 * (g1 + 1) * (g1 + 2) / 6
                                                             : g1 * g1 * g1;
      sdg_array_mesh_desc d{};
      d.element = am.element;
      d.polynomial_order = SC::kPolynomialOrder;
      d.geometry_order = g;
      d.node_count = static_cast<int64_t>(am.nodes.size());
      d.nodes = xyz.data();
      d.element_count = static_cast<int64_t>(am.elements.size()) / nvert;
      d.elements = am.elements.data();
      d.boundary_face_count = static_cast<int64_t>(am.boundary_physical.size());
      d.boundary_faces = am.boundary_faces.data();
      d.boundary_type = btype.data();
      d.boundary_physical = bphys.data();
      d.periodic_link_count = static_cast<int>(sizes.size());
      d.periodic_link_size = sizes.data();
      d.periodic_pairs = pairs.data();
      check(sdg_mesh_create_from_arrays(&d, &mesh_), nullptr);
      check(sdg_create(&problem_, mesh_, &ctx_), nullptr);
      return;
    }
    if (!mesh_file_.empty()) {
      std::vector<int> phys, types;
      for (const auto& b : boundary_types_) {
        phys.push_back(static_cast<int>(b.first));
        types.push_back(static_cast<int>(b.second));
      }
      sdg_gmsh_desc g{mesh_file_.c_str(), SC::kPolynomialOrder, static_cast<int>(phys.size()), phys.data(), types.data()};
      check(sdg_mesh_create_gmsh(&g, &mesh_), nullptr);
      sdg_mesh_info info{};
      check(sdg_mesh_get_info(mesh_, &info), nullptr);
      if (info.dimension != SC::kDimension) throw std::runtime_error("SubrosaDG: mesh file dimension mismatch");
      check(sdg_create(&problem_, mesh_, &ctx_), nullptr);
      return;
    }
    sdg_structured_desc d{};
    d.element = SC::kDimension == 1 ? SDG_ELEM_LINE : (SC::kDimension == 2 ? SDG_ELEM_QUADRANGLE : SDG_ELEM_HEXAHEDRON);
    d.polynomial_order = SC::kPolynomialOrder;
    for (int i = 0; i < 3; i++) {
      d.n[i] = mesh_desc_.cells[i];
      d.lo[i] = mesh_desc_.lower[i];
      d.hi[i] = mesh_desc_.upper[i];
    }
    for (int dir = 0; dir < SC::kDimension; dir++) {
      if (mesh_desc_.physical_index[2 * dir] < 0 && mesh_desc_.physical_index[2 * dir + 1] < 0) continue;
      d.bounded[dir] = 1;
      for (int side = 0; side < 2; side++) d.boundary[2 * dir + side] = boundaryType(mesh_desc_.physical_index[2 * dir + side]);
    }
    check(sdg_mesh_create_structured(&d, &mesh_), nullptr);
    check(sdg_create(&problem_, mesh_, &ctx_), nullptr);
  }

  int boundaryType(const Isize physical_index) const {
    for (const auto& b : boundary_types_)
      if (b.first == physical_index) return static_cast<int>(b.second);
    throw std::runtime_error("SubrosaDG: no boundary condition added for physical index " +
                             std::to_string(physical_index));
  }

  // boundary data at every boundary face point for time t (BoundaryCondition.cpp:29-50); the
  // side of an axis-aligned box face, hence its physical index, follows from its normal
  void updateBoundary(const Real time) {
    constexpr int D = SC::kDimension, NV = SC::kConservedVariableNumber;
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    if (info.boundary_faces == 0) return;
    const int64_t npts = info.boundary_faces * info.face_quadrature;
    std::vector<double> xb(static_cast<size_t>(npts * D)), nb(static_cast<size_t>(npts * D)),
        prim(static_cast<size_t>(npts * NV));
    std::vector<int32_t> phys(static_cast<size_t>(info.boundary_faces));
    check(sdg_mesh_export(mesh_, SDG_FIELD_BOUNDARY_FACE_COORD, xb.data(), npts * D), nullptr);
    check(sdg_mesh_export(mesh_, SDG_FIELD_BOUNDARY_FACE_NORMAL, nb.data(), npts * D), nullptr);
    check(sdg_mesh_export(mesh_, SDG_FIELD_BOUNDARY_FACE_PHYSICAL, phys.data(), info.boundary_faces), nullptr);
    for (int64_t p = 0; p < npts; p++) {
      std::array<Real, D> x;
      int side = 0;
      for (int dd = 0; dd < D; dd++) {
        x[dd] = xb[p * D + dd];
        if (std::fabs(nb[p * D + dd]) > 0.5) side = 2 * dd + (nb[p * D + dd] > 0.0 ? 1 : 0);
      }
      const Isize physical = mesh_file_.empty() && !use_arrays_ ? mesh_desc_.physical_index[side] : phys[p / info.face_quadrature];
      const auto v = boundary_condition_.calculatePrimitiveFromCoordinate(x, time, physical);
      for (int k = 0; k < NV; k++) prim[p * NV + k] = v[k];
    }
    check(sdg_set_boundary_primitive(ctx_, prim.data()), ctx_);
  }

  // SystemControl.cpp:159-195
  void solve() {
    if (!ctx_) synchronize();
    constexpr int D = SC::kDimension, NV = SC::kConservedVariableNumber;
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    const int64_t nq = info.quadrature;
    std::vector<double> xq(static_cast<size_t>(sdg_mesh_field_size(mesh_, SDG_FIELD_QUAD_COORD)));
    check(sdg_mesh_export(mesh_, SDG_FIELD_QUAD_COORD, xq.data(), static_cast<int64_t>(xq.size())), nullptr);
    if constexpr (SC::kInitialCondition == InitialConditionEnum::Function) {
      std::vector<double> prim(static_cast<size_t>(info.elements * nq * NV));
      for (int64_t p = 0; p < info.elements * nq; p++) {
        std::array<Real, D> x;
        for (int d = 0; d < D; d++) x[d] = xq[p * D + d];
        const auto v = initial_condition_.calculatePrimitiveFromCoordinate(x);
        for (int k = 0; k < NV; k++) prim[p * NV + k] = v[k];
      }
      check(sdg_set_initial_primitive(ctx_, prim.data()), ctx_);
    } else if constexpr (SC::kInitialCondition == InitialConditionEnum::SpecificFile) {
      check(sdg_raw_binary_read(ctx_, initial_condition_file_.c_str(), SDG_IC_SPECIFIC_FILE), ctx_);
    } else {  // LastStep: the raw file of the first iteration (SystemControl.cpp:150-154)
      check(sdg_raw_binary_read(ctx_, rawPath(iteration_start_).c_str(), SDG_IC_LAST_STEP), ctx_);
    }
    // SystemControl.cpp:163-165 and TimeIntegration.cpp:134-146: a time step set with setDeltaTime
    // is kept; otherwise LastStep takes the earlier run's from its log and the others compute the
    // CFL step
    if (delta_time_ == 0.0) {
      if (SC::kInitialCondition == InitialConditionEnum::LastStep && !output_directory_.empty())
        delta_time_ = deltaTimeFromErrorLog();
      else
        check(sdg_compute_dt(ctx_, cfl_, &delta_time_), ctx_);
    }
    updateBoundary(iteration_start_ * delta_time_);
    if (SC::kInitialCondition != InitialConditionEnum::LastStep) writeRaw(0);
    const bool log = !output_directory_.empty();
    if (log) openErrorLog();
    std::ofstream error_log;
    if (log) error_log.open(errorPath(), std::ios::app);
    const int interval = io_interval_ > 0 ? io_interval_ : iteration_end_;
    for (int i = iteration_start_ + 1; i <= iteration_end_; i++) {
      // stepSolver refreshes time-varying data at iteration_ * delta_time_, and iteration_ is the
      // previous step's number while step i runs (TimeIntegration.cpp:334-336, SystemControl.cpp:175-177)
      if (SC::kBoundaryTime == BoundaryTimeEnum::TimeVarying) updateBoundary((i - 1) * delta_time_);
      std::array<Real, NV> rel{};
      int nan = 0, taken = 0;
      check(sdg_step(ctx_, delta_time_, 1, rel.data(), &nan, &taken), ctx_);
      relative_error_history_.push_back(rel);
      if (log) error_log << errorLine(static_cast<Real>(i) * delta_time_, rel) << '\n' << std::flush;
      if (on_step_) on_step_(i, rel);
      if (nan) {  // all-NaN monitor: truncate the run (SystemControl.cpp:185-191)
        iteration_end_ = i;
        break;
      }
      if (interval > 0 && i % interval == 0) writeRaw(i);
    }
  }

  // Modal coefficients, per element N_b x NV with NV fastest (the reference's Eigen layout).
  std::vector<Real> state() const {
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    std::vector<Real> c(static_cast<size_t>(info.elements * info.basis * SC::kConservedVariableNumber));
    check(sdg_get_state(ctx_, c.data()), ctx_);
    return c;
  }
  // View::view (Paraview.cpp:280-313): every raw file of the run, read back, written as
  // <output_directory>/vtu/<prefix>_<step>.vtu (density, velocity, pressure, temperature, Mach
  // number, and the node artificial viscosity with shock capturing); the solver keeps its final state.
  void view() {
    if (output_directory_.empty()) throw std::runtime_error("SubrosaDG: view needs setViewConfig");
    std::filesystem::create_directories(output_directory_ + "/vtu");
    const std::vector<Real> final_state = state();
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    for (int step = iteration_start_; step <= iteration_end_; step++) {
      const std::string raw = rawPath(step);
      if (!std::filesystem::exists(raw)) continue;
      check(sdg_raw_binary_read(ctx_, raw.c_str(), SDG_IC_LAST_STEP), ctx_);
      const std::vector<Real> c = state();
      std::vector<double> payload;
      const double* node_values = nullptr;
      if constexpr (SC::kShockCapturing == ShockCapturingEnum::ArtificialViscosity) {
        int64_t bytes = 0;
        check(sdg_raw_file_read(raw.c_str(), nullptr, 0, &bytes), nullptr);
        payload.resize(static_cast<size_t>(bytes) / sizeof(double));
        check(sdg_raw_file_read(raw.c_str(), payload.data(), bytes, &bytes), nullptr);
        if (payload.size() >= static_cast<size_t>(info.nodes)) node_values = payload.data() + payload.size() - info.nodes;
      }
      const std::string out =
          output_directory_ + "/vtu/" + output_prefix_ + "_" + std::to_string(step) + ".vtu";
      check(sdg_write_vtu(mesh_, &problem_, c.data(), node_values, step, step * delta_time_, out.c_str()), nullptr);
    }
    check(sdg_set_state(ctx_, final_state.data()), ctx_);
  }

  // Node artificial viscosity of the last step (node_artificial_viscosity_, the reference's
  // ViewVariableEnum::ArtificialViscosity output), mesh nodes long.
  std::vector<Real> artificialViscosity() const {
    static_assert(SC::kShockCapturing == ShockCapturingEnum::ArtificialViscosity,
                  "artificialViscosity needs ShockCapturingEnum::ArtificialViscosity");
    sdg_mesh_info info{};
    check(sdg_mesh_get_info(mesh_, &info), nullptr);
    std::vector<Real> v(static_cast<size_t>(info.nodes));
    check(sdg_get_artificial_viscosity(ctx_, v.data()), ctx_);
    return v;
  }
};

}  // namespace SubrosaDG

#endif  // SUBROSA_DG_AMD_HPP_
