// sdg_vtu.cpp -- VTK XML unstructured-grid output of a state (the reference's View::stepView,
// src/View/Paraview.cpp:154-313, which writes each element as one VTK Lagrange cell through vtu11).
//
// Every element is one Lagrange cell of order P (VTK_LAGRANGE_CURVE / TRIANGLE / QUADRILATERAL /
// HEXAHEDRON) through its equispaced nodes of order P, listed in VTK's node order (vertices, edge
// nodes, face nodes, interior nodes); points are per element, so the DG discontinuities stay
// visible. Point data: the requested view variables (ViewVariableEnum, VariableConvertor.cpp:
// 756-882) evaluated from the modal state (and, for the gradient variables, the lifted gradient)
// at the nodes; in 2-D and 3-D Velocity, MachNumber and HeatFlux are 3-component vectors, as is
// Vorticity in 3-D (Paraview.cpp:43-96). Field data TMSTEP and TimeValue. ASCII, 17 significant
// digits.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "kit.hpp"
#include "sdg_internal.hpp"

namespace {

constexpr double kGamma = 1.4;  // PhysicalModel.cpp:45

int vtuFail(int code, const std::string& msg) {
  sdg::setGlobalError("sdg_write_view: " + msg);
  return code;
}

// VTK's node index of lattice point (i, j[, k]) of a Lagrange quadrilateral / hexahedron of order
// P (vertices, then the edges in VTK's edge order with their nodes in increasing parameter, then
// faces, then the interior, each lattice-ordered)
int quadNode(int i, int j, int P) {
  const bool ib = i == 0 || i == P, jb = j == 0 || j == P;
  if (ib && jb) return i ? (j ? 2 : 1) : (j ? 3 : 0);
  int off = 4;
  if (jb) return (i - 1) + (j ? (P - 1) + (P - 1) : 0) + off;
  if (ib) return (j - 1) + (i ? (P - 1) : 2 * (P - 1) + (P - 1)) + off;
  off += 4 * (P - 1);
  return off + (i - 1) + (P - 1) * (j - 1);
}
int hexNode(int i, int j, int k, int P) {
  const bool ib = i == 0 || i == P, jb = j == 0 || j == P, kb = k == 0 || k == P;
  const int nb = ib + jb + kb, m = P - 1;
  if (nb == 3) return (i ? (j ? 2 : 1) : (j ? 3 : 0)) + (k ? 4 : 0);
  int off = 8;
  if (nb == 2) {
    if (!ib) return (i - 1) + (j ? 2 * m : 0) + (k ? 4 * m : 0) + off;
    if (!jb) return (j - 1) + (i ? m : 3 * m) + (k ? 4 * m : 0) + off;
    off += 8 * m;
    return (k - 1) + m * (i ? (j ? 3 : 1) : (j ? 2 : 0)) + off;
  }
  off += 12 * m;
  if (nb == 1) {
    if (ib) return (j - 1) + m * (k - 1) + (i ? m * m : 0) + off;
    off += 2 * m * m;
    if (jb) return (i - 1) + m * (k - 1) + (j ? m * m : 0) + off;
    off += 2 * m * m;
    return (i - 1) + m * (j - 1) + (k ? m * m : 0) + off;
  }
  off += 6 * m * m;
  return off + (i - 1) + m * ((j - 1) + m * (k - 1));
}

// reference coordinates of the element's nodes in VTK Lagrange order, and the cell type
void lagrangeNodes(const sdg::RefElement& ref, std::vector<double>& xi, int& vtk_type) {
  const int P = ref.order, dim = ref.dim, n1 = P + 1;
  xi.clear();
  auto t = [&](int i) { return -1.0 + 2.0 * i / P; };
  if (ref.type == SDG_ELEM_TRIANGLE) {  // reference (0,0), (1,0), (0,1)
    std::vector<std::pair<int, int>> pts = {{0, 0}, {P, 0}, {0, P}};
    for (int k = 1; k < P; k++) pts.push_back({k, 0});
    for (int k = 1; k < P; k++) pts.push_back({P - k, k});
    for (int k = 1; k < P; k++) pts.push_back({0, P - k});
    if (P == 3) pts.push_back({1, 1});
    if (P == 4) pts.insert(pts.end(), {{1, 1}, {2, 1}, {1, 2}});  // the interior: an order-1 triangle
    for (const auto& p : pts) {
      xi.push_back(static_cast<double>(p.first) / P);
      xi.push_back(static_cast<double>(p.second) / P);
    }
    vtk_type = 69;  // VTK_LAGRANGE_TRIANGLE
    return;
  }
  if (ref.type == SDG_ELEM_TETRAHEDRON) {  // reference (0,0,0), (1,0,0), (0,1,0), (0,0,1)
    // barycentric lattice weights (a0, a1, a2, a3), sum P: vertices, the edges (0,1), (1,2), (2,0),
    // (0,3), (1,3), (2,3) in their direction, the faces (0,1,3), (1,2,3), (2,0,3), (0,2,1) (interior
    // points as a triangle of order P - 3 in the face's vertex order), then the interior (P = 4:
    // the centroid)
    auto add = [&](const int w[4]) {
      for (int d = 1; d <= 3; d++) xi.push_back(static_cast<double>(w[d]) / P);
    };
    for (int v = 0; v < 4; v++) {
      int w[4] = {0, 0, 0, 0};
      w[v] = P;
      add(w);
    }
    static const int kEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (const auto& ed : kEdge)
      for (int k = 1; k < P; k++) {
        int w[4] = {0, 0, 0, 0};
        w[ed[0]] = P - k;
        w[ed[1]] = k;
        add(w);
      }
    static const int kFace[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
    std::vector<std::pair<int, int>> inner;  // (i, j) lattice of the face interior
    if (P == 3) inner = {{1, 1}};
    if (P == 4) inner = {{1, 1}, {2, 1}, {1, 2}};
    for (const auto& f : kFace)
      for (const auto& q : inner) {
        int w[4] = {0, 0, 0, 0};
        w[f[0]] = P - q.first - q.second;
        w[f[1]] = q.first;
        w[f[2]] = q.second;
        add(w);
      }
    if (P == 4) {
      const int w[4] = {1, 1, 1, 1};
      add(w);
    }
    vtk_type = 71;  // VTK_LAGRANGE_TETRAHEDRON
    return;
  }
  if (dim == 1) {
    xi.push_back(t(0));
    xi.push_back(t(P));
    for (int i = 1; i < P; i++) xi.push_back(t(i));
    vtk_type = 68;  // VTK_LAGRANGE_CURVE
    return;
  }
  if (dim == 2) {
    xi.assign(static_cast<size_t>(n1) * n1 * 2, 0.0);
    for (int j = 0; j <= P; j++)
      for (int i = 0; i <= P; i++) {
        const int id = quadNode(i, j, P);
        xi[static_cast<size_t>(id) * 2] = t(i);
        xi[static_cast<size_t>(id) * 2 + 1] = t(j);
      }
    vtk_type = 70;  // VTK_LAGRANGE_QUADRILATERAL
    return;
  }
  xi.assign(static_cast<size_t>(n1) * n1 * n1 * 3, 0.0);
  for (int k = 0; k <= P; k++)
    for (int j = 0; j <= P; j++)
      for (int i = 0; i <= P; i++) {
        const int id = hexNode(i, j, k, P);
        xi[static_cast<size_t>(id) * 3] = t(i);
        xi[static_cast<size_t>(id) * 3 + 1] = t(j);
        xi[static_cast<size_t>(id) * 3 + 2] = t(k);
      }
  vtk_type = 72;  // VTK_LAGRANGE_HEXAHEDRON
}

const char* variableName(int v) {  // magic_enum names of ViewVariableEnum (Enum.cpp:210-233)
  static const char* kNames[] = {"Density",     "Velocity",   "Temperature", "Pressure",    "SoundSpeed",
                                 "MachNumber",  "Entropy",    "Vorticity",   "HeatFlux",    "ArtificialViscosity",
                                 "VelocityX",   "VelocityY",  "VelocityZ",   "MachNumberX", "MachNumberY",
                                 "MachNumberZ", "VorticityX", "VorticityY",  "VorticityZ",  "HeatFluxX",
                                 "HeatFluxY",   "HeatFluxZ"};
  return v >= 0 && v <= SDG_VIEW_HEAT_FLUX_Z ? kNames[v] : nullptr;
}

bool needsGradient(int v) {
  return v == SDG_VIEW_VORTICITY || v == SDG_VIEW_HEAT_FLUX || (v >= SDG_VIEW_VORTICITY_X && v <= SDG_VIEW_HEAT_FLUX_Z);
}

// components written for a variable (Paraview.cpp:43-57)
int componentsOf(int v, int dim) {
  if (dim >= 2 && (v == SDG_VIEW_VELOCITY || v == SDG_VIEW_MACH_NUMBER || v == SDG_VIEW_HEAT_FLUX ||
                   (v == SDG_VIEW_VORTICITY && dim == 3)))
    return 3;
  return 1;
}

// point state: conserved u, computational (rho, u.., e, p) and, with a gradient, the primitive
// gradient dv (velocity rows k*D+d = d u_k / d x_d) and d T / d x_d
struct NodeState {
  double rho = 0, vel[3] = {0, 0, 0}, e = 0, p = 0, c = 0, T = 0;
  double dvel[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dT[3] = {0, 0, 0};
  double av = 0;
};

// the scalar of view variable v (ViewVariable::get, VariableConvertor.cpp:756-882)
double scalarOf(int v, const NodeState& s, int dim) {
  const double speed2 = s.vel[0] * s.vel[0] + s.vel[1] * s.vel[1] + s.vel[2] * s.vel[2];
  auto du = [&](int k, int d) { return k < dim && d < dim ? s.dvel[k * dim + d] : 0.0; };
  switch (v) {
    case SDG_VIEW_DENSITY: return s.rho;
    case SDG_VIEW_VELOCITY: return std::sqrt(speed2);
    case SDG_VIEW_TEMPERATURE: return s.T;
    case SDG_VIEW_PRESSURE: return s.p;
    case SDG_VIEW_SOUND_SPEED: return s.c;
    case SDG_VIEW_MACH_NUMBER: return std::sqrt(speed2) / s.c;
    case SDG_VIEW_ENTROPY: return s.p / std::pow(s.rho, kGamma);
    case SDG_VIEW_VORTICITY: {
      if (dim == 2) return du(1, 0) - du(0, 1);
      const double wx = du(2, 1) - du(1, 2), wy = du(0, 2) - du(2, 0), wz = du(1, 0) - du(0, 1);
      return std::sqrt(wx * wx + wy * wy + wz * wz);
    }
    case SDG_VIEW_ARTIFICIAL_VISCOSITY: return s.av;
    case SDG_VIEW_VELOCITY_X: return s.vel[0];
    case SDG_VIEW_VELOCITY_Y: return s.vel[1];
    case SDG_VIEW_VELOCITY_Z: return s.vel[2];
    case SDG_VIEW_MACH_NUMBER_X: return s.vel[0] / s.c;
    case SDG_VIEW_MACH_NUMBER_Y: return s.vel[1] / s.c;
    case SDG_VIEW_MACH_NUMBER_Z: return s.vel[2] / s.c;
    case SDG_VIEW_VORTICITY_X: return du(2, 1) - du(1, 2);
    case SDG_VIEW_VORTICITY_Y: return du(0, 2) - du(2, 0);
    case SDG_VIEW_VORTICITY_Z: return du(1, 0) - du(0, 1);
    case SDG_VIEW_HEAT_FLUX_X: return s.dT[0];
    case SDG_VIEW_HEAT_FLUX_Y: return s.dT[1];
    case SDG_VIEW_HEAT_FLUX_Z: return s.dT[2];
    default: return 0.0;
  }
}

void writeArray(FILE* f, const char* type, const char* name, int comps, const std::vector<double>& v) {
  std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"ascii\">\n", type, name,
               comps);
  for (size_t i = 0; i < v.size(); i++) std::fprintf(f, "%.17g%c", v[i], (i + 1) % static_cast<size_t>(comps) ? ' ' : '\n');
  std::fprintf(f, "        </DataArray>\n");
}

}  // namespace

extern "C" {

int sdg_write_view(const sdg_mesh* mesh_handle, const sdg_problem_desc* problem, const double* coefficients,
                   const double* gradient, const double* node_values, const int* variables, int variable_count, int step,
                   double time, const char* path) {
  if (!mesh_handle || !problem || !coefficients || !path || (variable_count > 0 && !variables) || variable_count < 0)
    return vtuFail(SDG_E_INVALID, "null argument");
  const sdg::Mesh* mp = sdg::meshOf(mesh_handle);
  if (!mp) return vtuFail(SDG_E_UNSUPPORTED, "meshes of mixed element types are not supported");
  if (problem->abi_version != SDG_ABI_VERSION) return vtuFail(SDG_E_INVALID, "abi_version mismatch");
  if (!(problem->cv > 0.0)) return vtuFail(SDG_E_INVALID, "specific heats must be positive");
  const sdg::Mesh& m = *mp;
  const sdg::RefElement& ref = m.ref;
  if (ref.type == SDG_ELEM_PYRAMID)  // the Lagrange nodes include the apex, where the pyramid basis is 0/0
    return vtuFail(SDG_E_UNSUPPORTED, "pyramids are not written");
  const int D = ref.dim, nb = ref.nb, nv = D + 2, ng = nv * D;
  for (int i = 0; i < variable_count; i++) {
    if (!variableName(variables[i])) return vtuFail(SDG_E_INVALID, "unknown view variable");
    if (needsGradient(variables[i]) && !gradient)
      return vtuFail(SDG_E_INVALID, std::string(variableName(variables[i])) +
                                        " needs the lifted gradient (Navier-Stokes, sdg_get_gradient)");
    if (variables[i] == SDG_VIEW_ARTIFICIAL_VISCOSITY && !node_values)
      return vtuFail(SDG_E_INVALID, "ArtificialViscosity needs the node values");
  }
  const int64_t E = m.elements;
  std::vector<double> xi;
  int vtk_type = 0;
  lagrangeNodes(ref, xi, vtk_type);
  const int np = static_cast<int>(xi.size()) / D;
  // basis and P1 nodal functions at the nodes
  std::vector<double> B(static_cast<size_t>(np) * nb), N(static_cast<size_t>(np) * ref.nvert);
  for (int p = 0; p < np; p++) {
    sdg::basisValues(ref, &xi[static_cast<size_t>(p) * D], &B[static_cast<size_t>(p) * nb]);
    sdg::nodalValues(ref, &xi[static_cast<size_t>(p) * D], &N[static_cast<size_t>(p) * ref.nvert]);
  }
  const size_t npts = static_cast<size_t>(E) * np;
  std::vector<double> xyz(npts * 3, 0.0);
  std::vector<std::vector<double>> data(static_cast<size_t>(variable_count));
  for (int i = 0; i < variable_count; i++) data[i].assign(npts * componentsOf(variables[i], D), 0.0);
  const bool wc = problem->equation_of_state == SDG_EOS_WEAK_COMPRESSIBLE;
  const double c0 = problem->reference_sound_speed, rho0 = problem->reference_density;
  for (int64_t e = 0; e < E; e++)
    for (int p = 0; p < np; p++) {
      const size_t o = static_cast<size_t>(e) * np + p;
      double x[3] = {0.0, 0.0, 0.0};
      sdg::mapReferencePoint(m, e, &xi[static_cast<size_t>(p) * D], x);
      for (int d = 0; d < D; d++) xyz[o * 3 + d] = x[d];
      double u[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, g[15] = {};
      for (int b = 0; b < nb; b++) {
        const double bv = B[static_cast<size_t>(p) * nb + b];
        for (int v = 0; v < nv; v++) u[v] += bv * coefficients[(static_cast<size_t>(e) * nb + b) * nv + v];
        if (gradient)
          for (int r = 0; r < ng; r++) g[r] += bv * gradient[(static_cast<size_t>(e) * nb + b) * ng + r];
      }
      // conserved -> computational (VariableConvertor.cpp:315-339)
      NodeState s;
      s.rho = u[0];
      double v2 = 0.0;
      for (int d = 0; d < D; d++) {
        s.vel[d] = u[1 + d] / u[0];
        v2 += s.vel[d] * s.vel[d];
      }
      s.e = u[D + 1] / u[0] - v2 / 2.0;
      s.p = wc ? c0 * c0 * (u[0] - rho0) + 0.01 * rho0 * c0 * c0 : (kGamma - 1.0) * u[0] * s.e;
      s.c = wc ? c0 : std::sqrt(kGamma * s.p / u[0]);
      s.T = s.e / problem->cv;
      // conserved gradient -> primitive gradient (VariableConvertor.cpp:584-638): grad u_k =
      // (grad m_k - u_k grad rho) / rho, grad e = (grad(rho E) - (rho E / rho) grad rho) / rho - u . grad u
      if (gradient)
        for (int d = 0; d < D; d++) {
          const double drho = g[0 * D + d];
          double ude = 0.0;
          for (int k = 0; k < D; k++) {
            s.dvel[k * D + d] = (g[(1 + k) * D + d] - s.vel[k] * drho) / u[0];
            ude += s.vel[k] * s.dvel[k * D + d];
          }
          const double de = (g[(D + 1) * D + d] - u[D + 1] / u[0] * drho) / u[0] - ude;
          s.dT[d] = de / problem->cv;
        }
      if (node_values) {
        double a = 0.0;
        for (int j = 0; j < ref.nvert; j++)
          a += N[static_cast<size_t>(p) * ref.nvert + j] * node_values[m.node_tag[static_cast<size_t>(e) * ref.nvert + j]];
        s.av = a;
      }
      for (int i = 0; i < variable_count; i++) {
        const int v = variables[i];
        if (componentsOf(v, D) == 3) {
          const int base = v == SDG_VIEW_VELOCITY ? SDG_VIEW_VELOCITY_X
                                                  : (v == SDG_VIEW_MACH_NUMBER ? SDG_VIEW_MACH_NUMBER_X
                                                                               : (v == SDG_VIEW_HEAT_FLUX ? SDG_VIEW_HEAT_FLUX_X
                                                                                                          : SDG_VIEW_VORTICITY_X));
          for (int c = 0; c < 3; c++) data[i][o * 3 + c] = c < D ? scalarOf(base + c, s, D) : 0.0;
        } else if (D == 1 && (v == SDG_VIEW_VELOCITY || v == SDG_VIEW_MACH_NUMBER || v == SDG_VIEW_HEAT_FLUX)) {
          data[i][o] = scalarOf(v == SDG_VIEW_VELOCITY ? SDG_VIEW_VELOCITY_X
                                                       : (v == SDG_VIEW_MACH_NUMBER ? SDG_VIEW_MACH_NUMBER_X : SDG_VIEW_HEAT_FLUX_X),
                                s, D);
        } else {
          data[i][o] = scalarOf(v, s, D);
        }
      }
    }
  FILE* f = std::fopen(path, "w");
  if (!f) return vtuFail(SDG_E_INVALID, std::string("cannot open ") + path);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                  "header_type=\"UInt64\">\n  <UnstructuredGrid>\n    <FieldData>\n");
  std::fprintf(f, "      <DataArray type=\"Int32\" Name=\"TMSTEP\" NumberOfTuples=\"1\" format=\"ascii\">%d</DataArray>\n", step);
  std::fprintf(f, "      <DataArray type=\"Float64\" Name=\"TimeValue\" NumberOfTuples=\"1\" format=\"ascii\">%.17g</DataArray>\n",
               time);
  std::fprintf(f, "    </FieldData>\n    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%lld\">\n      <PointData>\n", npts,
               static_cast<long long>(E));
  for (int i = 0; i < variable_count; i++)
    writeArray(f, "Float64", variableName(variables[i]), componentsOf(variables[i], D), data[i]);
  std::fprintf(f, "      </PointData>\n      <Points>\n");
  writeArray(f, "Float64", "Points", 3, xyz);
  std::fprintf(f, "      </Points>\n      <Cells>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (int64_t e = 0; e < E; e++) {
    for (int k = 0; k < np; k++) std::fprintf(f, "%lld ", static_cast<long long>(e) * np + k);
    std::fputc('\n', f);
  }
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (int64_t c = 1; c <= E; c++) std::fprintf(f, "%lld\n", static_cast<long long>(c) * np);
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int64_t c = 0; c < E; c++) std::fprintf(f, "%d\n", vtk_type);
  std::fprintf(f, "        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::fclose(f) == 0;
  return ok ? SDG_OK : vtuFail(SDG_E_INVALID, std::string("write failed: ") + path);
}

int sdg_write_vtu(const sdg_mesh* mesh, const sdg_problem_desc* problem, const double* coefficients,
                  const double* node_values, int step, double time, const char* path) {
  const int vars[] = {SDG_VIEW_DENSITY, SDG_VIEW_VELOCITY, SDG_VIEW_PRESSURE, SDG_VIEW_TEMPERATURE, SDG_VIEW_MACH_NUMBER,
                      SDG_VIEW_ARTIFICIAL_VISCOSITY};
  return sdg_write_view(mesh, problem, coefficients, nullptr, node_values, vars, node_values ? 6 : 5, step, time, path);
}

}  // extern "C"
