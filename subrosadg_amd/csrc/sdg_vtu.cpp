// sdg_vtu.cpp -- VTK XML unstructured-grid output of a state (the reference's View::stepView,
// src/View/Paraview.cpp:280-313, which writes VTK Lagrange cells through vtu11).
//
// Every element is sampled at its equispaced points of order P -- (P+1)^D on quadrangles and
// hexahedra, (P+1)(P+2)/2 on triangles -- and written as the P^D linear sub-cells (P^2 triangles)
// through those points; points are per element, so the DG discontinuities stay visible. Point
// data: Density, Velocity (3 components), Pressure, Temperature, MachNumber and, when node values
// are given, ArtificialViscosity (their P1 interpolant, as the reference's node-based output);
// field data TMSTEP and TimeValue (Paraview.cpp:43-46). ASCII, 17 significant digits.
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
  sdg::setGlobalError("sdg_write_vtu: " + msg);
  return code;
}

// reference coordinates of the sample points and the linear sub-cells through them
void samplePattern(const sdg::RefElement& ref, std::vector<double>& xi, std::vector<int>& cells, int& cell_nodes,
                   int& vtk_type) {
  const int P = ref.order, dim = ref.dim, n1 = P + 1;
  xi.clear();
  cells.clear();
  if (ref.type == SDG_ELEM_TRIANGLE) {
    auto idx = [&](int i, int j) { return j * n1 - j * (j - 1) / 2 + i; };  // rows of decreasing length
    for (int j = 0; j <= P; j++)
      for (int i = 0; i + j <= P; i++) {
        xi.push_back(static_cast<double>(i) / P);
        xi.push_back(static_cast<double>(j) / P);
      }
    for (int j = 0; j < P; j++)
      for (int i = 0; i + j < P; i++) {
        cells.insert(cells.end(), {idx(i, j), idx(i + 1, j), idx(i, j + 1)});
        if (i + j < P - 1) cells.insert(cells.end(), {idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)});
      }
    cell_nodes = 3;
    vtk_type = 5;  // VTK_TRIANGLE
    return;
  }
  auto t = [&](int i) { return -1.0 + 2.0 * i / P; };
  if (dim == 1) {
    for (int i = 0; i <= P; i++) xi.push_back(t(i));
    for (int i = 0; i < P; i++) cells.insert(cells.end(), {i, i + 1});
    cell_nodes = 2;
    vtk_type = 3;  // VTK_LINE
    return;
  }
  if (dim == 2) {
    for (int j = 0; j <= P; j++)
      for (int i = 0; i <= P; i++) {
        xi.push_back(t(i));
        xi.push_back(t(j));
      }
    for (int j = 0; j < P; j++)
      for (int i = 0; i < P; i++) {
        const int a = i + n1 * j;
        cells.insert(cells.end(), {a, a + 1, a + 1 + n1, a + n1});
      }
    cell_nodes = 4;
    vtk_type = 9;  // VTK_QUAD
    return;
  }
  for (int k = 0; k <= P; k++)
    for (int j = 0; j <= P; j++)
      for (int i = 0; i <= P; i++) {
        xi.push_back(t(i));
        xi.push_back(t(j));
        xi.push_back(t(k));
      }
  const int s = n1 * n1;
  for (int k = 0; k < P; k++)
    for (int j = 0; j < P; j++)
      for (int i = 0; i < P; i++) {
        const int a = i + n1 * j + s * k;
        cells.insert(cells.end(), {a, a + 1, a + 1 + n1, a + n1, a + s, a + 1 + s, a + 1 + n1 + s, a + n1 + s});
      }
  cell_nodes = 8;
  vtk_type = 12;  // VTK_HEXAHEDRON
}

void writeArray(FILE* f, const char* type, const char* name, int comps, const std::vector<double>& v) {
  std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"ascii\">\n", type, name,
               comps);
  for (size_t i = 0; i < v.size(); i++) std::fprintf(f, "%.17g%c", v[i], (i + 1) % static_cast<size_t>(comps) ? ' ' : '\n');
  std::fprintf(f, "        </DataArray>\n");
}

}  // namespace

extern "C" {

int sdg_write_vtu(const sdg_mesh* mesh_handle, const sdg_problem_desc* problem, const double* coefficients,
                  const double* node_values, int step, double time, const char* path) {
  if (!mesh_handle || !problem || !coefficients || !path) return vtuFail(SDG_E_INVALID, "null argument");
  const sdg::Mesh* mp = sdg::meshOf(mesh_handle);
  if (!mp) return vtuFail(SDG_E_UNSUPPORTED, "meshes of mixed element types are not supported");
  if (problem->abi_version != SDG_ABI_VERSION) return vtuFail(SDG_E_INVALID, "abi_version mismatch");
  if (!(problem->cv > 0.0)) return vtuFail(SDG_E_INVALID, "specific heats must be positive");
  const sdg::Mesh& m = *mp;
  const sdg::RefElement& ref = m.ref;
  const int D = ref.dim, nb = ref.nb, nv = D + 2;
  const int64_t E = m.elements;
  std::vector<double> xi;
  std::vector<int> pattern;
  int cell_nodes = 0, vtk_type = 0;
  samplePattern(ref, xi, pattern, cell_nodes, vtk_type);
  const int np = static_cast<int>(xi.size()) / D, nc = static_cast<int>(pattern.size()) / cell_nodes;
  // basis and P1 nodal functions at the sample points
  std::vector<double> B(static_cast<size_t>(np) * nb), N(static_cast<size_t>(np) * ref.nvert);
  for (int p = 0; p < np; p++) {
    sdg::basisValues(ref, &xi[static_cast<size_t>(p) * D], &B[static_cast<size_t>(p) * nb]);
    sdg::nodalValues(ref, &xi[static_cast<size_t>(p) * D], &N[static_cast<size_t>(p) * ref.nvert]);
  }
  const size_t npts = static_cast<size_t>(E) * np;
  std::vector<double> xyz(npts * 3, 0.0), rho(npts), vel(npts * 3, 0.0), pres(npts), temp(npts), mach(npts), av;
  if (node_values) av.resize(npts);
  for (int64_t e = 0; e < E; e++)
    for (int p = 0; p < np; p++) {
      const size_t o = static_cast<size_t>(e) * np + p;
      double x[3] = {0.0, 0.0, 0.0};
      sdg::mapReferencePoint(m, e, &xi[static_cast<size_t>(p) * D], x);
      for (int d = 0; d < D; d++) xyz[o * 3 + d] = x[d];
      double u[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int b = 0; b < nb; b++)
        for (int v = 0; v < nv; v++) u[v] += B[static_cast<size_t>(p) * nb + b] * coefficients[(static_cast<size_t>(e) * nb + b) * nv + v];
      // conserved -> computational -> primitive (VariableConvertor.cpp:315-339, 118-139)
      double v2 = 0.0;
      for (int d = 0; d < D; d++) {
        vel[o * 3 + d] = u[1 + d] / u[0];
        v2 += vel[o * 3 + d] * vel[o * 3 + d];
      }
      const double ie = u[D + 1] / u[0] - v2 / 2.0;
      const double pr = problem->equation_of_state == SDG_EOS_IDEAL_GAS
                            ? (kGamma - 1.0) * u[0] * ie
                            : problem->reference_sound_speed * problem->reference_sound_speed * (u[0] - problem->reference_density) +
                                  0.01 * problem->reference_density * problem->reference_sound_speed * problem->reference_sound_speed;
      const double c = problem->equation_of_state == SDG_EOS_IDEAL_GAS ? std::sqrt(kGamma * pr / u[0])
                                                                       : problem->reference_sound_speed;
      rho[o] = u[0];
      pres[o] = pr;
      temp[o] = ie / problem->cv;
      mach[o] = std::sqrt(v2) / c;
      if (node_values) {
        double a = 0.0;
        for (int j = 0; j < ref.nvert; j++)
          a += N[static_cast<size_t>(p) * ref.nvert + j] * node_values[m.node_tag[static_cast<size_t>(e) * ref.nvert + j]];
        av[o] = a;
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
               static_cast<long long>(E) * nc);
  writeArray(f, "Float64", "Density", 1, rho);
  writeArray(f, "Float64", "Velocity", 3, vel);
  writeArray(f, "Float64", "Pressure", 1, pres);
  writeArray(f, "Float64", "Temperature", 1, temp);
  writeArray(f, "Float64", "MachNumber", 1, mach);
  if (node_values) writeArray(f, "Float64", "ArtificialViscosity", 1, av);
  std::fprintf(f, "      </PointData>\n      <Points>\n");
  writeArray(f, "Float64", "Points", 3, xyz);
  std::fprintf(f, "      </Points>\n      <Cells>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (int64_t e = 0; e < E; e++)
    for (int c = 0; c < nc; c++) {
      for (int k = 0; k < cell_nodes; k++)
        std::fprintf(f, "%lld ", static_cast<long long>(e) * np + pattern[static_cast<size_t>(c) * cell_nodes + k]);
      std::fputc('\n', f);
    }
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (int64_t c = 1; c <= E * nc; c++) std::fprintf(f, "%lld\n", static_cast<long long>(c) * cell_nodes);
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int64_t c = 0; c < E * nc; c++) std::fprintf(f, "%d\n", vtk_type);
  std::fprintf(f, "        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::fclose(f) == 0;
  return ok ? SDG_OK : vtuFail(SDG_E_INVALID, std::string("write failed: ") + path);
}

}  // extern "C"
