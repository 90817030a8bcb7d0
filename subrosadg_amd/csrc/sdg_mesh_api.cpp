// sdg_mesh_api.cpp -- mesh and operand-export part of the C-ABI (host only, no device work).
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "kit.hpp"
#include "sdg_internal.hpp"

struct sdg_mesh {
  sdg::Mesh mesh;         // single element type (all structured meshes, uniform split grids)
  sdg::MixedMesh mixed;   // split grids: per-element triangle / quadrangle description
  bool single = true;
  bool has_mixed = false;
};

namespace sdg {

thread_local std::string g_last_error;

void setGlobalError(const std::string& msg) { g_last_error = msg; }
const char* globalError() { return g_last_error.c_str(); }
const Mesh* meshOf(const sdg_mesh* m) { return m && m->single ? &m->mesh : nullptr; }
const MixedMesh* mixedOf(const sdg_mesh* m) { return m && m->has_mixed ? &m->mixed : nullptr; }

}  // namespace sdg

extern "C" {

int sdg_abi_version(void) { return SDG_ABI_VERSION; }

int sdg_mesh_create_gmsh(const sdg_gmsh_desc* desc, sdg_mesh** out) {
  if (!desc || !out || !desc->path || (desc->boundary_count > 0 && (!desc->physical_index || !desc->boundary))) {
    sdg::setGlobalError("sdg_mesh_create_gmsh: null argument");
    return SDG_E_INVALID;
  }
  *out = nullptr;
  sdg_mesh* m = new (std::nothrow) sdg_mesh();
  if (!m) {
    sdg::setGlobalError("sdg_mesh_create_gmsh: out of host memory");
    return SDG_E_NOMEM;
  }
  std::string err;
  try {
    std::map<int, int> bc;
    for (int i = 0; i < desc->boundary_count; i++) bc[desc->physical_index[i]] = desc->boundary[i];
    if (!sdg::readGmshMesh(desc->path, desc->polynomial_order, bc, m->mesh, err)) {
      delete m;
      sdg::setGlobalError("sdg_mesh_create_gmsh: " + err);
      return SDG_E_INVALID;
    }
  } catch (const std::exception& e) {
    delete m;
    sdg::setGlobalError(std::string("sdg_mesh_create_gmsh: ") + e.what());
    return SDG_E_NOMEM;
  }
  *out = m;
  return SDG_OK;
}

int sdg_mesh_create_from_arrays(const sdg_array_mesh_desc* d, sdg_mesh** out) {
  auto refuse = [](const std::string& msg) {
    sdg::setGlobalError("sdg_mesh_create_from_arrays: " + msg);
    return SDG_E_INVALID;
  };
  if (!d || !out) return refuse("null argument");
  *out = nullptr;
  const int type = d->element;
  int fvert = 0;
  if (type == SDG_ELEM_TRIANGLE || type == SDG_ELEM_QUADRANGLE) {
    fvert = 2;
  } else if (type == SDG_ELEM_TETRAHEDRON) {
    fvert = 3;
  } else if (type == SDG_ELEM_HEXAHEDRON || type == SDG_ELEM_PYRAMID) {
    fvert = 4;
  } else {
    sdg::setGlobalError("sdg_mesh_create_from_arrays: element must be triangle, quadrangle, tetrahedron, pyramid or "
                        "hexahedron");
    return SDG_E_UNSUPPORTED;
  }
  const int gorder = d->geometry_order > 0 ? d->geometry_order : 1;
  sdg::GeomMap gm;
  std::string gerr;
  if (!sdg::buildGeomMap(type, gorder, gm, gerr)) {
    sdg::setGlobalError("sdg_mesh_create_from_arrays: " + gerr);
    return SDG_E_UNSUPPORTED;
  }
  const int nvert = gm.n;  // nodes per element in `elements`
  if (d->node_count < 1 || !d->nodes || d->element_count < 1 || !d->elements) return refuse("no nodes or elements");
  if (d->boundary_face_count < 0 || (d->boundary_face_count > 0 && (!d->boundary_faces || !d->boundary_type)))
    return refuse("boundary faces without node lists or types");
  if (d->periodic_link_count < 0 || (d->periodic_link_count > 0 && (!d->periodic_link_size || !d->periodic_pairs)))
    return refuse("periodic links without pairs");
  for (int64_t i = 0; i < d->element_count * nvert; i++)
    if (d->elements[i] < 0 || d->elements[i] >= d->node_count) return refuse("element node out of range");
  for (int64_t i = 0; i < d->boundary_face_count * fvert; i++) {
    const bool pad = type == SDG_ELEM_PYRAMID && i % fvert == 3 && d->boundary_faces[i] == -1;  // a triangle
    if (!pad && (d->boundary_faces[i] < 0 || d->boundary_faces[i] >= d->node_count))
      return refuse("boundary face node out of range");
  }
  std::vector<std::map<int64_t, int64_t>> links(static_cast<size_t>(d->periodic_link_count));
  int64_t off = 0;
  for (int l = 0; l < d->periodic_link_count; l++) {
    if (d->periodic_link_size[l] < 0) return refuse("negative periodic link size");
    for (int64_t k = 0; k < d->periodic_link_size[l]; k++, off++) {
      const int64_t sl = d->periodic_pairs[2 * off], ms = d->periodic_pairs[2 * off + 1];
      if (sl < 0 || sl >= d->node_count || ms < 0 || ms >= d->node_count) return refuse("periodic node out of range");
      if (sl != ms) links[static_cast<size_t>(l)][sl] = ms;
    }
  }
  sdg_mesh* m = new (std::nothrow) sdg_mesh();
  if (!m) {
    sdg::setGlobalError("sdg_mesh_create_from_arrays: out of host memory");
    return SDG_E_NOMEM;
  }
  std::string err;
  try {
    std::vector<int> phys(static_cast<size_t>(d->boundary_face_count), -1);
    if (d->boundary_physical)
      for (int64_t i = 0; i < d->boundary_face_count; i++) phys[static_cast<size_t>(i)] = d->boundary_physical[i];
    if (!sdg::buildMeshFromConnectivity(type, d->polynomial_order, d->node_count, d->nodes, d->element_count,
                                        d->elements, d->boundary_face_count, d->boundary_faces, d->boundary_type,
                                        phys.data(), m->mesh, err, links.empty() ? nullptr : &links, gorder)) {
      delete m;
      sdg::setGlobalError("sdg_mesh_create_from_arrays: " + err);
      return SDG_E_INVALID;
    }
  } catch (const std::exception& e) {
    delete m;
    sdg::setGlobalError(std::string("sdg_mesh_create_from_arrays: ") + e.what());
    return SDG_E_NOMEM;
  }
  *out = m;
  return SDG_OK;
}

int sdg_mesh_create_structured(const sdg_structured_desc* desc, sdg_mesh** out) {
  if (!desc || !out) {
    sdg::setGlobalError("sdg_mesh_create_structured: null argument");
    return SDG_E_INVALID;
  }
  *out = nullptr;
  if (desc->element != SDG_ELEM_LINE && desc->element != SDG_ELEM_QUADRANGLE && desc->element != SDG_ELEM_HEXAHEDRON) {
    sdg::setGlobalError("sdg_mesh_create_structured: element must be line, quadrangle or hexahedron");
    return SDG_E_UNSUPPORTED;
  }
  sdg_mesh* m = new (std::nothrow) sdg_mesh();
  if (!m) {
    sdg::setGlobalError("sdg_mesh_create_structured: out of host memory");
    return SDG_E_NOMEM;
  }
  std::string err;
  try {
    int periodic[3], bc[6];
    for (int d = 0; d < 3; d++) periodic[d] = desc->bounded[d] ? 0 : 1;
    for (int i = 0; i < 6; i++) bc[i] = desc->boundary[i];
    if (!sdg::buildStructuredMesh(desc->element, desc->polynomial_order, desc->n, desc->lo, desc->hi, periodic, bc,
                                  m->mesh, err)) {
      delete m;
      sdg::setGlobalError("sdg_mesh_create_structured: " + err);
      return SDG_E_INVALID;
    }
  } catch (const std::bad_alloc&) {
    delete m;
    sdg::setGlobalError("sdg_mesh_create_structured: out of host memory");
    return SDG_E_NOMEM;
  }
  *out = m;
  return SDG_OK;
}

int sdg_mesh_create_split(const sdg_split_desc* desc, sdg_mesh** out) {
  if (!desc || !out) {
    sdg::setGlobalError("sdg_mesh_create_split: null argument");
    return SDG_E_INVALID;
  }
  *out = nullptr;
  if (desc->n[0] < 2 || desc->n[1] < 2 || static_cast<int64_t>(desc->n[0]) * desc->n[1] > (int64_t(1) << 28)) {
    sdg::setGlobalError("sdg_mesh_create_split: bad cell counts");
    return SDG_E_INVALID;
  }
  sdg_mesh* m = new (std::nothrow) sdg_mesh();
  if (!m) {
    sdg::setGlobalError("sdg_mesh_create_split: out of host memory");
    return SDG_E_NOMEM;
  }
  // split flags: Bernoulli(split_probability) per cell from a splitmix64 stream of the seed
  std::vector<uint8_t> split(static_cast<size_t>(desc->n[0]) * desc->n[1]);
  uint64_t state = desc->seed;
  for (auto& f : split) {
    state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    f = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0) < desc->split_probability ? 1 : 0;
  }
  std::string err;
  try {
    if (!sdg::buildSplitMesh(desc->polynomial_order, desc->n, desc->lo, desc->hi, split, m->mixed, err)) {
      delete m;
      sdg::setGlobalError("sdg_mesh_create_split: " + err);
      return SDG_E_INVALID;
    }
  } catch (const std::bad_alloc&) {
    delete m;
    sdg::setGlobalError("sdg_mesh_create_split: out of host memory");
    return SDG_E_NOMEM;
  }
  m->has_mixed = true;
  m->single = sdg::singleTypeView(m->mixed, m->mesh);
  *out = m;
  return SDG_OK;
}

void sdg_mesh_destroy(sdg_mesh* mesh) { delete mesh; }

int sdg_mesh_get_info(const sdg_mesh* mesh, sdg_mesh_info* info) {
  if (!mesh || !info) {
    sdg::setGlobalError("sdg_mesh_get_info: null argument");
    return SDG_E_INVALID;
  }
  if (!mesh->single) {  // triangles and quadrangles
    const sdg::MixedMesh& x = mesh->mixed;
    *info = sdg_mesh_info{};
    info->dimension = 2;
    info->element = SDG_ELEM_MIXED_TRIANGLE_QUADRANGLE;
    info->polynomial_order = x.order;
    info->elements = x.elements;
    info->faces = static_cast<int64_t>(x.faces.size());
    info->affine = 1;
    return SDG_OK;
  }
  const sdg::Mesh& m = mesh->mesh;
  info->dimension = m.ref.dim;
  info->element = m.ref.type;
  info->polynomial_order = m.ref.order;
  info->elements = m.elements;
  info->faces = static_cast<int64_t>(m.faces.size());
  info->basis = m.ref.nb;
  info->quadrature = m.ref.nq;
  info->faces_per_element = m.ref.nf;
  info->face_quadrature = m.ref.nfq;
  info->affine = m.affine ? 1 : 0;
  info->boundary_faces = static_cast<int64_t>(m.bfaces.size());
  info->nodes = m.nodes;
  info->vertices = m.ref.nvert;
  return SDG_OK;
}

int64_t sdg_mesh_field_size(const sdg_mesh* mesh, int field) {
  if (!mesh) return 0;
  if (field == SDG_FIELD_ELEMENT_TYPE) return mesh->has_mixed ? mesh->mixed.elements : mesh->mesh.elements;
  if (!mesh->single) {
    const sdg::MixedMesh& x = mesh->mixed;
    int64_t nq = 0, nb2 = 0;
    for (int64_t e = 0; e < x.elements; e++) {
      nq += x.refOf(e).nq;
      nb2 += static_cast<int64_t>(x.refOf(e).nb) * x.refOf(e).nb;
    }
    const int64_t F = static_cast<int64_t>(x.faces.size()), nfq = x.ref[0].nfq;  // line faces of both types
    switch (field) {
      case SDG_FIELD_QUAD_COORD: return nq * 2;
      case SDG_FIELD_JACOBIAN_DET_W: return nq;
      case SDG_FIELD_JACOBIAN_TIW: return nq * 4;
      case SDG_FIELD_MASS_INVERSE: return nb2;
      case SDG_FIELD_MIN_EDGE: return x.elements;
      case SDG_FIELD_FACE_NORMAL: return F * nfq * 2;
      case SDG_FIELD_FACE_JACOBIAN_W: return F * nfq;
      case SDG_FIELD_FACE_COORD: return F * nfq * 2;
      case SDG_FIELD_FACE_PARENT: return F * 2;
      case SDG_FIELD_FACE_LOCAL: return F * 2;
      case SDG_FIELD_FACE_SEQUENCE: return F * nfq;
      default: return 0;
    }
  }
  const sdg::Mesh& m = mesh->mesh;
  const int64_t E = m.elements, F = static_cast<int64_t>(m.faces.size()), Fb = static_cast<int64_t>(m.bfaces.size());
  const int64_t D = m.ref.dim, nb = m.ref.nb, nq = m.ref.nq, nf = m.ref.nf, nfq = m.ref.nfq;
  switch (field) {
    case SDG_FIELD_BASIS_VALUE: return nq * nb;
    case SDG_FIELD_BASIS_GRADIENT: return nq * D * nb;
    case SDG_FIELD_BASIS_FACE_VALUE: return nf * nfq * nb;
    case SDG_FIELD_BASIS_LSQ_INVERSE: return nb * nb;
    case SDG_FIELD_QUAD_WEIGHT: return nq;
    case SDG_FIELD_FACE_QUAD_WEIGHT: return nfq;
    case SDG_FIELD_QUAD_COORD: return E * nq * D;
    case SDG_FIELD_JACOBIAN_DET_W: return E * nq;
    case SDG_FIELD_JACOBIAN_TIW: return E * nq * D * D;
    case SDG_FIELD_MASS_INVERSE: return E * nb * nb;
    case SDG_FIELD_MIN_EDGE: return E;
    case SDG_FIELD_FACE_NORMAL: return F * nfq * D;
    case SDG_FIELD_FACE_JACOBIAN_W: return F * nfq;
    case SDG_FIELD_FACE_COORD: return F * nfq * D;
    case SDG_FIELD_FACE_PARENT: return F * 2;
    case SDG_FIELD_FACE_LOCAL: return F * 2;
    case SDG_FIELD_FACE_SEQUENCE: return F * nfq;
    case SDG_FIELD_FACE_ROTATION: return F;
    case SDG_FIELD_QUAD_POINT: return nq * D;
    case SDG_FIELD_FACE_QUAD_POINT: return nfq * (D - 1);
    case SDG_FIELD_BOUNDARY_FACE_COORD: return Fb * nfq * D;
    case SDG_FIELD_BOUNDARY_FACE_NORMAL: return Fb * nfq * D;
    case SDG_FIELD_BOUNDARY_FACE_JACOBIAN_W: return Fb * nfq;
    case SDG_FIELD_BOUNDARY_FACE_PARENT: return Fb * 2;
    case SDG_FIELD_BOUNDARY_FACE_TYPE: return Fb;
    case SDG_FIELD_INNER_RADIUS: return E;
    case SDG_FIELD_NODE_TAG: return E * mesh->mesh.ref.nvert;
    case SDG_FIELD_NODAL_VALUE: return nq * mesh->mesh.ref.nvert;
    case SDG_FIELD_NODAL_FACE_VALUE: return nf * nfq * mesh->mesh.ref.nvert;
    case SDG_FIELD_BOUNDARY_FACE_PHYSICAL: return Fb;
    default: return 0;
  }
}

int sdg_mesh_export(const sdg_mesh* mesh, int field, void* out, int64_t count) {
  const int64_t need = sdg_mesh_field_size(mesh, field);
  const bool boundary_field = (field >= SDG_FIELD_BOUNDARY_FACE_COORD && field <= SDG_FIELD_BOUNDARY_FACE_TYPE) ||
                              field == SDG_FIELD_BOUNDARY_FACE_PHYSICAL;
  if (mesh && boundary_field && need == 0) return SDG_OK;  // no boundary faces
  if (!mesh || !out || need == 0) {
    sdg::setGlobalError("sdg_mesh_export: null argument or unknown field");
    return SDG_E_INVALID;
  }
  if (count < need) {
    sdg::setGlobalError("sdg_mesh_export: output buffer too small");
    return SDG_E_INVALID;
  }
  if (field == SDG_FIELD_ELEMENT_TYPE) {
    int32_t* oi = static_cast<int32_t*>(out);
    if (mesh->has_mixed)
      for (int64_t e = 0; e < mesh->mixed.elements; e++) oi[e] = mesh->mixed.refOf(e).type;
    else
      for (int64_t e = 0; e < mesh->mesh.elements; e++) oi[e] = mesh->mesh.ref.type;
    return SDG_OK;
  }
  if (!mesh->single) {
    const sdg::MixedMesh& x = mesh->mixed;
    double* od = static_cast<double*>(out);
    int32_t* oi = static_cast<int32_t*>(out);
    const size_t nfq = static_cast<size_t>(x.ref[0].nfq);
    // per-element blocks concatenated in element order (each element's own N_q or N_b)
    if (field == SDG_FIELD_JACOBIAN_TIW || field == SDG_FIELD_MASS_INVERSE) {
      int64_t o = 0;
      for (int64_t e = 0; e < x.elements; e++) {
        const sdg::RefElement& r = x.refOf(e);
        if (field == SDG_FIELD_JACOBIAN_TIW) {
          sdg::affineGeometry(r, &x.jac[static_cast<size_t>(e) * 4], x.detj[e], nullptr, od + o);
          o += static_cast<int64_t>(r.nq) * 4;
        } else {
          sdg::affineMassInverse(r, x.detj[e], od + o);
          o += static_cast<int64_t>(r.nb) * r.nb;
        }
      }
      return SDG_OK;
    }
    for (size_t f = 0; field >= SDG_FIELD_FACE_NORMAL && field <= SDG_FIELD_FACE_SEQUENCE && f < x.faces.size(); f++) {
      const sdg::Face& fc = x.faces[f];
      switch (field) {
        case SDG_FIELD_FACE_NORMAL: std::memcpy(od + f * nfq * 2, fc.normal.data(), sizeof(double) * nfq * 2); break;
        case SDG_FIELD_FACE_JACOBIAN_W: std::memcpy(od + f * nfq, fc.jw.data(), sizeof(double) * nfq); break;
        case SDG_FIELD_FACE_COORD: std::memcpy(od + f * nfq * 2, fc.coord.data(), sizeof(double) * nfq * 2); break;
        case SDG_FIELD_FACE_PARENT:
          oi[2 * f] = fc.parent[0];
          oi[2 * f + 1] = fc.parent[1];
          break;
        case SDG_FIELD_FACE_LOCAL:
          oi[2 * f] = fc.local[0];
          oi[2 * f + 1] = fc.local[1];
          break;
        case SDG_FIELD_FACE_SEQUENCE: std::memcpy(oi + f * nfq, fc.seq.data(), sizeof(int32_t) * nfq); break;
        default: break;
      }
    }
    if (field >= SDG_FIELD_FACE_NORMAL && field <= SDG_FIELD_FACE_SEQUENCE) return SDG_OK;
    int64_t o = 0;
    for (int64_t e = 0; e < x.elements; e++) {
      const sdg::RefElement& r = x.refOf(e);
      for (int q = 0; q < r.nq; q++) {
        if (field == SDG_FIELD_QUAD_COORD) {
          double xq[2], jac[4];
          // affine map of the element's vertices
          const double* xi = &r.qpoint[static_cast<size_t>(q) * 2];
          double N[4], G[4][2];
          int nv = r.nvert;
          if (r.type == SDG_ELEM_TRIANGLE) {
            N[0] = 1.0 - xi[0] - xi[1];
            N[1] = xi[0];
            N[2] = xi[1];
          } else {
            for (int v = 0; v < 4; v++)
              N[v] = 0.25 * (1.0 + r.vertex[v * 2] * xi[0]) * (1.0 + r.vertex[v * 2 + 1] * xi[1]);
          }
          (void)G;
          (void)jac;
          xq[0] = xq[1] = 0.0;
          for (int v = 0; v < nv; v++)
            for (int d = 0; d < 2; d++) xq[d] += N[v] * x.vertex[static_cast<size_t>(e) * 8 + v * 2 + d];
          od[(o + q) * 2] = xq[0];
          od[(o + q) * 2 + 1] = xq[1];
        } else if (field == SDG_FIELD_JACOBIAN_DET_W) {
          od[o + q] = x.detj[e] * r.qweight[q];
        }
      }
      if (field == SDG_FIELD_MIN_EDGE) od[e] = x.min_edge[e];
      o += r.nq;
    }
    return SDG_OK;
  }
  const sdg::Mesh& m = mesh->mesh;
  const sdg::RefElement& r = m.ref;
  const int64_t E = m.elements;
  const int D = r.dim, nb = r.nb, nq = r.nq, nfq = r.nfq;
  double* od = static_cast<double*>(out);
  int32_t* oi = static_cast<int32_t*>(out);
  auto copy = [&](const std::vector<double>& v) { std::memcpy(od, v.data(), v.size() * sizeof(double)); };
  switch (field) {
    case SDG_FIELD_BASIS_VALUE: copy(r.V); break;
    case SDG_FIELD_BASIS_GRADIENT: copy(r.dV); break;
    case SDG_FIELD_BASIS_FACE_VALUE: copy(r.Vf); break;
    case SDG_FIELD_BASIS_LSQ_INVERSE: copy(r.lsq_inverse); break;
    case SDG_FIELD_QUAD_WEIGHT: copy(r.qweight); break;
    case SDG_FIELD_FACE_QUAD_WEIGHT: copy(r.fweight); break;
    case SDG_FIELD_QUAD_POINT: copy(r.qpoint); break;
    case SDG_FIELD_FACE_QUAD_POINT: copy(r.fpoint); break;
    case SDG_FIELD_QUAD_COORD:
      for (int64_t e = 0; e < E; e++) sdg::elementGeometry(m, e, nullptr, nullptr, od + e * nq * D);
      break;
    case SDG_FIELD_JACOBIAN_DET_W:
      for (int64_t e = 0; e < E; e++) sdg::elementGeometry(m, e, od + e * nq, nullptr, nullptr);
      break;
    case SDG_FIELD_JACOBIAN_TIW:
      for (int64_t e = 0; e < E; e++) sdg::elementGeometry(m, e, nullptr, od + e * nq * D * D, nullptr);
      break;
    case SDG_FIELD_MASS_INVERSE:
      for (int64_t e = 0; e < E; e++) sdg::elementMassInverse(m, e, od + e * nb * nb);
      break;
    case SDG_FIELD_MIN_EDGE: copy(m.min_edge); break;
    case SDG_FIELD_FACE_NORMAL:
      for (size_t f = 0; f < m.faces.size(); f++)
        std::memcpy(od + f * nfq * D, m.faces[f].normal.data(), sizeof(double) * nfq * D);
      break;
    case SDG_FIELD_FACE_JACOBIAN_W:
      for (size_t f = 0; f < m.faces.size(); f++) std::memcpy(od + f * nfq, m.faces[f].jw.data(), sizeof(double) * nfq);
      break;
    case SDG_FIELD_FACE_COORD:
      for (size_t f = 0; f < m.faces.size(); f++)
        std::memcpy(od + f * nfq * D, m.faces[f].coord.data(), sizeof(double) * nfq * D);
      break;
    case SDG_FIELD_FACE_PARENT:
      for (size_t f = 0; f < m.faces.size(); f++) {
        oi[2 * f] = m.faces[f].parent[0];
        oi[2 * f + 1] = m.faces[f].parent[1];
      }
      break;
    case SDG_FIELD_FACE_LOCAL:
      for (size_t f = 0; f < m.faces.size(); f++) {
        oi[2 * f] = m.faces[f].local[0];
        oi[2 * f + 1] = m.faces[f].local[1];
      }
      break;
    case SDG_FIELD_FACE_SEQUENCE:
      for (size_t f = 0; f < m.faces.size(); f++)
        std::memcpy(oi + f * nfq, m.faces[f].seq.data(), sizeof(int32_t) * nfq);
      break;
    case SDG_FIELD_FACE_ROTATION:
      for (size_t f = 0; f < m.faces.size(); f++) oi[f] = m.faces[f].rotation;
      break;
    case SDG_FIELD_BOUNDARY_FACE_COORD:
      for (size_t f = 0; f < m.bfaces.size(); f++)
        std::memcpy(od + f * nfq * D, m.bfaces[f].coord.data(), sizeof(double) * nfq * D);
      break;
    case SDG_FIELD_BOUNDARY_FACE_NORMAL:
      for (size_t f = 0; f < m.bfaces.size(); f++)
        std::memcpy(od + f * nfq * D, m.bfaces[f].normal.data(), sizeof(double) * nfq * D);
      break;
    case SDG_FIELD_BOUNDARY_FACE_JACOBIAN_W:
      for (size_t f = 0; f < m.bfaces.size(); f++) std::memcpy(od + f * nfq, m.bfaces[f].jw.data(), sizeof(double) * nfq);
      break;
    case SDG_FIELD_BOUNDARY_FACE_PARENT:
      for (size_t f = 0; f < m.bfaces.size(); f++) {
        oi[2 * f] = m.bfaces[f].parent[0];
        oi[2 * f + 1] = m.bfaces[f].local[0];
      }
      break;
    case SDG_FIELD_BOUNDARY_FACE_TYPE:
      for (size_t f = 0; f < m.bfaces.size(); f++) oi[f] = m.bfaces[f].bc;
      break;
    case SDG_FIELD_INNER_RADIUS: copy(m.inner_radius); break;
    case SDG_FIELD_NODE_TAG: std::memcpy(oi, m.node_tag.data(), m.node_tag.size() * sizeof(int32_t)); break;
    case SDG_FIELD_NODAL_VALUE: copy(r.nodal_value); break;
    case SDG_FIELD_NODAL_FACE_VALUE: copy(r.nodal_face_value); break;
    case SDG_FIELD_BOUNDARY_FACE_PHYSICAL:
      for (size_t f = 0; f < m.bfaces.size(); f++) oi[f] = m.bfaces[f].physical;
      break;
    default:
      sdg::setGlobalError("sdg_mesh_export: unknown field");
      return SDG_E_INVALID;
  }
  return SDG_OK;
}

}  // extern "C"
