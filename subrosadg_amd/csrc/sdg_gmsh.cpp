// sdg_gmsh.cpp -- reader of gmsh's MSH 4.1 ASCII format for meshes of one element type:
// the mesh-file half of the reference's Mesh<SC>::initializeMesh / readMeshElement
// (src/Mesh/ReadControl.cpp:286-335), which reads through the gmsh library; here the file is parsed
// directly and handed to the connectivity builder of the kit.
//
// Read: $MeshFormat (4.1, ASCII), $Entities (the physical tags of the boundary curves / surfaces),
// $Nodes, $Elements. The highest-dimensional elements are the mesh (triangles or quadrangles of
// order 1..5 in 2-D, tetrahedra or hexahedra of order 1 or 2, or first-order pyramids in 3-D; one
// type and order); their high-order nodes define the element map, so elements may be curved.
// Elements of one dimension
// lower carry the boundary faces (identified by their vertices), whose type comes from the physical
// group of their entity through the caller's map; $Periodic node correspondences (translations)
// pair the faces of periodic boundaries. Refused: binary files, other versions, prisms, high-order
// pyramids, serendipity elements, mixed element types.
#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "kit.hpp"
#include "sdg_internal.hpp"

namespace sdg {

namespace {

// The gmsh element types read here: element shape (enum sdg_element; 0 for a point), dimension,
// geometry order, nodes and vertices (the first nodes). The reference meshes with setOrder(P),
// so its files hold elements of order P (SimulationControl.cpp:26-31).
struct GmshType {
  int code, shape, dim, order, nodes, vertices;
};
constexpr GmshType kGmshTypes[] = {
    {15, 0, 0, 1, 1, 1},                                                       // point
    {1, 1, 1, 1, 2, 2},   {8, 1, 1, 2, 3, 2},   {26, 1, 1, 3, 4, 2},   {27, 1, 1, 4, 5, 2},   {28, 1, 1, 5, 6, 2},
    {2, 2, 2, 1, 3, 3},   {9, 2, 2, 2, 6, 3},   {21, 2, 2, 3, 10, 3},  {23, 2, 2, 4, 15, 3},  {25, 2, 2, 5, 21, 3},
    {3, 3, 2, 1, 4, 4},   {10, 3, 2, 2, 9, 4},  {36, 3, 2, 3, 16, 4},  {37, 3, 2, 4, 25, 4},  {38, 3, 2, 5, 36, 4},
    {4, 4, 3, 1, 4, 4},   {11, 4, 3, 2, 10, 4},
    {7, 5, 3, 1, 5, 5},
    {5, 6, 3, 1, 8, 8},   {12, 6, 3, 2, 27, 8},
};
const GmshType* gmshType(int code) {
  for (const auto& t : kGmshTypes)
    if (t.code == code) return &t;
  return nullptr;
}

bool skipTo(std::istream& in, const std::string& end) {
  std::string line;
  while (in >> line)
    if (line == end) return true;
  return false;
}

}  // namespace

bool readGmshMesh(const char* path, int order, const std::map<int, int>& bc_of_physical, Mesh& mesh,
                  std::string& error) {
  std::ifstream in(path);
  if (!in) {
    error = std::string("cannot open ") + path;
    return false;
  }
  std::map<std::pair<int, int>, int> physical;  // (dim, entity) -> first physical tag
  std::map<int64_t, int64_t> node_index;       // gmsh node tag -> 0-based index
  std::vector<double> xyz;
  struct Block {
    int dim, entity, type;
    std::vector<int64_t> nodes;  // gmsh node tags, gmshType(type)->nodes per element
  };
  std::vector<Block> blocks;
  std::vector<std::vector<std::pair<int64_t, int64_t>>> periodic_pairs;  // per link: slave, master tags
  bool have_format = false, have_nodes = false, have_elements = false;
  std::string tok;
  while (in >> tok) {
    if (tok == "$MeshFormat") {
      std::string version;
      int file_type = 0, data_size = 0;
      in >> version >> file_type >> data_size;
      if (version.rfind("4.1", 0) != 0 || file_type != 0) {
        error = "only the MSH 4.1 ASCII format is read";
        return false;
      }
      have_format = true;
      skipTo(in, "$EndMeshFormat");
    } else if (tok == "$Entities") {
      int64_t count[4];
      in >> count[0] >> count[1] >> count[2] >> count[3];
      for (int dim = 0; dim < 4; dim++)
        for (int64_t i = 0; i < count[dim]; i++) {
          int tag = 0;
          double box[6];
          in >> tag;
          const int nbox = dim == 0 ? 3 : 6;
          for (int k = 0; k < nbox; k++) in >> box[k];
          int64_t nphys = 0;
          in >> nphys;
          for (int64_t k = 0; k < nphys; k++) {
            int ph = 0;
            in >> ph;
            if (k == 0) physical[{dim, tag}] = ph;
          }
          if (dim > 0) {
            int64_t nbound = 0;
            in >> nbound;
            for (int64_t k = 0; k < nbound; k++) {
              int b = 0;
              in >> b;
            }
          }
        }
      if (!in) {
        error = "malformed $Entities section";
        return false;
      }
      skipTo(in, "$EndEntities");
    } else if (tok == "$Nodes") {
      int64_t nblocks = 0, nnodes = 0, mintag = 0, maxtag = 0;
      in >> nblocks >> nnodes >> mintag >> maxtag;
      xyz.reserve(static_cast<size_t>(nnodes) * 3);
      for (int64_t b = 0; b < nblocks; b++) {
        int dim = 0, entity = 0, parametric = 0;
        int64_t n = 0;
        in >> dim >> entity >> parametric >> n;
        if (parametric) {
          error = "parametric node coordinates are not read";
          return false;
        }
        std::vector<int64_t> tags(static_cast<size_t>(n));
        for (auto& t : tags) in >> t;
        for (int64_t i = 0; i < n; i++) {
          double x, y, z;
          in >> x >> y >> z;
          node_index[tags[static_cast<size_t>(i)]] = static_cast<int64_t>(xyz.size() / 3);
          xyz.insert(xyz.end(), {x, y, z});
        }
      }
      if (!in || static_cast<int64_t>(xyz.size() / 3) != nnodes) {
        error = "malformed $Nodes section";
        return false;
      }
      have_nodes = true;
      skipTo(in, "$EndNodes");
    } else if (tok == "$Elements") {
      int64_t nblocks = 0, nelem = 0, mintag = 0, maxtag = 0;
      in >> nblocks >> nelem >> mintag >> maxtag;
      for (int64_t b = 0; b < nblocks; b++) {
        Block blk;
        int64_t n = 0;
        in >> blk.dim >> blk.entity >> blk.type >> n;
        const GmshType* gt = gmshType(blk.type);
        if (!gt) {
          error = "element type " + std::to_string(blk.type) + " is not read (lines, triangles and quadrangles of "
                  "order 1..5, tetrahedra and hexahedra of order 1 or 2, first-order pyramids; no prisms or serendipity "
                  "elements)";
          return false;
        }
        const int nn = gt->nodes;
        blk.nodes.resize(static_cast<size_t>(n) * nn);
        for (int64_t i = 0; i < n; i++) {
          int64_t etag = 0;
          in >> etag;
          for (int k = 0; k < nn; k++) in >> blk.nodes[static_cast<size_t>(i) * nn + k];
        }
        blocks.push_back(std::move(blk));
      }
      if (!in) {
        error = "malformed $Elements section";
        return false;
      }
      have_elements = true;
      skipTo(in, "$EndElements");
    } else if (tok == "$Periodic") {  // node correspondences of translational periodicity
      int64_t links = 0;
      in >> links;
      for (int64_t l = 0; l < links; l++) {
        int dim_l = 0, tag_l = 0, master_l = 0;
        int64_t naffine = 0, npairs = 0;
        in >> dim_l >> tag_l >> master_l >> naffine;
        for (int64_t k = 0; k < naffine; k++) {
          double a = 0.0;
          in >> a;
        }
        in >> npairs;
        periodic_pairs.emplace_back();
        for (int64_t k = 0; k < npairs; k++) {
          int64_t slave = 0, mast = 0;
          in >> slave >> mast;
          periodic_pairs.back().push_back({slave, mast});
        }
      }
      if (!in) {
        error = "malformed $Periodic section";
        return false;
      }
      skipTo(in, "$EndPeriodic");
    } else if (tok.size() > 1 && tok[0] == '$' && tok.rfind("$End", 0) != 0) {
      skipTo(in, "$End" + tok.substr(1));  // sections not needed here ($PhysicalNames, ...)
    }
  }
  if (!have_format || !have_nodes || !have_elements) {
    error = "missing $MeshFormat, $Nodes or $Elements";
    return false;
  }
  int dim = 0;
  for (const auto& b : blocks) dim = std::max(dim, gmshType(b.type)->dim);
  if (dim < 2) {
    error = "no 2-D or 3-D elements";
    return false;
  }
  int type = -1, code = -1;
  std::vector<int64_t> conn, bconn;
  std::vector<int64_t> tri_faces;  // boundary faces that are triangles (padded to 4 nodes for pyramids)
  std::vector<int> btype, bphysical;
  std::vector<std::pair<int, int>> bnodes;  // per boundary block: nodes per element, physical group
  auto index = [&](int64_t tag, int64_t& out) {
    const auto it = node_index.find(tag);
    if (it == node_index.end()) return false;
    out = it->second;
    return true;
  };
  for (const auto& b : blocks) {
    const GmshType* gt = gmshType(b.type);
    const int d = gt->dim, nn = gt->nodes;
    if (d == dim) {
      if (code >= 0 && code != b.type) {
        error = "meshes of more than one element type or order are not read";
        return false;
      }
      type = gt->shape;
      code = b.type;
      for (int64_t tag : b.nodes) {
        int64_t i = 0;
        if (!index(tag, i)) {
          error = "element references an unknown node";
          return false;
        }
        conn.push_back(i);
      }
    } else if (d == dim - 1) {
      const auto ph = physical.find({d, b.entity});
      if (ph == physical.end()) continue;  // boundary elements outside any physical group
      const auto bc = bc_of_physical.find(ph->second);
      if (bc == bc_of_physical.end()) continue;
      bnodes.push_back({gt->shape, ph->second});
      for (size_t k = 0; k < b.nodes.size(); k += static_cast<size_t>(nn)) {
        if (gt->shape == 2) tri_faces.push_back(static_cast<int64_t>(btype.size()));
        for (int v = 0; v < gt->vertices; v++) {  // a face is identified by its vertices
          int64_t i = 0;
          if (!index(b.nodes[k + static_cast<size_t>(v)], i)) {
            error = "boundary element references an unknown node";
            return false;
          }
          bconn.push_back(i);
        }
        btype.push_back(bc->second);
        bphysical.push_back(ph->second);
      }
    }
  }
  // boundary elements must have the shape of the volume elements' faces: lines (triangles,
  // quadrangles), triangles (tetrahedra) or quadrangles (hexahedra); another shape would be read
  // as the wrong number of face vertices
  if (type >= 0) {
    const int face_shape = type == 6 ? 3 : (type == 4 ? 2 : 1);
    static const char* kShape[] = {"points", "lines", "triangles", "quadrangles"};
    for (const auto& bn : bnodes)
      if (bn.first != face_shape && !(type == 5 && (bn.first == 2 || bn.first == 3))) {
        error = "physical group " + std::to_string(bn.second) + " holds boundary " + kShape[bn.first] +
                "; faces of the volume elements are " + kShape[face_shape];
        return false;
      }
  }
  // one node map per periodic link (a corner node is a slave in more than one link)
  std::vector<std::map<int64_t, int64_t>> links;
  for (const auto& link : periodic_pairs) {
    links.emplace_back();
    for (const auto& pr : link) {
      int64_t si = 0, mi = 0;
      if (!index(pr.first, si) || !index(pr.second, mi)) {
        error = "$Periodic references an unknown node";
        return false;
      }
      if (si != mi) links.back()[si] = mi;
    }
  }
  if (type == 5 && !tri_faces.empty()) {  // pyramids: boundary rows of 4 nodes, triangles padded with -1
    std::vector<int64_t> padded;
    size_t o = 0, t = 0;
    for (int64_t f = 0; f < static_cast<int64_t>(btype.size()); f++) {
      const bool tri = t < tri_faces.size() && tri_faces[t] == f;
      const int n = tri ? 3 : 4;
      for (int v = 0; v < n; v++) padded.push_back(bconn[o + v]);
      if (tri) {
        padded.push_back(-1);
        t++;
      }
      o += n;
    }
    bconn.swap(padded);
  }
  const GmshType* gt = gmshType(code);
  return buildMeshFromConnectivity(type, order, static_cast<int64_t>(xyz.size() / 3), xyz.data(),
                                   static_cast<int64_t>(conn.size()) / gt->nodes, conn.data(),
                                   static_cast<int64_t>(btype.size()), bconn.data(), btype.data(), bphysical.data(),
                                   mesh, error, links.empty() ? nullptr : &links, gt->order);
}

}  // namespace sdg
