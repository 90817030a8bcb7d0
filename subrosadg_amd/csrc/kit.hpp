// kit.hpp -- host-side discretisation kit: quadrature rules, the hierarchical Lobatto modal
// basis, reference-element tables and structured periodic meshes with affine geometry.
//
// This replaces the gmsh-owned parts of the reference's setup
// (src/Mesh/Quadrature.cpp:28-54, src/Mesh/BasisFunction.cpp:136-230, src/Mesh/Geometry.cpp:29-168,
//  src/Mesh/Adjacency.cpp:37-495) with project-written code. Conventions kept from the reference:
//   * reference elements: quad [-1,1]^2, hex [-1,1]^3 with gmsh vertex numbering
//     (SimulationControl.cpp:218-241);
//   * local face vertex lists (SimulationControl.cpp:177-216);
//   * face points in parent coordinates = P1 map of the face rule through the face vertices
//     (BasisFunction.cpp:76-111);
//   * volume rule of order 2P, face rule of order 2P+1 (SimulationControl.cpp:275-283);
//   * normals outward from the left (first-seen) parent (Geometry.cpp:102-168);
//   * right-parent point index = seq[j] (SimulationControl.cpp:381-523), computed here by
//     matching physical face-point coordinates.
#ifndef SDG_KIT_HPP_
#define SDG_KIT_HPP_

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace sdg {

// Gauss-Legendre rule on [-1,1] with n points, ascending abscissae.
void gaussLegendre(int n, std::vector<double>& x, std::vector<double>& w);
// Legendre polynomial L_n and its derivative at x.
void legendre(int n, double x, double& value, double& derivative);
// Solin's 1-D Lobatto shape function l_k and its derivative on [-1,1]:
// l0 = (1-x)/2, l1 = (1+x)/2, l_k = (L_k - L_{k-2}) / sqrt(2(2k-1)) for k >= 2.
void lobatto(int k, double x, double& value, double& derivative);

// Reference-element tables for one (element type, order).
struct RefElement {
  int type = 0;       // sdg_element
  int dim = 0;        // D
  int order = 0;      // P
  int nb = 0;         // N_b
  int nq = 0;         // N_q
  int nf = 0;         // N_f
  int nfq = 0;        // N_fq
  int nvert = 0;      // vertices of the element
  int nfvert = 0;     // vertices per face
  std::vector<std::array<int, 3>> modes;   // per basis function: 1-D Lobatto indices (i,j,k)
  std::vector<double> qpoint;              // N_q x D reference volume points
  std::vector<double> qweight;             // N_q
  std::vector<double> fpoint;              // N_fq x (D-1) reference face points
  std::vector<double> fweight;             // N_fq
  std::vector<double> vertex;              // nvert x D reference vertices
  std::vector<int> face_vertex;            // N_f x nfvert local face vertex lists
  std::vector<double> V;                   // N_q x N_b        basis at volume points
  std::vector<double> dV;                  // (N_q*D) x N_b    reference gradient, row q*D+d
  std::vector<double> Vf;                  // (N_f*N_fq) x N_b basis at face points
  std::vector<double> face_param;          // N_f x N_fq x D   face points in parent coordinates
  std::vector<double> lsq_inverse;         // N_b x N_b        (V^T V)^-1
  std::vector<double> mass;                // N_b x N_b        sum_q V_qi V_qj w_q (reference mass)
  std::vector<double> mass_inverse;        // N_b x N_b
  // Cholesky factor of the reference mass, mass = L L^T (lower triangular), and L^-1. The test
  // functions psi = L^-1 phi are orthonormal on the reference element; the mixed-precision mode
  // forms residuals against them so that FP32 rounding is not amplified by the conditioning of
  // the mass matrix of the hierarchic basis.
  std::vector<double> mass_chol;           // N_b x N_b  L
  std::vector<double> mass_chol_inverse;   // N_b x N_b  L^-1
  // P1 Lagrange functions of the vertices (the reference's order-1 nodal basis on the basic
  // nodes, BasisFunction.cpp:161-207): at the volume points and at the face points
  std::vector<double> nodal_value;         // N_q x nvert
  std::vector<double> nodal_face_value;    // (N_f*N_fq) x nvert
  // per face: its vertex count (pyramids: 3 or 4, else nfvert), the face-rule point and weight of
  // each of its N_fq slots, and how many slots are real points. A pyramid's triangle faces carry
  // the triangle rule in their first slots and zero-weight copies of their first point in the
  // rest, so that every face has the N_fq = (P+1)^2 slots of the quadrangle base.
  std::vector<int> face_nv;                // N_f
  std::vector<int> face_nreal;             // N_f
  std::vector<double> fslot_point;         // N_f x N_fq x (D-1)
  std::vector<double> fslot_weight;        // N_f x N_fq
};

// N_b of the modal basis of an element type and order (the hierarchic basis of order P-1 is a
// prefix of order P's, so this is also where the modes above order P-1 start). Order 0: 1.
inline int basisSize(int type, int dim, int order) {
  const int n = order + 1;
  if (type == 2 /* SDG_ELEM_TRIANGLE */) return n * (n + 1) / 2;
  if (type == 4 /* SDG_ELEM_TETRAHEDRON */) return n * (n + 1) * (n + 2) / 6;
  if (type == 5 /* SDG_ELEM_PYRAMID */) return n * (n + 1) * (2 * n + 1) / 6;
  return dim == 1 ? n : (dim == 2 ? n * n : n * n * n);
}

// Build the reference tables; returns false for an unsupported (type, order).
bool buildRefElement(int type, int order, RefElement& ref, std::string& error);

// Dense inverse by Gauss-Jordan elimination with partial pivoting (row-major n x n).
bool invertMatrix(int n, const double* a, double* inverse);

// Element map of a gmsh element of geometry order g >= 1: the Lagrange functions through its
// geometry nodes, in gmsh's node order (vertices, then the interior nodes of the edges, faces and
// volume; ReadControl.cpp:286-335 reads them through gmsh, Geometry.cpp:44-67 evaluates the map).
struct GeomMap {
  int type = 0, dim = 0, order = 1, n = 0;  // n geometry nodes
  std::vector<double> node;                 // n x dim reference coordinates
  std::vector<std::array<int, 3>> expo;     // n monomial exponents spanning the map's space
  std::vector<double> ainv;                 // n x n: N_k(xi) = sum_m ainv[m n + k] xi^expo[m]
};
// false with a message for an unsupported (type, geometry order)
bool buildGeomMap(int type, int order, GeomMap& g, std::string& error);
// shape values N (n) and reference gradients dN (n x dim, dN[k dim + d] = dN_k / dxi_d) at xi
void geomShape(const GeomMap& g, const double* xi, double* N, double* dN);

struct Face {
  int32_t parent[2] = {0, 0};   // left, right element (boundary face: right = -1)
  int32_t local[2] = {0, 0};    // local face index in left, right
  int32_t rotation = 0;
  int32_t bc = -1;              // boundary faces: sdg_boundary type
  int32_t physical = -1;        // boundary faces read from a mesh file: the gmsh physical tag
  std::vector<int32_t> seq;     // N_fq: right point index of left point j
  std::vector<double> normal;   // N_fq x D, outward from the left parent
  std::vector<double> jw;       // N_fq: |J_f| w
  std::vector<double> coord;    // N_fq x D physical points (left parent's view)
};

struct Mesh {
  RefElement ref;
  int64_t elements = 0;
  std::vector<double> vertex;       // E x nvert x D physical vertex coordinates (unwrapped)
  std::vector<double> jac;          // E x D x D, jac[l*D+k] = d x_l / d xi_k (constant: affine)
  std::vector<double> detj;         // E
  std::vector<double> min_edge;     // E
  std::vector<Face> faces;          // interior faces
  std::vector<Face> bfaces;         // boundary faces (the element is the left parent)
  // element -> face lookup: per element and local face, face index and side (0 left, 1 right);
  // a boundary face b is stored as -2 - b
  std::vector<int32_t> elem_face;   // E x N_f
  std::vector<int8_t> elem_side;    // E x N_f
  bool affine = true;
  // curved (non-affine) meshes: the element map through each element's geometry nodes and their
  // physical coordinates (E x geom.n x D); jac / detj then hold the map at the element centre.
  // Empty for affine meshes.
  GeomMap geom;
  std::vector<double> gnode;
  // mesh nodes (the element vertices): per element the 0-based node index of each vertex, shared
  // by the elements that meet at a point. Periodic images are distinct nodes, as gmsh keeps the
  // nodes of a periodic side apart from their masters.
  std::vector<int32_t> node_tag;    // E x nvert
  int64_t nodes = 0;
  // inner radius per element: D |element| / |boundary| (the inscribed radius of a triangle, of a
  // square or cube, and of any polygon with an inscribed circle)
  std::vector<double> inner_radius; // E
};

// Structured mesh on an axis-aligned box; periodic[d] != 0 makes direction d periodic, otherwise
// its two sides are boundaries of type bc[2 d] (low) and bc[2 d + 1] (high). Returns false with
// a message on a bad descriptor.
bool buildStructuredMesh(int type, int order, const int n[3], const double lo[3], const double hi[3],
                         const int periodic[3], const int bc[6], Mesh& mesh, std::string& error);

// Mesh of one element type from node coordinates (nnode x 3) and element vertex lists (E x nvert,
// 0-based node indices, gmsh vertex order), with the boundary faces given by their node lists
// (nbface x nfvert, any order) and sdg_boundary types. Elements must be affine (triangles,
// parallelogram quadrangles, parallelepiped hexahedra, tetrahedra) or curved (below); an affine element
// with a negative Jacobian is mirrored. Faces pair by their node sets and are created at the first visit in element order,
// whose element is the left parent (Adjacency.cpp:61-183); right-parent points are matched by
// coordinates, boundary faces take the type of their boundary element. periodic_links (optional)
// are node maps slave -> master, one per periodic pair of boundaries: an unpaired face whose nodes
// are all slaves of one link pairs with the unpaired face on the master nodes, and the right
// parent's points are matched after the translation between the two faces (translational
// periodicity, Adjacency.cpp:185-328).
// geometry_order > 1: conn holds the element's gmsh geometry nodes of that order (vertices first);
// an element whose nodes do not all lie on the affine map of its vertices makes the mesh curved,
// as does a non-parallelogram quadrangle or non-parallelepiped hexahedron (then the bilinear /
// trilinear map of the vertices is the element map).
bool buildMeshFromConnectivity(int type, int order, int64_t nnode, const double* xyz, int64_t elements,
                               const int64_t* conn, int64_t nbface, const int64_t* bconn, const int* btype,
                               const int* bphysical, Mesh& mesh, std::string& error,
                               const std::vector<std::map<int64_t, int64_t>>* periodic_links = nullptr,
                               int geometry_order = 1);

// Triangles and quadrangles on a periodic structured n[0] x n[1] grid: cell (i, j) (i fastest) is a
// quadrangle, or, when split[cell] != 0, the two triangles (A, B, C), (A, C, D) of its corners
// A = (i, j), B = (i+1, j), C = (i+1, j+1), D = (i, j+1). Faces are matched geometrically
// (modulo the period) and created at their first visit, whose element is the left parent.
struct MixedMesh {
  int order = 0;
  RefElement ref[2];               // [0] triangle, [1] quadrangle
  int64_t elements = 0;
  std::vector<int8_t> etype;       // per element: 0 triangle, 1 quadrangle
  std::vector<double> vertex;      // E x 4 x 2 (a triangle uses the first 3)
  std::vector<double> jac;         // E x 2 x 2, jac[l*2+k] = d x_l / d xi_k
  std::vector<double> detj;        // E
  std::vector<double> min_edge;    // E
  std::vector<Face> faces;         // interior faces (all faces: the grid is periodic)
  std::vector<int32_t> elem_face;  // E x 4 (-1 past the element's face count)
  std::vector<int8_t> elem_side;   // E x 4
  const RefElement& refOf(int64_t e) const { return ref[etype[e]]; }
};

bool buildSplitMesh(int order, const int n[2], const double lo[2], const double hi[2], const std::vector<uint8_t>& split,
                    MixedMesh& mesh, std::string& error);
// The mesh as a single-type Mesh; false when it holds both element types.
bool singleTypeView(const MixedMesh& mm, Mesh& out);
// Node tags (vertices that coincide in physical space share a node) and inner radii from the
// element vertices and Jacobians of an affine mesh.
void assignNodes(Mesh& mesh);
// Modal basis values (N_b) at a reference point.
void basisValues(const RefElement& ref, const double* xi, double* value);
// Order-1 nodal functions of the vertices (nvert) at a reference point.
void nodalValues(const RefElement& ref, const double* xi, double* value);
// Physical point of reference point xi in element e.
void mapReferencePoint(const Mesh& mesh, int64_t e, const double* xi, double* x);

// Geometry derived per element and volume point (affine elements: constant Jacobian; curved
// elements: the Jacobian of the element map at each point).
//   detjw[q]           = |J| w_q
//   tiw[q*D*D + a+D*b] = ((J^T)^-1)(a,b) |J| w_q   (column-major D x D)
//   coord[q*D + d]     = physical point
void elementGeometry(const Mesh& mesh, int64_t e, double* detjw, double* tiw, double* coord);
// Inverse of the element mass matrix sum_q V_qi V_qj |J| w_q (row-major N_b x N_b).
void elementMassInverse(const Mesh& mesh, int64_t e, double* minv);
// The same two for an affine element given by its reference tables, Jacobian (jac[l*D+k] =
// d x_l / d xi_k) and determinant (elements of split triangle / quadrangle grids).
void affineGeometry(const RefElement& ref, const double* jac, double detj, double* detjw, double* tiw);
void affineMassInverse(const RefElement& ref, double detj, double* minv);

}  // namespace sdg

#endif  // SDG_KIT_HPP_
