// kit.cpp -- quadrature, Lobatto modal basis, reference tables and structured periodic meshes.
// See kit.hpp for the conventions kept from the reference.
#include "kit.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <utility>

namespace sdg {

namespace {
constexpr double kPi = 3.14159265358979323846;

// gmsh reference vertices (SimulationControl.cpp:218-241)
const double kQuadVertex[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
const double kHexVertex[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
// local face vertex lists (SimulationControl.cpp:177-216)
const int kQuadFace[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
const int kHexFace[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
// P1 face element vertices: line (-1),(1); quad (-1,-1),(1,-1),(1,1),(-1,1)
const double kLineFaceVertex[2] = {-1, 1};
const double kQuadFaceVertex[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};

// Tensor-product P1 shape function of a vertex with coordinates v (entries +-1) at xi, and its
// gradient.
void p1Shape(int dim, const double* v, const double* xi, double& n, double* grad) {
  double f[3];
  for (int d = 0; d < dim; d++) f[d] = 0.5 * (1.0 + v[d] * xi[d]);
  n = 1.0;
  for (int d = 0; d < dim; d++) n *= f[d];
  for (int d = 0; d < dim; d++) {
    double g = 0.5 * v[d];
    for (int e = 0; e < dim; e++)
      if (e != d) g *= f[e];
    grad[d] = g;
  }
}

// Symmetric quadrature rules on the reference triangle (0,0),(1,0),(0,1), weights summing to its
// measure 1/2, exact for polynomials of degree 2, 4, 6, 8 with 3, 6, 12, 16 points (the point
// counts of SimulationControl.cpp:275 for P = 1..4). Orbits in barycentric coordinates: S3 the
// centroid, S21(a) the 3 permutations of (a, a, 1-2a), S111(a, b) the 6 permutations of
// (a, b, 1-a-b). Parameters solved from the moment equations (Dunavant's orbit structure);
// exactness is checked in tests/test_kit.py.
struct TriOrbit {
  int kind;  // 0: S3, 1: S21, 2: S111
  double a, b, w;
};
const TriOrbit kTriRule2[] = {{1, 0.16666666666666666, 0.0, 0.16666666666666666}};
const TriOrbit kTriRule4[] = {{1, 0.44594849091596483, 0.0, 0.11169079483900568},
                              {1, 0.09157621350977076, 0.0, 0.05497587182766098}};
const TriOrbit kTriRule6[] = {{1, 0.24928674517091226, 0.0, 0.05839313786318872},
                              {1, 0.06308901449150205, 0.0, 0.025422453185103156},
                              {2, 0.05314504984481764, 0.3103524510337826, 0.041425537809187396}};
const TriOrbit kTriRule8[] = {{0, 0.0, 0.0, 0.0721578038388935},
                              {1, 0.459292588292723, 0.0, 0.0475458171336425},
                              {1, 0.17056930775176, 0.0, 0.051608685267359},
                              {1, 0.050547228317031, 0.0, 0.016229248811599},
                              {2, 0.008394777409958, 0.263112829634638, 0.0136151570872175}};

void triangleRule(int order, std::vector<double>& pts, std::vector<double>& w) {
  const TriOrbit* r = order == 1 ? kTriRule2 : order == 2 ? kTriRule4 : order == 3 ? kTriRule6 : kTriRule8;
  const int n = order == 1 ? 1 : order == 2 ? 2 : order == 3 ? 3 : 5;
  pts.clear();
  w.clear();
  auto add = [&](double l1, double l2, double wt) {  // (xi, eta) = (lambda_1, lambda_2)
    pts.push_back(l1);
    pts.push_back(l2);
    w.push_back(wt);
  };
  for (int i = 0; i < n; i++) {
    const TriOrbit& o = r[i];
    if (o.kind == 0) {
      add(1.0 / 3.0, 1.0 / 3.0, o.w);
    } else if (o.kind == 1) {
      const double c = 1.0 - 2.0 * o.a;
      add(o.a, c, o.w);
      add(c, o.a, o.w);
      add(o.a, o.a, o.w);
    } else {
      const double c = 1.0 - o.a - o.b;
      const double t[6][2] = {{o.b, c}, {c, o.b}, {o.a, c}, {c, o.a}, {o.a, o.b}, {o.b, o.a}};
      for (int k = 0; k < 6; k++) add(t[k][0], t[k][1], o.w);
    }
  }
}

// Legendre polynomials L_0..L_n and derivatives at x (any arithmetic type)
template <typename T>
void legendreSeries(int n, T x, T* L, T* dL) {
  L[0] = T(1.0);
  dL[0] = T(0.0);
  if (n == 0) return;
  L[1] = x;
  dL[1] = T(1.0);
  for (int k = 2; k <= n; k++) {
    L[k] = (T(2.0 * k - 1.0) * x * L[k - 1] - T(k - 1.0) * L[k - 2]) / T(double(k));
    dL[k] = dL[k - 2] + T(2.0 * k - 1.0) * L[k - 1];
  }
}

// Hierarchic modal basis of order P on the reference triangle (Solin's H1 construction spanning
// P_P, the space of BasisFunction.cpp's triangle basis): the barycentric vertex functions, then per
// degree k = 2..P the three edge functions lambda_a lambda_b kappa_{k-2}(lambda_b - lambda_a) with
// Solin's kernel kappa_{k-2}(x) = 4 l_k(x) / (1 - x^2), and the bubbles
// lambda_0 lambda_1 lambda_2 L_i(lambda_1 - lambda_0) L_j(lambda_2 - lambda_1), i + j = k - 3.
template <typename T>
void triangleBasis(int order, T xi, T eta, T* value) {
  const T lam[3] = {T(1.0) - xi - eta, xi, eta};
  static const int kEdge[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  T L[8], dL[8], L2[8], dL2[8];
  int b = 0;
  for (int v = 0; v < 3; v++) value[b++] = lam[v];
  for (int k = 2; k <= order; k++) {
    for (int e = 0; e < 3; e++) {
      const T l0 = lam[kEdge[e][0]], l1 = lam[kEdge[e][1]];
      legendreSeries(k - 1, l1 - l0, L, dL);
      // l_k = (L_k - L_{k-2}) / sqrt(2(2k-1)) = (2k-1)(x^2-1) L'_{k-1} / (k(k-1) sqrt(2(2k-1)))
      const T kappa = T(-4.0 * (2.0 * k - 1.0) / (k * (k - 1.0) * std::sqrt(2.0 * (2.0 * k - 1.0)))) * dL[k - 1];
      value[b++] = l0 * l1 * kappa;
    }
    if (k >= 3) {
      legendreSeries(k - 3, lam[1] - lam[0], L, dL);
      legendreSeries(k - 3, lam[2] - lam[1], L2, dL2);
      for (int i = 0; i <= k - 3; i++) value[b++] = lam[0] * lam[1] * lam[2] * L[i] * L2[k - 3 - i];
    }
  }
}

double det(int dim, const double* a) {  // row-major
  if (dim == 1) return a[0];
  if (dim == 2) return a[0] * a[3] - a[1] * a[2];
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
}  // namespace

void legendre(int n, double x, double& value, double& derivative) {
  if (n == 0) {
    value = 1.0;
    derivative = 0.0;
    return;
  }
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; k++) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  value = p1;
  // derivative from the recurrence (1-x^2) L_n' = n (L_{n-1} - x L_n), and the closed form at +-1
  if (std::fabs(1.0 - x * x) < 1e-300) {
    const double s = (x > 0 || (n % 2 == 1)) ? 1.0 : -1.0;
    derivative = s * 0.5 * n * (n + 1.0);
  } else {
    derivative = n * (p0 - x * p1) / (1.0 - x * x);
  }
}

void gaussLegendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    // i-th largest root, Newton from the Chebyshev-like guess
    double r = std::cos(kPi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double v, d;
      legendre(n, r, v, d);
      const double dr = v / d;
      r -= dr;
      if (std::fabs(dr) < 1e-17) break;
    }
    double v, d;
    legendre(n, r, v, d);
    x[n - 1 - i] = r;
    w[n - 1 - i] = 2.0 / ((1.0 - r * r) * d * d);
  }
}

void lobatto(int k, double x, double& value, double& derivative) {
  if (k == 0) {
    value = 0.5 * (1.0 - x);
    derivative = -0.5;
    return;
  }
  if (k == 1) {
    value = 0.5 * (1.0 + x);
    derivative = 0.5;
    return;
  }
  double lk, dk, lk2, dk2, lk1, dk1;
  legendre(k, x, lk, dk);
  legendre(k - 2, x, lk2, dk2);
  legendre(k - 1, x, lk1, dk1);
  const double s = std::sqrt(2.0 * (2.0 * k - 1.0));
  value = (lk - lk2) / s;
  derivative = (2.0 * k - 1.0) * lk1 / s;  // L_k' - L_{k-2}' = (2k-1) L_{k-1}
}

bool invertMatrix(int n, const double* a, double* inverse) {
  std::vector<double> m(static_cast<size_t>(n) * 2 * n, 0.0);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) m[i * 2 * n + j] = a[i * n + j];
    m[i * 2 * n + n + i] = 1.0;
  }
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(m[r * 2 * n + c]) > std::fabs(m[piv * 2 * n + c])) piv = r;
    if (std::fabs(m[piv * 2 * n + c]) < 1e-300) return false;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) std::swap(m[c * 2 * n + j], m[piv * 2 * n + j]);
    const double inv = 1.0 / m[c * 2 * n + c];
    for (int j = 0; j < 2 * n; j++) m[c * 2 * n + j] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = m[r * 2 * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; j++) m[r * 2 * n + j] -= f * m[c * 2 * n + j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inverse[i * n + j] = m[i * 2 * n + n + j];
  return true;
}

static void evalBasis(const RefElement& ref, const double* xi, double* value, double* grad /*N_b x D*/) {
  const int dim = ref.dim;
  if (ref.type == 2) {  // triangle: values, and gradients by complex-step differentiation
    triangleBasis<double>(ref.order, xi[0], xi[1], value);
    if (grad) {
      const double h = 1e-30;
      std::vector<std::complex<double>> cv(ref.nb);
      for (int d = 0; d < 2; d++) {
        const std::complex<double> x(xi[0], d == 0 ? h : 0.0), y(xi[1], d == 1 ? h : 0.0);
        triangleBasis<std::complex<double>>(ref.order, x, y, cv.data());
        for (int b = 0; b < ref.nb; b++) grad[b * 2 + d] = cv[b].imag() / h;
      }
    }
    return;
  }
  for (int b = 0; b < ref.nb; b++) {
    double l[3], dl[3];
    for (int d = 0; d < dim; d++) lobatto(ref.modes[b][d], xi[d], l[d], dl[d]);
    double v = 1.0;
    for (int d = 0; d < dim; d++) v *= l[d];
    value[b] = v;
    if (grad) {
      for (int d = 0; d < dim; d++) {
        double g = dl[d];
        for (int e = 0; e < dim; e++)
          if (e != d) g *= l[e];
        grad[b * dim + d] = g;
      }
    }
  }
}

// order-1 nodal functions of the vertices: barycentric on the triangle, tensor P1 otherwise
void nodalValues(const RefElement& ref, const double* xi, double* out) {
  if (ref.type == 2) {
    out[0] = 1.0 - xi[0] - xi[1];
    out[1] = xi[0];
    out[2] = xi[1];
    return;
  }
  double g[3];
  for (int v = 0; v < ref.nvert; v++) p1Shape(ref.dim, &ref.vertex[static_cast<size_t>(v) * ref.dim], xi, out[v], g);
}

void basisValues(const RefElement& ref, const double* xi, double* value) { evalBasis(ref, xi, value, nullptr); }

bool buildRefElement(int type, int order, RefElement& ref, std::string& error) {
  if (order < 1 || order > 5) {
    error = "polynomial order must be 1..5";
    return false;
  }
  ref = RefElement();
  ref.type = type;
  ref.order = order;
  const int n1 = order + 1;  // points per direction: Gauss order 2P -> P+1, order 2P+1 -> P+1
  std::vector<double> gx, gw;
  gaussLegendre(n1, gx, gw);
  if (type == 1) {  // line: faces are its two end points (normal -1 / +1, Geometry.cpp:102-112)
    ref.dim = 1;
    ref.nf = 2;
    ref.nvert = 2;
    ref.nfvert = 1;
    for (int i = 0; i <= order; i++) ref.modes.push_back({i, 0, 0});
    ref.vertex = {-1.0, 1.0};
    ref.face_vertex = {0, 1};
    for (int a = 0; a < n1; a++) {
      ref.qpoint.push_back(gx[a]);
      ref.qweight.push_back(gw[a]);
    }
    ref.fweight.push_back(1.0);  // one point per face; fpoint stays empty (0-D face coordinates)
  } else if (type == 3) {  // quadrangle
    ref.dim = 2;
    ref.nf = 4;
    ref.nvert = 4;
    ref.nfvert = 2;
    for (int j = 0; j <= order; j++)
      for (int i = 0; i <= order; i++) ref.modes.push_back({i, j, 0});
    for (int v = 0; v < 4; v++)
      for (int d = 0; d < 2; d++) ref.vertex.push_back(kQuadVertex[v][d]);
    for (int f = 0; f < 4; f++)
      for (int v = 0; v < 2; v++) ref.face_vertex.push_back(kQuadFace[f][v]);
    for (int b = 0; b < n1; b++)
      for (int a = 0; a < n1; a++) {
        ref.qpoint.push_back(gx[a]);
        ref.qpoint.push_back(gx[b]);
        ref.qweight.push_back(gw[a] * gw[b]);
      }
    for (int a = 0; a < n1; a++) {
      ref.fpoint.push_back(gx[a]);
      ref.fweight.push_back(gw[a]);
    }
  } else if (type == 6) {  // hexahedron
    ref.dim = 3;
    ref.nf = 6;
    ref.nvert = 8;
    ref.nfvert = 4;
    for (int k = 0; k <= order; k++)
      for (int j = 0; j <= order; j++)
        for (int i = 0; i <= order; i++) ref.modes.push_back({i, j, k});
    for (int v = 0; v < 8; v++)
      for (int d = 0; d < 3; d++) ref.vertex.push_back(kHexVertex[v][d]);
    for (int f = 0; f < 6; f++)
      for (int v = 0; v < 4; v++) ref.face_vertex.push_back(kHexFace[f][v]);
    for (int c = 0; c < n1; c++)
      for (int b = 0; b < n1; b++)
        for (int a = 0; a < n1; a++) {
          ref.qpoint.push_back(gx[a]);
          ref.qpoint.push_back(gx[b]);
          ref.qpoint.push_back(gx[c]);
          ref.qweight.push_back(gw[a] * gw[b] * gw[c]);
        }
    // face points with the second face coordinate fastest: with this order the right-parent
    // permutations equal the rotation tables of SimulationControl.cpp:454-520
    for (int a = 0; a < n1; a++)
      for (int b = 0; b < n1; b++) {
        ref.fpoint.push_back(gx[a]);
        ref.fpoint.push_back(gx[b]);
        ref.fweight.push_back(gw[a] * gw[b]);
      }
  } else if (type == 2) {  // triangle
    if (order > 4) {
      error = "triangles are implemented for P = 1..4";
      return false;
    }
    ref.dim = 2;
    ref.nf = 3;
    ref.nvert = 3;
    ref.nfvert = 2;
    const double tv[3][2] = {{0, 0}, {1, 0}, {0, 1}};  // SimulationControl.cpp:218-241
    const int tf[3][2] = {{0, 1}, {1, 2}, {2, 0}};      // SimulationControl.cpp:182-186
    for (int v = 0; v < 3; v++)
      for (int d = 0; d < 2; d++) ref.vertex.push_back(tv[v][d]);
    for (int f = 0; f < 3; f++)
      for (int v = 0; v < 2; v++) ref.face_vertex.push_back(tf[f][v]);
    triangleRule(order, ref.qpoint, ref.qweight);
    for (int a = 0; a < n1; a++) {
      ref.fpoint.push_back(gx[a]);
      ref.fweight.push_back(gw[a]);
    }
    ref.modes.assign(static_cast<size_t>((order + 1) * (order + 2) / 2), {0, 0, 0});
  } else {
    error = "unsupported element type (line=1, triangle=2, quadrangle=3 and hexahedron=6 are implemented)";
    return false;
  }
  // hierarchical ordering: sort modes by (max index, then slowest-to-fastest index) so that the
  // first N_b(P-1) functions span the order P-1 space
  if (type != 2)
  std::stable_sort(ref.modes.begin(), ref.modes.end(), [](const std::array<int, 3>& a, const std::array<int, 3>& b) {
    const int ma = std::max(a[0], std::max(a[1], a[2]));
    const int mb = std::max(b[0], std::max(b[1], b[2]));
    if (ma != mb) return ma < mb;
    if (a[2] != b[2]) return a[2] < b[2];
    if (a[1] != b[1]) return a[1] < b[1];
    return a[0] < b[0];
  });
  const int dim = ref.dim;
  ref.nb = static_cast<int>(ref.modes.size());
  ref.nq = static_cast<int>(ref.qweight.size());
  ref.nfq = static_cast<int>(ref.fweight.size());
  const int nb = ref.nb, nq = ref.nq, nf = ref.nf, nfq = ref.nfq;
  ref.V.assign(static_cast<size_t>(nq) * nb, 0.0);
  ref.dV.assign(static_cast<size_t>(nq) * dim * nb, 0.0);
  std::vector<double> val(nb), grad(static_cast<size_t>(nb) * dim);
  for (int q = 0; q < nq; q++) {
    evalBasis(ref, &ref.qpoint[static_cast<size_t>(q) * dim], val.data(), grad.data());
    for (int b = 0; b < nb; b++) {
      ref.V[static_cast<size_t>(q) * nb + b] = val[b];
      for (int d = 0; d < dim; d++) ref.dV[(static_cast<size_t>(q) * dim + d) * nb + b] = grad[b * dim + d];
    }
  }
  // face points in parent coordinates: P1 map of the face rule through the face vertices
  ref.face_param.assign(static_cast<size_t>(nf) * nfq * dim, 0.0);
  ref.Vf.assign(static_cast<size_t>(nf) * nfq * nb, 0.0);
  for (int f = 0; f < nf; f++) {
    for (int j = 0; j < nfq; j++) {
      double xi[3] = {0, 0, 0};
      for (int v = 0; v < ref.nfvert; v++) {
        double n = 1.0, g[2];
        if (dim == 1) {
          n = 1.0;  // the face is the vertex itself
        } else if (dim == 2) {
          const double fv[1] = {kLineFaceVertex[v]};
          p1Shape(1, fv, &ref.fpoint[j], n, g);
        } else {
          p1Shape(2, kQuadFaceVertex[v], &ref.fpoint[static_cast<size_t>(j) * 2], n, g);
        }
        const int lv = ref.face_vertex[f * ref.nfvert + v];
        for (int d = 0; d < dim; d++) xi[d] += n * ref.vertex[lv * dim + d];
      }
      for (int d = 0; d < dim; d++) ref.face_param[(static_cast<size_t>(f) * nfq + j) * dim + d] = xi[d];
      evalBasis(ref, xi, val.data(), nullptr);
      for (int b = 0; b < nb; b++) ref.Vf[(static_cast<size_t>(f) * nfq + j) * nb + b] = val[b];
    }
  }
  std::vector<double> vtv(static_cast<size_t>(nb) * nb, 0.0);
  ref.mass.assign(static_cast<size_t>(nb) * nb, 0.0);
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < nb; j++) {
      double s = 0.0, m = 0.0;
      for (int q = 0; q < nq; q++) {
        const double p = ref.V[static_cast<size_t>(q) * nb + i] * ref.V[static_cast<size_t>(q) * nb + j];
        s += p;
        m += p * ref.qweight[q];
      }
      vtv[static_cast<size_t>(i) * nb + j] = s;
      ref.mass[static_cast<size_t>(i) * nb + j] = m;
    }
  ref.lsq_inverse.assign(static_cast<size_t>(nb) * nb, 0.0);
  ref.mass_inverse.assign(static_cast<size_t>(nb) * nb, 0.0);
  if (!invertMatrix(nb, vtv.data(), ref.lsq_inverse.data()) ||
      !invertMatrix(nb, ref.mass.data(), ref.mass_inverse.data())) {
    error = "singular reference matrix";
    return false;
  }
  // mass = L L^T (Cholesky-Banachiewicz), then L^-1 by forward substitution
  ref.mass_chol.assign(static_cast<size_t>(nb) * nb, 0.0);
  ref.mass_chol_inverse.assign(static_cast<size_t>(nb) * nb, 0.0);
  double* Lc = ref.mass_chol.data();
  for (int i = 0; i < nb; i++)
    for (int j = 0; j <= i; j++) {
      double acc = ref.mass[static_cast<size_t>(i) * nb + j];
      for (int k = 0; k < j; k++) acc -= Lc[i * nb + k] * Lc[j * nb + k];
      if (i == j) {
        if (!(acc > 0.0)) {
          error = "reference mass matrix is not positive definite";
          return false;
        }
        Lc[i * nb + i] = std::sqrt(acc);
      } else {
        Lc[i * nb + j] = acc / Lc[j * nb + j];
      }
    }
  double* Li = ref.mass_chol_inverse.data();
  for (int c = 0; c < nb; c++)
    for (int i = c; i < nb; i++) {
      double acc = i == c ? 1.0 : 0.0;
      for (int k = c; k < i; k++) acc -= Lc[i * nb + k] * Li[k * nb + c];
      Li[i * nb + c] = acc / Lc[i * nb + i];
    }
  auto nodal = [&](const double* xi, double* out) { nodalValues(ref, xi, out); };
  ref.nodal_value.assign(static_cast<size_t>(nq) * ref.nvert, 0.0);
  for (int q = 0; q < nq; q++) nodal(&ref.qpoint[static_cast<size_t>(q) * dim], &ref.nodal_value[static_cast<size_t>(q) * ref.nvert]);
  ref.nodal_face_value.assign(static_cast<size_t>(nf) * nfq * ref.nvert, 0.0);
  for (int s = 0; s < nf * nfq; s++)
    nodal(&ref.face_param[static_cast<size_t>(s) * dim], &ref.nodal_face_value[static_cast<size_t>(s) * ref.nvert]);
  return true;
}

namespace {
// physical point of reference point xi in element e, and the Jacobian J(l,k) = dx_l/dxi_k
// physical point of reference point xi for an element of type ref with vertices xv (nvert x D),
// and the Jacobian J(l,k) = dx_l/dxi_k
void mapPointRef(const RefElement& ref, const double* xv_all, const double* xi, double* x, double* jac) {
  const int dim = ref.dim;
  for (int d = 0; d < dim; d++) x[d] = 0.0;
  if (jac)
    for (int i = 0; i < dim * dim; i++) jac[i] = 0.0;
  for (int v = 0; v < ref.nvert; v++) {
    double n, g[3];
    if (ref.type == 2) {  // barycentric P1 shape functions
      const double N[3] = {1.0 - xi[0] - xi[1], xi[0], xi[1]};
      const double G[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
      n = N[v];
      g[0] = G[v][0];
      g[1] = G[v][1];
    } else {
      p1Shape(dim, &ref.vertex[static_cast<size_t>(v) * dim], xi, n, g);
    }
    const double* xv = xv_all + static_cast<size_t>(v) * dim;
    for (int l = 0; l < dim; l++) {
      x[l] += n * xv[l];
      if (jac)
        for (int k = 0; k < dim; k++) jac[l * dim + k] += xv[l] * g[k];
    }
  }
}
void mapPoint(const Mesh& mesh, int64_t e, const double* xi, double* x, double* jac) {
  mapPointRef(mesh.ref, &mesh.vertex[static_cast<size_t>(e) * mesh.ref.nvert * mesh.ref.dim], xi, x, jac);
}
// normals (outward from element e through its local face s) and |J_f| w at the face points, from
// the element's affine face map (Geometry.cpp:69-86,102-168)
void faceNormalsRef(const RefElement& ref, const double* J, int s, Face& face) {
  const int dim = ref.dim, nfq = ref.nfq;
  face.normal.assign(static_cast<size_t>(nfq) * dim, 0.0);
  face.jw.assign(nfq, 0.0);
  if (dim == 1) {  // end points: outward -1 at the first vertex, +1 at the second, weight 1
    face.normal[0] = s == 0 ? -1.0 : 1.0;
    face.jw[0] = ref.fweight[0];
    return;
  }
  for (int p = 0; p < nfq; p++) {
    double ts[3] = {0, 0, 0}, tt[3] = {0, 0, 0};  // d xi / d s, d xi / d t
    for (int v = 0; v < ref.nfvert; v++) {
      double nval, g[2];
      if (dim == 2) {
        const double fv[1] = {kLineFaceVertex[v]};
        p1Shape(1, fv, &ref.fpoint[p], nval, g);
      } else {
        p1Shape(2, kQuadFaceVertex[v], &ref.fpoint[static_cast<size_t>(p) * 2], nval, g);
      }
      const int lv = ref.face_vertex[s * ref.nfvert + v];
      for (int d = 0; d < dim; d++) {
        ts[d] += g[0] * ref.vertex[lv * dim + d];
        if (dim == 3) tt[d] += g[1] * ref.vertex[lv * dim + d];
      }
    }
    double Ts[3] = {0, 0, 0}, Tt[3] = {0, 0, 0};
    for (int l = 0; l < dim; l++)
      for (int kk = 0; kk < dim; kk++) {
        Ts[l] += J[l * dim + kk] * ts[kk];
        Tt[l] += J[l * dim + kk] * tt[kk];
      }
    double nv[3];
    if (dim == 2) {
      nv[0] = Ts[1];
      nv[1] = -Ts[0];
    } else {
      nv[0] = Ts[1] * Tt[2] - Ts[2] * Tt[1];
      nv[1] = Ts[2] * Tt[0] - Ts[0] * Tt[2];
      nv[2] = Ts[0] * Tt[1] - Ts[1] * Tt[0];
    }
    double norm = 0.0;
    for (int d = 0; d < dim; d++) norm += nv[d] * nv[d];
    norm = std::sqrt(norm);
    for (int d = 0; d < dim; d++) face.normal[static_cast<size_t>(p) * dim + d] = nv[d] / norm;
    face.jw[p] = norm * ref.fweight[p];
  }
}
void faceNormals(const Mesh& mesh, int64_t e, int s, Face& face) {
  faceNormalsRef(mesh.ref, &mesh.jac[static_cast<size_t>(e) * mesh.ref.dim * mesh.ref.dim], s, face);
}
}  // namespace

bool buildStructuredMesh(int type, int order, const int n[3], const double lo[3], const double hi[3],
                         const int periodic[3], const int bc[6], Mesh& mesh, std::string& error) {
  mesh = Mesh();
  if (!buildRefElement(type, order, mesh.ref, error)) return false;
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim;
  int nc[3] = {n[0], dim >= 2 ? n[1] : 1, dim == 3 ? n[2] : 1};
  for (int d = 0; d < dim; d++) {
    if (nc[d] < (periodic[d] ? 2 : 1)) {
      error = "a periodic direction needs at least 2 cells, a bounded one at least 1";
      return false;
    }
    if (!periodic[d] && (bc[2 * d] < 0 || bc[2 * d] > 5 || bc[2 * d + 1] < 0 || bc[2 * d + 1] > 5)) {
      error = "a non-periodic side needs a boundary condition type (0..5)";
      return false;
    }
    if (!(hi[d] > lo[d])) {
      error = "box upper corner must exceed the lower corner";
      return false;
    }
  }
  double h[3], period[3];
  for (int d = 0; d < dim; d++) {
    period[d] = hi[d] - lo[d];
    h[d] = period[d] / nc[d];
  }
  const int64_t E = static_cast<int64_t>(nc[0]) * nc[1] * nc[2];
  if (E > (int64_t(1) << 30)) {
    error = "too many elements";
    return false;
  }
  mesh.elements = E;
  mesh.vertex.assign(static_cast<size_t>(E) * ref.nvert * dim, 0.0);
  auto cellIndex = [&](int i, int j, int k) -> int64_t {
    i = (i % nc[0] + nc[0]) % nc[0];
    j = (j % nc[1] + nc[1]) % nc[1];
    k = (k % nc[2] + nc[2]) % nc[2];
    return i + static_cast<int64_t>(nc[0]) * (j + static_cast<int64_t>(nc[1]) * k);
  };
  auto vertexId = [&](int i, int j, int k) -> int64_t {  // periodic global vertex id
    return cellIndex(i, j, k);
  };
  std::vector<int64_t> vid(static_cast<size_t>(E) * ref.nvert);
  for (int k = 0; k < nc[2]; k++)
    for (int j = 0; j < nc[1]; j++)
      for (int i = 0; i < nc[0]; i++) {
        const int64_t e = cellIndex(i, j, k);
        for (int v = 0; v < ref.nvert; v++) {
          const double* rv = &ref.vertex[static_cast<size_t>(v) * dim];
          const int off[3] = {rv[0] > 0 ? 1 : 0, dim >= 2 && rv[1] > 0 ? 1 : 0, dim == 3 && rv[2] > 0 ? 1 : 0};
          const int idx[3] = {i + off[0], j + off[1], k + off[2]};
          for (int d = 0; d < dim; d++) mesh.vertex[(static_cast<size_t>(e) * ref.nvert + v) * dim + d] = lo[d] + idx[d] * h[d];
          vid[static_cast<size_t>(e) * ref.nvert + v] = vertexId(idx[0], idx[1], idx[2]);
        }
      }
  // geometry: affine elements (rectangular cells), Jacobian at the centre
  mesh.jac.assign(static_cast<size_t>(E) * dim * dim, 0.0);
  mesh.detj.assign(E, 0.0);
  mesh.min_edge.assign(E, 0.0);
  static const int kQuadEdge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int kHexEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                      {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  for (int64_t e = 0; e < E; e++) {
    const double xi0[3] = {0, 0, 0};
    double x[3];
    mapPoint(mesh, e, xi0, x, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
    mesh.detj[e] = det(dim, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
    if (!(mesh.detj[e] > 0.0)) {
      error = "non-positive Jacobian";
      return false;
    }
    double me = std::numeric_limits<double>::max();
    static const int kLineEdge[1][2] = {{0, 1}};
    const int ne = dim == 1 ? 1 : (dim == 2 ? 4 : 12);
    for (int g = 0; g < ne; g++) {
      const int a = dim == 1 ? kLineEdge[g][0] : (dim == 2 ? kQuadEdge[g][0] : kHexEdge[g][0]);
      const int b = dim == 1 ? kLineEdge[g][1] : (dim == 2 ? kQuadEdge[g][1] : kHexEdge[g][1]);
      double s = 0.0;
      for (int d = 0; d < dim; d++) {
        const double t = mesh.vertex[(static_cast<size_t>(e) * ref.nvert + a) * dim + d] -
                         mesh.vertex[(static_cast<size_t>(e) * ref.nvert + b) * dim + d];
        s += t * t;
      }
      me = std::min(me, std::sqrt(s));
    }
    mesh.min_edge[e] = me;
  }
  // faces: neighbour across each local face, created on first visit (the visitor is the left
  // parent), as the reference's first-seen rule (Adjacency.cpp:61-183)
  static const int kQuadNb[4][4] = {{0, -1, 0, 2}, {1, 0, 0, 3}, {0, 1, 0, 0}, {-1, 0, 0, 1}};  // di,dj,dk,face
  static const int kHexNb[6][4] = {{0, 0, -1, 5}, {0, -1, 0, 4}, {-1, 0, 0, 3},
                                   {1, 0, 0, 2},  {0, 1, 0, 1},  {0, 0, 1, 0}};
  const int nf = ref.nf, nfq = ref.nfq;
  mesh.elem_face.assign(static_cast<size_t>(E) * nf, -1);
  mesh.elem_side.assign(static_cast<size_t>(E) * nf, 0);
  std::vector<double> xl(static_cast<size_t>(nfq) * dim), xr(static_cast<size_t>(nfq) * dim);
  for (int k = 0; k < nc[2]; k++)
    for (int j = 0; j < nc[1]; j++)
      for (int i = 0; i < nc[0]; i++) {
        const int64_t e = cellIndex(i, j, k);
        for (int s = 0; s < nf; s++) {
          if (mesh.elem_face[static_cast<size_t>(e) * nf + s] != -1) continue;
          static const int kLineNb[2][4] = {{-1, 0, 0, 1}, {1, 0, 0, 0}};
          const int* nbr = dim == 1 ? kLineNb[s] : (dim == 2 ? kQuadNb[s] : kHexNb[s]);
          const int ijk[3] = {i + nbr[0], j + nbr[1], k + nbr[2]};
          int side_out = -1;  // boundary side 2 d + (0 low, 1 high) when the neighbour is outside
          for (int d = 0; d < dim; d++)
            if (!periodic[d] && (ijk[d] < 0 || ijk[d] >= nc[d])) side_out = 2 * d + (ijk[d] < 0 ? 0 : 1);
          if (side_out >= 0) {
            Face face;
            face.parent[0] = static_cast<int32_t>(e);
            face.parent[1] = -1;
            face.local[0] = s;
            face.local[1] = -1;
            face.bc = bc[side_out];
            face.seq.resize(nfq);
            for (int p = 0; p < nfq; p++) {
              face.seq[p] = p;
              mapPoint(mesh, e, &ref.face_param[(static_cast<size_t>(s) * nfq + p) * dim], &xl[p * dim], nullptr);
            }
            face.coord.assign(xl.begin(), xl.end());
            faceNormals(mesh, e, s, face);
            mesh.elem_face[static_cast<size_t>(e) * nf + s] = -2 - static_cast<int32_t>(mesh.bfaces.size());
            mesh.elem_side[static_cast<size_t>(e) * nf + s] = 0;
            mesh.bfaces.push_back(std::move(face));
            continue;
          }
          const int64_t r = cellIndex(ijk[0], ijk[1], ijk[2]);
          const int rs = nbr[3];
          Face face;
          face.parent[0] = static_cast<int32_t>(e);
          face.parent[1] = static_cast<int32_t>(r);
          face.local[0] = s;
          face.local[1] = rs;
          // physical face points from both sides, matched modulo the period
          for (int p = 0; p < nfq; p++) {
            mapPoint(mesh, e, &ref.face_param[(static_cast<size_t>(s) * nfq + p) * dim], &xl[p * dim], nullptr);
            mapPoint(mesh, r, &ref.face_param[(static_cast<size_t>(rs) * nfq + p) * dim], &xr[p * dim], nullptr);
          }
          face.seq.assign(nfq, -1);
          for (int p = 0; p < nfq; p++) {
            double best = std::numeric_limits<double>::max();
            int bi = -1;
            for (int qq = 0; qq < nfq; qq++) {
              double s2 = 0.0;
              for (int d = 0; d < dim; d++) {
                double t = xl[p * dim + d] - xr[qq * dim + d];
                t -= period[d] * std::round(t / period[d]);
                s2 += t * t;
              }
              if (s2 < best) {
                best = s2;
                bi = qq;
              }
            }
            if (best > 1e-20 * (1.0 + h[0] * h[0])) {
              error = "face points of the two parents do not match";
              return false;
            }
            face.seq[p] = bi;
          }
          // rotation: position of the left's first face vertex in the right's face vertex list
          face.rotation = 0;
          if (dim == 3) {
            const int64_t first = vid[static_cast<size_t>(e) * ref.nvert + ref.face_vertex[s * ref.nfvert]];
            for (int v = 0; v < ref.nfvert; v++)
              if (vid[static_cast<size_t>(r) * ref.nvert + ref.face_vertex[rs * ref.nfvert + v]] == first) face.rotation = v;
          }
          face.coord.assign(xl.begin(), xl.end());
          faceNormals(mesh, e, s, face);
          const int32_t fid = static_cast<int32_t>(mesh.faces.size());
          mesh.faces.push_back(std::move(face));
          mesh.elem_face[static_cast<size_t>(e) * nf + s] = fid;
          mesh.elem_side[static_cast<size_t>(e) * nf + s] = 0;
          if (mesh.elem_face[static_cast<size_t>(r) * nf + rs] != -1) {
            error = "face recorded more than twice";
            return false;
          }
          mesh.elem_face[static_cast<size_t>(r) * nf + rs] = fid;
          mesh.elem_side[static_cast<size_t>(r) * nf + rs] = 1;
        }
      }
  mesh.affine = true;
  assignNodes(mesh);
  return true;
}

bool buildMeshFromConnectivity(int type, int order, int64_t nnode, const double* xyz, int64_t elements,
                               const int64_t* conn, int64_t nbface, const int64_t* bconn, const int* btype,
                               const int* bphysical, Mesh& mesh, std::string& error,
                               const std::vector<std::map<int64_t, int64_t>>* periodic_links) {
  mesh = Mesh();
  if (type != 2 && type != 3 && type != 6) {
    error = "element type must be triangle, quadrangle or hexahedron";
    return false;
  }
  if (!buildRefElement(type, order, mesh.ref, error)) return false;
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim, nv = ref.nvert, nf = ref.nf, nfq = ref.nfq, nfv = ref.nfvert;
  if (elements < 1 || elements > (int64_t(1) << 30)) {
    error = "element count out of range";
    return false;
  }
  const int64_t E = elements;
  mesh.elements = E;
  std::vector<int64_t> cv(conn, conn + E * nv);
  for (int64_t i = 0; i < E * nv; i++)
    if (cv[i] < 0 || cv[i] >= nnode) {
      error = "element node index out of range";
      return false;
    }
  mesh.vertex.assign(static_cast<size_t>(E) * nv * dim, 0.0);
  mesh.jac.assign(static_cast<size_t>(E) * dim * dim, 0.0);
  mesh.detj.assign(E, 0.0);
  mesh.min_edge.assign(E, 0.0);
  static const int kMirror[3][8] = {{0, 2, 1}, {0, 3, 2, 1}, {0, 3, 2, 1, 4, 7, 6, 5}};
  const int* mirror = kMirror[type == 2 ? 0 : (type == 3 ? 1 : 2)];
  const double xi_c[3] = {type == 2 ? 1.0 / 3.0 : 0.0, type == 2 ? 1.0 / 3.0 : 0.0, 0.0};
  static const int kHexEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                      {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  double xmax = 0.0;
  for (int64_t e = 0; e < E; e++) {
    for (int pass = 0; pass < 2; pass++) {
      for (int v = 0; v < nv; v++)
        for (int d = 0; d < dim; d++)
          mesh.vertex[(static_cast<size_t>(e) * nv + v) * dim + d] = xyz[cv[static_cast<size_t>(e) * nv + v] * 3 + d];
      double x[3];
      mapPoint(mesh, e, xi_c, x, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
      mesh.detj[e] = det(dim, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
      if (mesh.detj[e] > 0.0) break;
      if (pass == 1 || mesh.detj[e] == 0.0) {
        error = "degenerate element";
        return false;
      }
      int64_t tmp[8];
      for (int v = 0; v < nv; v++) tmp[v] = cv[static_cast<size_t>(e) * nv + mirror[v]];
      for (int v = 0; v < nv; v++) cv[static_cast<size_t>(e) * nv + v] = tmp[v];
    }
    // affine: the map through the centre and its Jacobian reproduces every vertex
    double xc[3];
    mapPoint(mesh, e, xi_c, xc, nullptr);
    double me = std::numeric_limits<double>::max(), span = 0.0;
    const int ne = dim == 2 ? nf : 12;
    for (int g = 0; g < ne; g++) {
      const int a = dim == 2 ? ref.face_vertex[g * 2] : kHexEdge[g][0];
      const int b = dim == 2 ? ref.face_vertex[g * 2 + 1] : kHexEdge[g][1];
      double s2 = 0.0;
      for (int d = 0; d < dim; d++) {
        const double t = mesh.vertex[(static_cast<size_t>(e) * nv + a) * dim + d] - mesh.vertex[(static_cast<size_t>(e) * nv + b) * dim + d];
        s2 += t * t;
      }
      me = std::min(me, std::sqrt(s2));
      span = std::max(span, std::sqrt(s2));
    }
    mesh.min_edge[e] = me;
    const double* J = &mesh.jac[static_cast<size_t>(e) * dim * dim];
    for (int v = 0; v < nv; v++) {
      double err2 = 0.0;
      for (int l = 0; l < dim; l++) {
        double x = xc[l];
        for (int k = 0; k < dim; k++) x += J[l * dim + k] * (ref.vertex[static_cast<size_t>(v) * dim + k] - xi_c[k]);
        const double t = x - mesh.vertex[(static_cast<size_t>(e) * nv + v) * dim + l];
        err2 += t * t;
        xmax = std::max(xmax, std::fabs(mesh.vertex[(static_cast<size_t>(e) * nv + v) * dim + l]));
      }
      if (err2 > 1e-20 * span * span) {
        error = "non-affine element (the engine needs triangles, parallelograms or parallelepipeds)";
        return false;
      }
    }
  }
  // faces keyed by their sorted node indices
  auto keyOf = [&](const int64_t* nodes) {
    std::array<int64_t, 4> k = {-1, -1, -1, -1};
    for (int v = 0; v < nfv; v++) k[v] = nodes[v];
    std::sort(k.begin(), k.begin() + nfv);
    return k;
  };
  std::map<std::array<int64_t, 4>, std::vector<std::pair<int64_t, int>>> visits;
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      int64_t fn[4];
      for (int v = 0; v < nfv; v++) fn[v] = cv[static_cast<size_t>(e) * nv + ref.face_vertex[s * nfv + v]];
      visits[keyOf(fn)].push_back({e, s});
    }
  std::map<std::array<int64_t, 4>, std::pair<int, int>> boundary_type;  // type, physical tag
  for (int64_t b = 0; b < nbface; b++) {
    if (btype[b] < 0 || btype[b] > 5) {
      error = "boundary type must be 0..5";
      return false;
    }
    boundary_type[keyOf(&bconn[static_cast<size_t>(b) * nfv])] = {btype[b], bphysical ? bphysical[b] : -1};
  }
  // periodic partners: an unpaired face on the slave nodes of a link and the unpaired face on
  // their masters; link_of records the link for the rotation's node comparison
  std::map<std::pair<int64_t, int>, std::pair<std::pair<int64_t, int>, int>> partner;
  if (periodic_links) {
    for (const auto& kv : visits) {
      if (kv.second.size() != 1) continue;
      const auto [e, s] = kv.second[0];
      for (size_t l = 0; l < periodic_links->size(); l++) {
        const auto& link = (*periodic_links)[l];
        int64_t mapped[4];
        bool slave = true;
        for (int v = 0; v < nfv && slave; v++) {
          const auto it = link.find(cv[static_cast<size_t>(e) * nv + ref.face_vertex[s * nfv + v]]);
          slave = it != link.end();
          if (slave) mapped[v] = it->second;
        }
        if (!slave) continue;
        const auto m = visits.find(keyOf(mapped));
        if (m == visits.end() || m->second.size() != 1) {
          error = "a periodic face has no unpaired master face";
          return false;
        }
        partner[{e, s}] = {m->second[0], static_cast<int>(l)};
        partner[m->second[0]] = {{e, s}, static_cast<int>(l)};
        break;
      }
    }
  }
  mesh.elem_face.assign(static_cast<size_t>(E) * nf, -1);
  mesh.elem_side.assign(static_cast<size_t>(E) * nf, 0);
  std::vector<double> xl(static_cast<size_t>(nfq) * dim), xr(static_cast<size_t>(nfq) * dim);
  const double tol = 1e-20 * (1.0 + xmax * xmax);
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      if (mesh.elem_face[static_cast<size_t>(e) * nf + s] != -1) continue;
      int64_t fn[4];
      for (int v = 0; v < nfv; v++) fn[v] = cv[static_cast<size_t>(e) * nv + ref.face_vertex[s * nfv + v]];
      const auto key = keyOf(fn);
      std::vector<std::pair<int64_t, int>> list = visits[key];
      const auto pp = partner.find({e, s});
      if (pp != partner.end()) list.push_back(pp->second.first);
      for (int p = 0; p < nfq; p++)
        mapPoint(mesh, e, &ref.face_param[(static_cast<size_t>(s) * nfq + p) * dim], &xl[p * dim], nullptr);
      if (list.size() == 1) {
        const auto it = boundary_type.find(key);
        if (it == boundary_type.end()) {
          error = "a boundary face has no boundary condition (no boundary element of a mapped physical group)";
          return false;
        }
        Face face;
        face.parent[0] = static_cast<int32_t>(e);
        face.parent[1] = -1;
        face.local[0] = s;
        face.local[1] = -1;
        face.bc = it->second.first;
        face.physical = it->second.second;
        face.seq.resize(nfq);
        for (int p = 0; p < nfq; p++) face.seq[p] = p;
        face.coord.assign(xl.begin(), xl.end());
        faceNormals(mesh, e, s, face);
        mesh.elem_face[static_cast<size_t>(e) * nf + s] = -2 - static_cast<int32_t>(mesh.bfaces.size());
        mesh.bfaces.push_back(std::move(face));
        continue;
      }
      if (list.size() != 2) {
        error = "a face is shared by more than two elements";
        return false;
      }
      const int64_t r = list[1].first;
      const int rs = list[1].second;
      Face face;
      face.parent[0] = static_cast<int32_t>(e);
      face.parent[1] = static_cast<int32_t>(r);
      face.local[0] = s;
      face.local[1] = rs;
      for (int p = 0; p < nfq; p++)
        mapPoint(mesh, r, &ref.face_param[(static_cast<size_t>(rs) * nfq + p) * dim], &xr[p * dim], nullptr);
      // translation between the two faces (zero unless the face is periodic): difference of the
      // face-vertex centroids
      double shift[3] = {0.0, 0.0, 0.0};
      for (int v = 0; v < nfv; v++)
        for (int d = 0; d < dim; d++)
          shift[d] += (xyz[fn[v] * 3 + d] - xyz[cv[static_cast<size_t>(r) * nv + ref.face_vertex[rs * nfv + v]] * 3 + d]) / nfv;
      face.seq.assign(nfq, -1);
      for (int p = 0; p < nfq; p++) {
        double best = std::numeric_limits<double>::max();
        int bi = -1;
        for (int qq = 0; qq < nfq; qq++) {
          double s2 = 0.0;
          for (int d = 0; d < dim; d++) {
            const double t = xl[p * dim + d] - (xr[qq * dim + d] + shift[d]);
            s2 += t * t;
          }
          if (s2 < best) {
            best = s2;
            bi = qq;
          }
        }
        if (best > tol) {
          error = "face points of the two parents do not match";
          return false;
        }
        face.seq[p] = bi;
      }
      face.rotation = 0;
      if (dim == 3) {  // the left's first face vertex in the right's list (through the periodic map)
        auto image = [&](int64_t n) {
          if (pp == partner.end()) return n;
          const auto& link = (*periodic_links)[static_cast<size_t>(pp->second.second)];
          const auto it = link.find(n);
          return it == link.end() ? n : it->second;
        };
        const int64_t first = image(fn[0]);
        for (int v = 0; v < nfv; v++)
          if (image(cv[static_cast<size_t>(r) * nv + ref.face_vertex[rs * nfv + v]]) == first) face.rotation = v;
      }
      face.coord.assign(xl.begin(), xl.end());
      faceNormals(mesh, e, s, face);
      const int32_t fid = static_cast<int32_t>(mesh.faces.size());
      mesh.faces.push_back(std::move(face));
      mesh.elem_face[static_cast<size_t>(e) * nf + s] = fid;
      mesh.elem_side[static_cast<size_t>(e) * nf + s] = 0;
      mesh.elem_face[static_cast<size_t>(r) * nf + rs] = fid;
      mesh.elem_side[static_cast<size_t>(r) * nf + rs] = 1;
    }
  mesh.affine = true;
  assignNodes(mesh);
  return true;
}

bool buildSplitMesh(int order, const int n[2], const double lo[2], const double hi[2], const std::vector<uint8_t>& split,
                    MixedMesh& mesh, std::string& error) {
  mesh = MixedMesh();
  mesh.order = order;
  if (!buildRefElement(2, order, mesh.ref[0], error) || !buildRefElement(3, order, mesh.ref[1], error)) return false;
  for (int d = 0; d < 2; d++) {
    if (n[d] < 2) {
      error = "a periodic direction needs at least 2 cells";
      return false;
    }
    if (!(hi[d] > lo[d])) {
      error = "box upper corner must exceed the lower corner";
      return false;
    }
  }
  if (split.size() != static_cast<size_t>(n[0]) * n[1]) {
    error = "one split flag per cell is required";
    return false;
  }
  const double period[2] = {hi[0] - lo[0], hi[1] - lo[1]};
  const double h[2] = {period[0] / n[0], period[1] / n[1]};
  // elements
  for (int j = 0; j < n[1]; j++)
    for (int i = 0; i < n[0]; i++) {
      const double A[2] = {lo[0] + i * h[0], lo[1] + j * h[1]}, B[2] = {A[0] + h[0], A[1]},
                   C[2] = {A[0] + h[0], A[1] + h[1]}, Dv[2] = {A[0], A[1] + h[1]};
      auto push = [&](int8_t t, std::initializer_list<const double*> vs) {
        mesh.etype.push_back(t);
        int k = 0;
        for (const double* v : vs) {
          mesh.vertex.push_back(v[0]);
          mesh.vertex.push_back(v[1]);
          k++;
        }
        for (; k < 4; k++) {
          mesh.vertex.push_back(0.0);
          mesh.vertex.push_back(0.0);
        }
      };
      if (split[static_cast<size_t>(j) * n[0] + i]) {
        push(0, {A, B, C});
        push(0, {A, C, Dv});
      } else {
        push(1, {A, B, C, Dv});
      }
    }
  const int64_t E = static_cast<int64_t>(mesh.etype.size());
  mesh.elements = E;
  mesh.jac.assign(static_cast<size_t>(E) * 4, 0.0);
  mesh.detj.assign(E, 0.0);
  mesh.min_edge.assign(E, 0.0);
  for (int64_t e = 0; e < E; e++) {
    const RefElement& ref = mesh.refOf(e);
    const double xi0[2] = {ref.type == 2 ? 1.0 / 3.0 : 0.0, ref.type == 2 ? 1.0 / 3.0 : 0.0};
    double x[2];
    mapPointRef(ref, &mesh.vertex[static_cast<size_t>(e) * 8], xi0, x, &mesh.jac[static_cast<size_t>(e) * 4]);
    mesh.detj[e] = det(2, &mesh.jac[static_cast<size_t>(e) * 4]);
    if (!(mesh.detj[e] > 0.0)) {
      error = "non-positive Jacobian";
      return false;
    }
    double me = std::numeric_limits<double>::max();
    for (int a = 0; a < ref.nvert; a++) {
      const int b = (a + 1) % ref.nvert;
      const double dx = mesh.vertex[static_cast<size_t>(e) * 8 + 2 * a] - mesh.vertex[static_cast<size_t>(e) * 8 + 2 * b];
      const double dy = mesh.vertex[static_cast<size_t>(e) * 8 + 2 * a + 1] - mesh.vertex[static_cast<size_t>(e) * 8 + 2 * b + 1];
      me = std::min(me, std::sqrt(dx * dx + dy * dy));
    }
    mesh.min_edge[e] = me;
  }
  // faces: edges keyed by their midpoint modulo the period (on a grid of quarter cells)
  auto key = [&](int64_t e, int s) {
    const RefElement& ref = mesh.refOf(e);
    const int a = ref.face_vertex[s * 2], b = ref.face_vertex[s * 2 + 1];
    int64_t k[2];
    for (int d = 0; d < 2; d++) {
      const double m = 0.5 * (mesh.vertex[static_cast<size_t>(e) * 8 + 2 * a + d] + mesh.vertex[static_cast<size_t>(e) * 8 + 2 * b + d]);
      const int64_t q = static_cast<int64_t>(std::llround(4.0 * (m - lo[d]) / h[d]));
      const int64_t per = 4 * static_cast<int64_t>(n[d]);
      k[d] = ((q % per) + per) % per;
    }
    return k[0] + k[1] * 4 * static_cast<int64_t>(n[0]);
  };
  std::unordered_map<int64_t, std::vector<std::pair<int64_t, int>>> edges;
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < mesh.refOf(e).nf; s++) edges[key(e, s)].push_back({e, s});
  mesh.elem_face.assign(static_cast<size_t>(E) * 4, -1);
  mesh.elem_side.assign(static_cast<size_t>(E) * 4, 0);
  const int nfq = order + 1;
  std::vector<double> xl(static_cast<size_t>(nfq) * 2), xr(static_cast<size_t>(nfq) * 2);
  for (int64_t e = 0; e < E; e++) {
    const RefElement& ref = mesh.refOf(e);
    for (int s = 0; s < ref.nf; s++) {
      if (mesh.elem_face[static_cast<size_t>(e) * 4 + s] >= 0) continue;
      const auto& list = edges[key(e, s)];
      if (list.size() != 2) {
        error = "an edge is not shared by exactly two elements";
        return false;
      }
      const auto other = list[0].first == e && list[0].second == s ? list[1] : list[0];
      const int64_t r = other.first;
      const int rs = other.second;
      const RefElement& rref = mesh.refOf(r);
      Face face;
      face.parent[0] = static_cast<int32_t>(e);
      face.parent[1] = static_cast<int32_t>(r);
      face.local[0] = s;
      face.local[1] = rs;
      for (int p = 0; p < nfq; p++) {
        mapPointRef(ref, &mesh.vertex[static_cast<size_t>(e) * 8], &ref.face_param[(static_cast<size_t>(s) * nfq + p) * 2],
                    &xl[p * 2], nullptr);
        mapPointRef(rref, &mesh.vertex[static_cast<size_t>(r) * 8], &rref.face_param[(static_cast<size_t>(rs) * nfq + p) * 2],
                    &xr[p * 2], nullptr);
      }
      face.seq.assign(nfq, -1);
      for (int p = 0; p < nfq; p++) {
        double best = std::numeric_limits<double>::max();
        int bi = -1;
        for (int qq = 0; qq < nfq; qq++) {
          double s2 = 0.0;
          for (int d = 0; d < 2; d++) {
            double t = xl[p * 2 + d] - xr[qq * 2 + d];
            t -= period[d] * std::round(t / period[d]);
            s2 += t * t;
          }
          if (s2 < best) {
            best = s2;
            bi = qq;
          }
        }
        if (best > 1e-20 * (1.0 + h[0] * h[0])) {
          error = "face points of the two parents do not match";
          return false;
        }
        face.seq[p] = bi;
      }
      face.coord.assign(xl.begin(), xl.end());
      faceNormalsRef(ref, &mesh.jac[static_cast<size_t>(e) * 4], s, face);
      const int32_t fid = static_cast<int32_t>(mesh.faces.size());
      mesh.faces.push_back(std::move(face));
      mesh.elem_face[static_cast<size_t>(e) * 4 + s] = fid;
      mesh.elem_side[static_cast<size_t>(e) * 4 + s] = 0;
      mesh.elem_face[static_cast<size_t>(r) * 4 + rs] = fid;
      mesh.elem_side[static_cast<size_t>(r) * 4 + rs] = 1;
    }
  }
  return true;
}

bool singleTypeView(const MixedMesh& mm, Mesh& out) {
  const int64_t E = mm.elements;
  if (E == 0) return false;
  for (int64_t e = 1; e < E; e++)
    if (mm.etype[e] != mm.etype[0]) return false;
  out = Mesh();
  out.ref = mm.ref[mm.etype[0]];
  const int nv = out.ref.nvert, nf = out.ref.nf;
  out.elements = E;
  out.vertex.resize(static_cast<size_t>(E) * nv * 2);
  for (int64_t e = 0; e < E; e++)
    for (int i = 0; i < nv * 2; i++) out.vertex[static_cast<size_t>(e) * nv * 2 + i] = mm.vertex[static_cast<size_t>(e) * 8 + i];
  out.jac = mm.jac;
  out.detj = mm.detj;
  out.min_edge = mm.min_edge;
  out.faces = mm.faces;
  out.elem_face.resize(static_cast<size_t>(E) * nf);
  out.elem_side.resize(static_cast<size_t>(E) * nf);
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      out.elem_face[static_cast<size_t>(e) * nf + s] = mm.elem_face[static_cast<size_t>(e) * 4 + s];
      out.elem_side[static_cast<size_t>(e) * nf + s] = mm.elem_side[static_cast<size_t>(e) * 4 + s];
    }
  out.affine = true;
  assignNodes(out);
  return true;
}

void mapReferencePoint(const Mesh& mesh, int64_t e, const double* xi, double* x) { mapPoint(mesh, e, xi, x, nullptr); }

void assignNodes(Mesh& mesh) {
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim, nv = ref.nvert;
  const int64_t E = mesh.elements;
  // vertices coincide when their coordinates agree to a millionth of the smallest edge
  double hmin = std::numeric_limits<double>::max();
  for (int64_t e = 0; e < E; e++) hmin = std::min(hmin, mesh.min_edge[e]);
  const double step = 1e-6 * hmin;
  std::map<std::array<int64_t, 3>, int32_t> index;
  mesh.node_tag.assign(static_cast<size_t>(E) * nv, 0);
  for (int64_t e = 0; e < E; e++)
    for (int v = 0; v < nv; v++) {
      std::array<int64_t, 3> key = {0, 0, 0};
      for (int d = 0; d < dim; d++)
        key[d] = static_cast<int64_t>(std::llround(mesh.vertex[(static_cast<size_t>(e) * nv + v) * dim + d] / step));
      auto it = index.find(key);
      if (it == index.end()) it = index.emplace(key, static_cast<int32_t>(index.size())).first;
      mesh.node_tag[static_cast<size_t>(e) * nv + v] = it->second;
    }
  mesh.nodes = static_cast<int64_t>(index.size());
  // inner radius D |element| / |boundary|; reference measures: triangle 1/2, quadrangle 4, hexahedron 8
  const double measure = ref.type == 2 ? 0.5 : (dim == 1 ? 2.0 : (dim == 2 ? 4.0 : 8.0));
  mesh.inner_radius.assign(E, 0.0);
  for (int64_t e = 0; e < E; e++) {
    const double* x = &mesh.vertex[static_cast<size_t>(e) * nv * dim];
    double surface = 0.0;
    for (int f = 0; f < ref.nf; f++) {
      const int* fv = &ref.face_vertex[static_cast<size_t>(f) * ref.nfvert];
      if (dim == 1) {
        surface += 1.0;  // an end point
      } else if (dim == 2) {
        double s2 = 0.0;
        for (int d = 0; d < 2; d++) {
          const double t = x[fv[1] * 2 + d] - x[fv[0] * 2 + d];
          s2 += t * t;
        }
        surface += std::sqrt(s2);
      } else {  // planar parallelogram face a, b, c, d: |(b - a) x (d - a)|
        double u[3], w[3];
        for (int d = 0; d < 3; d++) {
          u[d] = x[fv[1] * 3 + d] - x[fv[0] * 3 + d];
          w[d] = x[fv[3] * 3 + d] - x[fv[0] * 3 + d];
        }
        const double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        surface += std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      }
    }
    mesh.inner_radius[e] = dim * (measure * mesh.detj[e]) / surface;
  }
}

void affineGeometry(const RefElement& ref, const double* J, double dj, double* detjw, double* tiw) {
  const int dim = ref.dim;
  // Jt(k,l) = J(l,k); G = inverse(Jt) |J| w, column-major: tiw[a + D b] = inv(Jt)(a,b)
  double jt[9], jti[9];
  for (int k = 0; k < dim; k++)
    for (int l = 0; l < dim; l++) jt[k * dim + l] = J[l * dim + k];
  invertMatrix(dim, jt, jti);
  for (int q = 0; q < ref.nq; q++) {
    const double jw = dj * ref.qweight[q];
    if (detjw) detjw[q] = jw;
    if (tiw)
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) tiw[static_cast<size_t>(q) * dim * dim + a + dim * b] = jti[a * dim + b] * jw;
  }
}

void elementGeometry(const Mesh& mesh, int64_t e, double* detjw, double* tiw, double* coord) {
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim;
  affineGeometry(ref, &mesh.jac[static_cast<size_t>(e) * dim * dim], mesh.detj[e], detjw, tiw);
  if (coord)
    for (int q = 0; q < ref.nq; q++)
      mapPoint(mesh, e, &ref.qpoint[static_cast<size_t>(q) * dim], &coord[static_cast<size_t>(q) * dim], nullptr);
}

void elementMassInverse(const Mesh& mesh, int64_t e, double* minv) { affineMassInverse(mesh.ref, mesh.detj[e], minv); }

void affineMassInverse(const RefElement& ref, double dj, double* minv) {
  const int nb = ref.nb;
  std::vector<double> m(static_cast<size_t>(nb) * nb, 0.0);
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < nb; j++) {
      double s = 0.0;
      for (int q = 0; q < ref.nq; q++)
        s += ref.V[static_cast<size_t>(q) * nb + i] * (ref.V[static_cast<size_t>(q) * nb + j] * (dj * ref.qweight[q]));
      m[static_cast<size_t>(i) * nb + j] = s;
    }
  invertMatrix(nb, m.data(), minv);
}

}  // namespace sdg
