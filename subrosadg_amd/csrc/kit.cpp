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
// P1 face element vertices: line (-1),(1); quad (-1,-1),(1,-1),(1,1),(-1,1); triangle (0,0),(1,0),(0,1)
const double kLineFaceVertex[2] = {-1, 1};
const double kQuadFaceVertex[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
// tetrahedron: gmsh reference vertices and local face vertex lists (SimulationControl.cpp:193-198),
// each face ordered so that (v1 - v0) x (v2 - v0) points out of the element
const double kTetVertex[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
const int kTetFace[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
// pyramid: gmsh reference vertices (base [-1,1]^2 at z = 0, apex (0,0,1)) and local faces
// (SimulationControl.cpp:199-204): four triangles, then the quadrangle base; -1 pads a triangle
const double kPyrVertex[5][3] = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {0, 0, 1}};
const int kPyrFace[5][4] = {{0, 1, 4, -1}, {3, 0, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 3, 2, 1}};
const int kPyrEdge[8][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 4}, {2, 4}, {3, 4}};
// element edges (vertex pairs) of the 2-D and 3-D elements, for minimum edge lengths
const int kTriEdge[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int kQuadEdge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
const int kHexEdge12[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

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

// Symmetric rules on the reference triangle of degree 3, 5, 7, 9 with 4, 7, 13, 19 points: the face
// rules (order 2P+1, SimulationControl.cpp:275-283) of tetrahedra P = 1..4, with the orbit types of
// the volume rules above; parameters solved from the moment equations (Dunavant's orbit structures)
// in 50-digit arithmetic, exactness checked in tests/test_kit_tet.py.
const TriOrbit kTriRule3[] = {{0, 0.0, 0.0, -0.28125}, {1, 0.2, 0.0, 0.26041666666666667}};
// degree 5: Radon's rule, a = (6 +- sqrt 15) / 21, w = (155 +- sqrt 15) / 2400, centroid 9/80
const TriOrbit kTriRule5[] = {{0, 0.0, 0.0, 0.1125},
                              {1, 0.47014206410511509, 0.0, 0.066197076394253090},
                              {1, 0.10128650732345634, 0.0, 0.062969590272413576}};
const TriOrbit kTriRule7[] = {{0, 0.0, 0.0, -0.074785022233840875},
                              {1, 0.26034596607903983, 0.0, 0.087807628716603906},
                              {1, 0.065130102902215812, 0.0, 0.026673617804419246},
                              {2, 0.048690315425316412, 0.31286549600487386, 0.03855688044512857}};
const TriOrbit kTriRule9[] = {{0, 0.0, 0.0, 0.048567898141399417},
                              {1, 0.48968251919873763, 0.0, 0.015667350113569535},
                              {1, 0.43708959149293664, 0.0, 0.03891377050238714},
                              {1, 0.18820353561903273, 0.0, 0.039823869463605127},
                              {1, 0.04472951339445271, 0.0, 0.012788837829349016},
                              {2, 0.036838412054736284, 0.2219629891607657, 0.021641769688644689}};

void expandTriRule(const TriOrbit* r, int n, std::vector<double>& pts, std::vector<double>& w) {
  pts.clear();
  w.clear();
  auto add = [&](double l1, double l2, double wt) {
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
      add(o.a, o.a, o.w);
      add(o.a, c, o.w);
      add(c, o.a, o.w);
    } else {
      const double c = 1.0 - o.a - o.b;
      const double t[6][2] = {{o.b, c}, {c, o.a}, {o.a, o.b}, {o.b, o.a}, {c, o.b}, {o.a, c}};
      for (int k = 0; k < 6; k++) add(t[k][0], t[k][1], o.w);
    }
  }
}

// face rule of a tetrahedron of order P: degree 2P+1 on the reference triangle
void triangleFaceRule(int order, std::vector<double>& pts, std::vector<double>& w) {
  if (order == 1) expandTriRule(kTriRule3, 2, pts, w);
  else if (order == 2) expandTriRule(kTriRule5, 3, pts, w);
  else if (order == 3) expandTriRule(kTriRule7, 4, pts, w);
  else expandTriRule(kTriRule9, 6, pts, w);
}

// Grundmann-Moeller rule of degree 2s+1 on the reference simplex of dimension n (the origin and
// the unit vectors; A. Grundmann, H. M. Moeller, SIAM J. Numer. Anal. 15 (1978)): for i = 0..s and
// every composition beta of s - i into n + 1 parts, the point with barycentric coordinates
// (2 beta_j + 1) / (d + n - 2i), weight proportional to (-1)^i (d + n - 2i)^d / (i! (d + n - i)!),
// d = 2s + 1; weights scaled to the simplex volume 1/n!. Volume rules of tetrahedra (s = P).
void grundmannMoeller(int n, int s, std::vector<double>& pts, std::vector<double>& w) {
  pts.clear();
  w.clear();
  const int d = 2 * s + 1;
  auto fact = [](int k) {
    double f = 1.0;
    for (int i = 2; i <= k; i++) f *= i;
    return f;
  };
  std::vector<int> beta(static_cast<size_t>(n) + 1, 0);
  for (int i = 0; i <= s; i++) {
    const int m = s - i;
    const double den = d + n - 2 * i;
    const double coef = (i % 2 ? -1.0 : 1.0) * std::pow(den, d) / (fact(i) * fact(d + n - i));
    // compositions of m into n + 1 parts, beta_1..beta_n enumerated, beta_0 the rest
    std::fill(beta.begin(), beta.end(), 0);
    while (true) {
      int used = 0;
      for (int j = 1; j <= n; j++) used += beta[j];
      if (used <= m) {
        beta[0] = m - used;
        for (int j = 1; j <= n; j++) pts.push_back((2.0 * beta[j] + 1.0) / den);
        w.push_back(coef);
      }
      int j = 1;
      while (j <= n) {
        if (++beta[j] <= m) break;
        beta[j] = 0;
        j++;
      }
      if (j > n) break;
    }
  }
  double sum = 0.0;
  for (double x : w) sum += x;
  const double vol = 1.0 / fact(n);
  for (double& x : w) x *= vol / sum;
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

// Hierarchic modal basis of order P on the reference tetrahedron spanning P_P, built like the
// triangle's: the 4 barycentric vertex functions, then per degree k = 2..P the 6 edge functions
// lambda_a lambda_b kappa_{k-2}(lambda_b - lambda_a), the face functions
// lambda_a lambda_b lambda_c L_i(lambda_b - lambda_a) L_j(lambda_c - lambda_b) (i + j = k - 3) of
// the 4 faces, and the bubbles lambda_0 lambda_1 lambda_2 lambda_3 L_i L_j L_l of
// (lambda_1 - lambda_0, lambda_2 - lambda_1, lambda_3 - lambda_2) with i + j + l = k - 4:
// 4 + 6 (P-1) + 2 (P-1)(P-2) + (P-1)(P-2)(P-3)/6 = (P+1)(P+2)(P+3)/6 functions. The edge, face and
// bubble functions are scaled by 4, 27 and 256 (the reciprocal maxima of lambda_a lambda_b,
// lambda_a lambda_b lambda_c and the product of all four), which keeps the reference mass matrix
// of P4 at a condition number near 1e4 instead of 3e6. The reference takes gmsh's H1Legendre basis
// of the same space (BasisFunction.cpp:58-73).
template <typename T>
void tetBasis(int order, T x, T y, T z, T* value) {
  const T lam[4] = {T(1.0) - x - y - z, x, y, z};
  static const int kEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  static const int kFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
  T L[8], dL[8], L2[8], dL2[8], L3[8], dL3[8];
  int b = 0;
  for (int v = 0; v < 4; v++) value[b++] = lam[v];
  for (int k = 2; k <= order; k++) {
    for (int e = 0; e < 6; e++) {
      const T l0 = lam[kEdge[e][0]], l1 = lam[kEdge[e][1]];
      legendreSeries(k - 1, l1 - l0, L, dL);
      const T kappa = T(-4.0 * (2.0 * k - 1.0) / (k * (k - 1.0) * std::sqrt(2.0 * (2.0 * k - 1.0)))) * dL[k - 1];
      value[b++] = T(4.0) * l0 * l1 * kappa;
    }
    if (k >= 3)
      for (int f = 0; f < 4; f++) {
        const T la = lam[kFace[f][0]], lb = lam[kFace[f][1]], lc = lam[kFace[f][2]];
        legendreSeries(k - 3, lb - la, L, dL);
        legendreSeries(k - 3, lc - lb, L2, dL2);
        for (int i = 0; i <= k - 3; i++) value[b++] = T(27.0) * la * lb * lc * L[i] * L2[k - 3 - i];
      }
    if (k >= 4) {
      legendreSeries(k - 4, lam[1] - lam[0], L, dL);
      legendreSeries(k - 4, lam[2] - lam[1], L2, dL2);
      legendreSeries(k - 4, lam[3] - lam[2], L3, dL3);
      for (int i = 0; i <= k - 4; i++)
        for (int j = 0; i + j <= k - 4; j++)
          value[b++] = T(256.0) * lam[0] * lam[1] * lam[2] * lam[3] * L[i] * L2[j] * L3[k - 4 - i - j];
    }
  }
}

// Gauss-Jacobi rule of n points for the weight (1 - t)^alpha on [-1, 1] (beta = 0): the nodes are the
// eigenvalues of the Jacobi matrix of the monic recurrence and the weights mu0 times the squared
// first eigenvector components (Golub-Welsch), by cyclic Jacobi rotations of the small matrix.
void gaussJacobi(int n, double alpha, std::vector<double>& x, std::vector<double>& w) {
  std::vector<double> A(static_cast<size_t>(n) * n, 0.0), Q(static_cast<size_t>(n) * n, 0.0);
  const double beta = 0.0;
  for (int k = 0; k < n; k++) {
    const double s = 2.0 * k + alpha + beta;
    A[k * n + k] = (beta * beta - alpha * alpha) / (s * (s + 2.0));
    if (k > 0) {
      const double num = 4.0 * k * (k + alpha) * (k + beta) * (k + alpha + beta);
      const double den = s * s * (s + 1.0) * (s - 1.0);
      A[k * n + k - 1] = A[(k - 1) * n + k] = std::sqrt(num / den);
    }
    Q[k * n + k] = 1.0;
  }
  if (n >= 1) A[0] = (beta - alpha) / (alpha + beta + 2.0);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-34) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; k++) {  // A <- A G
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - sn * akq;
          A[k * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // A <- G^T A
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - sn * aqk;
          A[q * n + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {  // Q <- Q G
          const double qkp = Q[k * n + p], qkq = Q[k * n + q];
          Q[k * n + p] = c * qkp - sn * qkq;
          Q[k * n + q] = sn * qkp + c * qkq;
        }
      }
  }
  const double mu0 = std::pow(2.0, alpha + beta + 1.0) * std::tgamma(alpha + 1.0) * std::tgamma(beta + 1.0) /
                     std::tgamma(alpha + beta + 2.0);
  std::vector<std::pair<double, double>> nw;
  for (int k = 0; k < n; k++) nw.push_back({A[k * n + k], mu0 * Q[k] * Q[k]});
  std::sort(nw.begin(), nw.end());
  x.clear();
  w.clear();
  for (const auto& p : nw) {
    x.push_back(p.first);
    w.push_back(p.second);
  }
}

// Jacobi polynomials P_k^(alpha, 0)(t), k = 0..n (three-term recurrence)
template <typename T>
void jacobiSeries(int n, double alpha, T t, T* P) {
  P[0] = T(1.0);
  if (n == 0) return;
  P[1] = (T(alpha + 2.0) * t + T(alpha)) / T(2.0);
  for (int k = 2; k <= n; k++) {
    const double a = k + alpha;  // beta = 0
    const double c1 = 2.0 * k * a * (2.0 * k + alpha - 2.0);
    const double c2 = (2.0 * k + alpha - 1.0) * (2.0 * k + alpha) * (2.0 * k + alpha - 2.0);
    const double c3 = (2.0 * k + alpha - 1.0) * alpha * alpha;
    const double c4 = 2.0 * (k + alpha - 1.0) * (k - 1.0) * (2.0 * k + alpha);
    P[k] = ((T(c2) * t + T(c3)) * P[k - 1] - T(c4) * P[k - 2]) / T(c1);
  }
}

// Modal basis of order P on the reference pyramid: the space of gmsh's pyramid elements (Bergot's
// functions, (P+1)(P+2)(2P+3)/6 of them; the reference takes gmsh's Lagrange pyramid basis,
// BasisFunction.cpp:62-67), phi_ijk = L_i(x/(1-z)) L_j(y/(1-z)) (1-z)^m P_k^(2m+2,0)(2z-1), m = max(i, j),
// 0 <= k <= P - m; ordered by m + k, so the order P-1 basis is a prefix. Rational in z; the rules
// never reach the apex.
template <typename T>
void pyramidBasis(int order, T x, T y, T z, T* value) {
  const T s = T(1.0) - z;
  const T xt = x / s, yt = y / s;
  T Lx[8], dLx[8], Ly[8], dLy[8], J[8];
  legendreSeries(order, xt, Lx, dLx);
  legendreSeries(order, yt, Ly, dLy);
  int b = 0;
  for (int d = 0; d <= order; d++)
    for (int m = 0; m <= d; m++) {
      const int k = d - m;
      jacobiSeries(k, 2.0 * m + 2.0, T(2.0) * z - T(1.0), J);
      T sm = T(1.0);
      for (int i = 0; i < m; i++) sm = sm * s;
      const T radial = sm * J[k];
      for (int j = 0; j <= m; j++) value[b++] = Lx[m] * Ly[j] * radial;
      for (int i = 0; i < m; i++) value[b++] = Lx[i] * Ly[m] * radial;
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
  if (ref.type == 5) {  // pyramid: values, and gradients by complex-step differentiation
    pyramidBasis<double>(ref.order, xi[0], xi[1], xi[2], value);
    if (grad) {
      const double h = 1e-30;
      std::vector<std::complex<double>> cv(ref.nb);
      for (int d = 0; d < 3; d++) {
        const std::complex<double> x(xi[0], d == 0 ? h : 0.0), y(xi[1], d == 1 ? h : 0.0), z(xi[2], d == 2 ? h : 0.0);
        pyramidBasis<std::complex<double>>(ref.order, x, y, z, cv.data());
        for (int b = 0; b < ref.nb; b++) grad[b * 3 + d] = cv[b].imag() / h;
      }
    }
    return;
  }
  if (ref.type == 4) {  // tetrahedron: values, and gradients by complex-step differentiation
    tetBasis<double>(ref.order, xi[0], xi[1], xi[2], value);
    if (grad) {
      const double h = 1e-30;
      std::vector<std::complex<double>> cv(ref.nb);
      for (int d = 0; d < 3; d++) {
        const std::complex<double> x(xi[0], d == 0 ? h : 0.0), y(xi[1], d == 1 ? h : 0.0), z(xi[2], d == 2 ? h : 0.0);
        tetBasis<std::complex<double>>(ref.order, x, y, z, cv.data());
        for (int b = 0; b < ref.nb; b++) grad[b * 3 + d] = cv[b].imag() / h;
      }
    }
    return;
  }
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
  if (ref.type == 5) {  // the rational P1 pyramid functions: base (1 +- x)(1 +- y)/4 - z/4 + +-x y z/(4(1 - z)), apex z
    const double x = xi[0], y = xi[1], z = xi[2], r = z < 1.0 ? x * y * z / (1.0 - z) : 0.0;
    for (int v = 0; v < 4; v++) {
      const double a = kPyrVertex[v][0], b = kPyrVertex[v][1];
      out[v] = 0.25 * ((1.0 + a * x) * (1.0 + b * y) - z + a * b * r);
    }
    out[4] = z;
    return;
  }
  if (ref.type == 4) {
    out[0] = 1.0 - xi[0] - xi[1] - xi[2];
    out[1] = xi[0];
    out[2] = xi[1];
    out[3] = xi[2];
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
  } else if (type == 4) {  // tetrahedron
    if (order > 4) {
      error = "tetrahedra are implemented for P = 1..4";
      return false;
    }
    ref.dim = 3;
    ref.nf = 4;
    ref.nvert = 4;
    ref.nfvert = 3;
    for (int v = 0; v < 4; v++)
      for (int d = 0; d < 3; d++) ref.vertex.push_back(kTetVertex[v][d]);
    for (int f = 0; f < 4; f++)
      for (int v = 0; v < 3; v++) ref.face_vertex.push_back(kTetFace[f][v]);
    grundmannMoeller(3, order, ref.qpoint, ref.qweight);
    triangleFaceRule(order, ref.fpoint, ref.fweight);
    ref.modes.assign(static_cast<size_t>((order + 1) * (order + 2) * (order + 3) / 6), {0, 0, 0});
  } else if (type == 5) {  // pyramid
    if (order > 4) {
      error = "pyramids are implemented for P = 1..4";
      return false;
    }
    ref.dim = 3;
    ref.nf = 5;
    ref.nvert = 5;
    ref.nfvert = 4;  // the largest face: the quadrangle base
    for (int v = 0; v < 5; v++)
      for (int d = 0; d < 3; d++) ref.vertex.push_back(kPyrVertex[v][d]);
    for (int f = 0; f < 5; f++)
      for (int v = 0; v < 4; v++) ref.face_vertex.push_back(kPyrFace[f][v]);
    // volume rule of order 2P: collapsed Gauss -- x = a (1 - z), y = b (1 - z) with Gauss-Legendre
    // in a, b and Gauss-Jacobi (weight (1 - z)^2) in z, (P+1)^3 points (the reference's counts,
    // SimulationControl.cpp:265); exact for the mass matrix of the space above
    std::vector<double> jx, jw;
    gaussJacobi(n1, 2.0, jx, jw);
    for (int c = 0; c < n1; c++)
      for (int b = 0; b < n1; b++)
        for (int a = 0; a < n1; a++) {
          const double z = 0.5 * (1.0 + jx[c]);
          ref.qpoint.push_back(gx[a] * (1.0 - z));
          ref.qpoint.push_back(gx[b] * (1.0 - z));
          ref.qpoint.push_back(z);
          ref.qweight.push_back(gw[a] * gw[b] * jw[c] / 8.0);
        }
    // the face rule of the record (exports): the base's Gauss rule, second coordinate fastest
    for (int a = 0; a < n1; a++)
      for (int b = 0; b < n1; b++) {
        ref.fpoint.push_back(gx[a]);
        ref.fpoint.push_back(gx[b]);
        ref.fweight.push_back(gw[a] * gw[b]);
      }
    ref.modes.assign(static_cast<size_t>(basisSize(5, 3, order)), {0, 0, 0});
  } else {
    error = "unsupported element type (line=1, triangle=2, quadrangle=3, tetrahedron=4, pyramid=5 and hexahedron=6 "
            "are implemented)";
    return false;
  }
  // hierarchical ordering: sort modes by (max index, then slowest-to-fastest index) so that the
  // first N_b(P-1) functions span the order P-1 space
  if (type != 2 && type != 4 && type != 5)
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
  // per face: vertex count, slot points and weights (pyramids: the triangle rule padded with
  // zero-weight copies of its first point)
  const int fd = dim > 1 ? dim - 1 : 0;
  ref.face_nv.assign(nf, ref.nfvert);
  ref.face_nreal.assign(nf, nfq);
  ref.fslot_point.assign(static_cast<size_t>(nf) * nfq * fd, 0.0);
  ref.fslot_weight.assign(static_cast<size_t>(nf) * nfq, 0.0);
  std::vector<double> tp, tw;
  if (type == 5) triangleFaceRule(order, tp, tw);
  for (int f = 0; f < nf; f++) {
    const bool tri5 = type == 5 && f < 4;
    if (tri5) {
      ref.face_nv[f] = 3;
      ref.face_nreal[f] = static_cast<int>(tw.size());
    }
    for (int j = 0; j < nfq; j++) {
      const int src = tri5 ? (j < static_cast<int>(tw.size()) ? j : 0) : j;
      for (int d = 0; d < fd; d++)
        ref.fslot_point[(static_cast<size_t>(f) * nfq + j) * fd + d] = tri5 ? tp[static_cast<size_t>(src) * 2 + d] : ref.fpoint[static_cast<size_t>(j) * fd + d];
      ref.fslot_weight[static_cast<size_t>(f) * nfq + j] = tri5 ? (j < static_cast<int>(tw.size()) ? tw[src] : 0.0) : ref.fweight[j];
    }
  }
  // face points in parent coordinates: P1 map of the face rule through the face vertices
  ref.face_param.assign(static_cast<size_t>(nf) * nfq * dim, 0.0);
  ref.Vf.assign(static_cast<size_t>(nf) * nfq * nb, 0.0);
  for (int f = 0; f < nf; f++) {
    for (int j = 0; j < nfq; j++) {
      double xi[3] = {0, 0, 0};
      const double* fp = fd ? &ref.fslot_point[(static_cast<size_t>(f) * nfq + j) * fd] : nullptr;
      for (int v = 0; v < ref.face_nv[f]; v++) {
        double n = 1.0, g[2];
        if (dim == 1) {
          n = 1.0;  // the face is the vertex itself
        } else if (dim == 2) {
          const double fv[1] = {kLineFaceVertex[v]};
          p1Shape(1, fv, fp, n, g);
        } else if (ref.face_nv[f] == 3) {  // triangle face: barycentric P1 functions
          n = v == 0 ? 1.0 - fp[0] - fp[1] : fp[v - 1];
        } else {
          p1Shape(2, kQuadFaceVertex[v], fp, n, g);
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
  if (ref.type == 5) {  // pyramid with a parallelogram base: x = c + a xi + b eta + d zeta (affine)
    const double* v = xv_all;
    for (int l = 0; l < 3; l++) {
      const double c = 0.5 * (v[l] + v[2 * 3 + l]);
      const double a = 0.5 * (v[3 + l] - v[l]), b = 0.5 * (v[3 * 3 + l] - v[l]), d = v[4 * 3 + l] - c;
      x[l] = c + a * xi[0] + b * xi[1] + d * xi[2];
      if (jac) {
        jac[l * 3 + 0] = a;
        jac[l * 3 + 1] = b;
        jac[l * 3 + 2] = d;
      }
    }
    return;
  }
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
    } else if (ref.type == 4) {
      n = v == 0 ? 1.0 - xi[0] - xi[1] - xi[2] : xi[v - 1];
      for (int d = 0; d < 3; d++) g[d] = v == 0 ? -1.0 : (d == v - 1 ? 1.0 : 0.0);
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
  if (mesh.gnode.empty()) {
    mapPointRef(mesh.ref, &mesh.vertex[static_cast<size_t>(e) * mesh.ref.nvert * mesh.ref.dim], xi, x, jac);
    return;
  }
  // curved element: x = sum_k N_k(xi) X_k, J(l, k) = sum_n X_n,l dN_n / dxi_k
  const GeomMap& gm = mesh.geom;
  const int dim = mesh.ref.dim, n = gm.n;
  double N[64], dN[64 * 3];
  geomShape(gm, xi, N, dN);
  const double* X = &mesh.gnode[static_cast<size_t>(e) * n * dim];
  for (int l = 0; l < dim; l++) {
    double acc = 0.0;
    for (int k = 0; k < n; k++) acc += N[k] * X[k * dim + l];
    x[l] = acc;
    if (jac)
      for (int kk = 0; kk < dim; kk++) {
        double j = 0.0;
        for (int k = 0; k < n; k++) j += X[k * dim + l] * dN[k * dim + kk];
        jac[l * dim + kk] = j;
      }
  }
}
// normals (outward from element e through its local face s) and |J_f| w at the face points, from
// the element's affine face map (Geometry.cpp:69-86,102-168)
void faceNormalsRef(const RefElement& ref, const double* Jall, int jstride, int s, Face& face) {
  const int dim = ref.dim, nfq = ref.nfq;
  face.normal.assign(static_cast<size_t>(nfq) * dim, 0.0);
  face.jw.assign(nfq, 0.0);
  if (dim == 1) {  // end points: outward -1 at the first vertex, +1 at the second, weight 1
    face.normal[0] = s == 0 ? -1.0 : 1.0;
    face.jw[0] = ref.fweight[0];
    return;
  }
  const int fd = dim - 1;
  for (int p = 0; p < nfq; p++) {
    const double* J = Jall + static_cast<size_t>(p) * jstride;
    const double* fp = &ref.fslot_point[(static_cast<size_t>(s) * nfq + p) * fd];
    double ts[3] = {0, 0, 0}, tt[3] = {0, 0, 0};  // d xi / d s, d xi / d t
    for (int v = 0; v < ref.face_nv[s]; v++) {
      double nval, g[2];
      if (dim == 2) {
        const double fv[1] = {kLineFaceVertex[v]};
        p1Shape(1, fv, fp, nval, g);
      } else if (ref.face_nv[s] == 3) {  // triangle face: gradients of the barycentric P1 functions
        g[0] = v == 0 ? -1.0 : (v == 1 ? 1.0 : 0.0);
        g[1] = v == 0 ? -1.0 : (v == 2 ? 1.0 : 0.0);
      } else {
        p1Shape(2, kQuadFaceVertex[v], fp, nval, g);
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
    face.jw[p] = norm * ref.fslot_weight[static_cast<size_t>(s) * nfq + p];
  }
}
void faceNormals(const Mesh& mesh, int64_t e, int s, Face& face) {
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim;
  if (mesh.gnode.empty() || dim == 1) {
    faceNormalsRef(ref, &mesh.jac[static_cast<size_t>(e) * dim * dim], 0, s, face);
    return;
  }
  // curved element: the Jacobian of the element map at each face point
  std::vector<double> J(static_cast<size_t>(ref.nfq) * dim * dim);
  double x[3];
  for (int p = 0; p < ref.nfq; p++)
    mapPoint(mesh, e, &ref.face_param[(static_cast<size_t>(s) * ref.nfq + p) * dim], x, &J[static_cast<size_t>(p) * dim * dim]);
  faceNormalsRef(ref, J.data(), dim * dim, s, face);
}
}  // namespace

namespace {
// lattice points of a triangle of order g in gmsh's order: the vertices (0,0), (g,0), (0,g), the
// interior points of the edges (0,1), (1,2), (2,0) in the edge's direction, then the interior as a
// triangle of order g - 3 shifted by (1, 1)
void triangleLattice(int g, std::vector<std::array<int, 3>>& out) {
  if (g < 0) return;
  if (g == 0) {
    out.push_back({0, 0, 0});
    return;
  }
  out.push_back({0, 0, 0});
  out.push_back({g, 0, 0});
  out.push_back({0, g, 0});
  for (int k = 1; k < g; k++) out.push_back({k, 0, 0});
  for (int k = 1; k < g; k++) out.push_back({g - k, k, 0});
  for (int k = 1; k < g; k++) out.push_back({0, g - k, 0});
  std::vector<std::array<int, 3>> inner;
  triangleLattice(g - 3, inner);
  for (const auto& q : inner) out.push_back({q[0] + 1, q[1] + 1, 0});
}
// the same for a quadrangle: vertices (0,0), (g,0), (g,g), (0,g), edges (0,1), (1,2), (2,3), (3,0),
// then the interior as a quadrangle of order g - 2 shifted by (1, 1)
void quadrangleLattice(int g, std::vector<std::array<int, 3>>& out) {
  if (g < 0) return;
  if (g == 0) {
    out.push_back({0, 0, 0});
    return;
  }
  out.push_back({0, 0, 0});
  out.push_back({g, 0, 0});
  out.push_back({g, g, 0});
  out.push_back({0, g, 0});
  for (int k = 1; k < g; k++) out.push_back({k, 0, 0});
  for (int k = 1; k < g; k++) out.push_back({g, k, 0});
  for (int k = 1; k < g; k++) out.push_back({g - k, g, 0});
  for (int k = 1; k < g; k++) out.push_back({0, g - k, 0});
  std::vector<std::array<int, 3>> inner;
  quadrangleLattice(g - 2, inner);
  for (const auto& q : inner) out.push_back({q[0] + 1, q[1] + 1, 0});
}
}  // namespace

bool buildGeomMap(int type, int order, GeomMap& g, std::string& error) {
  g = GeomMap();
  g.type = type;
  g.order = order;
  if (order < 1) {
    error = "geometry order must be >= 1";
    return false;
  }
  std::vector<double>& x = g.node;
  if (type == 2 || type == 3) {  // lattice nodes
    g.dim = 2;
    if (order > 5) {
      error = "triangles and quadrangles of geometry order 1..5 are read";
      return false;
    }
    std::vector<std::array<int, 3>> lat;
    if (type == 2) triangleLattice(order, lat);
    else quadrangleLattice(order, lat);
    for (const auto& q : lat)
      for (int d = 0; d < 2; d++) x.push_back(type == 2 ? double(q[d]) / order : -1.0 + 2.0 * q[d] / order);
  } else if (type == 6 || type == 4) {
    g.dim = 3;
    if (order > 2) {
      error = "hexahedra and tetrahedra of geometry order 1 or 2 are read";
      return false;
    }
    const int nv = type == 6 ? 8 : 4;
    for (int v = 0; v < nv; v++)
      for (int d = 0; d < 3; d++) x.push_back(type == 6 ? kHexVertex[v][d] : kTetVertex[v][d]);
    if (order == 2) {
      auto mid = [&](const std::vector<int>& vs) {
        for (int d = 0; d < 3; d++) {
          double c = 0.0;
          for (int v : vs) c += type == 6 ? kHexVertex[v][d] : kTetVertex[v][d];
          x.push_back(c / vs.size());
        }
      };
      if (type == 6) {  // 27-node hexahedron: 12 edge, 6 face and the centre nodes
        static const int kE[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                      {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
        for (const auto& e : kE) mid({e[0], e[1]});
        for (const auto& f : kHexFace) mid({f[0], f[1], f[2], f[3]});
        mid({0, 1, 2, 3, 4, 5, 6, 7});
      } else {  // 10-node tetrahedron: edge midpoints
        static const int kE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
        for (const auto& e : kE) mid({e[0], e[1]});
      }
    }
  } else if (type == 5) {  // pyramids: first order, affine (parallelogram base) only
    g.dim = 3;
    if (order != 1) {
      error = "pyramids of geometry order 1 are read";
      return false;
    }
    for (int v = 0; v < 5; v++)
      for (int d = 0; d < 3; d++) x.push_back(kPyrVertex[v][d]);
    g.n = 5;
    return true;  // no interpolation space: the element map is the affine one of mapPointRef
  } else {
    error = "curved elements are read for triangles, quadrangles, tetrahedra and hexahedra";
    return false;
  }
  const int dim = g.dim;
  g.n = static_cast<int>(x.size()) / dim;
  // monomials: total degree <= order on simplices, degree <= order per coordinate otherwise
  const bool simplex = type == 2 || type == 4;
  for (int c = 0; c <= (dim == 3 ? order : 0); c++)
    for (int b = 0; b <= order; b++)
      for (int a = 0; a <= order; a++)
        if (!simplex || a + b + c <= order) g.expo.push_back({a, b, c});
  if (static_cast<int>(g.expo.size()) != g.n) {
    error = "geometry node count does not match the interpolation space";
    return false;
  }
  const int n = g.n;
  std::vector<double> A(static_cast<size_t>(n) * n);
  for (int i = 0; i < n; i++)
    for (int m = 0; m < n; m++) {
      double v = 1.0;
      for (int d = 0; d < dim; d++) v *= std::pow(x[static_cast<size_t>(i) * dim + d], g.expo[m][d]);
      A[static_cast<size_t>(i) * n + m] = v;
    }
  g.ainv.assign(static_cast<size_t>(n) * n, 0.0);
  if (!invertMatrix(n, A.data(), g.ainv.data())) {
    error = "singular geometry interpolation";
    return false;
  }
  return true;
}

void geomShape(const GeomMap& g, const double* xi, double* N, double* dN) {
  const int n = g.n, dim = g.dim;
  double mono[64], dmono[64 * 3];
  for (int m = 0; m < n; m++) {
    double v = 1.0;
    for (int d = 0; d < dim; d++) v *= std::pow(xi[d], g.expo[m][d]);
    mono[m] = v;
    for (int k = 0; k < dim; k++) {
      const int ek = g.expo[m][k];
      double dv = ek == 0 ? 0.0 : ek * std::pow(xi[k], ek - 1);
      for (int d = 0; d < dim; d++)
        if (d != k) dv *= std::pow(xi[d], g.expo[m][d]);
      dmono[m * dim + k] = dv;
    }
  }
  for (int k = 0; k < n; k++) {
    double v = 0.0, dv[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < n; m++) {
      const double a = g.ainv[static_cast<size_t>(m) * n + k];
      v += a * mono[m];
      for (int d = 0; d < dim; d++) dv[d] += a * dmono[m * dim + d];
    }
    N[k] = v;
    if (dN)
      for (int d = 0; d < dim; d++) dN[k * dim + d] = dv[d];
  }
}

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
                               const std::vector<std::map<int64_t, int64_t>>* periodic_links, int geometry_order) {
  mesh = Mesh();
  if (type != 2 && type != 3 && type != 4 && type != 5 && type != 6) {
    error = "element type must be triangle, quadrangle, tetrahedron, pyramid or hexahedron";
    return false;
  }
  if (!buildRefElement(type, order, mesh.ref, error)) return false;
  GeomMap gm;
  if (!buildGeomMap(type, geometry_order, gm, error)) return false;
  const RefElement& ref = mesh.ref;
  const int dim = ref.dim, nv = ref.nvert, nf = ref.nf, nfq = ref.nfq, nfv = ref.nfvert, ngn = gm.n;
  if (elements < 1 || elements > (int64_t(1) << 30)) {
    error = "element count out of range";
    return false;
  }
  const int64_t E = elements;
  mesh.elements = E;
  for (int64_t i = 0; i < E * ngn; i++)
    if (conn[i] < 0 || conn[i] >= nnode) {
      error = "element node index out of range";
      return false;
    }
  std::vector<int64_t> cn(conn, conn + E * ngn);  // element nodes (mirrored elements are reordered)
  std::vector<int64_t> cv(static_cast<size_t>(E) * nv);  // vertices: the first nodes of each element
  for (int64_t e = 0; e < E; e++)
    for (int v = 0; v < nv; v++) cv[static_cast<size_t>(e) * nv + v] = cn[static_cast<size_t>(e) * ngn + v];
  // mirror of an inverted element: the map with the first two reference coordinates swapped covers
  // the same cell with the opposite orientation; node k of the mirrored list is the old node at the
  // swapped position of node k (on the vertices: (0,2,1), (0,3,2,1), (0,2,1,3), (0,3,2,1,4,7,6,5))
  std::vector<int> gmirror(ngn, -1);
  for (int k = 0; k < ngn; k++) {
    double sw[3];
    for (int d = 0; d < dim; d++) sw[d] = gm.node[static_cast<size_t>(k) * dim + d];
    std::swap(sw[0], sw[1]);
    for (int j = 0; j < ngn; j++) {
      double d2 = 0.0;
      for (int d = 0; d < dim; d++) d2 += std::pow(gm.node[static_cast<size_t>(j) * dim + d] - sw[d], 2);
      if (d2 < 1e-20) gmirror[k] = j;
    }
  }
  mesh.vertex.assign(static_cast<size_t>(E) * nv * dim, 0.0);
  mesh.jac.assign(static_cast<size_t>(E) * dim * dim, 0.0);
  mesh.detj.assign(E, 0.0);
  mesh.min_edge.assign(E, 0.0);
  const double c3 = type == 2 ? 1.0 / 3.0 : (type == 4 ? 0.25 : 0.0);
  const double xi_c[3] = {c3, c3, type == 4 || type == 5 ? 0.25 : 0.0};
  const int(*edges)[2] = type == 2 ? kTriEdge
                         : type == 3 ? kQuadEdge
                         : type == 4 ? kTetEdge
                         : type == 5 ? kPyrEdge
                                     : kHexEdge12;
  const int nedge = type == 2 ? 3 : (type == 3 ? 4 : (type == 4 ? 6 : (type == 5 ? 8 : 12)));
  double xmax = 0.0;
  bool curved = false;
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
      std::vector<int64_t> tmp(cn.begin() + e * ngn, cn.begin() + (e + 1) * ngn);
      for (int k = 0; k < ngn; k++) cn[static_cast<size_t>(e) * ngn + k] = tmp[static_cast<size_t>(gmirror[k])];
      for (int v = 0; v < nv; v++) cv[static_cast<size_t>(e) * nv + v] = cn[static_cast<size_t>(e) * ngn + v];
    }
    double xc[3];
    mapPoint(mesh, e, xi_c, xc, nullptr);
    double me = std::numeric_limits<double>::max(), span = 0.0;
    for (int g = 0; g < nedge; g++) {
      const int a = edges[g][0], b = edges[g][1];
      double s2 = 0.0;
      for (int d = 0; d < dim; d++) {
        const double t = mesh.vertex[(static_cast<size_t>(e) * nv + a) * dim + d] - mesh.vertex[(static_cast<size_t>(e) * nv + b) * dim + d];
        s2 += t * t;
      }
      me = std::min(me, std::sqrt(s2));
      span = std::max(span, std::sqrt(s2));
    }
    mesh.min_edge[e] = me;
    // affine: the map through the centre and its Jacobian reproduces every geometry node
    const double* J = &mesh.jac[static_cast<size_t>(e) * dim * dim];
    for (int k = 0; k < ngn && !curved; k++) {
      const int64_t node = cn[static_cast<size_t>(e) * ngn + k];
      double err2 = 0.0;
      for (int l = 0; l < dim; l++) {
        double x = xc[l];
        for (int kk = 0; kk < dim; kk++) x += J[l * dim + kk] * (gm.node[static_cast<size_t>(k) * dim + kk] - xi_c[kk]);
        const double t = x - xyz[node * 3 + l];
        err2 += t * t;
        xmax = std::max(xmax, std::fabs(xyz[node * 3 + l]));
      }
      if (err2 > 1e-20 * span * span) curved = true;
    }
  }
  if (curved && type == 5) {
    error = "a pyramid is not affine (its base must be a parallelogram)";
    return false;
  }
  if (curved) {
    // every element through its geometry nodes (the reference's per-point Jacobians,
    // Geometry.cpp:44-67, and per-element mass matrices, Geometry.cpp:88-100)
    mesh.affine = false;
    mesh.geom = gm;
    mesh.gnode.assign(static_cast<size_t>(E) * ngn * dim, 0.0);
    for (int64_t e = 0; e < E; e++)
      for (int k = 0; k < ngn; k++)
        for (int d = 0; d < dim; d++) mesh.gnode[(static_cast<size_t>(e) * ngn + k) * dim + d] = xyz[cn[static_cast<size_t>(e) * ngn + k] * 3 + d];
    std::vector<double> jq(static_cast<size_t>(dim) * dim);
    for (int64_t e = 0; e < E; e++) {
      double x[3];
      mapPoint(mesh, e, xi_c, x, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
      mesh.detj[e] = det(dim, &mesh.jac[static_cast<size_t>(e) * dim * dim]);
      for (int q = 0; q < ref.nq; q++) {
        mapPoint(mesh, e, &ref.qpoint[static_cast<size_t>(q) * dim], x, jq.data());
        if (!(det(dim, jq.data()) > 0.0)) {
          error = "a curved element has a non-positive Jacobian at a quadrature point";
          return false;
        }
      }
    }
  }
  // faces keyed by their sorted node indices
  // (a face of fewer vertices than nfv -- a pyramid's triangle -- is padded with -1)
  auto keyOf = [&](const int64_t* nodes) {
    std::array<int64_t, 4> k = {-1, -1, -1, -1};
    for (int v = 0; v < nfv; v++) k[v] = nodes[v] < 0 ? -1 : nodes[v];
    std::sort(k.begin(), k.begin() + nfv);
    return k;
  };
  auto faceNodes = [&](int64_t e, int s, int64_t* fn) {
    for (int v = 0; v < 4; v++) fn[v] = -1;
    for (int v = 0; v < ref.face_nv[s]; v++) fn[v] = cv[static_cast<size_t>(e) * nv + ref.face_vertex[s * nfv + v]];
  };
  std::map<std::array<int64_t, 4>, std::vector<std::pair<int64_t, int>>> visits;
  for (int64_t e = 0; e < E; e++)
    for (int s = 0; s < nf; s++) {
      int64_t fn[4];
      faceNodes(e, s, fn);
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
        for (int v = 0; v < 4; v++) mapped[v] = -1;
        for (int v = 0; v < ref.face_nv[s] && slave; v++) {
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
      faceNodes(e, s, fn);
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
      const int fnv = ref.face_nv[s], nreal = ref.face_nreal[s];
      for (int v = 0; v < fnv; v++)
        for (int d = 0; d < dim; d++)
          shift[d] += (xyz[fn[v] * 3 + d] - xyz[cv[static_cast<size_t>(r) * nv + ref.face_vertex[rs * nfv + v]] * 3 + d]) / fnv;
      face.seq.assign(nfq, -1);
      for (int p = nreal; p < nfq; p++) face.seq[p] = p;  // zero-weight padding slots
      for (int p = 0; p < nreal; p++) {
        double best = std::numeric_limits<double>::max();
        int bi = -1;
        for (int qq = 0; qq < nreal; qq++) {
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
        for (int v = 0; v < ref.face_nv[rs]; v++)
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
      faceNormalsRef(ref, &mesh.jac[static_cast<size_t>(e) * 4], 0, s, face);
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
  // inner radius D |element| / |boundary|; reference volumes: triangle 1/2, quadrangle 4,
  // tetrahedron 1/6, pyramid 4/3, hexahedron 8. Curved elements: the volume and the face areas by
  // quadrature.
  const double measure = ref.type == 2   ? 0.5
                         : ref.type == 4 ? 1.0 / 6.0
                         : ref.type == 5 ? 4.0 / 3.0
                                         : (dim == 1 ? 2.0 : (dim == 2 ? 4.0 : 8.0));
  mesh.inner_radius.assign(E, 0.0);
  for (int64_t e = 0; e < E; e++) {
    const double* x = &mesh.vertex[static_cast<size_t>(e) * nv * dim];
    if (!mesh.gnode.empty()) {
      std::vector<double> jw(static_cast<size_t>(ref.nq));
      elementGeometry(mesh, e, jw.data(), nullptr, nullptr);
      double volume = 0.0, surface = 0.0;
      for (double v : jw) volume += v;
      for (int f = 0; f < ref.nf; f++) {
        const int32_t id = mesh.elem_face[static_cast<size_t>(e) * ref.nf + f];
        const Face& face = id >= 0 ? mesh.faces[id] : mesh.bfaces[-2 - id];
        for (double v : face.jw) surface += v;
      }
      mesh.inner_radius[e] = dim * volume / surface;
      continue;
    }
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
      } else {  // planar parallelogram face a, b, c, d: |(b - a) x (d - a)|; triangle a, b, c: half |(b - a) x (c - a)|
        const bool tri = ref.face_nv[f] == 3;
        double u[3], w[3];
        for (int d = 0; d < 3; d++) {
          u[d] = x[fv[1] * 3 + d] - x[fv[0] * 3 + d];
          w[d] = x[fv[tri ? 2 : 3] * 3 + d] - x[fv[0] * 3 + d];
        }
        const double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        surface += (tri ? 0.5 : 1.0) * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
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
  if (!mesh.gnode.empty()) {  // curved: the Jacobian at each point (Geometry.cpp:44-67)
    double x[3], J[9], jt[9], jti[9];
    for (int q = 0; q < ref.nq; q++) {
      mapPoint(mesh, e, &ref.qpoint[static_cast<size_t>(q) * dim], x, J);
      const double jw = det(dim, J) * ref.qweight[q];
      if (detjw) detjw[q] = jw;
      if (tiw) {
        for (int k = 0; k < dim; k++)
          for (int l = 0; l < dim; l++) jt[k * dim + l] = J[l * dim + k];
        invertMatrix(dim, jt, jti);
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++) tiw[static_cast<size_t>(q) * dim * dim + a + dim * b] = jti[a * dim + b] * jw;
      }
      if (coord)
        for (int d = 0; d < dim; d++) coord[static_cast<size_t>(q) * dim + d] = x[d];
    }
    return;
  }
  affineGeometry(ref, &mesh.jac[static_cast<size_t>(e) * dim * dim], mesh.detj[e], detjw, tiw);
  if (coord)
    for (int q = 0; q < ref.nq; q++)
      mapPoint(mesh, e, &ref.qpoint[static_cast<size_t>(q) * dim], &coord[static_cast<size_t>(q) * dim], nullptr);
}

void elementMassInverse(const Mesh& mesh, int64_t e, double* minv) {
  if (mesh.gnode.empty()) {
    affineMassInverse(mesh.ref, mesh.detj[e], minv);
    return;
  }
  // curved: (sum_q V_qi V_qj |J_q| w_q)^-1 (Geometry.cpp:88-100)
  const RefElement& ref = mesh.ref;
  const int nb = ref.nb;
  std::vector<double> jw(static_cast<size_t>(ref.nq)), m(static_cast<size_t>(nb) * nb, 0.0);
  elementGeometry(mesh, e, jw.data(), nullptr, nullptr);
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < nb; j++) {
      double acc = 0.0;
      for (int q = 0; q < ref.nq; q++)
        acc += ref.V[static_cast<size_t>(q) * nb + i] * (ref.V[static_cast<size_t>(q) * nb + j] * jw[q]);
      m[static_cast<size_t>(i) * nb + j] = acc;
    }
  invertMatrix(nb, m.data(), minv);
}

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
