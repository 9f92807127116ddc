// Host-side setup for the self_amd engine: Legendre quadrature, Lagrange matrices and
// mesh connectivity.  These run once per model on the CPU (as in the reference) and are
// the product's own implementation; device kernels consume their output.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <vector>

#include "selfamd_common.hpp"
#include "../../include/self_amd.h"

namespace selfamd {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char *last_error() { return g_err; }

namespace {

constexpr double kEps = 2.220446049250313080847e-16;  // epsilon(1.0_real64)
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;

// Legendre polynomial L_n(x) and derivative by the three-term recurrence
// (src/SELF_Quadrature.f90:342-383).
void legendre(int n, double x, double &l, double &dl) {
  if (n == 0) { l = 1.0; dl = 0.0; return; }
  if (n == 1) { l = x; dl = 1.0; return; }
  double l2 = 1.0, l1 = x, d2 = 0.0, d1 = 1.0;
  for (int k = 2; k <= n; k++) {
    l = ((2.0 * k - 1.0) * x * l1 - (k - 1.0) * l2) / (double)k;
    dl = d2 + (2.0 * k - 1.0) * l1;
    l2 = l1; l1 = l; d2 = d1; d1 = dl;
  }
}

// q = L_{n+1} - L_{n-1}, its derivative, and L_n (src/SELF_Quadrature.f90:385-418).
void legendre_q(int n, double x, double &q, double &qp, double &ln) {
  double l2 = 1.0, l1 = x, d2 = 0.0, d1 = 1.0, l = 0, dl = 0;
  for (int k = 2; k <= n; k++) {
    l = ((2.0 * k - 1.0) * x * l1 - (k - 1.0) * l2) / (double)k;
    dl = d2 + (2.0 * k - 1.0) * l1;
    l2 = l1; l1 = l; d2 = d1; d1 = dl;
  }
  const double lp = ((2.0 * (n + 1) - 1.0) * x * l - (double)n * l2) / (double)(n + 1);
  const double dlp = d2 + (2.0 * (n + 1) - 1.0) * l1;
  q = lp - l2;
  qp = dlp - d2;
  ln = l;
}

// Gauss nodes by Newton iteration on L_{N+1} (src/SELF_Quadrature.f90:201-262).  The
// reference seeds node j with -cos((2j+1)pi/(2N+1)), iterates j = 0..(N+1)/2 and mirrors
// each result into N-j; its exit test |delta| <= eps*x only fires for x > 0.  The same
// sweep is kept here because it fixes the node ORDER: at N = 3 the j = 2 Newton run jumps
// across the origin, so the reference's nodes are (-0.861, +0.340, -0.340, +0.861).
void gauss(int N, double *x, double *w) {
  if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
  if (N == 1) {
    x[0] = -std::sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = w[0];
    return;
  }
  double l, dl;
  for (int j = 0; j <= (N + 1) / 2; j++) {
    double xj = -std::cos((2.0 * j + 1.0) * kPi / (2.0 * N + 1.0));
    for (int it = 0; it < 500; it++) {
      legendre(N + 1, xj, l, dl);
      const double delta = -l / dl;
      xj += delta;
      if (std::fabs(delta) <= kEps * xj) break;
    }
    legendre(N + 1, xj, l, dl);
    w[j] = 2.0 / ((1.0 - xj * xj) * dl * dl);
    x[j] = xj;
    w[N - j] = w[j];
    x[N - j] = -xj;
  }
  if (N % 2 == 0) {
    legendre(N + 1, 0.0, l, dl);
    x[N / 2] = 0.0;
    w[N / 2] = 2.0 / (dl * dl);
  }
}

// Gauss-Lobatto nodes (src/SELF_Quadrature.f90:270-334).
void gauss_lobatto(int N, double *x, double *w) {
  x[0] = -1.0; x[N] = 1.0;
  if (N == 1) { w[0] = 1.0; w[1] = 1.0; return; }
  w[0] = 2.0 / ((double)N * (N + 1.0));
  w[N] = w[0];
  double q, qp, ln;
  for (int j = 1; j <= (N + 1) / 2 - 1; j++) {
    double xj = -std::cos((j + 0.25) * kPi / N - 3.0 / (8.0 * N * kPi * (j + 0.25)));
    for (int it = 0; it < 500; it++) {
      legendre_q(N, xj, q, qp, ln);
      const double delta = -q / qp;
      xj += delta;
      if (std::fabs(delta) <= kEps * xj) break;
    }
    legendre_q(N, xj, q, qp, ln);
    w[j] = 2.0 / ((double)N * (N + 1.0) * ln * ln);
    x[j] = xj;
    w[N - j] = w[j];
    x[N - j] = -xj;
  }
  if (N % 2 == 0) {
    legendre_q(N, 0.0, q, qp, ln);
    x[N / 2] = 0.0;
    w[N / 2] = 2.0 / ((double)N * (N + 1.0) * ln * ln);
  }
}

bool almost_equal(double a, double b) {  // src/SELF_SupportRoutines.f90:76-96
  if (a == 0.0 || b == 0.0) return std::fabs(a - b) <= kEps;
  return std::fabs(a - b) <= kEps * std::fabs(a) || std::fabs(a - b) <= kEps * std::fabs(b);
}

void points(int n, int type, double *x, double *w) {
  if (type == SELFAMD_GAUSS) {
    gauss(n, x, w);
  } else if (type == SELFAMD_GAUSS_LOBATTO) {
    gauss_lobatto(n, x, w);
  } else {  // uniform (src/SELF_SupportRoutines.f90:502-520)
    const double dx = 2.0 / (double)n;
    for (int i = 0; i <= n; i++) { x[i] = -1.0 + dx * (double)i; w[i] = 2.0 / (double)n; }
  }
}

// Barycentric Lagrange basis at s (src/SELF_Lagrange.f90:2721-2774).
void basis_at(int N, const double *cp, const double *bw, double s, double *out) {
  bool match = false;
  for (int j = 0; j <= N; j++) {
    out[j] = 0.0;
    if (almost_equal(s, cp[j])) { out[j] = 1.0; match = true; }
  }
  if (match) return;
  double sum = 0.0;
  for (int j = 0; j <= N; j++) { out[j] = bw[j] / (s - cp[j]); sum += out[j]; }
  for (int j = 0; j <= N; j++) out[j] /= sum;
}

}  // namespace
}  // namespace selfamd

using namespace selfamd;

extern "C" {

const char *selfamd_last_error(void) { return last_error(); }

int selfamd_version(void) { return SELFAMD_VERSION; }

int selfamd_lagrange_init(int N, int controlNodeType, int M, int targetNodeType,
                          double *controlPoints, double *targetPoints, double *bWeights,
                          double *qWeights, double *iMatrix, double *dMatrix, double *dgMatrix,
                          double *bMatrix) {
  SELFAMD_REQUIRE(N >= 1 && N <= 64 && M >= 1 && M <= 64, "lagrange_init: N=%d M=%d out of range", N, M);
  const int P = N + 1;
  std::vector<double> qt(M + 1);
  points(N, controlNodeType, controlPoints, qWeights);
  points(M, targetNodeType, targetPoints, qt.data());
  // Barycentric weights w_k = 1/prod_{j!=k}(x_k - x_j), accumulated pairwise in the
  // reference's order (src/SELF_Lagrange.f90:2542-2570).
  for (int i = 0; i <= N; i++) bWeights[i] = 1.0;
  for (int j = 1; j <= N; j++)
    for (int i = 0; i < j; i++) {
      bWeights[i] *= controlPoints[i] - controlPoints[j];
      bWeights[j] *= controlPoints[j] - controlPoints[i];
    }
  for (int j = 0; j <= N; j++) bWeights[j] = 1.0 / bWeights[j];
  // Interpolation matrix, stored iMatrix(control, target) (src/SELF_Lagrange.f90:2582-2643).
  for (int r = 0; r <= M; r++) basis_at(N, controlPoints, bWeights, targetPoints[r], iMatrix + (size_t)P * r);
  // Derivative matrices (src/SELF_Lagrange.f90:2656-2709): D(r,c) = l'_c(x_r) with the
  // negative-sum diagonal; stored transposed as dMatrix(c, r).  The weak-form matrix is
  // dgMatrix(r, c) = -D(r, c) * w_r / w_c.
  std::vector<double> D((size_t)P * P);
  for (int r = 0; r <= N; r++) {
    double diag = 0.0;
    for (int c = 0; c <= N; c++) {
      if (c == r) continue;
      const double v = bWeights[c] / (bWeights[r] * (controlPoints[r] - controlPoints[c]));
      D[r + (size_t)P * c] = v;
      diag -= v;
    }
    D[r + (size_t)P * r] = diag;
  }
  for (int r = 0; r <= N; r++)
    for (int c = 0; c <= N; c++) {
      dMatrix[r + (size_t)P * c] = D[c + (size_t)P * r];
      dgMatrix[r + (size_t)P * c] = -D[r + (size_t)P * c] * qWeights[r] / qWeights[c];
    }
  basis_at(N, controlPoints, bWeights, -1.0, bMatrix);
  basis_at(N, controlPoints, bWeights, 1.0, bMatrix + P);
  return kOk;
}

// Structured nx-by-ny quad mesh on [x0,x1]x[y0,y1] with corner nodes (nGeo = 1) in the
// HOPr/CGNS conventions of src/SELF_Mesh.f90:27-91.  Elements are numbered in raster
// order e = ix + nx*iy (stored 1-based).  Periodic directions are encoded as interior
// sides (bcid = 0, neighbour across the wrap) and the wrap corner nodes share global ids,
// so the corner-id flip test of RecalculateFlip yields flip = 0 everywhere.
int selfamd_mesh2d_structured(int nx, int ny, double x0, double x1, double y0, double y1,
                              int periodicX, int periodicY, int boundaryBcid, int32_t *sideInfo,
                              double *nodeCoords, int32_t *globalNodeIDs) {
  SELFAMD_REQUIRE(nx >= 1 && ny >= 1, "mesh2d_structured: nx=%d ny=%d", nx, ny);
  SELFAMD_REQUIRE(!periodicX || nx >= 2, "mesh2d_structured: periodic x needs nx >= 2");
  SELFAMD_REQUIRE(!periodicY || ny >= 2, "mesh2d_structured: periodic y needs ny >= 2");
  const int nnx = periodicX ? nx : nx + 1;  // distinct node columns
  auto node_id = [&](int ix, int iy) {
    if (periodicX) ix %= nx;
    if (periodicY) iy %= ny;
    return 1 + ix + nnx * iy;
  };
  const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
  // Side ids: one per unique face; the lower-numbered element owns the positive id
  // (HOPr's master/slave sign convention).  Horizontal faces first, then vertical.
  const int hx = nx, hy = periodicY ? ny : ny + 1;   // horizontal faces (S/N sides)
  const int vx = periodicX ? nx : nx + 1, vy = ny;   // vertical faces (E/W sides)
  auto hface = [&](int ix, int iy) { return 1 + ix + hx * (periodicY ? iy % ny : iy); };
  auto vface = [&](int ix, int iy) { return 1 + hx * hy + (periodicX ? ix % nx : ix) + vx * iy; };
  (void)vy;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      const int e = ix + nx * iy;
      const int cx[4] = {ix, ix + 1, ix + 1, ix}, cy[4] = {iy, iy, iy + 1, iy + 1};
      for (int c = 0; c < 4; c++) {
        // corner (i,j) positions of the nGeo = 1 element: CN1 (0,0) CN2 (1,0) CN3 (1,1) CN4 (0,1)
        const int gi = (c == 1 || c == 2) ? 1 : 0, gj = (c >= 2) ? 1 : 0;
        const size_t o = gi + 2 * (gj + 2 * (size_t)e);
        nodeCoords[2 * o] = x0 + dx * cx[c];
        nodeCoords[2 * o + 1] = y0 + dy * cy[c];
        globalNodeIDs[o] = node_id(cx[c], cy[c]);
      }
      // neighbours: S, E, N, W; neighbour local side is the opposite side
      int nbr[4], fid[4];
      const bool hasS = iy > 0 || periodicY, hasN = iy < ny - 1 || periodicY;
      const bool hasW = ix > 0 || periodicX, hasE = ix < nx - 1 || periodicX;
      nbr[0] = hasS ? 1 + ix + nx * ((iy + ny - 1) % ny) : 0;
      nbr[1] = hasE ? 1 + (ix + 1) % nx + nx * iy : 0;
      nbr[2] = hasN ? 1 + ix + nx * ((iy + 1) % ny) : 0;
      nbr[3] = hasW ? 1 + (ix + nx - 1) % nx + nx * iy : 0;
      fid[0] = hface(ix, iy);
      fid[2] = hface(ix, iy + 1);
      fid[3] = vface(ix, iy);
      fid[1] = vface(ix + 1, iy);
      const int opp[4] = {3, 4, 1, 2};
      for (int s = 0; s < 4; s++) {
        int32_t *si = sideInfo + 5 * (s + 4 * (size_t)e);
        si[0] = 2;  // side type: linear line segment between two corner nodes
        const bool master = nbr[s] == 0 || (e + 1) <= nbr[s];
        si[1] = master ? fid[s] : -fid[s];
        si[2] = nbr[s];
        si[3] = nbr[s] ? 10 * opp[s] : 0;
        si[4] = nbr[s] ? 0 : boundaryBcid;
      }
    }
  return selfamd_mesh2d_recalculate_flip(nx * ny, 1, sideInfo, globalNodeIDs);
}

// Flip recomputation from corner node ids (src/SELF_Mesh.f90:1018-1174, single rank):
// a side matches its neighbour when the ordered corner-id pairs agree.
int selfamd_mesh2d_recalculate_flip(int nElem, int nGeo, int32_t *sideInfo, const int32_t *globalNodeIDs) {
  SELFAMD_REQUIRE(nElem >= 1 && nGeo >= 1, "recalculate_flip: nElem=%d nGeo=%d", nElem, nGeo);
  const int G = nGeo + 1;
  const int ci[4] = {0, nGeo, nGeo, 0}, cj[4] = {0, 0, nGeo, nGeo};
  const int smap[4][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}};  // CGNS side -> corners
  std::vector<int32_t> newv((size_t)4 * nElem);
  for (int e = 0; e < nElem; e++)
    for (int s = 0; s < 4; s++) {
      const int32_t *si = sideInfo + 5 * (s + 4 * (size_t)e);
      newv[s + 4 * (size_t)e] = si[3];
      if (si[4] != 0) continue;
      const int e2 = si[2] - 1, s2 = si[3] / 10 - 1;
      SELFAMD_REQUIRE(e2 >= 0 && e2 < nElem && s2 >= 0 && s2 < 4,
                      "recalculate_flip: bad neighbour (elem %d side %d)", e + 1, s + 1);
      int diff = 0;
      for (int l = 0; l < 2; l++) {
        const int a = smap[s][l], b = smap[s2][l];
        diff += std::abs(globalNodeIDs[ci[a] + G * (cj[a] + G * (size_t)e)] -
                         globalNodeIDs[ci[b] + G * (cj[b] + G * (size_t)e2)]);
      }
      newv[s + 4 * (size_t)e] = 10 * (s2 + 1) + (diff == 0 ? 0 : 1);
    }
  for (size_t k = 0; k < newv.size(); k++) sideInfo[5 * k + 3] = newv[k];
  return kOk;
}

int selfamd_mesh2d_reset_boundary_condition_type(int nElem, int32_t *sideInfo, int bcid) {
  for (size_t k = 0; k < (size_t)4 * nElem; k++)
    if (sideInfo[5 * k + 2] == 0) sideInfo[5 * k + 4] = bcid;
  return kOk;
}

// Extruded 3-D HOPr arrays -> 2-D mesh (src/SELF_Mesh.f90:962-1007): keep the bottom node
// layer and sides 2..5 of each hexahedron, shift the neighbour local side down by one
// (sideInfo(4) - 10, so boundary sides carry -10) and recompute the flips.
int selfamd_mesh2d_from_hopr(int nElem, int nGeo, const int32_t *sideInfo3D, const double *nodeCoords3D,
                             const int32_t *globalNodeIDs3D, int32_t *sideInfo, double *nodeCoords,
                             int32_t *globalNodeIDs) {
  SELFAMD_REQUIRE(nElem >= 1 && nGeo >= 1, "mesh2d_from_hopr: nElem=%d nGeo=%d", nElem, nGeo);
  const int G = nGeo + 1;
  for (int e = 0; e < nElem; e++)
    for (int j = 0; j < G; j++)
      for (int i = 0; i < G; i++) {
        const size_t src = i + (size_t)G * (j + (size_t)G * (G * (size_t)e));
        const size_t dst = i + (size_t)G * (j + (size_t)G * e);
        nodeCoords[2 * dst] = nodeCoords3D[3 * src];
        nodeCoords[2 * dst + 1] = nodeCoords3D[3 * src + 1];
        globalNodeIDs[dst] = globalNodeIDs3D[src];
      }
  for (int e = 0; e < nElem; e++)
    for (int s = 0; s < 4; s++) {
      const int32_t *src = sideInfo3D + 5 * ((size_t)s + 1 + 6 * (size_t)e);
      int32_t *dst = sideInfo + 5 * (s + 4 * (size_t)e);
      for (int k = 0; k < 5; k++) dst[k] = src[k];
      dst[3] -= 10;
    }
  return selfamd_mesh2d_recalculate_flip(nElem, nGeo, sideInfo, globalNodeIDs);
}

}  // extern "C"
