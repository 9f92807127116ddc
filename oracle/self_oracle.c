/*
 * self_oracle.c -- CPU restatement of the SELF 2D DGSEM tendency + RK3 path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load this library; the product never does.
 *
 * Each routine follows one reference routine, cited file:line relative to the reference
 * checkout, keeping its loop order, accumulation order and quirks.  The element loops of
 * each pass may run under OpenMP (or_set_threads); passes are element-independent, so the
 * threaded result is bit-identical to the serial one.
 *
 * Parity pinning (what checks each group; tests/test_oracle*.py):
 *   quadrature/Lagrange  -> numpy leggauss closed forms, D*1=0, D*x=1, exactness, dg identity
 *   interp/divergence    -> polynomial exactness, weak==strong for interpolated normal flux
 *   connectivity         -> HOPr mesh fixtures derived from the etc/mesh HOPr files (h5dump)
 *   side exchange        -> brute-force permutation identity on structured meshes
 *   Riemann solvers      -> consistency F*(q,q,n) = F(q).n nScale and conservation
 *                           F*(L,R,n) = -F*(R,L,-n) against the physical fluxes; the SW and
 *                           Euler dissipation coefficients against eigenvalues of
 *                           finite-difference flux Jacobians; LSW = exact Godunov flux
 *   wall BC              -> reflection u - 2(u.n)n for arbitrary unit normals (Circle mesh),
 *                           mass conservation on walled meshes
 *   volume flux, source  -> the tendency of a smooth state converges to -div F + S of the
 *                           PDE; closed-form tendency of a uniform state (Euler, quirk Q3)
 *   tendency             -> mass conservation, lake at rest, free-stream preservation
 *   SW velocity lag (Q1) -> first-order temporal convergence of the nonlinear RK3 step
 *   RK3                  -> third-order convergence of the linear semi-discrete system
 * No function on the tendency path is left unpinned.
 */
#include "self_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P_ (N + 1)
#define SC(i, j, v, e) ((i) + P_ * ((j) + P_ * ((v) + (size_t)nVar * (e))))
#define VE(d, i, j, v, e) ((d) + 2 * SC(i, j, v, e))
#define TE(r, c, i, j, v, e) ((r) + 2 * (c) + 4 * SC(i, j, v, e))
#define SCB(i, v, s, e) ((i) + P_ * ((v) + nVar * ((s) + 4 * (size_t)(e))))
#define VEB(d, i, v, s, e) ((d) + 2 * SCB(i, v, s, e))
#define TEB(r, c, i, v, s, e) ((r) + 2 * (c) + 4 * SCB(i, v, s, e))
#define SI(k, s, e) ((k) + 5 * ((s) + 4 * (size_t)(e)))
#define MAT(a, b) ((a) + P_ * (b))

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1)")

static const double TOL = 2.220446049250313080847e-16; /* epsilon(1.0_real64) */
static const double PI = 3.141592653589793115997963468544185161590576171875;

/* ===================================================================================== */
/* a1: quadrature (src/SELF_Quadrature.f90)                                              */
/* ===================================================================================== */

/* LegendrePolynomial, src/SELF_Quadrature.f90:342-383 (Kopriva Alg. 22) */
void or_legendre_polynomial(int N, double x, double *lAtX, double *dLdxAtX) {
  if (N == 0) { *lAtX = 1.0; *dLdxAtX = 0.0; return; }
  if (N == 1) { *lAtX = x; *dLdxAtX = 1.0; return; }
  double lnm2 = 1.0, lnm1 = x, dlnm2 = 0.0, dlnm1 = 1.0, l = 0, dl = 0;
  for (int i = 2; i <= N; i++) {
    l = ((2.0 * i - 1.0) * x * lnm1 - ((double)i - 1.0) * lnm2) / ((double)i);
    dl = dlnm2 + (2.0 * i - 1.0) * lnm1;
    lnm2 = lnm1; lnm1 = l; dlnm2 = dlnm1; dlnm1 = dl;
  }
  *lAtX = l; *dLdxAtX = dl;
}

/* LegendreGauss, src/SELF_Quadrature.f90:201-262 (Kopriva Alg. 23).
 * Quirk Q5: the j loop runs to (N+1)/2 inclusive and mirrors; the Newton exit test
 * abs(delta) <= TOL*x is never true for x < 0, so those nodes take all newtonMax=500
 * iterations.  The N==1 branch in the reference reads the (unset) output array; the
 * evident intent (mirror of node/weight 0) is restated here. */
void or_legendre_gauss(int N, double *nodes, double *weights) {
  double *xl = (double *)malloc(sizeof(double) * (N + 1));
  double *wl = (double *)malloc(sizeof(double) * (N + 1));
  double lN1, dlN1, delta;
  if (N == 0) {
    xl[0] = 0.0; wl[0] = 2.0;
  } else if (N == 1) {
    xl[0] = -sqrt(1.0 / 3.0); wl[0] = 1.0; xl[1] = -xl[0]; wl[1] = wl[0];
  } else {
    for (int j = 0; j <= (N + 1) / 2; j++) {
      xl[j] = -cos((2.0 * j + 1.0) * PI / (2.0 * N + 1.0));
      for (int k = 1; k <= 500; k++) {
        or_legendre_polynomial(N + 1, xl[j], &lN1, &dlN1);
        delta = -lN1 / dlN1;
        xl[j] = xl[j] + delta;
        if (fabs(delta) <= TOL * xl[j]) break;
      }
      or_legendre_polynomial(N + 1, xl[j], &lN1, &dlN1);
      wl[j] = 2.0 / ((1.0 - xl[j] * xl[j]) * dlN1 * dlN1);
      wl[N - j] = wl[j];
      xl[N - j] = -xl[j];
    }
  }
  if (fmod((double)N, 2.0) == 0.0) {
    or_legendre_polynomial(N + 1, 0.0, &lN1, &dlN1);
    xl[N / 2] = 0.0;
    wl[N / 2] = 2.0 / (dlN1 * dlN1);
  }
  for (int j = 0; j <= N; j++) { nodes[j] = xl[j]; weights[j] = wl[j]; }
  free(xl); free(wl);
}

/* LegendreQandL, src/SELF_Quadrature.f90:385-418 (Kopriva Alg. 24) */
static void legendre_q_and_l(int N, double x, double *q, double *qprime, double *lN) {
  double lNm2 = 1.0, lNm1 = x, dlNm2 = 0.0, dlNm1 = 1.0, l = 0, dl = 0, lN1, dlN1;
  for (int i = 2; i <= N; i++) {
    l = ((2.0 * i - 1.0) * x * lNm1 - (i - 1.0) * lNm2) / ((double)i);
    dl = dlNm2 + (2.0 * i - 1.0) * lNm1;
    lNm2 = lNm1; lNm1 = l; dlNm2 = dlNm1; dlNm1 = dl;
  }
  lN1 = ((2.0 * (N + 1) - 1.0) * x * l - ((double)(N + 1) - 1.0) * lNm2) / ((double)(N + 1));
  dlN1 = dlNm2 + (2.0 * (N + 1) - 1.0) * lNm1;
  *q = lN1 - lNm2;
  *qprime = dlN1 - dlNm2;
  *lN = l;
}

/* LegendreGaussLobatto, src/SELF_Quadrature.f90:270-334 (Kopriva Alg. 25) */
void or_legendre_gauss_lobatto(int N, double *nodes, double *weights) {
  double q, qp, lN, delta;
  if (N == 1) {
    nodes[0] = -1.0; weights[0] = 1.0; nodes[1] = 1.0; weights[1] = 1.0;
  } else {
    nodes[0] = -1.0;
    weights[0] = 2.0 / ((double)N * ((double)N + 1.0));
    nodes[N] = 1.0;
    weights[N] = weights[0];
    for (int j = 1; j <= (N + 1) / 2 - 1; j++) {
      nodes[j] = -cos((j + 0.25) * PI / N - 3.0 / (8.0 * N * PI * (j + 0.25)));
      for (int k = 1; k <= 500; k++) {
        legendre_q_and_l(N, nodes[j], &q, &qp, &lN);
        delta = -q / qp;
        nodes[j] = nodes[j] + delta;
        if (fabs(delta) <= TOL * nodes[j]) break;
      }
      legendre_q_and_l(N, nodes[j], &q, &qp, &lN);
      weights[j] = 2.0 / ((double)N * ((double)N + 1.0) * lN * lN);
      weights[N - j] = weights[j];
      nodes[N - j] = -nodes[j];
    }
  }
  if (fmod((double)N, 2.0) == 0.0) {
    legendre_q_and_l(N, 0.0, &q, &qp, &lN);
    nodes[N / 2] = 0.0;
    weights[N / 2] = 2.0 / ((double)N * ((double)N + 1.0) * lN * lN);
  }
}

/* AlmostEqual_r64, src/SELF_SupportRoutines.f90:76-96 */
static int almost_equal(double a, double b) {
  if (a == 0.0 || b == 0.0) return fabs(a - b) <= TOL;
  return (fabs(a - b) <= TOL * fabs(a)) || (fabs(a - b) <= TOL * fabs(b));
}

/* UniformPoints, src/SELF_SupportRoutines.f90:502-520 */
static void uniform_points(double a, double b, int n, double *x) {
  double dx = (b - a) / (double)n;
  for (int i = 0; i <= n; i++) x[i] = a + dx * (double)i;
}

/* CalculateLagrangePolynomials, src/SELF_Lagrange.f90:2721-2774 */
static void lagrange_polynomials(int N, const double *cp, const double *bw, double sE, double *lAtS) {
  int match = 0;
  for (int j = 0; j <= N; j++) {
    lAtS[j] = 0.0;
    if (almost_equal(sE, cp[j])) { lAtS[j] = 1.0; match = 1; }
  }
  if (match) return;
  double t1 = 0.0;
  for (int j = 0; j <= N; j++) { double t2 = bw[j] / (sE - cp[j]); lAtS[j] = t2; t1 += t2; }
  for (int j = 0; j <= N; j++) lAtS[j] = lAtS[j] / t1;
}

/* Init_Lagrange, src/SELF_Lagrange.f90:521-615, with
 * CalculateBarycentricWeights :2542-2570, CalculateInterpolationMatrix :2582-2643,
 * CalculateDerivativeMatrix :2656-2709. */
void or_lagrange_init(int N, int ctrlType, int M, int tgtType, double *cp, double *tp,
                      double *bw, double *qw, double *iMatrix, double *dMatrix,
                      double *dgMatrix, double *bMatrix) {
  double *qtmp = (double *)malloc(sizeof(double) * (M + 1));
  if (ctrlType == OR_GAUSS) or_legendre_gauss(N, cp, qw);
  else if (ctrlType == OR_GAUSS_LOBATTO) or_legendre_gauss_lobatto(N, cp, qw);
  else { uniform_points(-1.0, 1.0, N, cp); for (int i = 0; i <= N; i++) qw[i] = 2.0 / (double)N; }
  if (tgtType == OR_GAUSS) or_legendre_gauss(M, tp, qtmp);
  else if (tgtType == OR_GAUSS_LOBATTO) or_legendre_gauss_lobatto(M, tp, qtmp);
  else uniform_points(-1.0, 1.0, M, tp);
  free(qtmp);
  /* barycentric weights */
  for (int i = 0; i <= N; i++) bw[i] = 1.0;
  for (int j = 1; j <= N; j++)
    for (int i = 0; i <= j - 1; i++) {
      bw[i] = bw[i] * (cp[i] - cp[j]);
      bw[j] = bw[j] * (cp[j] - cp[i]);
    }
  for (int j = 0; j <= N; j++) bw[j] = 1.0 / bw[j];
  /* interpolation matrix: iMatrix%hostData(col,row) = I(row,col) */
  for (int row = 0; row <= M; row++) {
    int hasMatch = 0;
    for (int col = 0; col <= N; col++) {
      iMatrix[col + (N + 1) * row] = 0.0;
      if (almost_equal(tp[row], cp[col])) { hasMatch = 1; iMatrix[col + (N + 1) * row] = 1.0; }
    }
    if (!hasMatch) {
      double t1 = 0.0;
      for (int col = 0; col <= N; col++) {
        double t2 = bw[col] / (tp[row] - cp[col]);
        iMatrix[col + (N + 1) * row] = t2;
        t1 += t2;
      }
      for (int col = 0; col <= N; col++) iMatrix[col + (N + 1) * row] = iMatrix[col + (N + 1) * row] / t1;
    }
  }
  /* derivative matrices; dmat(row,col) = l'_col(x_row); stored transposed */
  double *dmat = (double *)malloc(sizeof(double) * (N + 1) * (N + 1));
#define DM(r, c) dmat[(r) + (N + 1) * (c)]
  for (int row = 0; row <= N; row++) {
    DM(row, row) = 0.0;
    for (int col = 0; col <= N; col++) {
      if (col != row) {
        DM(row, col) = bw[col] / (bw[row] * (cp[row] - cp[col]));
        DM(row, row) = DM(row, row) - DM(row, col);
      }
    }
  }
  for (int row = 0; row <= N; row++)
    for (int col = 0; col <= N; col++) {
      /* dgmat(row,col) = -dmat(col,row)*qw(col)/qw(row); hostData(row,col) = mat(col,row) */
      dMatrix[row + (N + 1) * col] = DM(col, row);
      dgMatrix[row + (N + 1) * col] = -DM(row, col) * qw[row] / qw[col];
    }
#undef DM
  free(dmat);
  lagrange_polynomials(N, cp, bw, -1.0, bMatrix);
  lagrange_polynomials(N, cp, bw, 1.0, bMatrix + (N + 1));
}

/* ===================================================================================== */
/* a2, a3, a12, a16: Lagrange 2D operators (src/SELF_Lagrange.f90)                         */
/* ===================================================================================== */

/* ScalarBoundaryInterp_2D_cpu, src/SELF_Lagrange.f90:2241-2270 */
void or_scalar_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int i = 0; i <= N; i++) {
        double s = 0, ea = 0, no = 0, w = 0;
        for (int ii = 0; ii <= N; ii++) {
          s += bM[MAT(ii, 0)] * f[SC(i, ii, v, e)];
          ea += bM[MAT(ii, 1)] * f[SC(ii, i, v, e)];
          no += bM[MAT(ii, 1)] * f[SC(i, ii, v, e)];
          w += bM[MAT(ii, 0)] * f[SC(ii, i, v, e)];
        }
        fb[SCB(i, v, 0, e)] = s; fb[SCB(i, v, 1, e)] = ea;
        fb[SCB(i, v, 2, e)] = no; fb[SCB(i, v, 3, e)] = w;
      }
}

/* VectorBoundaryInterp_2D_cpu, src/SELF_Lagrange.f90:2284-2316 */
void or_vector_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int i = 0; i <= N; i++) {
        double b[2][4] = {{0}};
        for (int ii = 0; ii <= N; ii++)
          for (int d = 0; d < 2; d++) {
            b[d][0] += bM[MAT(ii, 0)] * f[VE(d, i, ii, v, e)];
            b[d][1] += bM[MAT(ii, 1)] * f[VE(d, ii, i, v, e)];
            b[d][2] += bM[MAT(ii, 1)] * f[VE(d, i, ii, v, e)];
            b[d][3] += bM[MAT(ii, 0)] * f[VE(d, ii, i, v, e)];
          }
        for (int d = 0; d < 2; d++)
          for (int s = 0; s < 4; s++) fb[VEB(d, i, v, s, e)] = b[d][s];
      }
}

/* TensorBoundaryInterp_2D_cpu, src/SELF_Lagrange.f90:2330-2365 */
void or_tensor_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int i = 0; i <= N; i++) {
        double b[2][2][4] = {{{0}}};
        for (int ii = 0; ii <= N; ii++)
          for (int c = 0; c < 2; c++)
            for (int r = 0; r < 2; r++) {
              b[r][c][0] += bM[MAT(ii, 0)] * f[TE(r, c, i, ii, v, e)];
              b[r][c][1] += bM[MAT(ii, 1)] * f[TE(r, c, ii, i, v, e)];
              b[r][c][2] += bM[MAT(ii, 1)] * f[TE(r, c, i, ii, v, e)];
              b[r][c][3] += bM[MAT(ii, 0)] * f[TE(r, c, ii, i, v, e)];
            }
        for (int c = 0; c < 2; c++)
          for (int r = 0; r < 2; r++)
            for (int s = 0; s < 4; s++) fb[TEB(r, c, i, v, s, e)] = b[r][c][s];
      }
}

/* ScalarGridInterp_2D_cpu, src/SELF_Lagrange.f90:741-789 (iM is iMatrix(0:N,0:M)) */
void or_scalar_grid_interp_2d(int N, int M, const double *iM, const double *f, double *fNew, int nVar, int nEl) {
  const int Q = M + 1;
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= M; j++)
        for (int i = 0; i <= M; i++) {
          double fij = 0.0;
          for (int jj = 0; jj <= N; jj++) {
            double fi = 0.0;
            for (int ii = 0; ii <= N; ii++) fi += f[SC(ii, jj, v, e)] * iM[ii + P_ * i];
            fij += fi * iM[jj + P_ * j];
          }
          fNew[i + Q * (j + Q * (v + (size_t)nVar * e))] = fij;
        }
}

/* VectorGridInterp_2D_cpu, src/SELF_Lagrange.f90:824-876 */
void or_vector_grid_interp_2d(int N, int M, const double *iM, const double *f, double *fNew, int nVar, int nEl) {
  const int Q = M + 1;
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= M; j++)
        for (int i = 0; i <= M; i++) {
          size_t o = 2 * (i + Q * (j + Q * (v + (size_t)nVar * e)));
          fNew[o] = 0.0; fNew[o + 1] = 0.0;
          for (int jj = 0; jj <= N; jj++) {
            double fi0 = 0.0, fi1 = 0.0;
            for (int ii = 0; ii <= N; ii++) {
              fi0 += f[VE(0, ii, jj, v, e)] * iM[ii + P_ * i];
              fi1 += f[VE(1, ii, jj, v, e)] * iM[ii + P_ * i];
            }
            fNew[o] += fi0 * iM[jj + P_ * j];
            fNew[o + 1] += fi1 * iM[jj + P_ * j];
          }
        }
}

/* VectorGradient_2D_cpu, src/SELF_Lagrange.f90:1424-1463 */
void or_vector_gradient_2d(int N, const double *dM, const double *f, double *gradF, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          double g11 = 0, g21 = 0, g12 = 0, g22 = 0;
          for (int ii = 0; ii <= N; ii++) {
            g11 += dM[MAT(ii, i)] * f[VE(0, ii, j, v, e)];
            g21 += dM[MAT(ii, i)] * f[VE(1, ii, j, v, e)];
            g12 += dM[MAT(ii, j)] * f[VE(0, i, ii, v, e)];
            g22 += dM[MAT(ii, j)] * f[VE(1, i, ii, v, e)];
          }
          gradF[TE(0, 0, i, j, v, e)] = g11; gradF[TE(1, 0, i, j, v, e)] = g21;
          gradF[TE(0, 1, i, j, v, e)] = g12; gradF[TE(1, 1, i, j, v, e)] = g22;
        }
}

/* VectorDivergence_2D_cpu (strong form), src/SELF_Lagrange.f90:1549-1574 */
void or_vector_divergence_2d(int N, const double *dM, const double *f, double *dF, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          double d = 0.0;
          for (int ii = 0; ii <= N; ii++)
            d = d + dM[MAT(ii, i)] * f[VE(0, ii, j, v, e)] + dM[MAT(ii, j)] * f[VE(1, i, ii, v, e)];
          dF[SC(i, j, v, e)] = d;
        }
}

/* VectorDGDivergence_2D_cpu (weak form), src/SELF_Lagrange.f90:1589-1625 */
void or_vector_dg_divergence_2d(int N, const double *dgM, const double *bM, const double *qW,
                                const double *f, const double *bF, double *dF, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          double d = 0.0;
          for (int ii = 0; ii <= N; ii++)
            d = d + dgM[MAT(ii, i)] * f[VE(0, ii, j, v, e)] + dgM[MAT(ii, j)] * f[VE(1, i, ii, v, e)];
          d = d + (bM[MAT(i, 1)] * bF[SCB(j, v, 1, e)] + bM[MAT(i, 0)] * bF[SCB(j, v, 3, e)]) / qW[i] +
              (bM[MAT(j, 1)] * bF[SCB(i, v, 2, e)] + bM[MAT(j, 0)] * bF[SCB(i, v, 0, e)]) / qW[j];
          dF[SC(i, j, v, e)] = d;
        }
}

/* Determinant_Tensor2D, src/SELF_Data.f90:1570-1607 */
void or_determinant_tensor_2d(int N, const double *t, double *det, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++)
          det[SC(i, j, v, e)] = t[TE(0, 0, i, j, v, e)] * t[TE(1, 1, i, j, v, e)] -
                                t[TE(0, 1, i, j, v, e)] * t[TE(1, 0, i, j, v, e)];
}

/* ===================================================================================== */
/* a18: mesh connectivity (src/SELF_Mesh.f90)                                              */
/* ===================================================================================== */

/* RecalculateFlip_Mesh2D, src/SELF_Mesh.f90:1018-1174 (serial: no decomposition).
 * CGNSCornerMap / CGNSSideMap from Init_Mesh2D :636-645; CompareArray
 * (src/SELF_SupportRoutines.f90:404-425) is an ordered element-wise comparison. */
void or_recalculate_flip(int nElem, int nGeo, int32_t *sideInfo, const int32_t *gnid) {
  const int cornerI[4] = {0, nGeo, nGeo, 0}, cornerJ[4] = {0, 0, nGeo, nGeo};
  const int sideMap[4][2] = {{1, 2}, {2, 3}, {4, 3}, {1, 4}};
  const int G = nGeo + 1;
  int *nid1 = (int *)calloc((size_t)8 * nElem, sizeof(int));
  int *nid2 = (int *)calloc((size_t)8 * nElem, sizeof(int));
  for (int e1 = 0; e1 < nElem; e1++)
    for (int s1 = 0; s1 < 4; s1++) {
      int e2 = sideInfo[SI(2, s1, e1)];
      int s2 = sideInfo[SI(3, s1, e1)] / 10;
      int bcid = sideInfo[SI(4, s1, e1)];
      if (bcid == 0) {
        for (int l = 0; l < 2; l++) {
          int c1 = sideMap[s1][l] - 1, c2 = sideMap[s2 - 1][l] - 1;
          nid1[l + 2 * (s1 + 4 * e1)] = gnid[cornerI[c1] + G * (cornerJ[c1] + G * e1)];
          nid2[l + 2 * (s1 + 4 * e1)] = gnid[cornerI[c2] + G * (cornerJ[c2] + G * (e2 - 1))];
        }
      }
    }
  for (int e1 = 0; e1 < nElem; e1++)
    for (int s1 = 0; s1 < 4; s1++) {
      int s2 = sideInfo[SI(3, s1, e1)] / 10;
      int bcid = sideInfo[SI(4, s1, e1)];
      if (bcid == 0) {
        int diff = 0;
        for (int l = 0; l < 2; l++) diff += abs(nid1[l + 2 * (s1 + 4 * e1)] - nid2[l + 2 * (s1 + 4 * e1)]);
        sideInfo[SI(3, s1, e1)] = diff == 0 ? 10 * s2 : 10 * s2 + 1;
      }
    }
  free(nid1); free(nid2);
}

/* ResetBoundaryConditionType_Mesh2D, src/SELF_Mesh.f90:695-719 */
void or_reset_boundary_condition_type(int nElem, int32_t *sideInfo, int bcid) {
  for (int e = 0; e < nElem; e++)
    for (int s = 0; s < 4; s++)
      if (sideInfo[SI(2, s, e)] == 0) sideInfo[SI(4, s, e)] = bcid;
}

/* Read_HOPr_Mesh2D 3-D -> 2-D conversion, src/SELF_Mesh.f90:962-1007, then RecalculateFlip */
void or_hopr_to_mesh2d(int nElem, int nGeo, const int32_t *si3, const double *nc3, const int32_t *gn3,
                       int32_t *sideInfo, double *nodeCoords, int32_t *gnid) {
  const int G = nGeo + 1;
  for (int e = 0; e < nElem; e++)
    for (int j = 0; j <= nGeo; j++)
      for (int i = 0; i <= nGeo; i++) {
        int nid = i + G * (j + G * (G * e)); /* zero-based form of i+1+(nGeo+1)*(j+(nGeo+1)*((nGeo+1)*(eid-1))) */
        size_t o = i + G * (j + (size_t)G * e);
        nodeCoords[2 * o] = nc3[3 * (size_t)nid];
        nodeCoords[2 * o + 1] = nc3[3 * (size_t)nid + 1];
        gnid[o] = gn3[nid];
      }
  for (int e = 0; e < nElem; e++)
    for (int l = 0; l < 4; l++) {
      int iSide = l + 1 + 6 * e; /* zero-based form of lsid+1+6*(eid-1) */
      for (int k = 0; k < 5; k++) sideInfo[SI(k, l, e)] = si3[5 * iSide + k];
      sideInfo[SI(3, l, e)] = sideInfo[SI(3, l, e)] - 10;
    }
  or_recalculate_flip(nElem, nGeo, sideInfo, gnid);
}

/* ===================================================================================== */
/* a16: geometry (src/SELF_Geometry.f90)                                                   */
/* ===================================================================================== */

/* GenerateFromMesh_SEMQuad :393-435, CalculateMetricTerms_SEMQuad :528-552,
 * CalculateContravariantBasis_SEMQuad :437-526 */
or_geometry *or_geometry_from_mesh(int N, int ctrlType, int nGeo, int nElem, const double *nodeCoords) {
  or_geometry *g = (or_geometry *)calloc(1, sizeof(or_geometry));
  const int P = N + 1, G = nGeo + 1;
  const size_t nn = (size_t)P * P * nElem, nb = (size_t)P * 4 * nElem;
  g->N = N; g->nElem = nElem;
  g->x = calloc(2 * nn, sizeof(double)); g->xb = calloc(2 * nb, sizeof(double));
  g->dxds = calloc(4 * nn, sizeof(double)); g->dxdsb = calloc(4 * nb, sizeof(double));
  g->dsdx = calloc(4 * nn, sizeof(double)); g->dsdxb = calloc(4 * nb, sizeof(double));
  g->nHatb = calloc(2 * nb, sizeof(double)); g->nScaleb = calloc(nb, sizeof(double));
  g->J = calloc(nn, sizeof(double)); g->Jb = calloc(nb, sizeof(double));
  /* meshToModel: control = mesh nodes (uniform, nGeo), target = model nodes */
  double *cp = malloc(sizeof(double) * G), *tp = malloc(sizeof(double) * P), *bw = malloc(sizeof(double) * G),
         *qw = malloc(sizeof(double) * G), *iM = malloc(sizeof(double) * G * P), *dM = malloc(sizeof(double) * G * G),
         *dgM = malloc(sizeof(double) * G * G), *bM = malloc(sizeof(double) * G * 2);
  or_lagrange_init(nGeo, OR_UNIFORM, N, ctrlType, cp, tp, bw, qw, iM, dM, dgM, bM);
  or_vector_grid_interp_2d(nGeo, N, iM, nodeCoords, g->x, 1, nElem);
  free(cp); free(tp); free(bw); free(qw); free(iM); free(dM); free(dgM); free(bM);
  /* model Lagrange (control points of the model) */
  double *mcp = malloc(sizeof(double) * P), *mtp = malloc(sizeof(double) * P), *mbw = malloc(sizeof(double) * P),
         *mqw = malloc(sizeof(double) * P), *miM = malloc(sizeof(double) * P * P), *mdM = malloc(sizeof(double) * P * P),
         *mdgM = malloc(sizeof(double) * P * P), *mbM = malloc(sizeof(double) * P * 2);
  or_lagrange_init(N, ctrlType, N, ctrlType, mcp, mtp, mbw, mqw, miM, mdM, mdgM, mbM);
  or_vector_boundary_interp_2d(N, mbM, g->x, g->xb, 1, nElem);
  or_vector_gradient_2d(N, mdM, g->x, g->dxds, 1, nElem);
  or_tensor_boundary_interp_2d(N, mbM, g->dxds, g->dxdsb, 1, nElem);
  or_determinant_tensor_2d(N, g->dxds, g->J, 1, nElem);
  or_scalar_boundary_interp_2d(N, mbM, g->J, g->Jb, 1, nElem);
  /* contravariant basis */
  const int nVar = 1;
  for (int e = 0; e < nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        g->dsdx[TE(0, 0, i, j, 0, e)] = g->dxds[TE(1, 1, i, j, 0, e)];
        g->dsdx[TE(1, 0, i, j, 0, e)] = -g->dxds[TE(0, 1, i, j, 0, e)];
        g->dsdx[TE(0, 1, i, j, 0, e)] = -g->dxds[TE(1, 0, i, j, 0, e)];
        g->dsdx[TE(1, 1, i, j, 0, e)] = g->dxds[TE(0, 0, i, j, 0, e)];
      }
  or_tensor_boundary_interp_2d(N, mbM, g->dsdx, g->dsdxb, 1, nElem);
  for (int e = 0; e < nElem; e++)
    for (int k = 0; k < 4; k++)
      for (int i = 0; i <= N; i++) {
        double jb = g->Jb[SCB(i, 0, k, e)];
        double sg = (jb >= 0.0) ? 1.0 : -1.0; /* SIGN(1.0, J) (J never -0.0 here) */
        double fac = (k == 1 || k == 2) ? sg : -sg;
        int c = (k == 0 || k == 2) ? 1 : 0; /* S,N use contravariant vector 2; E,W vector 1 */
        double a0 = g->dsdxb[TEB(0, c, i, 0, k, e)], a1 = g->dsdxb[TEB(1, c, i, 0, k, e)];
        double mag = sqrt(a0 * a0 + a1 * a1);
        g->nScaleb[SCB(i, 0, k, e)] = mag;
        g->nHatb[VEB(0, i, 0, k, e)] = fac * a0 / mag;
        g->nHatb[VEB(1, i, 0, k, e)] = fac * a1 / mag;
        for (int cc = 0; cc < 2; cc++)
          for (int r = 0; r < 2; r++) g->dsdxb[TEB(r, cc, i, 0, k, e)] = g->dsdxb[TEB(r, cc, i, 0, k, e)] * fac;
      }
  free(mcp); free(mtp); free(mbw); free(mqw); free(miM); free(mdM); free(mdgM); free(mbM);
  return g;
}

void or_geometry_free(or_geometry *g) {
  if (!g) return;
  free(g->x); free(g->xb); free(g->dxds); free(g->dxdsb); free(g->dsdx); free(g->dsdxb);
  free(g->nHatb); free(g->nScaleb); free(g->J); free(g->Jb); free(g);
}

/* CovariantArcMin_SEMQuad, src/SELF_Geometry.f90:554-582 */
double or_covariant_arc_min(const or_geometry *g) {
  const int N = g->N, nVar = 1;
  double mn = INFINITY;
  for (int e = 0; e < g->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        double d11 = g->dxds[TE(0, 0, i, j, 0, e)], d12 = g->dxds[TE(0, 1, i, j, 0, e)];
        double d21 = g->dxds[TE(1, 0, i, j, 0, e)], d22 = g->dxds[TE(1, 1, i, j, 0, e)];
        double dx = sqrt(d11 * d11 + d12 * d12), dy = sqrt(d21 * d21 + d22 * d22);
        double ds = 2.0 * fmin(dx, dy) / ((double)N * (double)N);
        if (ds < mn) mn = ds;
      }
  return mn;
}

/* ===================================================================================== */
/* a4, a5, a11, a12, a17: mapped operators (src/SELF_MappedData.f90)                       */
/* ===================================================================================== */

/* SideExchange_MappedScalar2D CPU branch, src/SELF_MappedData.f90:834-917 (single rank) */
void or_side_exchange_scalar_2d(int N, int nVar, int nEl, const int32_t *si, const double *bnd, double *ext) {
  OMP_FOR
  for (int e1 = 0; e1 < nEl; e1++)
    for (int s1 = 0; s1 < 4; s1++) {
      int e2 = si[SI(2, s1, e1)] - 1;
      int s2 = si[SI(3, s1, e1)] / 10;
      int flip = si[SI(3, s1, e1)] - s2 * 10;
      int bcid = si[SI(4, s1, e1)];
      if (bcid != 0) continue;
      for (int v = 0; v < nVar; v++)
        for (int i1 = 0; i1 <= N; i1++) {
          int i2 = flip == 0 ? i1 : N - i1;
          if (flip == 0 || flip == 1) ext[SCB(i1, v, s1, e1)] = bnd[SCB(i2, v, s2 - 1, e2)];
        }
    }
}

/* SideExchange_MappedVector2D CPU branch, src/SELF_MappedData.f90:1870-1958 */
void or_side_exchange_vector_2d(int N, int nVar, int nEl, const int32_t *si, const double *bnd, double *ext) {
  OMP_FOR
  for (int e1 = 0; e1 < nEl; e1++)
    for (int s1 = 0; s1 < 4; s1++) {
      int e2 = si[SI(2, s1, e1)] - 1;
      int s2 = si[SI(3, s1, e1)] / 10;
      int flip = si[SI(3, s1, e1)] - s2 * 10;
      int bcid = si[SI(4, s1, e1)];
      if (bcid != 0) continue;
      for (int v = 0; v < nVar; v++)
        for (int i1 = 0; i1 <= N; i1++) {
          int i2 = flip == 0 ? i1 : N - i1;
          if (flip == 0 || flip == 1) {
            ext[VEB(0, i1, v, s1, e1)] = bnd[VEB(0, i2, v, s2 - 1, e2)];
            ext[VEB(1, i1, v, s1, e1)] = bnd[VEB(1, i2, v, s2 - 1, e2)];
          }
        }
    }
}

/* GradientSF_MappedScalar2D CPU branch, src/SELF_MappedData.f90:1080-1149 */
void or_gradient_sf_2d(int N, const double *dM, const double *dsdx, const double *J, const double *f,
                       double *gradF, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          double gx = 0.0, gy = 0.0;
          for (int ii = 0; ii <= N; ii++) {
            /* geometry arrays have nVar = 1: index them with their own layout */
            size_t g1 = (size_t)4 * (ii + P_ * (j + (size_t)P_ * e));
            size_t g2 = (size_t)4 * (i + P_ * (ii + (size_t)P_ * e));
            double f1 = f[SC(ii, j, v, e)] * dsdx[g1 + 0];
            double f2 = f[SC(i, ii, v, e)] * dsdx[g2 + 2];
            gx = gx + dM[MAT(ii, i)] * f1 + dM[MAT(ii, j)] * f2;
            f1 = f[SC(ii, j, v, e)] * dsdx[g1 + 1];
            f2 = f[SC(i, ii, v, e)] * dsdx[g2 + 3];
            gy = gy + dM[MAT(ii, i)] * f1 + dM[MAT(ii, j)] * f2;
          }
          double jac = J[i + P_ * (j + (size_t)P_ * e)];
          gradF[VE(0, i, j, v, e)] = gx / jac;
          gradF[VE(1, i, j, v, e)] = gy / jac;
        }
}

/* JacobianWeight_MappedScalar2D CPU branch, src/SELF_MappedData.f90:1235-1268 */
void or_jacobian_weight_2d(int N, const double *J, double *f, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) f[SC(i, j, v, e)] = f[SC(i, j, v, e)] / J[i + P_ * (j + (size_t)P_ * e)];
}

/* ContravariantProjection_MappedVector2D CPU branch, src/SELF_MappedData.f90:2220-2269 */
void or_contravariant_projection_2d(int N, const double *dsdx, double *f, int nVar, int nEl) {
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          size_t g = (size_t)4 * (i + P_ * (j + (size_t)P_ * e));
          double Fx = f[VE(0, i, j, v, e)], Fy = f[VE(1, i, j, v, e)];
          f[VE(0, i, j, v, e)] = dsdx[g + 0] * Fx + dsdx[g + 1] * Fy;
          f[VE(1, i, j, v, e)] = dsdx[g + 2] * Fx + dsdx[g + 3] * Fy;
        }
}

/* ===================================================================================== */
/* Models                                                                                  */
/* ===================================================================================== */

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

or_model *or_model_create(int kind, int N, int nElem, const int32_t *sideInfo, const or_geometry *geom) {
  or_model *m = (or_model *)calloc(1, sizeof(or_model));
  const int P = N + 1;
  m->kind = kind; m->N = N; m->nElem = nElem;
  m->nVar = kind == OR_MODEL_EULER ? 4 : 3;
  m->sideInfo = sideInfo; m->geom = geom;
  const int nVar = m->nVar;
  const size_t nn = (size_t)P * P * nElem, nb = (size_t)P * 4 * nElem;
  m->qWeights = zalloc(P); m->dMatrix = zalloc(P * P); m->dgMatrix = zalloc(P * P); m->bMatrix = zalloc(2 * P);
  double *cp = zalloc(P), *tp = zalloc(P), *bw = zalloc(P), *iM = zalloc(P * P);
  or_lagrange_init(N, OR_GAUSS, N, OR_GAUSS, cp, tp, bw, m->qWeights, iM, m->dMatrix, m->dgMatrix, m->bMatrix);
  free(cp); free(tp); free(bw); free(iM);
  m->sol = zalloc(nn * nVar); m->solB = zalloc(nb * nVar); m->solE = zalloc(nb * nVar);
  m->workSol = zalloc(nn * nVar); m->dSdt = zalloc(nn * nVar); m->source = zalloc(nn * nVar);
  m->fluxDiv = zalloc(nn * nVar); m->flux = zalloc(2 * nn * nVar); m->fluxBN = zalloc(nb * nVar);
  m->vel = zalloc(2 * nn); m->velB = zalloc(2 * nb); m->velE = zalloc(2 * nb);
  m->g = 1.0; m->fCoriConst = 0.0;
  m->topo = zalloc(nn); m->topoB = zalloc(nb); m->gradTopo = zalloc(2 * nn); m->fCori = zalloc(nn);
  if (kind == OR_MODEL_EULER) {
    m->prim = zalloc(nn * 4); m->primB = zalloc(nb * 4); m->primE = zalloc(nb * 4);
    m->diag = zalloc(nn * 4); m->diagB = zalloc(nb * 4); m->diagE = zalloc(nb * 4);
    m->ent = zalloc(nn * 4); m->entB = zalloc(nb * 4); m->entE = zalloc(nb * 4);
    m->env = zalloc(nn * 2); m->envGrad = zalloc(2 * nn * 2);
    m->presSolB = zalloc(nb * 4); m->presPrimB = zalloc(nb * 4); m->presDiagB = zalloc(nb * 4);
    m->gamma = 1.4; m->Rgas = 1.0;
  }
  return m;
}

void or_model_free(or_model *m) {
  if (!m) return;
  double *p[] = {m->qWeights, m->dMatrix, m->dgMatrix, m->bMatrix, m->sol, m->solB, m->solE, m->workSol,
                 m->dSdt, m->source, m->fluxDiv, m->flux, m->fluxBN, m->vel, m->velB, m->velE, m->topo,
                 m->topoB, m->gradTopo, m->fCori, m->prim, m->primB, m->primE, m->diag, m->diagB, m->diagE,
                 m->ent, m->entB, m->entE, m->env, m->envGrad, m->presSolB, m->presPrimB, m->presDiagB};
  for (size_t k = 0; k < sizeof(p) / sizeof(p[0]); k++) free(p[k]);
  free(m);
}

/* SetSolution (src/SELF_Model2D.f90:313-332): interior then BoundaryInterp */
void or_model_set_solution(or_model *m, const double *interior) {
  const int N = m->N;
  memcpy(m->sol, interior, sizeof(double) * (size_t)(N + 1) * (N + 1) * m->nVar * m->nElem);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->sol, m->solB, m->nVar, m->nElem);
}

/* SetTopography (src/models/SELF_ShallowWater.f90:241-262): H, H%BoundaryInterp, GradientSF */
void or_sw_set_topography(or_model *m, const double *topo) {
  const int N = m->N;
  memcpy(m->topo, topo, sizeof(double) * (size_t)(N + 1) * (N + 1) * m->nElem);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->topo, m->topoB, 1, m->nElem);
  or_gradient_sf_2d(N, m->dMatrix, m->geom->dsdx, m->geom->J, m->topo, m->gradTopo, 1, m->nElem);
}

/* SetBathymetry (src/models/SELF_LinearShallowWater.f90:251-296) */
void or_lsw_set_bathymetry(or_model *m, const double *H) {
  const int N = m->N;
  memcpy(m->topo, H, sizeof(double) * (size_t)(N + 1) * (N + 1) * m->nElem);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->topo, m->topoB, 1, m->nElem);
}

/* SetCoriolis (src/models/SELF_LinearShallowWater.f90:218-249) */
void or_lsw_set_coriolis(or_model *m, const double *f) {
  const int N = m->N;
  memcpy(m->fCori, f, sizeof(double) * (size_t)(N + 1) * (N + 1) * m->nElem);
}

/* SetGravityFromChar (src/models/SELF_CompressibleIdealGas2D.f90:608-630): environmentals
 * var 1 = potential, then GradientSF over both environmentals variables. */
void or_euler_set_gravity(or_model *m, const double *phi) {
  const int N = m->N, nVar = 2;
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) m->env[SC(i, j, 0, e)] = phi[i + (N + 1) * (j + (size_t)(N + 1) * e)];
  or_gradient_sf_2d(N, m->dMatrix, m->geom->dsdx, m->geom->J, m->env, m->envGrad, 2, m->nElem);
}

/* SetDragFromConstant / FromChar (src/models/SELF_CompressibleIdealGas2D.f90:632-672): var 2 */
void or_euler_set_drag(or_model *m, const double *cd) {
  const int N = m->N, nVar = 2;
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) m->env[SC(i, j, 1, e)] = cd[i + (N + 1) * (j + (size_t)(N + 1) * e)];
}

/* CalculateDiagnostics, src/models/SELF_CompressibleIdealGas2D.f90:984-1127 */
void or_euler_diagnostics(or_model *m) {
  const int N = m->N, nVar = 4;
  const double gam = m->gamma, R = m->Rgas;
  OMP_FOR
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        double rhoU = m->sol[SC(i, j, 0, e)], rhoV = m->sol[SC(i, j, 1, e)];
        double rho = m->sol[SC(i, j, 2, e)], rhoE = m->sol[SC(i, j, 3, e)];
        double u = rhoU / rho, v = rhoV / rho;
        double rhoKE = 0.5 * (rhoU * u + rhoV * v);
        double p = (gam - 1.0) * (rhoE - rhoKE);
        m->diag[SC(i, j, 0, e)] = rhoKE;
        m->diag[SC(i, j, 1, e)] = rhoE + p;
        m->diag[SC(i, j, 2, e)] = sqrt(gam * p / rho);
        m->diag[SC(i, j, 3, e)] = (2.0 / 3.0) * ((rhoE - rhoKE) / rho) / R;
      }
}

/* ConservativeToPrimitive, src/models/SELF_CompressibleIdealGas2D.f90:1129-1170 */
void or_euler_primitive(or_model *m) {
  const int N = m->N, nVar = 4;
  const double gam = m->gamma;
  OMP_FOR
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        double rhoU = m->sol[SC(i, j, 0, e)], rhoV = m->sol[SC(i, j, 1, e)];
        double rho = m->sol[SC(i, j, 2, e)], rhoE = m->sol[SC(i, j, 3, e)];
        double u = rhoU / rho, v = rhoV / rho;
        double rhoKE = 0.5 * (rhoU * u + rhoV * v);
        double p = (gam - 1.0) * (rhoE - rhoKE);
        m->prim[SC(i, j, 0, e)] = u; m->prim[SC(i, j, 1, e)] = v;
        m->prim[SC(i, j, 2, e)] = rho; m->prim[SC(i, j, 3, e)] = p;
      }
}

/* ConservativeToEntropy, src/models/SELF_CompressibleIdealGas2D.f90:1172-1215 */
void or_euler_entropy_vars(or_model *m) {
  const int N = m->N, nVar = 4;
  const double gam = m->gamma;
  OMP_FOR
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        double rhoU = m->sol[SC(i, j, 0, e)], rhoV = m->sol[SC(i, j, 1, e)];
        double rho = m->sol[SC(i, j, 2, e)], E = m->sol[SC(i, j, 3, e)];
        double u = rhoU / rho, v = rhoV / rho;
        double KE = 0.5 * (rhoU * u + rhoV * v);
        double p = (gam - 1.0) * (E - KE);
        double s = log(p) - gam * log(rho);
        m->ent[SC(i, j, 0, e)] = u * rho / p;
        m->ent[SC(i, j, 1, e)] = v * rho / p;
        m->ent[SC(i, j, 2, e)] = (gam - s) / (gam - 1.0) - KE / p;
        m->ent[SC(i, j, 3, e)] = -rho / p;
      }
}

/* Prescribed state := current state (SetPrescribedSolutionFromSolution,
 * src/models/SELF_CompressibleIdealGas2D.f90:694-740) */
void or_euler_set_prescribed(or_model *m) {
  const int N = m->N;
  or_euler_diagnostics(m);
  or_euler_primitive(m);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->sol, m->presSolB, 4, m->nElem);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->prim, m->presPrimB, 4, m->nElem);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->diag, m->presDiagB, 4, m->nElem);
}

/* PreTendency: SW src/models/SELF_ShallowWater.f90:190-239 (quirk Q1: velocity traces from the
 * solution%boundary left by the previous stage); Euler :955-982; LSW: base no-op. */
void or_pre_tendency(or_model *m) {
  const int N = m->N, nEl = m->nElem;
  if (m->kind == OR_MODEL_SW) {
    int nVar = 3;
    OMP_FOR
    for (int e = 0; e < nEl; e++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          double H = m->sol[SC(i, j, 2, e)];
          m->vel[2 * (i + (N + 1) * (j + (size_t)(N + 1) * e))] = m->sol[SC(i, j, 0, e)] / H;
          m->vel[2 * (i + (N + 1) * (j + (size_t)(N + 1) * e)) + 1] = m->sol[SC(i, j, 1, e)] / H;
        }
    OMP_FOR
    for (int e = 0; e < nEl; e++)
      for (int s = 0; s < 4; s++)
        for (int i = 0; i <= N; i++) {
          double H = m->solB[SCB(i, 2, s, e)];
          size_t o = 2 * (i + (N + 1) * (s + 4 * (size_t)e));
          m->velB[o] = m->solB[SCB(i, 0, s, e)] / H;
          m->velB[o + 1] = m->solB[SCB(i, 1, s, e)] / H;
        }
    or_side_exchange_vector_2d(N, 1, nEl, m->sideInfo, m->velB, m->velE);
  } else if (m->kind == OR_MODEL_EULER) {
    or_euler_diagnostics(m);
    or_euler_primitive(m);
    or_euler_entropy_vars(m);
    or_scalar_boundary_interp_2d(N, m->bMatrix, m->prim, m->primB, 4, nEl);
    or_scalar_boundary_interp_2d(N, m->bMatrix, m->diag, m->diagB, 4, nEl);
    or_scalar_boundary_interp_2d(N, m->bMatrix, m->ent, m->entB, 4, nEl);
    or_side_exchange_scalar_2d(N, 4, nEl, m->sideInfo, m->primB, m->primE);
    or_side_exchange_scalar_2d(N, 4, nEl, m->sideInfo, m->diagB, m->diagE);
    or_side_exchange_scalar_2d(N, 4, nEl, m->sideInfo, m->entB, m->entE);
  }
}

/* SetBoundaryCondition: LSW src/models/SELF_LinearShallowWater.f90:298-355;
 * SW src/models/SELF_ShallowWater.f90:284-343 (quirk Q2: radiation ext H = topography trace);
 * Euler src/models/SELF_CompressibleIdealGas2D.f90:1217-1323. */
void or_set_boundary_condition(or_model *m) {
  const int N = m->N, nEl = m->nElem, nVar = m->nVar;
  const int32_t *si = m->sideInfo;
  const double *nh = m->geom->nHatb;
  for (int e = 0; e < nEl; e++)
    for (int s = 0; s < 4; s++) {
      int bcid = si[SI(4, s, e)], e2 = si[SI(2, s, e)];
      if (e2 != 0) continue;
      for (int i = 0; i <= N; i++) {
        size_t gb = 2 * (i + (N + 1) * (s + 4 * (size_t)e));
        double nx = nh[gb], ny = nh[gb + 1];
        if (m->kind == OR_MODEL_EULER) {
          if (bcid == OR_BC_NONORMALFLOW) {
            double u = m->solB[SCB(i, 0, s, e)], v = m->solB[SCB(i, 1, s, e)];
            m->solE[SCB(i, 0, s, e)] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
            m->solE[SCB(i, 1, s, e)] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
            m->solE[SCB(i, 2, s, e)] = m->solB[SCB(i, 2, s, e)];
            m->solE[SCB(i, 3, s, e)] = m->solB[SCB(i, 3, s, e)];
            u = m->primB[SCB(i, 0, s, e)]; v = m->primB[SCB(i, 1, s, e)];
            m->primE[SCB(i, 0, s, e)] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
            m->primE[SCB(i, 1, s, e)] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
            m->primE[SCB(i, 2, s, e)] = m->primB[SCB(i, 2, s, e)];
            m->primE[SCB(i, 3, s, e)] = m->primB[SCB(i, 3, s, e)];
            for (int d = 0; d < 4; d++) m->diagE[SCB(i, d, s, e)] = m->diagB[SCB(i, d, s, e)];
          } else if (bcid == OR_BC_PRESCRIBED || bcid == OR_BC_RADIATION) {
            for (int d = 0; d < 4; d++) {
              m->solE[SCB(i, d, s, e)] = m->presSolB[SCB(i, d, s, e)];
              m->primE[SCB(i, d, s, e)] = m->presPrimB[SCB(i, d, s, e)];
              m->diagE[SCB(i, d, s, e)] = m->presDiagB[SCB(i, d, s, e)];
            }
          }
        } else {
          if (bcid == OR_BC_NONORMALFLOW) {
            double u = m->solB[SCB(i, 0, s, e)], v = m->solB[SCB(i, 1, s, e)];
            m->solE[SCB(i, 0, s, e)] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
            m->solE[SCB(i, 1, s, e)] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
            m->solE[SCB(i, 2, s, e)] = m->solB[SCB(i, 2, s, e)];
          } else { /* radiation and default */
            m->solE[SCB(i, 0, s, e)] = 0.0;
            m->solE[SCB(i, 1, s, e)] = 0.0;
            m->solE[SCB(i, 2, s, e)] = m->kind == OR_MODEL_SW ? m->topoB[i + (N + 1) * (s + 4 * (size_t)e)] : 0.0;
          }
        }
      }
    }
}

/* Source: LSW src/models/SELF_LinearShallowWater.f90:413-446; SW SELF_ShallowWater.f90:345-384;
 * Euler SELF_CompressibleIdealGas2D.f90:1325-1366 (quirk Q3: energy source -rhou*gx-rhou*gy). */
void or_source(or_model *m) {
  const int N = m->N, nEl = m->nElem, nVar = m->nVar;
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        size_t n1 = i + (N + 1) * (j + (size_t)(N + 1) * e);
        if (m->kind == OR_MODEL_LSW) {
          m->source[SC(i, j, 0, e)] = m->fCori[n1] * m->sol[SC(i, j, 1, e)];
          m->source[SC(i, j, 1, e)] = -m->fCori[n1] * m->sol[SC(i, j, 0, e)];
          m->source[SC(i, j, 2, e)] = 0.0;
        } else if (m->kind == OR_MODEL_SW) {
          m->source[SC(i, j, 0, e)] = -m->fCoriConst * m->sol[SC(i, j, 1, e)] +
                                      m->g * m->sol[SC(i, j, 2, e)] * m->gradTopo[2 * n1];
          m->source[SC(i, j, 1, e)] = m->fCoriConst * m->sol[SC(i, j, 0, e)] +
                                      m->g * m->sol[SC(i, j, 2, e)] * m->gradTopo[2 * n1 + 1];
          m->source[SC(i, j, 2, e)] = 0.0;
        } else {
          double rhou = m->sol[SC(i, j, 0, e)], rhov = m->sol[SC(i, j, 1, e)], rho = m->sol[SC(i, j, 2, e)];
          /* environmentalsGradient(1:2,i,j,1,iEl) and environmentals(i,j,2,iEl); nVar = 2 layout */
          size_t g1 = i + (N + 1) * (j + (size_t)(N + 1) * (2 * (size_t)e));
          size_t c2 = i + (N + 1) * (j + (size_t)(N + 1) * (1 + 2 * (size_t)e));
          double gx = m->envGrad[2 * g1], gy = m->envGrad[2 * g1 + 1];
          double Cd = m->env[c2];
          m->source[SC(i, j, 0, e)] = -rho * gx - Cd * rhou;
          m->source[SC(i, j, 1, e)] = -rho * gy - Cd * rhov;
          m->source[SC(i, j, 2, e)] = 0.0;
          m->source[SC(i, j, 3, e)] = -rhou * gx - rhou * gy;
        }
      }
}

/* RiemannSolver: LSW src/models/SELF_LinearShallowWater.f90:500-557 (quirk Q7, characteristic
 * upwind); SW SELF_ShallowWater.f90:451-529 (Rusanov); Euler NaiveLLF :1455-1542. */
void or_riemann(or_model *m) {
  const int N = m->N, nEl = m->nElem, nVar = m->nVar;
  const double *nh = m->geom->nHatb, *ns = m->geom->nScaleb;
  const double g = m->g;
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int s = 0; s < 4; s++)
      for (int i = 0; i <= N; i++) {
        size_t b1 = i + (N + 1) * (s + 4 * (size_t)e);
        double nx = nh[2 * b1], ny = nh[2 * b1 + 1], nmag = ns[b1];
        if (m->kind == OR_MODEL_LSW) {
          double c = sqrt(g * m->topoB[b1]);
          double unL = m->solB[SCB(i, 0, s, e)] * nx + m->solB[SCB(i, 1, s, e)] * ny;
          double unR = m->solE[SCB(i, 0, s, e)] * nx + m->solE[SCB(i, 1, s, e)] * ny;
          double etaL = m->solB[SCB(i, 2, s, e)], etaR = m->solE[SCB(i, 2, s, e)];
          double wL = 0.5 * (unL / g + etaL / c);
          double wR = 0.5 * (unR / g - etaR / c);
          m->fluxBN[SCB(i, 0, s, e)] = g * c * (wL - wR) * nx * nmag;
          m->fluxBN[SCB(i, 1, s, e)] = g * c * (wL - wR) * ny * nmag;
          m->fluxBN[SCB(i, 2, s, e)] = c * c * (wL + wR) * nmag;
        } else if (m->kind == OR_MODEL_SW) {
          double unL = m->velB[2 * b1] * nx + m->velB[2 * b1 + 1] * ny;
          double unR = m->velE[2 * b1] * nx + m->velE[2 * b1 + 1] * ny;
          double cL = sqrt(g * m->solB[SCB(i, 2, s, e)]);
          double cR = sqrt(g * m->solE[SCB(i, 2, s, e)]);
          double HuL = m->solB[SCB(i, 0, s, e)], HuR = m->solE[SCB(i, 0, s, e)];
          double HvL = m->solB[SCB(i, 1, s, e)], HvR = m->solE[SCB(i, 1, s, e)];
          double HL = m->solB[SCB(i, 2, s, e)], HR = m->solE[SCB(i, 2, s, e)];
          double fL0 = HuL * unL + g * HL * HL * 0.5 * nx;
          double fL1 = HvL * unL + g * HL * HL * 0.5 * ny;
          double fL2 = HuL * nx + HvL * ny;
          double fR0 = HuR * unR + g * HR * HR * 0.5 * nx;
          double fR1 = HvR * unR + g * HR * HR * 0.5 * ny;
          double fR2 = HuR * nx + HvR * ny;
          double a = fmax(fmax(fabs(unL + cL), fabs(unR + cR)), fmax(fabs(unL - cL), fabs(unR - cR)));
          m->fluxBN[SCB(i, 0, s, e)] = 0.5 * (fL0 + fR0 + a * (HuL - HuR)) * nmag;
          m->fluxBN[SCB(i, 1, s, e)] = 0.5 * (fL1 + fR1 + a * (HvL - HvR)) * nmag;
          m->fluxBN[SCB(i, 2, s, e)] = 0.5 * (fL2 + fR2 + a * (HL - HR)) * nmag;
        } else {
          double unL = m->primB[SCB(i, 0, s, e)] * nx + m->primB[SCB(i, 1, s, e)] * ny;
          double unR = m->primE[SCB(i, 0, s, e)] * nx + m->primE[SCB(i, 1, s, e)] * ny;
          double fL[4], fR[4], jump[4];
          fL[0] = unL * m->solB[SCB(i, 0, s, e)] + m->primB[SCB(i, 3, s, e)] * nx;
          fL[1] = unL * m->solB[SCB(i, 1, s, e)] + m->primB[SCB(i, 3, s, e)] * ny;
          fL[2] = m->solB[SCB(i, 0, s, e)] * nx + m->solB[SCB(i, 1, s, e)] * ny;
          fL[3] = unL * m->diagB[SCB(i, 1, s, e)];
          fR[0] = unR * m->solE[SCB(i, 0, s, e)] + m->primE[SCB(i, 3, s, e)] * nx;
          fR[1] = unR * m->solE[SCB(i, 1, s, e)] + m->primE[SCB(i, 3, s, e)] * ny;
          fR[2] = m->solE[SCB(i, 0, s, e)] * nx + m->solE[SCB(i, 1, s, e)] * ny;
          fR[3] = unR * m->diagE[SCB(i, 1, s, e)];
          for (int d = 0; d < 4; d++) jump[d] = m->solB[SCB(i, d, s, e)] - m->solE[SCB(i, d, s, e)];
          double cL = m->diagB[SCB(i, 2, s, e)], cR = m->diagE[SCB(i, 2, s, e)];
          double a = fmax(fabs(unL), fabs(unR)) + fmax(fabs(cL), fabs(cR));
          for (int d = 0; d < 4; d++) m->fluxBN[SCB(i, d, s, e)] = 0.5 * (fL[d] + fR[d] + a * jump[d]) * nmag;
        }
      }
}

/* Flux: LSW src/models/SELF_LinearShallowWater.f90:448-498; SW SELF_ShallowWater.f90:386-449;
 * Euler SELF_CompressibleIdealGas2D.f90:1368-1440 */
void or_flux(or_model *m) {
  const int N = m->N, nEl = m->nElem, nVar = m->nVar;
  const double g = m->g;
  OMP_FOR
  for (int e = 0; e < nEl; e++)
    for (int v = 0; v < nVar; v++)
      for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
          size_t n1 = i + (N + 1) * (j + (size_t)(N + 1) * e);
          double f0 = 0, f1 = 0;
          if (m->kind == OR_MODEL_LSW) {
            if (v == 0) { f0 = g * m->sol[SC(i, j, 2, e)]; f1 = 0.0; }
            else if (v == 1) { f0 = 0.0; f1 = g * m->sol[SC(i, j, 2, e)]; }
            else { f0 = m->topo[n1] * m->sol[SC(i, j, 0, e)]; f1 = m->topo[n1] * m->sol[SC(i, j, 1, e)]; }
          } else if (m->kind == OR_MODEL_SW) {
            double u = m->vel[2 * n1], vv = m->vel[2 * n1 + 1], H = m->sol[SC(i, j, 2, e)];
            if (v == 0) { f0 = u * m->sol[SC(i, j, 0, e)] + 0.5 * g * H * H; f1 = vv * m->sol[SC(i, j, 0, e)]; }
            else if (v == 1) { f0 = u * m->sol[SC(i, j, 1, e)]; f1 = vv * m->sol[SC(i, j, 1, e)] + 0.5 * g * H * H; }
            else { f0 = m->sol[SC(i, j, 0, e)]; f1 = m->sol[SC(i, j, 1, e)]; }
          } else {
            double u = m->prim[SC(i, j, 0, e)], vv = m->prim[SC(i, j, 1, e)], p = m->prim[SC(i, j, 3, e)];
            if (v == 0) { f0 = u * m->sol[SC(i, j, 0, e)] + p; f1 = vv * m->sol[SC(i, j, 0, e)]; }
            else if (v == 1) { f0 = u * m->sol[SC(i, j, 1, e)]; f1 = vv * m->sol[SC(i, j, 1, e)] + p; }
            else if (v == 2) { f0 = m->sol[SC(i, j, 0, e)]; f1 = m->sol[SC(i, j, 1, e)]; }
            else { f0 = u * m->diag[SC(i, j, 1, e)]; f1 = vv * m->diag[SC(i, j, 1, e)]; }
          }
          m->flux[VE(0, i, j, v, e)] = f0;
          m->flux[VE(1, i, j, v, e)] = f1;
        }
}

/* CalculateTendency_Model2D, src/SELF_Model2D.f90:816-859 */
void or_calculate_tendency(or_model *m) {
  const int N = m->N, nEl = m->nElem, nVar = m->nVar;
  or_pre_tendency(m);
  or_scalar_boundary_interp_2d(N, m->bMatrix, m->sol, m->solB, nVar, nEl);
  or_side_exchange_scalar_2d(N, nVar, nEl, m->sideInfo, m->solB, m->solE);
  or_set_boundary_condition(m);
  or_source(m);
  or_riemann(m);
  or_flux(m);
  or_contravariant_projection_2d(N, m->geom->dsdx, m->flux, nVar, nEl);
  or_vector_dg_divergence_2d(N, m->dgMatrix, m->bMatrix, m->qWeights, m->flux, m->fluxBN, m->fluxDiv, nVar, nEl);
  or_jacobian_weight_2d(N, m->geom->J, m->fluxDiv, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  OMP_FOR
  for (size_t k = 0; k < n; k++) m->dSdt[k] = m->source[k] - m->fluxDiv[k];
}

/* Williamson low-storage RK3 coefficients, src/SELF_Model.f90:25-27 */
static const double rk3_a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double rk3_b[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
static const double rk3_g[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

/* UpdateGRK3_Model2D, src/SELF_Model2D.f90:710-752 */
void or_update_grk3(or_model *m, int stage) {
  const size_t n = (size_t)(m->N + 1) * (m->N + 1) * m->nVar * m->nElem;
  const double a = rk3_a[stage], gdt = rk3_g[stage];
  OMP_FOR
  for (size_t k = 0; k < n; k++) {
    m->workSol[k] = a * m->workSol[k] + m->dSdt[k];
    m->sol[k] = m->sol[k] + gdt * m->dt * m->workSol[k];
  }
}

/* LowStorageRK3_timeIntegrator, src/SELF_Model.f90:704-732, driven for nSteps full steps
 * (tn = t + nSteps*dt, so dt = MIN(dtLim, tRemain) is dt for every step). */
void or_rk3_steps(or_model *m, int nSteps, double dt) {
  m->dt = dt;
  for (int n = 0; n < nSteps; n++) {
    double t0 = m->t;
    for (int s = 0; s < 3; s++) {
      or_calculate_tendency(m);
      or_update_grk3(m, s);
      m->t = t0 + rk3_b[s] * m->dt;
    }
    m->t = t0 + m->dt;
  }
}

/* CalculateEntropy: SW src/models/SELF_ShallowWater.f90:149-188; LSW
 * SELF_LinearShallowWater.f90:171-216; Euler SELF_CompressibleIdealGas2D.f90:782-823 */
double or_entropy(const or_model *m) {
  const int N = m->N, nVar = m->nVar;
  const double *J = m->geom->J, *w = m->qWeights;
  double ent = 0.0;
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++) {
        size_t n1 = i + (N + 1) * (j + (size_t)(N + 1) * e);
        double jac = J[n1], wi = w[i], wj = w[j];
        if (m->kind == OR_MODEL_SW) {
          double Hu = m->sol[SC(i, j, 0, e)], Hv = m->sol[SC(i, j, 1, e)], H = m->sol[SC(i, j, 2, e)];
          double b = m->topo[n1];
          ent = ent + (0.5 * (Hu * Hu / H + Hv * Hv / H) + 0.5 * m->g * H * H - m->g * H * b) * jac * wi * wj;
        } else if (m->kind == OR_MODEL_LSW) {
          double u = m->sol[SC(i, j, 0, e)], v = m->sol[SC(i, j, 1, e)], eta = m->sol[SC(i, j, 2, e)];
          double H = m->topo[n1];
          ent = ent + 0.5 * (H * (u * u + v * v) + m->g * eta * eta) * jac * wi * wj;
        } else {
          /* Euler: pressure from primitive(4) as left by the last PreTendency */
          double rho = m->sol[SC(i, j, 2, e)], P = m->prim[SC(i, j, 3, e)];
          ent = ent + rho * (log(P) - m->gamma * log(rho)) / (m->gamma - 1.0) * wi * wj * jac;
        }
      }
  return ent;
}

/* Mass: sum w_i w_j J f(var) (no reference routine; Integral_MappedScalar2D pattern,
 * src/SELF_MappedData.f90:1270-1311, with the variable as an argument) */
double or_mass(const or_model *m, int var) {
  const int N = m->N, nVar = m->nVar;
  double s = 0.0;
  for (int e = 0; e < m->nElem; e++)
    for (int j = 0; j <= N; j++)
      for (int i = 0; i <= N; i++)
        s = s + m->sol[SC(i, j, var, e)] * m->qWeights[i] * m->qWeights[j] *
                    m->geom->J[i + (N + 1) * (j + (size_t)(N + 1) * e)];
  return s;
}

/* ===================================================================================== */
/* Accessors for the Python test harness                                                 */
/* ===================================================================================== */
double *or_model_field(or_model *m, const char *name) {
  struct { const char *n; double *p; } t[] = {
      {"sol", m->sol}, {"solB", m->solB}, {"solE", m->solE}, {"workSol", m->workSol}, {"dSdt", m->dSdt},
      {"source", m->source}, {"fluxDiv", m->fluxDiv}, {"flux", m->flux}, {"fluxBN", m->fluxBN},
      {"vel", m->vel}, {"velB", m->velB}, {"velE", m->velE}, {"topo", m->topo}, {"topoB", m->topoB},
      {"gradTopo", m->gradTopo}, {"fCori", m->fCori}, {"prim", m->prim}, {"primB", m->primB},
      {"primE", m->primE}, {"diag", m->diag}, {"diagB", m->diagB}, {"diagE", m->diagE}, {"ent", m->ent},
      {"entB", m->entB}, {"entE", m->entE}, {"env", m->env}, {"envGrad", m->envGrad},
      {"presSolB", m->presSolB}, {"presPrimB", m->presPrimB}, {"presDiagB", m->presDiagB},
      {"qWeights", m->qWeights}, {"dMatrix", m->dMatrix}, {"dgMatrix", m->dgMatrix}, {"bMatrix", m->bMatrix}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); k++)
    if (strcmp(t[k].n, name) == 0) return t[k].p;
  return NULL;
}

void or_model_set_scalar(or_model *m, const char *name, double v) {
  if (!strcmp(name, "g")) m->g = v;
  else if (!strcmp(name, "fCori")) m->fCoriConst = v;
  else if (!strcmp(name, "gamma")) m->gamma = v;
  else if (!strcmp(name, "R")) m->Rgas = v;
  else if (!strcmp(name, "t")) m->t = v;
  else if (!strcmp(name, "dt")) m->dt = v;
}

double or_model_get_scalar(or_model *m, const char *name) {
  if (!strcmp(name, "g")) return m->g;
  if (!strcmp(name, "fCori")) return m->fCoriConst;
  if (!strcmp(name, "gamma")) return m->gamma;
  if (!strcmp(name, "R")) return m->Rgas;
  if (!strcmp(name, "t")) return m->t;
  if (!strcmp(name, "dt")) return m->dt;
  return NAN;
}

double *or_geometry_field(or_geometry *g, const char *name) {
  struct { const char *n; double *p; } t[] = {
      {"x", g->x}, {"xb", g->xb}, {"dxds", g->dxds}, {"dxdsb", g->dxdsb}, {"dsdx", g->dsdx},
      {"dsdxb", g->dsdxb}, {"nHatb", g->nHatb}, {"nScaleb", g->nScaleb}, {"J", g->J}, {"Jb", g->Jb}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); k++)
    if (strcmp(t[k].n, name) == 0) return t[k].p;
  return NULL;
}
