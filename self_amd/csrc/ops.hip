// Per-operator HIP kernels: the reference's operator set (Lagrange interpolation and
// derivatives, side exchange, mapped projection/gradient/Jacobian weight, RK register
// update) as standalone device kernels.  They serve the operator-level API (one call per
// reference method), the geometry setup and the curvilinear fallback; the hot path is the
// fused stage kernel in dg2d_fused.hip.
//
// Mapping: one work-item per output value, 256-wide blocks, the fastest index of the
// output (the node index i) on the lane index so every load/store of a variable's P*P
// block is coalesced.  Each work-item accumulates its contraction in registers and
// writes once (the reference kernels accumulate into global memory inside the loop).
#include "selfamd_common.hpp"
#include "../../include/self_amd.h"

using namespace selfamd;

namespace {

constexpr unsigned kBlock = 256;

// ---------------------------------------------------------------------------------
// Lagrange operators
// ---------------------------------------------------------------------------------

// ScalarBoundaryInterp_2D (src/SELF_Lagrange.f90:2241-2270).  Work-item = (i, v, s, e):
// S/N sides contract along j, E/W along i.  ncomp > 1 handles vector/tensor components
// stored innermost (component c of node (i,j) at c + ncomp*SC).
__global__ void k_boundary_interp(const double *__restrict__ bM, const double *__restrict__ f,
                                  double *__restrict__ fb, int N, int nVar, int nEl, int ncomp) {
  const size_t P = N + 1;
  const size_t total = (size_t)ncomp * P * nVar * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int c = r % ncomp; r /= ncomp;
    const int i = r % P; r /= P;
    const int v = r % nVar; r /= nVar;
    const int s = r % 4;
    const size_t e = r / 4;
    const double *b = bM + (s == 1 || s == 2 ? P : 0);
    double acc = 0.0;
    for (int ii = 0; ii <= N; ii++) {
      const size_t node = (s == 0 || s == 2) ? sc_idx(i, ii, v, e, P, nVar) : sc_idx(ii, i, v, e, P, nVar);
      acc += b[ii] * f[c + ncomp * node];
    }
    fb[t] = acc;
  }
}

// GridInterp_2D (src/SELF_Lagrange.f90:741-876): fNew(i,j) = sum_jj (sum_ii f(ii,jj) I(ii,i)) I(jj,j)
__global__ void k_grid_interp(const double *__restrict__ iM, const double *__restrict__ f,
                              double *__restrict__ fNew, int N, int M, int nVar, int nEl, int ncomp) {
  const size_t P = N + 1, Q = M + 1;
  const size_t total = (size_t)ncomp * Q * Q * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int c = r % ncomp; r /= ncomp;
    const int i = r % Q; r /= Q;
    const int j = r % Q; r /= Q;
    const int v = r % nVar;
    const size_t e = r / nVar;
    double fij = 0.0;
    for (int jj = 0; jj <= N; jj++) {
      double fi = 0.0;
      for (int ii = 0; ii <= N; ii++) fi += f[c + ncomp * sc_idx(ii, jj, v, e, P, nVar)] * iM[ii + P * i];
      fij += fi * iM[jj + P * j];
    }
    fNew[t] = fij;
  }
}

// VectorGradient_2D (src/SELF_Lagrange.f90:1424-1463): gradF(r,c) = d f(r) / d xi_c
__global__ void k_vector_gradient(const double *__restrict__ dM, const double *__restrict__ f,
                                  double *__restrict__ g, int N, int nVar, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * P * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int i = r % P; r /= P;
    const int j = r % P; r /= P;
    const int v = r % nVar;
    const size_t e = r / nVar;
    double g11 = 0, g21 = 0, g12 = 0, g22 = 0;
    for (int ii = 0; ii <= N; ii++) {
      const double di = dM[ii + P * i], dj = dM[ii + P * j];
      const size_t a = 2 * sc_idx(ii, j, v, e, P, nVar), b = 2 * sc_idx(i, ii, v, e, P, nVar);
      g11 += di * f[a];
      g21 += di * f[a + 1];
      g12 += dj * f[b];
      g22 += dj * f[b + 1];
    }
    g[4 * t + 0] = g11; g[4 * t + 1] = g21; g[4 * t + 2] = g12; g[4 * t + 3] = g22;
  }
}

// VectorDivergence_2D (strong, :1549-1574) and VectorDGDivergence_2D (weak, :1589-1625).
template <bool WEAK>
__global__ void k_divergence(const double *__restrict__ dM, const double *__restrict__ bM,
                             const double *__restrict__ qW, const double *__restrict__ f,
                             const double *__restrict__ bF, double *__restrict__ dF, int N, int nVar, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * P * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int i = r % P; r /= P;
    const int j = r % P; r /= P;
    const int v = r % nVar;
    const size_t e = r / nVar;
    double d = 0.0;
    for (int ii = 0; ii <= N; ii++)
      d = d + dM[ii + P * i] * f[2 * sc_idx(ii, j, v, e, P, nVar)] + dM[ii + P * j] * f[2 * sc_idx(i, ii, v, e, P, nVar) + 1];
    if (WEAK) {
      d = d + (bM[i + P] * bF[scb_idx(j, v, 1, e, P, nVar)] + bM[i] * bF[scb_idx(j, v, 3, e, P, nVar)]) / qW[i] +
          (bM[j + P] * bF[scb_idx(i, v, 2, e, P, nVar)] + bM[j] * bF[scb_idx(i, v, 0, e, P, nVar)]) / qW[j];
    }
    dF[t] = d;
  }
}

// Determinant_Tensor2D (src/SELF_Data.f90:1570-1607)
__global__ void k_determinant(const double *__restrict__ tsr, double *__restrict__ det, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    det[t] = tsr[4 * t] * tsr[4 * t + 3] - tsr[4 * t + 2] * tsr[4 * t + 1];
}

// ---------------------------------------------------------------------------------
// Mapped operators
// ---------------------------------------------------------------------------------

// SideExchange (src/SELF_MappedData.f90:834-917, 1870-1958): work-item = (c, i, v, s, e)
__global__ void k_side_exchange(const double *__restrict__ bnd, double *__restrict__ ext,
                                const int32_t *__restrict__ si, int N, int nVar, int nEl, int ncomp) {
  const size_t P = N + 1;
  const size_t total = (size_t)ncomp * P * nVar * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int c = r % ncomp; r /= ncomp;
    const int i1 = r % P; r /= P;
    const int v = r % nVar; r /= nVar;
    const int s1 = r % 4;
    const size_t e1 = r / 4;
    const int32_t *s = si + 5 * (s1 + 4 * e1);
    if (s[4] != 0) continue;
    const int s2 = s[3] / 10, flip = s[3] - 10 * s2;
    if (flip != 0 && flip != 1) continue;
    if (s[2] < 1 || s[2] > nEl || s2 < 1 || s2 > 4) continue;  // malformed side: never read out of range
    const int i2 = flip == 0 ? i1 : N - i1;
    ext[t] = bnd[c + ncomp * scb_idx(i2, v, s2 - 1, (size_t)(s[2] - 1), P, nVar)];
  }
}

// GradientSF_MappedScalar2D (src/SELF_MappedData.f90:1080-1149)
__global__ void k_gradient_sf(const double *__restrict__ dM, const double *__restrict__ dsdx,
                              const double *__restrict__ J, const double *__restrict__ f,
                              double *__restrict__ gradF, int N, int nVar, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * P * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t r = t;
    const int i = r % P; r /= P;
    const int j = r % P; r /= P;
    const int v = r % nVar;
    const size_t e = r / nVar;
    double gx = 0.0, gy = 0.0;
    for (int ii = 0; ii <= N; ii++) {
      const size_t g1 = 4 * (ii + P * (j + P * e)), g2 = 4 * (i + P * (ii + P * e));
      const double fa = f[sc_idx(ii, j, v, e, P, nVar)], fb = f[sc_idx(i, ii, v, e, P, nVar)];
      gx = gx + dM[ii + P * i] * (fa * dsdx[g1 + 0]) + dM[ii + P * j] * (fb * dsdx[g2 + 2]);
      gy = gy + dM[ii + P * i] * (fa * dsdx[g1 + 1]) + dM[ii + P * j] * (fb * dsdx[g2 + 3]);
    }
    const double jac = J[i + P * (j + P * e)];
    gradF[2 * t] = gx / jac;
    gradF[2 * t + 1] = gy / jac;
  }
}

// JacobianWeight (src/SELF_MappedData.f90:1235-1268)
__global__ void k_jacobian_weight(const double *__restrict__ J, double *__restrict__ f, int N, int nVar, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / (PP * nVar);
    f[t] = f[t] / J[node + PP * e];
  }
}

// ContravariantProjection (src/SELF_MappedData.f90:2220-2269)
__global__ void k_contravariant_projection(const double *__restrict__ dsdx, double *__restrict__ f, int N, int nVar,
                                           int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nVar * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / (PP * nVar);
    const double *g = dsdx + 4 * (node + PP * e);
    const double Fx = f[2 * t], Fy = f[2 * t + 1];
    f[2 * t] = g[0] * Fx + g[1] * Fy;
    f[2 * t + 1] = g[2] * Fx + g[3] * Fy;
  }
}

// UpdateGRK_Model2D (src/SELF_Model2D.f90:710-752)
__global__ void k_update_grk(double *__restrict__ G, double *__restrict__ S, const double *__restrict__ dSdt,
                             double a, double gdt, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const double g = a * G[t] + dSdt[t];
    G[t] = g;
    S[t] = S[t] + gdt * g;
  }
}

// dSdt = source - fluxDivergence (src/SELF_Model2D.f90:843-855)
__global__ void k_dsdt(const double *__restrict__ fd, const double *__restrict__ src, double *__restrict__ d, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    d[t] = src[t] - fd[t];
}

// Contravariant basis, outward normals and face scale (src/SELF_Geometry.f90:437-526):
// work-item = (i, s, e) on the boundary, after dsdx has been interpolated to the faces.
__global__ void k_geom_normals(double *__restrict__ dsdxb, const double *__restrict__ Jb, double *__restrict__ nHat,
                               double *__restrict__ nScale, int N, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int s = (t / P) % 4;
    const double sg = Jb[t] >= 0.0 ? 1.0 : -1.0;
    const double fac = (s == 1 || s == 2) ? sg : -sg;
    const int c = (s == 0 || s == 2) ? 1 : 0;  // S/N: contravariant vector 2; E/W: vector 1
    double *g = dsdxb + 4 * t;
    const double a0 = g[2 * c], a1 = g[2 * c + 1];
    const double mag = sqrt(a0 * a0 + a1 * a1);
    nScale[t] = mag;
    nHat[2 * t] = fac * a0 / mag;
    nHat[2 * t + 1] = fac * a1 / mag;
    for (int k = 0; k < 4; k++) g[k] = g[k] * fac;
  }
}

__global__ void k_dsdx_from_dxds(const double *__restrict__ dxds, double *__restrict__ dsdx, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const double *a = dxds + 4 * t;
    double *b = dsdx + 4 * t;
    b[0] = a[3];   // dsdx(1,1) =  dxds(2,2)
    b[1] = -a[2];  // dsdx(2,1) = -dxds(1,2)
    b[2] = -a[1];  // dsdx(1,2) = -dxds(2,1)
    b[3] = a[0];   // dsdx(2,2) =  dxds(1,1)
  }
}

// ---------------------------------------------------------------------------------
// Physics operators
// ---------------------------------------------------------------------------------

// SW PreTendency velocity (src/models/SELF_ShallowWater.f90:206-235)
__global__ void k_sw_velocity(const double *__restrict__ sol, const double *__restrict__ solB, double *__restrict__ vel,
                              double *__restrict__ velB, int N, int nEl) {
  const size_t P = N + 1, PP = P * P;
  const size_t nInt = PP * nEl, nBnd = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nInt + nBnd; t += (size_t)gridDim.x * blockDim.x) {
    if (t < nInt) {
      const size_t node = t % PP, e = t / PP;
      const double H = sol[node + PP * (2 + 3 * e)];
      vel[2 * t] = sol[node + PP * (3 * e)] / H;
      vel[2 * t + 1] = sol[node + PP * (1 + 3 * e)] / H;
    } else {
      const size_t b = t - nInt;
      const size_t i = b % P, se = b / P;  // se = s + 4e
      const double H = solB[i + P * (2 + 3 * se)];
      velB[2 * b] = solB[i + P * (3 * se)] / H;
      velB[2 * b + 1] = solB[i + P * (1 + 3 * se)] / H;
    }
  }
}

// Boundary conditions on physical sides (e2 == 0), work-item = (i, s, e).
// kind: 1 LSW (src/models/SELF_LinearShallowWater.f90:298-355), 2 SW (quirk: radiation
// sets the external thickness to the topography trace, SELF_ShallowWater.f90:284-343).
__global__ void k_sw_bc(const double *__restrict__ solB, double *__restrict__ solE, const double *__restrict__ topoB,
                        const double *__restrict__ nHat, const int32_t *__restrict__ si, int N, int nEl, int kind) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % P, se = t / P;
    const int32_t *s = si + 5 * se;
    if (s[2] != 0) continue;
    const double nx = nHat[2 * t], ny = nHat[2 * t + 1];
    const size_t o0 = i + P * (3 * se), o1 = o0 + P, o2 = o0 + 2 * P;
    if (s[4] == SELFAMD_BC_NONORMALFLOW) {
      const double u = solB[o0], v = solB[o1];
      solE[o0] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
      solE[o1] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
      solE[o2] = solB[o2];
    } else {
      solE[o0] = 0.0;
      solE[o1] = 0.0;
      solE[o2] = kind == 2 ? topoB[t] : 0.0;
    }
  }
}

// SW source (src/models/SELF_ShallowWater.f90:345-384)
__global__ void k_sw_source(double *__restrict__ src, const double *__restrict__ sol, const double *__restrict__ gradH,
                            double g, double f, int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 3 * e, o1 = o0 + PP, o2 = o1 + PP;
    src[o0] = -f * sol[o1] + g * sol[o2] * gradH[2 * t];
    src[o1] = f * sol[o0] + g * sol[o2] * gradH[2 * t + 1];
    src[o2] = 0.0;
  }
}

// SW flux (src/models/SELF_ShallowWater.f90:386-449)
__global__ void k_sw_flux(double *__restrict__ flux, const double *__restrict__ sol, const double *__restrict__ vel,
                          double g, int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 3 * e, o1 = o0 + PP, o2 = o1 + PP;
    const double u = vel[2 * t], v = vel[2 * t + 1], Hu = sol[o0], Hv = sol[o1], H = sol[o2];
    flux[2 * o0] = u * Hu + 0.5 * g * H * H;
    flux[2 * o0 + 1] = v * Hu;
    flux[2 * o1] = u * Hv;
    flux[2 * o1 + 1] = v * Hv + 0.5 * g * H * H;
    flux[2 * o2] = Hu;
    flux[2 * o2 + 1] = Hv;
  }
}

// SW Rusanov flux (src/models/SELF_ShallowWater.f90:451-529)
__global__ void k_sw_riemann(double *__restrict__ fbn, const double *__restrict__ solB, const double *__restrict__ solE,
                             const double *__restrict__ velB, const double *__restrict__ velE,
                             const double *__restrict__ nHat, const double *__restrict__ nScale, double g, int N,
                             int nEl) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % P, se = t / P;
    const double nx = nHat[2 * t], ny = nHat[2 * t + 1], nmag = nScale[t];
    const double unL = velB[2 * t] * nx + velB[2 * t + 1] * ny;
    const double unR = velE[2 * t] * nx + velE[2 * t + 1] * ny;
    const size_t o0 = i + P * (3 * se), o1 = o0 + P, o2 = o0 + 2 * P;
    const double HuL = solB[o0], HuR = solE[o0], HvL = solB[o1], HvR = solE[o1], HL = solB[o2], HR = solE[o2];
    const double cL = sqrt(g * HL), cR = sqrt(g * HR);
    const double fL0 = HuL * unL + g * HL * HL * 0.5 * nx, fL1 = HvL * unL + g * HL * HL * 0.5 * ny;
    const double fL2 = HuL * nx + HvL * ny;
    const double fR0 = HuR * unR + g * HR * HR * 0.5 * nx, fR1 = HvR * unR + g * HR * HR * 0.5 * ny;
    const double fR2 = HuR * nx + HvR * ny;
    const double a = fmax(fmax(fabs(unL + cL), fabs(unR + cR)), fmax(fabs(unL - cL), fabs(unR - cR)));
    fbn[o0] = 0.5 * (fL0 + fR0 + a * (HuL - HuR)) * nmag;
    fbn[o1] = 0.5 * (fL1 + fR1 + a * (HvL - HvR)) * nmag;
    fbn[o2] = 0.5 * (fL2 + fR2 + a * (HL - HR)) * nmag;
  }
}

// LSW source (src/models/SELF_LinearShallowWater.f90:413-446)
__global__ void k_lsw_source(double *__restrict__ src, const double *__restrict__ sol, const double *__restrict__ fc,
                             int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 3 * e, o1 = o0 + PP, o2 = o1 + PP;
    src[o0] = fc[t] * sol[o1];
    src[o1] = -fc[t] * sol[o0];
    src[o2] = 0.0;
  }
}

// LSW flux (src/models/SELF_LinearShallowWater.f90:448-498)
__global__ void k_lsw_flux(double *__restrict__ flux, const double *__restrict__ sol, const double *__restrict__ H,
                           double g, int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 3 * e, o1 = o0 + PP, o2 = o1 + PP;
    flux[2 * o0] = g * sol[o2];
    flux[2 * o0 + 1] = 0.0;
    flux[2 * o1] = 0.0;
    flux[2 * o1 + 1] = g * sol[o2];
    flux[2 * o2] = H[t] * sol[o0];
    flux[2 * o2 + 1] = H[t] * sol[o1];
  }
}

// LSW characteristic upwind flux (src/models/SELF_LinearShallowWater.f90:500-557)
__global__ void k_lsw_riemann(double *__restrict__ fbn, const double *__restrict__ solB, const double *__restrict__ solE,
                              const double *__restrict__ Hb, const double *__restrict__ nHat,
                              const double *__restrict__ nScale, double g, int N, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % P, se = t / P;
    const double nx = nHat[2 * t], ny = nHat[2 * t + 1], nmag = nScale[t];
    const double c = sqrt(g * Hb[t]);
    const size_t o0 = i + P * (3 * se), o1 = o0 + P, o2 = o0 + 2 * P;
    const double unL = solB[o0] * nx + solB[o1] * ny, unR = solE[o0] * nx + solE[o1] * ny;
    const double wL = 0.5 * (unL / g + solB[o2] / c), wR = 0.5 * (unR / g - solE[o2] / c);
    fbn[o0] = g * c * (wL - wR) * nx * nmag;
    fbn[o1] = g * c * (wL - wR) * ny * nmag;
    fbn[o2] = c * c * (wL + wR) * nmag;
  }
}

// Euler node-local diagnostics (src/models/SELF_CompressibleIdealGas2D.f90:984-1215).
// what: 0 diagnostics (KE, rho*E+p, c, T), 1 primitive (u, v, rho, p), 2 entropy variables.
__global__ void k_euler_nodal(const double *__restrict__ sol, double *__restrict__ out, double gam, double R, int N,
                              int nEl, int what) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 4 * e, o1 = o0 + PP, o2 = o1 + PP, o3 = o2 + PP;
    const double rhoU = sol[o0], rhoV = sol[o1], rho = sol[o2], rhoE = sol[o3];
    const double u = rhoU / rho, v = rhoV / rho;
    const double ke = 0.5 * (rhoU * u + rhoV * v);
    const double p = (gam - 1.0) * (rhoE - ke);
    if (what == 0) {
      out[o0] = ke;
      out[o1] = rhoE + p;
      out[o2] = sqrt(gam * p / rho);
      out[o3] = (2.0 / 3.0) * ((rhoE - ke) / rho) / R;
    } else if (what == 1) {
      out[o0] = u; out[o1] = v; out[o2] = rho; out[o3] = p;
    } else {
      const double s = log(p) - gam * log(rho);
      out[o0] = u * rho / p;
      out[o1] = v * rho / p;
      out[o2] = (gam - s) / (gam - 1.0) - ke / p;
      out[o3] = -rho / p;
    }
  }
}

// Euler boundary conditions (src/models/SELF_CompressibleIdealGas2D.f90:1217-1323)
__global__ void k_euler_bc(const double *__restrict__ solB, double *__restrict__ solE, const double *__restrict__ pSolB,
                           const double *__restrict__ primB, double *__restrict__ primE,
                           const double *__restrict__ pPrimB, const double *__restrict__ diagB,
                           double *__restrict__ diagE, const double *__restrict__ pDiagB,
                           const double *__restrict__ nHat, const int32_t *__restrict__ si, int N, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % P, se = t / P;
    const int32_t *s = si + 5 * se;
    if (s[2] != 0) continue;
    const size_t o[4] = {i + P * (4 * se), i + P * (1 + 4 * se), i + P * (2 + 4 * se), i + P * (3 + 4 * se)};
    if (s[4] == SELFAMD_BC_NONORMALFLOW) {
      const double nx = nHat[2 * t], ny = nHat[2 * t + 1];
      double u = solB[o[0]], v = solB[o[1]];
      solE[o[0]] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
      solE[o[1]] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
      solE[o[2]] = solB[o[2]];
      solE[o[3]] = solB[o[3]];
      u = primB[o[0]]; v = primB[o[1]];
      primE[o[0]] = (ny * ny - nx * nx) * u - 2.0 * nx * ny * v;
      primE[o[1]] = (nx * nx - ny * ny) * v - 2.0 * nx * ny * u;
      primE[o[2]] = primB[o[2]];
      primE[o[3]] = primB[o[3]];
      for (int k = 0; k < 4; k++) diagE[o[k]] = diagB[o[k]];
    } else if (s[4] == SELFAMD_BC_PRESCRIBED || s[4] == SELFAMD_BC_RADIATION) {
      for (int k = 0; k < 4; k++) {
        solE[o[k]] = pSolB[o[k]];
        primE[o[k]] = pPrimB[o[k]];
        diagE[o[k]] = pDiagB[o[k]];
      }
    }
  }
}

// Euler source, drag included (src/models/SELF_CompressibleIdealGas2D.f90:1325-1366;
// quirk: the energy source is -rhou*gx - rhou*gy).
__global__ void k_euler_source(double *__restrict__ src, const double *__restrict__ sol, const double *__restrict__ env,
                               const double *__restrict__ envGrad, int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 4 * e, o1 = o0 + PP, o2 = o1 + PP, o3 = o2 + PP;
    const double rhou = sol[o0], rhov = sol[o1], rho = sol[o2];
    const double gx = envGrad[2 * (node + PP * 2 * e)], gy = envGrad[2 * (node + PP * 2 * e) + 1];
    const double Cd = env[node + PP * (1 + 2 * e)];
    src[o0] = -rho * gx - Cd * rhou;
    src[o1] = -rho * gy - Cd * rhov;
    src[o2] = 0.0;
    src[o3] = -rhou * gx - rhou * gy;
  }
}

// Euler flux (src/models/SELF_CompressibleIdealGas2D.f90:1368-1440)
__global__ void k_euler_flux(double *__restrict__ flux, const double *__restrict__ sol, const double *__restrict__ prim,
                             const double *__restrict__ diag, int N, int nEl) {
  const size_t PP = (size_t)(N + 1) * (N + 1);
  const size_t total = PP * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t node = t % PP, e = t / PP;
    const size_t o0 = node + PP * 4 * e, o1 = o0 + PP, o2 = o1 + PP, o3 = o2 + PP;
    const double u = prim[o0], v = prim[o1], p = prim[o3], Hn = diag[o1];
    flux[2 * o0] = u * sol[o0] + p;
    flux[2 * o0 + 1] = v * sol[o0];
    flux[2 * o1] = u * sol[o1];
    flux[2 * o1 + 1] = v * sol[o1] + p;
    flux[2 * o2] = sol[o0];
    flux[2 * o2 + 1] = sol[o1];
    flux[2 * o3] = u * Hn;
    flux[2 * o3 + 1] = v * Hn;
  }
}

// Euler naive local Lax-Friedrichs flux (src/models/SELF_CompressibleIdealGas2D.f90:1455-1542)
__global__ void k_euler_riemann(double *__restrict__ fbn, const double *__restrict__ solB,
                                const double *__restrict__ solE, const double *__restrict__ primB,
                                const double *__restrict__ primE, const double *__restrict__ diagB,
                                const double *__restrict__ diagE, const double *__restrict__ nHat,
                                const double *__restrict__ nScale, int N, int nEl) {
  const size_t P = N + 1;
  const size_t total = P * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % P, se = t / P;
    const double nx = nHat[2 * t], ny = nHat[2 * t + 1], nmag = nScale[t];
    const size_t o[4] = {i + P * (4 * se), i + P * (1 + 4 * se), i + P * (2 + 4 * se), i + P * (3 + 4 * se)};
    const double unL = primB[o[0]] * nx + primB[o[1]] * ny, unR = primE[o[0]] * nx + primE[o[1]] * ny;
    double fL[4], fR[4];
    fL[0] = unL * solB[o[0]] + primB[o[3]] * nx;
    fL[1] = unL * solB[o[1]] + primB[o[3]] * ny;
    fL[2] = solB[o[0]] * nx + solB[o[1]] * ny;
    fL[3] = unL * diagB[o[1]];
    fR[0] = unR * solE[o[0]] + primE[o[3]] * nx;
    fR[1] = unR * solE[o[1]] + primE[o[3]] * ny;
    fR[2] = solE[o[0]] * nx + solE[o[1]] * ny;
    fR[3] = unR * diagE[o[1]];
    const double cL = diagB[o[2]], cR = diagE[o[2]];
    const double a = fmax(fabs(unL), fabs(unR)) + fmax(fabs(cL), fabs(cR));
    for (int k = 0; k < 4; k++) fbn[o[k]] = 0.5 * (fL[k] + fR[k] + a * (solB[o[k]] - solE[o[k]])) * nmag;
  }
}

// Streaming copy dst = src (16-byte vector loads/stores): calibrates HBM counters and the
// achievable bandwidth against kernels of known traffic.  Each work-item issues four
// independent loads before its stores, so that enough bytes are in flight per CU to cover
// the HBM latency, and every access is non-temporal (nothing is reused): 5.2 TB/s read +
// write over 1 GiB on MI355X, against 4.5 TB/s with cached accesses.
__device__ inline double2 nt_load2(const double2 *p) {
  return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ inline void nt_store2(double2 *p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (; t + 3 * stride < n2; t += 4 * stride) {
    const double2 x0 = nt_load2(src + t), x1 = nt_load2(src + t + stride);
    const double2 x2 = nt_load2(src + t + 2 * stride), x3 = nt_load2(src + t + 3 * stride);
    nt_store2(dst + t, x0);
    nt_store2(dst + t + stride, x1);
    nt_store2(dst + t + 2 * stride, x2);
    nt_store2(dst + t + 3 * stride, x3);
  }
  for (; t < n2; t += stride) nt_store2(dst + t, nt_load2(src + t));
}

inline unsigned blocks_for(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > (1u << 20)) b = 1u << 20;  // grid-stride loops cover the rest
  return (unsigned)(b ? b : 1);
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define CHECK_DIMS(N, nVar, nEl)                                                                  \
  SELFAMD_REQUIRE((N) >= 1 && (N) <= kMaxN && (nVar) >= 1 && (nEl) >= 1,                          \
                  "%s: bad sizes N=%d nVar=%d nEl=%d", __func__, (int)(N), (int)(nVar), (int)(nEl))

}  // namespace

extern "C" {

int selfamd_scalar_boundary_interp_2d(const double *bM, const double *f, double *fb, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * nVar * 4 * nEl;
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(n), kBlock, 0, S(s), bM, f, fb, N, nVar, nEl, 1);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_vector_boundary_interp_2d(const double *bM, const double *f, double *fb, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)2 * (N + 1) * nVar * 4 * nEl;
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(n), kBlock, 0, S(s), bM, f, fb, N, nVar, nEl, 2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_tensor_boundary_interp_2d(const double *bM, const double *f, double *fb, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)4 * (N + 1) * nVar * 4 * nEl;
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(n), kBlock, 0, S(s), bM, f, fb, N, nVar, nEl, 4);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_scalar_grid_interp_2d(const double *iM, const double *f, double *fNew, int N, int M, int nVar, int nEl,
                                  void *s) {
  CHECK_DIMS(N, nVar, nEl);
  SELFAMD_REQUIRE(M >= 1 && M <= 64, "scalar_grid_interp_2d: M=%d", M);
  const size_t n = (size_t)(M + 1) * (M + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_grid_interp, blocks_for(n), kBlock, 0, S(s), iM, f, fNew, N, M, nVar, nEl, 1);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_vector_grid_interp_2d(const double *iM, const double *f, double *fNew, int N, int M, int nVar, int nEl,
                                  void *s) {
  CHECK_DIMS(N, nVar, nEl);
  SELFAMD_REQUIRE(M >= 1 && M <= 64, "vector_grid_interp_2d: M=%d", M);
  const size_t n = (size_t)2 * (M + 1) * (M + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_grid_interp, blocks_for(n), kBlock, 0, S(s), iM, f, fNew, N, M, nVar, nEl, 2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_vector_gradient_2d(const double *dM, const double *f, double *g, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_vector_gradient, blocks_for(n), kBlock, 0, S(s), dM, f, g, N, nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_vector_divergence_2d(const double *dM, const double *f, double *dF, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_divergence<false>, blocks_for(n), kBlock, 0, S(s), dM, nullptr, nullptr, f, nullptr, dF, N,
                     nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_vector_dg_divergence_2d(const double *dgM, const double *bM, const double *qW, const double *f,
                                    const double *bF, double *dF, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_divergence<true>, blocks_for(n), kBlock, 0, S(s), dgM, bM, qW, f, bF, dF, N, nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_determinant_tensor_2d(const double *t, double *det, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_determinant, blocks_for(n), kBlock, 0, S(s), t, det, n);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_side_exchange_scalar_2d(const double *b, double *ext, const int32_t *si, int N, int nVar, int nEl,
                                    void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * nVar * 4 * nEl;
  hipLaunchKernelGGL(k_side_exchange, blocks_for(n), kBlock, 0, S(s), b, ext, si, N, nVar, nEl, 1);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_side_exchange_vector_2d(const double *b, double *ext, const int32_t *si, int N, int nVar, int nEl,
                                    void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)2 * (N + 1) * nVar * 4 * nEl;
  hipLaunchKernelGGL(k_side_exchange, blocks_for(n), kBlock, 0, S(s), b, ext, si, N, nVar, nEl, 2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_gradient_sf_2d(const double *dM, const double *dsdx, const double *J, const double *f, double *gradF, int N,
                           int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_gradient_sf, blocks_for(n), kBlock, 0, S(s), dM, dsdx, J, f, gradF, N, nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_jacobian_weight_2d(const double *J, double *f, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_jacobian_weight, blocks_for(n), kBlock, 0, S(s), J, f, N, nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_contravariant_projection_2d(const double *dsdx, double *f, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_contravariant_projection, blocks_for(n), kBlock, 0, S(s), dsdx, f, N, nVar, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_update_grk(double *G, double *Sv, const double *dSdt, double a, double g, double dt, int N, int nVar,
                       int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_update_grk, blocks_for(n), kBlock, 0, S(s), G, Sv, dSdt, a, g * dt, n);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_copy_f64(const double *src, double *dst, size_t n, void *s) {
  SELFAMD_REQUIRE(src && dst && n % 2 == 0, "copy_f64: null pointer or odd length %zu", n);
  const size_t n2 = n / 2;
  unsigned blocks = (unsigned)((n2 + kBlock - 1) / kBlock);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy, blocks ? blocks : 1, kBlock, 0, S(s), reinterpret_cast<const double2 *>(src),
                     reinterpret_cast<double2 *>(dst), n2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}

int selfamd_calculate_dsdt(const double *fd, const double *src, double *d, int N, int nVar, int nEl, void *s) {
  CHECK_DIMS(N, nVar, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nVar * nEl;
  hipLaunchKernelGGL(k_dsdt, blocks_for(n), kBlock, 0, S(s), fd, src, d, n);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}

int selfamd_semquad_generate(int N, int ctrl, int nGeo, int nEl, const double *nodeCoords, double *x, double *xb,
                             double *dxds, double *dxdsb, double *dsdx, double *dsdxb, double *nHatb, double *nScaleb,
                             double *J, double *Jb, void *s) {
  CHECK_DIMS(N, 1, nEl);
  SELFAMD_REQUIRE(nGeo >= 1 && nGeo <= kMaxN, "semquad_generate: nGeo=%d", nGeo);
  const int P = N + 1, G = nGeo + 1;
  // meshToModel: control = uniform mesh nodes, target = model nodes; model matrices.
  double cp[kMaxN + 1], tp[kMaxN + 1], bw[kMaxN + 1], qw[kMaxN + 1], iMg[(kMaxN + 1) * (kMaxN + 1)],
      dMg[(kMaxN + 1) * (kMaxN + 1)], dgMg[(kMaxN + 1) * (kMaxN + 1)], bMg[2 * (kMaxN + 1)];
  int rc = selfamd_lagrange_init(nGeo, SELFAMD_UNIFORM, N, ctrl, cp, tp, bw, qw, iMg, dMg, dgMg, bMg);
  if (rc) return rc;
  double mcp[kMaxN + 1], mtp[kMaxN + 1], mbw[kMaxN + 1], mqw[kMaxN + 1], miM[(kMaxN + 1) * (kMaxN + 1)],
      mdM[(kMaxN + 1) * (kMaxN + 1)], mdgM[(kMaxN + 1) * (kMaxN + 1)], mbM[2 * (kMaxN + 1)];
  rc = selfamd_lagrange_init(N, ctrl, N, ctrl, mcp, mtp, mbw, mqw, miM, mdM, mdgM, mbM);
  if (rc) return rc;
  double *dbuf = nullptr;  // device copies: iMatrix (G x P), dMatrix (P x P), bMatrix (2P)
  SELFAMD_HIP(hipMalloc(&dbuf, sizeof(double) * (G * P + P * P + 2 * P)));
  double *d_iM = dbuf, *d_dM = dbuf + G * P, *d_bM = d_dM + P * P;
  hipStream_t st = S(s);
  SELFAMD_HIP(hipMemcpyAsync(d_iM, iMg, sizeof(double) * G * P, hipMemcpyHostToDevice, st));
  SELFAMD_HIP(hipMemcpyAsync(d_dM, mdM, sizeof(double) * P * P, hipMemcpyHostToDevice, st));
  SELFAMD_HIP(hipMemcpyAsync(d_bM, mbM, sizeof(double) * 2 * P, hipMemcpyHostToDevice, st));
  const size_t nn = (size_t)P * P * nEl, nb = (size_t)P * 4 * nEl;
  hipLaunchKernelGGL(k_grid_interp, blocks_for(2 * nn), kBlock, 0, st, d_iM, nodeCoords, x, nGeo, N, 1, nEl, 2);
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(2 * nb), kBlock, 0, st, d_bM, x, xb, N, 1, nEl, 2);
  hipLaunchKernelGGL(k_vector_gradient, blocks_for(nn), kBlock, 0, st, d_dM, x, dxds, N, 1, nEl);
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(4 * nb), kBlock, 0, st, d_bM, dxds, dxdsb, N, 1, nEl, 4);
  hipLaunchKernelGGL(k_determinant, blocks_for(nn), kBlock, 0, st, dxds, J, nn);
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(nb), kBlock, 0, st, d_bM, J, Jb, N, 1, nEl, 1);
  hipLaunchKernelGGL(k_dsdx_from_dxds, blocks_for(nn), kBlock, 0, st, dxds, dsdx, nn);
  hipLaunchKernelGGL(k_boundary_interp, blocks_for(4 * nb), kBlock, 0, st, d_bM, dsdx, dsdxb, N, 1, nEl, 4);
  hipLaunchKernelGGL(k_geom_normals, blocks_for(nb), kBlock, 0, st, dsdxb, Jb, nHatb, nScaleb, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  SELFAMD_HIP(hipStreamSynchronize(st));
  SELFAMD_HIP(hipFree(dbuf));
  return kOk;
}

int selfamd_sw2d_velocity(const double *sol, const double *solB, double *vel, double *velB, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl + (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_sw_velocity, blocks_for(n), kBlock, 0, S(s), sol, solB, vel, velB, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_sw2d_set_boundary_condition(const double *solB, double *solE, const double *topoB, const double *nHat,
                                        const int32_t *si, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_sw_bc, blocks_for(n), kBlock, 0, S(s), solB, solE, topoB, nHat, si, N, nEl, 2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_sw2d_source(double *src, const double *sol, const double *gradH, double g, double f, int N, int nEl,
                        void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_sw_source, blocks_for(n), kBlock, 0, S(s), src, sol, gradH, g, f, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_sw2d_flux(double *flux, const double *sol, const double *vel, double g, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_sw_flux, blocks_for(n), kBlock, 0, S(s), flux, sol, vel, g, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_sw2d_riemann(double *fbn, const double *solB, const double *solE, const double *velB, const double *velE,
                         const double *nHat, const double *nScale, double g, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_sw_riemann, blocks_for(n), kBlock, 0, S(s), fbn, solB, solE, velB, velE, nHat, nScale, g, N,
                     nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_lsw2d_set_boundary_condition(const double *solB, double *solE, const double *nHat, const int32_t *si, int N,
                                         int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_sw_bc, blocks_for(n), kBlock, 0, S(s), solB, solE, nullptr, nHat, si, N, nEl, 1);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_lsw2d_source(double *src, const double *sol, const double *fc, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_lsw_source, blocks_for(n), kBlock, 0, S(s), src, sol, fc, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_lsw2d_flux(double *flux, const double *sol, const double *H, double g, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_lsw_flux, blocks_for(n), kBlock, 0, S(s), flux, sol, H, g, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_lsw2d_riemann(double *fbn, const double *solB, const double *solE, const double *Hb, const double *nHat,
                          const double *nScale, double g, int N, int nEl, void *s) {
  CHECK_DIMS(N, 3, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_lsw_riemann, blocks_for(n), kBlock, 0, S(s), fbn, solB, solE, Hb, nHat, nScale, g, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_diagnostics(const double *sol, double *diag, double gam, double R, int N, int nEl, void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_euler_nodal, blocks_for(n), kBlock, 0, S(s), sol, diag, gam, R, N, nEl, 0);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_primitive(const double *sol, double *prim, double gam, int N, int nEl, void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_euler_nodal, blocks_for(n), kBlock, 0, S(s), sol, prim, gam, 1.0, N, nEl, 1);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_entropy_vars(const double *sol, double *ent, double gam, int N, int nEl, void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_euler_nodal, blocks_for(n), kBlock, 0, S(s), sol, ent, gam, 1.0, N, nEl, 2);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_set_boundary_condition(const double *solB, double *solE, const double *pSolB, const double *primB,
                                           double *primE, const double *pPrimB, const double *diagB, double *diagE,
                                           const double *pDiagB, const double *nHat, const int32_t *si, int N, int nEl,
                                           void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_euler_bc, blocks_for(n), kBlock, 0, S(s), solB, solE, pSolB, primB, primE, pPrimB, diagB, diagE,
                     pDiagB, nHat, si, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_source(double *src, const double *sol, const double *env, const double *envGrad, int N, int nEl,
                           void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_euler_source, blocks_for(n), kBlock, 0, S(s), src, sol, env, envGrad, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_flux(double *flux, const double *sol, const double *prim, const double *diag, int N, int nEl,
                         void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * (N + 1) * nEl;
  hipLaunchKernelGGL(k_euler_flux, blocks_for(n), kBlock, 0, S(s), flux, sol, prim, diag, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}
int selfamd_euler2d_riemann(double *fbn, const double *solB, const double *solE, const double *primB,
                            const double *primE, const double *diagB, const double *diagE, const double *nHat,
                            const double *nScale, int N, int nEl, void *s) {
  CHECK_DIMS(N, 4, nEl);
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  hipLaunchKernelGGL(k_euler_riemann, blocks_for(n), kBlock, 0, S(s), fbn, solB, solE, primB, primE, diagB, diagE,
                     nHat, nScale, N, nEl);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}

}  // extern "C"
