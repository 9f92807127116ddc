// Adjoint (vector-Jacobian product) of the nonlinear shallow-water tendency on affine
// periodic meshes, for training the learned closure through the solver.  No reference
// counterpart: the reference has no adjoint.
//
// The tendency is the one the fused stage kernel evaluates (CalculateTendency_Model2D,
// src/SELF_Model2D.f90:816-859, with the SW plug-ins of src/models/SELF_ShallowWater.f90):
//   dSdt = source(S) - J^-1 [ weak divergence of the contravariant flux F(S)
//                             + surface terms of the Rusanov flux F*(qL, qR) ]
// with qL the element's own face traces of S, qR the neighbour's, and the lagged face normal
// velocities un (the SW PreTendency lag, src/models/SELF_ShallowWater.f90:222-237), formed from
// the face traces of the previous stage's input state Sp.
// Given a cotangent lambda of dSdt, mu = (d dSdt / dS)^T lambda is formed by two kernels:
//   k_sw_vjp_local   per element: the transposed weak divergence, the flux Jacobians, the
//                    source, and the transposed Riemann flux; the part belonging to the own
//                    traces goes back to the element's nodes, the part belonging to the
//                    neighbour's traces (chiR) is stored per face node in the trace layout;
//   k_sw_vjp_gather  per node: the chiR the neighbours stored for this element's traces,
//                    spread back through the boundary interpolation;
// and, when asked, the lag path muP = (d dSdt / d Sp)^T lambda (k_sw_vjp_lag) from the
// cotangents of un that k_sw_vjp_local also stores.
// Layouts are the engine's: state (0:N,0:N,3,nEl), traces (0:N,3,4,nEl), lagged u.n of each
// element's own sides (0:N,4,nEl), 20-double affine element records, neighbour codes
// (e2 << 3) | (s2 << 1) | flip, matrices dgM(P*P) D(P*P) bM(2P) qW(P) 1/qW(P) with entry
// (c, r) of a P x P matrix at c + P r.
#include "sw_adjoint.hpp"

namespace selfamd {
namespace {

constexpr int kVjpThreads = 256;

template <int N>
__global__ void __launch_bounds__(kVjpThreads)
    k_sw_vjp_local(const double *__restrict__ S, const double *__restrict__ tr, const double *__restrict__ un,
                   const double *__restrict__ aux, const double *__restrict__ geom, const int32_t *__restrict__ nbr,
                   const double *__restrict__ mats, double g, double fCori, double auxConst,
                   const double *__restrict__ lam, double *__restrict__ mu, double *__restrict__ chiR,
                   double *__restrict__ chiUnL, double *__restrict__ chiUnR, int nEl) {
  constexpr int NP = N + 1, NN = NP * NP, EPB = kVjpThreads / NN, NF = 4 * NP;
  __shared__ double sK[3][EPB][NN];    // kappa = -lambda / J, cotangent of the bracket
  __shared__ double sAux[EPB][NN];     // static field (topography)
  __shared__ double sChi[3][EPB][NF];  // cotangent of the own face traces
  __shared__ double sM[2 * NN + 4 * NP];
  const int tid = threadIdx.x, el = tid / NN, node = tid % NN, i = node % NP, j = node / NP;
  const int e = blockIdx.x * EPB + el;
  const bool active = el < EPB && e < nEl;
  for (int k = tid; k < 2 * NN + 4 * NP; k += kVjpThreads) sM[k] = mats[k];
  const double *dgM = sM, *Dm = sM + NN, *b0 = sM + 2 * NN, *b1 = b0 + NP, *iw = sM + 2 * NN + 3 * NP;
  double s[3] = {1.0, 0.0, 0.0}, lm[3] = {0.0, 0.0, 0.0}, d[4] = {0, 0, 0, 0}, invJ = 0.0;
  if (active) {
    const double *gr = geom + (size_t)e * kGeomStride;
    for (int k = 0; k < 4; k++) d[k] = gr[k];
    invJ = gr[6];
    for (int v = 0; v < 3; v++) {
      s[v] = S[((size_t)e * 3 + v) * NN + node];
      lm[v] = lam[((size_t)e * 3 + v) * NN + node];
      sK[v][el][node] = -lm[v] * invJ;
    }
    sAux[el][node] = aux ? aux[(size_t)e * NN + node] : auxConst;
  }
  __syncthreads();

  // ---- nodes: transposed volume terms, flux Jacobians, source ----
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  if (active) {
    // forward: sum_ii dgM(ii + P i) F1(ii, j) + sum_ii dgM(ii + P j) F2(i, ii)
    double p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
    for (int a = 0; a < NP; a++) {
      const double ai = dgM[i + NP * a], aj = dgM[j + NP * a];
      for (int v = 0; v < 3; v++) {
        p1[v] += ai * sK[v][el][a + NP * j];
        p2[v] += aj * sK[v][el][i + NP * a];
      }
    }
    // F1 = d0 fx + d1 fy, F2 = d2 fx + d3 fy
    double fxb[3], fyb[3];
    for (int v = 0; v < 3; v++) {
      fxb[v] = d[0] * p1[v] + d[2] * p2[v];
      fyb[v] = d[1] * p1[v] + d[3] * p2[v];
    }
    const double Hu = s[0], Hv = s[1], H = s[2], u = Hu / H, w = Hv / H;
    // fx = (u Hu + g H^2/2, u Hv, Hu), fy = (w Hu, w Hv + g H^2/2, Hv)
    m0 = fxb[0] * 2.0 * u + fxb[1] * w + fxb[2] + fyb[0] * w;
    m1 = fxb[1] * u + fyb[0] * u + fyb[1] * 2.0 * w + fyb[2];
    m2 = fxb[0] * (g * H - u * u) - fxb[1] * u * w - fyb[0] * u * w + fyb[1] * (g * H - w * w);
    // source (-f Hv + g H gx, f Hu + g H gy, 0), gx, gy the affine GradientSF of the static field
    double dxi = 0.0, deta = 0.0;
    for (int a = 0; a < NP; a++) {
      dxi += Dm[a + NP * i] * sAux[el][a + NP * j];
      deta += Dm[a + NP * j] * sAux[el][i + NP * a];
    }
    const double gx = (d[0] * dxi + d[2] * deta) * invJ, gy = (d[1] * dxi + d[3] * deta) * invJ;
    m0 += fCori * lm[1];
    m1 -= fCori * lm[0];
    m2 += g * (gx * lm[0] + gy * lm[1]);
  }

  // ---- face nodes: transposed surface terms and Rusanov flux ----
  for (int ft = tid; ft < EPB * NF; ft += kVjpThreads) {
    const int fel = ft / NF, fnode = ft % NF, side = fnode / NP, fi = fnode % NP;
    const int fe = blockIdx.x * EPB + fel;
    if (fe >= nEl) continue;
    // forward surface term: b1(i)/w_i F*(E, j) + b0(i)/w_i F*(W, j) + b1(j)/w_j F*(N, i) + b0(j)/w_j F*(S, i)
    const double *bl = (side == 1 || side == 2) ? b1 : b0;
    const bool alongJ = side == 0 || side == 2;  // S, N: column i = fi; E, W: row j = fi
    double psi[3] = {0, 0, 0};
    for (int a = 0; a < NP; a++) {
      const double wgt = bl[a] * iw[a];
      for (int v = 0; v < 3; v++) psi[v] += wgt * (alongJ ? sK[v][fel][fi + NP * a] : sK[v][fel][a + NP * fi]);
    }
    const double *gr = geom + (size_t)fe * kGeomStride;
    const double nx = gr[8 + side], ny = gr[12 + side], nmag = gr[16 + side];
    const int code = nbr[4 * (size_t)fe + side];
    const int flip = code & 1, s2 = (code >> 1) & 3;
    const size_t e2 = (size_t)(code >> 3);
    const double *tl = tr + (size_t)fi + NP * (size_t)3 * (side + 4 * (size_t)fe);
    const double *trr = tr + (size_t)(flip ? N - fi : fi) + NP * (size_t)3 * (s2 + 4 * e2);
    const double HuL = tl[0], HvL = tl[NP], HL = tl[2 * NP];
    const double HuR = trr[0], HvR = trr[NP], HR = trr[2 * NP];
    const double unL = un[(size_t)fi + NP * (side + 4 * (size_t)fe)];
    const double unR = -un[(size_t)(flip ? N - fi : fi) + NP * (s2 + 4 * e2)];
    const double cL = sqrt(g * HL), cR = sqrt(g * HR);
    const double c0 = fabs(unL + cL), c1 = fabs(unR + cR), c2 = fabs(unL - cL), c3 = fabs(unR - cR);
    const double a = fmax(fmax(c0, c1), fmax(c2, c3));
    // derivative of the active branch of the maximum
    double daL = 0.0, daR = 0.0;
    const double dcL = 0.5 * g / cL, dcR = 0.5 * g / cR;
    if (a == c0) daL = (unL + cL >= 0.0 ? 1.0 : -1.0) * dcL;
    else if (a == c1) daR = (unR + cR >= 0.0 ? 1.0 : -1.0) * dcR;
    else if (a == c2) daL = -(unL - cL >= 0.0 ? 1.0 : -1.0) * dcL;
    else daR = -(unR - cR >= 0.0 ? 1.0 : -1.0) * dcR;
    // F*_v = (fL_v + fR_v + a (qL_v - qR_v)) nmag / 2
    const double q0 = 0.5 * nmag * psi[0], q1 = 0.5 * nmag * psi[1], q2 = 0.5 * nmag * psi[2];
    const double jump = q0 * (HuL - HuR) + q1 * (HvL - HvR) + q2 * (HL - HR);
    const double chiL[3] = {q0 * (unL + a) + q2 * nx, q1 * (unL + a) + q2 * ny,
                            q0 * g * HL * nx + q1 * g * HL * ny + q2 * a + jump * daL};
    const double chiRv[3] = {q0 * (unR - a) + q2 * nx, q1 * (unR - a) + q2 * ny,
                             q0 * g * HR * nx + q1 * g * HR * ny - q2 * a + jump * daR};
    for (int v = 0; v < 3; v++) {
      sChi[v][fel][fnode] = chiL[v];
      chiR[(size_t)fi + NP * ((size_t)v + 3 * (side + 4 * (size_t)fe))] = chiRv[v];
    }
    if (chiUnL) {
      // cotangents of the lagged normal velocities: fL0 = HuL unL + .., fL1 = HvL unL + .., and a
      double aL = 0.0, aR = 0.0;
      if (a == c0) aL = unL + cL >= 0.0 ? 1.0 : -1.0;
      else if (a == c1) aR = unR + cR >= 0.0 ? 1.0 : -1.0;
      else if (a == c2) aL = unL - cL >= 0.0 ? 1.0 : -1.0;
      else aR = unR - cR >= 0.0 ? 1.0 : -1.0;
      const size_t fb = (size_t)fi + NP * (side + 4 * (size_t)fe);
      chiUnL[fb] = q0 * HuL + q1 * HvL + jump * aL;
      chiUnR[fb] = q0 * HuR + q1 * HvR + jump * aR;
    }
  }
  __syncthreads();

  // ---- nodes: own traces qL(S, i) = sum_j b0(j) q(i, j), qL(N, i) with b1, qL(W, j) = sum_i
  // b0(i) q(i, j), qL(E, j) with b1; q = S for SW ----
  if (active) {
    const double t[3] = {m0, m1, m2};
    for (int v = 0; v < 3; v++) {
      const double add = b0[j] * sChi[v][el][0 * NP + i] + b1[j] * sChi[v][el][2 * NP + i] +
                         b1[i] * sChi[v][el][1 * NP + j] + b0[i] * sChi[v][el][3 * NP + j];
      mu[((size_t)e * 3 + v) * NN + node] = t[v] + add;
    }
  }
}

// mu += the neighbours' cotangents of this element's traces, through the boundary interpolation
template <int N>
__global__ void __launch_bounds__(kVjpThreads)
    k_sw_vjp_gather(const int32_t *__restrict__ nbr, const double *__restrict__ mats,
                    const double *__restrict__ chiR, double *__restrict__ mu, int nEl) {
  constexpr int NP = N + 1, NN = NP * NP;
  const size_t t = blockIdx.x * (size_t)kVjpThreads + threadIdx.x;
  if (t >= (size_t)nEl * NN) return;
  const size_t e = t / NN;
  const int node = (int)(t % NN), i = node % NP, j = node / NP;
  const double *b0 = mats + 2 * NN, *b1 = b0 + NP;
  double add[3] = {0, 0, 0};
  for (int side = 0; side < 4; side++) {
    const int code = nbr[4 * e + side];
    const int flip = code & 1, s2 = (code >> 1) & 3;
    const size_t e2 = (size_t)(code >> 3);
    const int fi = (side == 0 || side == 2) ? i : j;
    const double wgt = side == 0 ? b0[j] : side == 2 ? b1[j] : side == 1 ? b1[i] : b0[i];
    const double *c = chiR + (size_t)(flip ? N - fi : fi) + NP * (size_t)3 * (s2 + 4 * e2);
    for (int v = 0; v < 3; v++) add[v] += wgt * c[(size_t)v * NP];
  }
  for (int v = 0; v < 3; v++) mu[(e * 3 + v) * NN + node] += add[v];
}

// Lag path: stage s reads the lagged velocities un = (Hu/H, Hv/H).n of the face traces of the
// previous stage's input state Sp (own sides; the outer side's is the neighbour's, negated).
// muP = (d dSdt / d Sp)^T lambda through them: the cotangent of each own un collects chiUnL of
// this element and -chiUnR of the neighbour that reads it, then goes to the traces of Sp and,
// through the boundary interpolation, to Sp's nodes.
template <int N>
__global__ void __launch_bounds__(kVjpThreads)
    k_sw_vjp_lag(const double *__restrict__ trP, const double *__restrict__ geom, const int32_t *__restrict__ nbr,
                 const double *__restrict__ mats, const double *__restrict__ chiUnL,
                 const double *__restrict__ chiUnR, double *__restrict__ muP, int nEl) {
  constexpr int NP = N + 1, NN = NP * NP;
  const size_t t = blockIdx.x * (size_t)kVjpThreads + threadIdx.x;
  if (t >= (size_t)nEl * NN) return;
  const size_t e = t / NN;
  const int node = (int)(t % NN), i = node % NP, j = node / NP;
  const double *b0 = mats + 2 * NN, *b1 = b0 + NP;
  const double *gr = geom + e * kGeomStride;
  double add[3] = {0, 0, 0};
  for (int side = 0; side < 4; side++) {
    const int code = nbr[4 * e + side];
    const int flip = code & 1, s2 = (code >> 1) & 3;
    const size_t e2 = (size_t)(code >> 3);
    const int fi = (side == 0 || side == 2) ? i : j;
    const double wgt = side == 0 ? b0[j] : side == 2 ? b1[j] : side == 1 ? b1[i] : b0[i];
    const double xi = chiUnL[(size_t)fi + NP * (side + 4 * e)] -
                      chiUnR[(size_t)(flip ? N - fi : fi) + NP * (s2 + 4 * e2)];
    const double *tl = trP + (size_t)fi + NP * (size_t)3 * (side + 4 * e);
    const double Hu = tl[0], Hv = tl[NP], H = tl[2 * NP], nx = gr[8 + side], ny = gr[12 + side];
    add[0] += wgt * xi * nx / H;
    add[1] += wgt * xi * ny / H;
    add[2] -= wgt * xi * (Hu * nx + Hv * ny) / (H * H);
  }
  for (int v = 0; v < 3; v++) muP[(e * 3 + v) * NN + node] = add[v];
}

}  // namespace

hipError_t sw_tendency_vjp(int N, int nEl, const double *S, const double *traces, const double *un,
                           const double *aux, const double *geom, const int32_t *nbr, const double *mats, double g,
                           double fCori, double auxConst, const double *lam, double *mu, double *chiR,
                           const double *tracesPrev, double *muPrev, double *chiUn, hipStream_t st) {
  const size_t nodes = (size_t)nEl * (N + 1) * (N + 1);
  double *chiUnL = muPrev ? chiUn : nullptr;
  double *chiUnR = muPrev ? chiUn + (size_t)(N + 1) * 4 * nEl : nullptr;
  const unsigned gBlocks = (unsigned)((nodes + kVjpThreads - 1) / kVjpThreads);
  switch (N) {
#define CASE(NN_)                                                                                          \
  case NN_: {                                                                                              \
    constexpr int EPB = kVjpThreads / ((NN_ + 1) * (NN_ + 1));                                            \
    hipLaunchKernelGGL((k_sw_vjp_local<NN_>), dim3((nEl + EPB - 1) / EPB), kVjpThreads, 0, st, S, traces, un, aux, \
                       geom, nbr, mats, g, fCori, auxConst, lam, mu, chiR, chiUnL, chiUnR, nEl);          \
    hipError_t e = hipGetLastError();                                                                     \
    if (e != hipSuccess) return e;                                                                         \
    hipLaunchKernelGGL((k_sw_vjp_gather<NN_>), dim3(gBlocks), kVjpThreads, 0, st, nbr, mats, chiR, mu, nEl); \
    if (muPrev) {                                                                                          \
      if ((e = hipGetLastError()) != hipSuccess) return e;                                                 \
      hipLaunchKernelGGL((k_sw_vjp_lag<NN_>), dim3(gBlocks), kVjpThreads, 0, st, tracesPrev, geom, nbr, mats, \
                         chiUnL, chiUnR, muPrev, nEl);                                                     \
    }                                                                                                      \
    return hipGetLastError();                                                                              \
  }
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace selfamd
