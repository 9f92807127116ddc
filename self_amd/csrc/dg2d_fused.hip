// Fused DGSEM tendency + low-storage RK3 stage kernels and the Model2D engine.
//
// One launch per RK stage does the whole of CalculateTendency_Model2D + UpdateGRK3
// (src/SELF_Model2D.f90:710-752, 816-859) for a block of elements:
//   phase 0  node work-items: load S (and G), physical flux -> contravariant flux (LDS),
//            nodal trace quantities q(S) (LDS)
//   phase 1  face work-items: own traces from LDS (bMatrix contraction), neighbour traces
//            from the previous launch's trace buffer (the only cross-element read),
//            boundary conditions, Riemann flux * nScale (LDS)
//   phase 2  node work-items: weak DG divergence (dgMatrix + bMatrix/qWeights), / J,
//            source, dSdt, G = a G + dSdt, S += (g dt) G; write S, G; q(S_new) -> LDS
//   phase 3  face work-items: traces of S_new -> the other trace buffer
// so per stage each element reads S, G, its static fields and its neighbours' traces and
// writes S, G and its own traces: the algorithmic minimum for a nodal DG stage with
// Gauss points (no collocated face nodes).  Trace buffers alternate between launches.
//
// Work mapping: blocks of 64, 128 or 256 work-items (the tile, autotuned per degree), EPB =
// THREADS/(N+1)^2 elements per block; at N = 7 one wave64 owns one element, lane = i + 8 j,
// so every variable is one 512-byte coalesced access and the per-element geometry record is
// wave-uniform.
//
// Geometry: the fast path takes affine elements (metric terms constant per element, true
// for the structured and block meshes); each element carries one 20-double record.
#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "closure.hpp"
#include "sw_adjoint.hpp"
#include "../../include/self_amd.h"

using namespace selfamd;

namespace {

constexpr int kThreads = 256;
constexpr int kNbrOrphan = -0x40000000;
// forward_step calls of at most this many steps with a step size other than the captured
// graphs' are launched stage by stage instead of recapturing
constexpr int kDirectSteps = 2;

// Williamson low-storage RK3 (src/SELF_Model.f90:25-27)
constexpr double kRk3A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double kRk3G[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

struct Params {
  double g, fCori, gamma, Rgas;
  double auxConst, stat2Const;  // value of the static fields when no array is given
};

struct FusedArgs {
  double *S, *G, *dSdt;
  const double *traceIn;
  double *traceOut;
  // SW lagged velocity traces, read from the previous stage's buffer and written to the other
  // one for the next stage.  Affine mode: u.n of each face node, own side only, (0:N,4,nEl);
  // the outer side's value is the neighbour's own, negated (the two sides' normals are
  // opposite).  Nodal mode: the vector (u, v) of each face node, (2,0:N,4,nEl), and both normal
  // velocities are formed with this side's normal, as the reference does, because on curved
  // elements interpolated below the mesh degree the two sides' normals differ.
  const double *unIn;
  double *unOut;
  const double *aux;        // SW topography / LSW bathymetry / Euler potential, or null
  const double *stat2;      // LSW Coriolis / Euler drag, or null
  const double *presTrace;  // Euler prescribed traces (0:N,9,4,nEl), or null
  const double *geom;       // per-element records (affine mode)
  const double *ndsdx, *nJ, *nnHat, *nnScale;  // per-node metrics (nodal mode), SEMQuad shapes
  const int32_t *nbr;       // per (side, element) neighbour code
  const double *mats;       // dgM(P*P) D(P*P) bM(2P) qW(P) 1/qW(P)
  int nEl;
  double rkA, rkGdt;        // stage coefficients a_m and g_m*dt
  double rkG;               // g_m alone: batched runs take dt from their element records
  int perElemDt;            // 1: g_m * dt_e with dt_e in the element record (batched runs)
  int writeG;               // 0 at the last RK3 stage: the next step's first stage (a_1 = 0) never reads G
  int reverse;              // 1: blocks take the elements in reverse order (alternate stages)
  const int32_t *elems;     // the launch's elements (a band), or null: elements 0 .. nEl-1
  const double *corr;       // learned tendency correction (interior layout), or null
  const double *corrShift;  // global mean of the mass-variable correction (one double)
  // MODE 1 only, or null: the contravariant flux (2,0:N,0:N,nVar,nEl) and its weak divergence
  // divided by J (0:N,0:N,nVar,nEl), the flux and fluxDivergence registers of
  // CalculateTendency_Model2D (src/SELF_Model2D.f90:816-859)
  double *fluxOut, *fluxDivOut;
  int anyBoundary;          // 1 when some element side is a physical boundary (neighbour code < 0)
  Params p;
};

// ------------------------------------------------------------------------------
// Physics policies
// ------------------------------------------------------------------------------

// Nonlinear shallow water (src/models/SELF_ShallowWater.f90): state (Hu, Hv, H), traces
// are the state traces, aux = topography (gradient in the source, trace in radiation BC).
struct PhysSW {
  static constexpr int kKind = SELFAMD_MODEL_SW;
  static constexpr int NV = 3, NT = 3;
  static constexpr bool kAuxGrad = true, kAuxTrace = true, kLag = true;
  static constexpr bool kAuxTraceInFlux = false;  // the topography trace enters the radiation BC only
  __device__ static void volume(const double *s, double, const Params &p, double *fx, double *fy, double *q) {
    const double Hu = s[0], Hv = s[1], H = s[2];
    const double invH = 1.0 / H;
    const double u = Hu * invH, v = Hv * invH;  // PreTendency velocity (src/models/SELF_ShallowWater.f90:206-218)
    fx[0] = u * Hu + 0.5 * p.g * H * H;
    fy[0] = v * Hu;
    fx[1] = u * Hv;
    fy[1] = v * Hv + 0.5 * p.g * H * H;
    fx[2] = Hu;
    fy[2] = Hv;
    q[0] = Hu; q[1] = Hv; q[2] = H;
  }
  __device__ static void nodal(const double *s, double *q) { q[0] = s[0]; q[1] = s[1]; q[2] = s[2]; }
  __device__ static void source(const double *s, double gx, double gy, double, const Params &p, double *src) {
    src[0] = -p.fCori * s[1] + p.g * s[2] * gx;
    src[1] = p.fCori * s[0] + p.g * s[2] * gy;
    src[2] = 0.0;
  }
  // SetBoundaryCondition_ShallowWater (src/models/SELF_ShallowWater.f90:284-343)
  __device__ static void bc(const double *qL, double auxB, double nx, double ny, int bcid, const double *, bool, double *qR) {
    if (bcid == SELFAMD_BC_NONORMALFLOW) {
      qR[0] = (ny * ny - nx * nx) * qL[0] - 2.0 * nx * ny * qL[1];
      qR[1] = (nx * nx - ny * ny) * qL[1] - 2.0 * nx * ny * qL[0];
      qR[2] = qL[2];
    } else {
      qR[0] = 0.0; qR[1] = 0.0; qR[2] = auxB;
    }
  }
  // RiemannSolver_ShallowWater (src/models/SELF_ShallowWater.f90:451-529); unL/unR are the
  // lagged normal velocities.
  __device__ static void riemann(const double *qL, const double *qR, double nx, double ny, double nmag, double,
                                 double unL, double unR, const Params &p, double *f) {
    const double g = p.g;
    const double HuL = qL[0], HvL = qL[1], HL = qL[2], HuR = qR[0], HvR = qR[1], HR = qR[2];
    const double cL = sqrt(g * HL), cR = sqrt(g * HR);
    const double fL0 = HuL * unL + g * HL * HL * 0.5 * nx, fL1 = HvL * unL + g * HL * HL * 0.5 * ny;
    const double fL2 = HuL * nx + HvL * ny;
    const double fR0 = HuR * unR + g * HR * HR * 0.5 * nx, fR1 = HvR * unR + g * HR * HR * 0.5 * ny;
    const double fR2 = HuR * nx + HvR * ny;
    const double a = fmax(fmax(fabs(unL + cL), fabs(unR + cR)), fmax(fabs(unL - cL), fabs(unR - cR)));
    f[0] = 0.5 * (fL0 + fR0 + a * (HuL - HuR)) * nmag;
    f[1] = 0.5 * (fL1 + fR1 + a * (HvL - HvR)) * nmag;
    f[2] = 0.5 * (fL2 + fR2 + a * (HL - HR)) * nmag;
  }
};

// Linear shallow water (src/models/SELF_LinearShallowWater.f90): state (u, v, eta),
// aux = bathymetry H (flux and trace for c), stat2 = Coriolis f (source).
struct PhysLSW {
  static constexpr int kKind = SELFAMD_MODEL_LSW;
  static constexpr int NV = 3, NT = 3;
  static constexpr bool kAuxGrad = false, kAuxTrace = true, kLag = false;
  static constexpr bool kAuxTraceInFlux = true;  // the Riemann flux takes c = sqrt(g H) of the bathymetry trace
  __device__ static void volume(const double *s, double H, const Params &p, double *fx, double *fy, double *q) {
    fx[0] = p.g * s[2]; fy[0] = 0.0;
    fx[1] = 0.0;        fy[1] = p.g * s[2];
    fx[2] = H * s[0];   fy[2] = H * s[1];
    q[0] = s[0]; q[1] = s[1]; q[2] = s[2];
  }
  __device__ static void nodal(const double *s, double *q) { q[0] = s[0]; q[1] = s[1]; q[2] = s[2]; }
  __device__ static void source(const double *s, double, double, double f, const Params &, double *src) {
    src[0] = f * s[1];
    src[1] = -f * s[0];
    src[2] = 0.0;
  }
  __device__ static void bc(const double *qL, double, double nx, double ny, int bcid, const double *, bool, double *qR) {
    if (bcid == SELFAMD_BC_NONORMALFLOW) {
      qR[0] = (ny * ny - nx * nx) * qL[0] - 2.0 * nx * ny * qL[1];
      qR[1] = (nx * nx - ny * ny) * qL[1] - 2.0 * nx * ny * qL[0];
      qR[2] = qL[2];
    } else {
      qR[0] = 0.0; qR[1] = 0.0; qR[2] = 0.0;
    }
  }
  // characteristic upwind flux (src/models/SELF_LinearShallowWater.f90:500-557)
  __device__ static void riemann(const double *qL, const double *qR, double nx, double ny, double nmag, double Hb,
                                 double, double, const Params &p, double *f) {
    const double g = p.g;
    const double c = sqrt(g * Hb);
    const double unL = qL[0] * nx + qL[1] * ny, unR = qR[0] * nx + qR[1] * ny;
    const double wL = 0.5 * (unL / g + qL[2] / c), wR = 0.5 * (unR / g - qR[2] / c);
    f[0] = g * c * (wL - wR) * nx * nmag;
    f[1] = g * c * (wL - wR) * ny * nmag;
    f[2] = c * c * (wL + wR) * nmag;
  }
};

// Compressible ideal gas (src/models/SELF_CompressibleIdealGas2D.f90): state
// (rho u, rho v, rho, rho E); traces: state(4), primitive u, v, p, enthalpy rho E + p,
// sound speed c (each interpolated from its nodal values, as PreTendency does);
// aux = gravity potential (gradient in the source), stat2 = drag coefficient.
struct PhysEuler {
  static constexpr int kKind = SELFAMD_MODEL_EULER;
  static constexpr int NV = 4, NT = 9;
  static constexpr bool kAuxGrad = true, kAuxTrace = false, kLag = false;
  static constexpr bool kAuxTraceInFlux = false;
  __device__ static void nodal_p(const double *s, const Params &p, double &u, double &v, double &pr, double &ke,
                                 double &invRho) {
    invRho = 1.0 / s[2];
    u = s[0] * invRho;
    v = s[1] * invRho;
    ke = 0.5 * (s[0] * u + s[1] * v);
    pr = (p.gamma - 1.0) * (s[3] - ke);
  }
  __device__ static void nodal_q(const double *s, const Params &p, double *q) {
    double u, v, pr, ke, invRho;
    nodal_p(s, p, u, v, pr, ke, invRho);
    q[0] = s[0]; q[1] = s[1]; q[2] = s[2]; q[3] = s[3];
    q[4] = u; q[5] = v; q[6] = pr;
    q[7] = s[3] + pr;                        // diagnostics(2): enthalpy
    q[8] = sqrt(p.gamma * pr * invRho);      // diagnostics(3): sound speed
  }
  __device__ static void volume(const double *s, double, const Params &p, double *fx, double *fy, double *q) {
    nodal_q(s, p, q);
    const double u = q[4], v = q[5], pr = q[6], Hn = q[7];
    fx[0] = u * s[0] + pr; fy[0] = v * s[0];
    fx[1] = u * s[1];      fy[1] = v * s[1] + pr;
    fx[2] = s[0];          fy[2] = s[1];
    fx[3] = u * Hn;        fy[3] = v * Hn;
  }
  __device__ static void source(const double *s, double gx, double gy, double Cd, const Params &, double *src) {
    src[0] = -s[2] * gx - Cd * s[0];
    src[1] = -s[2] * gy - Cd * s[1];
    src[2] = 0.0;
    src[3] = -s[0] * gx - s[0] * gy;  // quirk: energy source uses rho*u twice
  }
  // SetBoundaryCondition_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:1217-1323)
  __device__ static void bc(const double *qL, double, double nx, double ny, int bcid, const double *pres, bool hasPres,
                            double *qR) {
    if (bcid == SELFAMD_BC_NONORMALFLOW) {
      qR[0] = (ny * ny - nx * nx) * qL[0] - 2.0 * nx * ny * qL[1];
      qR[1] = (nx * nx - ny * ny) * qL[1] - 2.0 * nx * ny * qL[0];
      qR[2] = qL[2]; qR[3] = qL[3];
      qR[4] = (ny * ny - nx * nx) * qL[4] - 2.0 * nx * ny * qL[5];
      qR[5] = (nx * nx - ny * ny) * qL[5] - 2.0 * nx * ny * qL[4];
      qR[6] = qL[6]; qR[7] = qL[7]; qR[8] = qL[8];
    } else if ((bcid == SELFAMD_BC_PRESCRIBED || bcid == SELFAMD_BC_RADIATION) && hasPres) {
      for (int k = 0; k < 9; k++) qR[k] = pres[k];
    } else {
      for (int k = 0; k < 9; k++) qR[k] = 0.0;
    }
  }
  // NaiveLLF (src/models/SELF_CompressibleIdealGas2D.f90:1455-1542)
  __device__ static void riemann(const double *qL, const double *qR, double nx, double ny, double nmag, double,
                                 double, double, const Params &, double *f) {
    const double unL = qL[4] * nx + qL[5] * ny, unR = qR[4] * nx + qR[5] * ny;
    const double fL0 = unL * qL[0] + qL[6] * nx, fL1 = unL * qL[1] + qL[6] * ny;
    const double fL2 = qL[0] * nx + qL[1] * ny, fL3 = unL * qL[7];
    const double fR0 = unR * qR[0] + qR[6] * nx, fR1 = unR * qR[1] + qR[6] * ny;
    const double fR2 = qR[0] * nx + qR[1] * ny, fR3 = unR * qR[7];
    const double a = fmax(fabs(unL), fabs(unR)) + fmax(fabs(qL[8]), fabs(qR[8]));
    f[0] = 0.5 * (fL0 + fR0 + a * (qL[0] - qR[0])) * nmag;
    f[1] = 0.5 * (fL1 + fR1 + a * (qL[1] - qR[1])) * nmag;
    f[2] = 0.5 * (fL2 + fR2 + a * (qL[2] - qR[2])) * nmag;
    f[3] = 0.5 * (fL3 + fR3 + a * (qL[3] - qR[3])) * nmag;
  }
  __device__ static void nodal(const double *s, double *q) {}  // unused: nodal_q needs params
};

template <class Phys>
__device__ inline void phys_nodal(const double *s, const Params &p, double *q) {
  if constexpr (Phys::NT == 9) PhysEuler::nodal_q(s, p, q);
  else Phys::nodal(s, q);
}

// Sum of x over the lane pair l, l ^ 32 of a wavefront, returned to both lanes.  The
// exchange is v_permlane32_swap (gfx950), a VALU operation: it does not use the LDS pipe that
// a shuffle through ds_bpermute would occupy.  Both lanes get the same sum x_l + x_(l+32).
// Lagged SW normal velocities (the layouts are described with FusedArgs::unIn): the own side's
// and the outer side's u.n at face-node slot fb (the neighbour's slot for the outer side).
template <bool NODAL>
__device__ inline double lag_own(const double *un, size_t fb, double nx, double ny) {
  if constexpr (NODAL) return un[2 * fb] * nx + un[2 * fb + 1] * ny;
  return un[fb];
}
// The outer side's value is formed in two steps: lag_outer_raw at the load (affine: the stored
// value as it is, so that no arithmetic waits for the load there) and outer_un where it is used.
template <bool NODAL>
__device__ inline double lag_outer_raw(const double *un, size_t fbNeighbour, double nx, double ny) {
  if constexpr (NODAL) return un[2 * fbNeighbour] * nx + un[2 * fbNeighbour + 1] * ny;
  return un[fbNeighbour];
}
template <bool NODAL>
__device__ inline double outer_un(double raw) {
  if constexpr (NODAL) return raw;
  return -raw;
}

// Ordering point between kernel phases.  When every element's node lanes and face lanes lie in
// one wavefront (WAVE), a phase only reads LDS values its own wavefront wrote, and a wavefront's
// LDS accesses execute in program order: a code-motion barrier suffices.  A workgroup barrier
// would also make each wavefront wait for its outstanding global stores (the workgroup-scope
// release) and for the block's slowest wavefront.
template <bool WAVE>
__device__ inline void phase_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

__device__ inline double pair_sum32(double x) {
  const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
  const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  // a[0], b[0]: the value of lane l mod 32; a[1], b[1]: the value of lane l mod 32 + 32
  return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
}

// ------------------------------------------------------------------------------
// The fused kernel
// ------------------------------------------------------------------------------
// MODE 0: RK stage (update S, G, write next traces); MODE 1: tendency only (write dSdt).
// NODAL: metric terms per node and per face node exactly as SEMQuad stores them (curved
// elements, and the reference's own per-node metric values); otherwise one affine record
// per element.  THREADS: block size (the tile shape); EPB = THREADS/(N+1)^2 elements per
// block.  With THREADS = 64 at N = 7 a block is one wavefront owning one element, so its
// barriers never wait on other waves.
//
// Every global load the stage needs is issued before any arithmetic: the element's state,
// RK register and static field on the node lanes, and (on the face lanes) the neighbour
// code, the neighbour's traces and the lagged normal velocities.  The gather of neighbour
// traces is the only dependent load and it overlaps the volume-flux work.
//
// Tile parameters: THREADS work-items per block (THREADS / (N+1)^2 elements), WPE the
// occupancy the register allocator must allow (waves per SIMD; 1 = unconstrained), XCD an
// element order in which the blocks one XCD receives cover a contiguous run of elements.
// With a.reverse the elements are visited last to first: successive stages alternate
// direction, so a stage starts on the elements whose data the previous stage touched last.
//
// LDS of one tile (block) of the stage kernel, in doubles: the node arrays sQ, sF1, sF2, the
// static field sAux, the nodal-mode metric products sGb, the face fluxes sBF and the matrices
// sM, each starting on a 16-byte boundary.
template <class Phys, int N, bool NODAL, int THREADS>
struct TileLds {
  static constexpr int NP = N + 1, NN = NP * NP, EPB = THREADS / NN, NF = 4 * NP;
  static constexpr int NV = Phys::NV, NT = Phys::NT;
  static constexpr bool USE_AUX = Phys::kAuxGrad || Phys::kAuxTrace;
  static constexpr bool GRADN = NODAL && Phys::kAuxGrad;
  // At N = 5 two pad doubles per element array (38 instead of 36) shift the elements of a
  // tile against each other in the banks: 70.0 against 71.1 us per SW launch at 256^2 (row
  // strides of 7 were slower, 72.3-72.7 us).  Matrices hold entry (c, r) at c + MRS r, with two
  // pad doubles per 8-double row (they are written once per tile).
  static constexpr int MRS = NP % 8 == 0 ? NP + 2 : NP, NNS = NP == 6 ? NN + 2 : NN;
  static constexpr int even(int n) { return (n + 1) & ~1; }
  static constexpr int Q = 0;
  static constexpr int F1 = Q + even(NT * EPB * NNS);
  static constexpr int F2 = F1 + even(NV * EPB * NNS);
  static constexpr int AUX = F2 + even(NV * EPB * NNS);
  static constexpr int GB = AUX + even(USE_AUX ? EPB * NNS : 1);
  static constexpr int BF = GB + even(GRADN ? 4 * EPB * NNS : 1);
  static constexpr int M = BF + even(NV * EPB * NF);
  static constexpr int SIZE = M + even(2 * NP * MRS + 4 * NP);
  // face-lane arrangement (see the stage kernel): every face node has its own lane; two lanes per
  // face node; N = 5 face lanes padded to 32 per element; every element's node and face lanes in
  // one wavefront
  static constexpr bool ONE_FACE_PASS = EPB * NF <= THREADS;
  static constexpr bool SPLIT = ONE_FACE_PASS && 2 * EPB * NF <= THREADS && NF <= 32;
  static constexpr bool FPAD = !SPLIT && NF == 24 && 32 * EPB <= THREADS;
  static constexpr bool WAVE = (NN == 64 && SPLIT) || (NN == 16 && NF == 16 && ONE_FACE_PASS && !SPLIT && !FPAD);
};

// The tile's interpolation and derivative matrices and quadrature weights into its LDS (dgM, D
// in rows of MRS, bM(2P), qW, 1/qW), from mats = [dgM, D, bM, qW, 1/qW] as SetupMatrices stores
// them.  Every work-item of the block takes part.
template <class Phys, int N, bool NODAL, int THREADS>
__device__ __forceinline__ int tile_matrix_slot(int k) {
  constexpr int NP = N + 1, NN = NP * NP, MRS = TileLds<Phys, N, NODAL, THREADS>::MRS;
  const int r = k % NN;
  return k < 2 * NN ? (k / NN) * NP * MRS + (r / NP) * MRS + r % NP : 2 * NP * MRS + k - 2 * NN;
}
template <class Phys, int N, bool NODAL, int THREADS>
__device__ __forceinline__ void load_tile_matrices(const double *mats, double *lds, int tid) {
  double *sM = lds + TileLds<Phys, N, NODAL, THREADS>::M;
  for (int k = tid; k < 2 * (N + 1) * (N + 1) + 4 * (N + 1); k += THREADS)
    sM[tile_matrix_slot<Phys, N, NODAL, THREADS>(k)] = mats[k];
}

// One tile of the stage kernel: tile blk of the launch's element list (after any reordering),
// with the tile's LDS at lds, run by work-item tid of the block.  The __global__ wrappers
// below map blocks to tiles.  PRELOADED: the block has already put the matrices into the LDS
// (load_tile_matrices) and made them visible, so the tile neither loads them nor needs a block
// barrier where the element's own wavefront suffices.
template <class Phys, int N, int MODE, bool NODAL, int THREADS, bool PRELOADED = false>
__device__ __forceinline__ void fused_tile(const FusedArgs &a, int blk, double *lds, const int tid) {
  using L = TileLds<Phys, N, NODAL, THREADS>;
  constexpr int NP = N + 1, NN = NP * NP, EPB = THREADS / NN, NF = 4 * NP;
  constexpr int NV = Phys::NV, NT = Phys::NT;
  constexpr bool USE_AUX = Phys::kAuxGrad || Phys::kAuxTrace;
  constexpr bool ONE_FACE_PASS = L::ONE_FACE_PASS;  // every face node has its own lane
  // two lanes per face node (lanes l and l ^ 32 of a wavefront): each interpolates half of
  // the face line, the halves are summed across the pair
  constexpr bool SPLIT = L::SPLIT;
  static_assert(EPB >= 1, "block too small for one element");
  // LDS layouts.  Node arrays hold node (i, j) at slot(i, j).  With 8 x 8 nodes a row is 64
  // bytes, and a column access (fixed i, one row per lane group) would find rows 0, 2, 4, 6 in
  // the same banks; permuting the columns of row j by i ^ (j & 6) spreads them over the banks,
  // keeps the pairs (2m, 2m + 1) adjacent for 16-byte reads and keeps every row in place, so
  // row-wise writes stay conflict-free.
  constexpr bool SWZ = NP == 8;
  constexpr int MRS = L::MRS, NNS = L::NNS;
  auto slot = [](int i_, int j_) { return SWZ ? NP * j_ + (i_ ^ (j_ & 6)) : i_ + NP * j_; };
  auto sQ = reinterpret_cast<double (*)[EPB][NNS]>(lds + L::Q);
  auto sF1 = reinterpret_cast<double (*)[EPB][NNS]>(lds + L::F1);
  auto sF2 = reinterpret_cast<double (*)[EPB][NNS]>(lds + L::F2);
  auto sAux = reinterpret_cast<double (*)[USE_AUX ? NNS : 1]>(lds + L::AUX);
  // nodal mode: static field times the node's contravariant metric, for GradientSF
  constexpr bool GRADN = L::GRADN;
  auto sGb = reinterpret_cast<double (*)[GRADN ? EPB : 1][GRADN ? NNS : 1]>(lds + L::GB);
  auto sBF = reinterpret_cast<double (*)[EPB][NF]>(lds + L::BF);
  double *sM = lds + L::M;  // dgM, D (rows of MRS), bM(2P), qW, 1/qW

  const int el = tid / NN, node = tid % NN;
  const int i = node % NP, j = node / NP;
  const int nl = slot(i, j);  // the node's LDS slot
  const int pe = blk * EPB + el;  // position in the launch's element list
  bool nodeActive = el < EPB && pe < a.nEl;
  int e = a.elems ? (nodeActive ? a.elems[pe] : 0) : pe;
  // one element per wavefront: its index, whether it is active, and every address and record
  // value derived from it are uniform, so they live in scalar registers and the branches on
  // them are scalar
  if constexpr (NN == 64) {
    e = __builtin_amdgcn_readfirstlane(e);
    nodeActive = __builtin_amdgcn_readfirstlane((int)nodeActive) != 0;
  }
  // ---- face lanes ----
  // SPLIT: wavefront w covers the face nodes of element w (NF <= 32 slots), lanes l and l + 32
  // share face node l; otherwise face node = thread index
  // FPAD (N = 5, tiles of 128 or 256): each element's 24 face nodes take 32 lanes, south and
  // north faces in lanes 0-11 and east and west in lanes 16-27 of the group, so that a 16-lane
  // LDS group reads either columns or rows of the element, never both (no bank conflicts)
  constexpr bool FPAD = L::FPAD;
  const int fhalf = SPLIT ? (tid >> 5) & 1 : 0;
  const int fslot = SPLIT ? (tid & 31) : FPAD ? (tid & 15) : 0;
  const int fidx = SPLIT ? (tid >> 6) * NF + fslot : FPAD ? (tid >> 5) * NF + (fslot % 12) : tid;
  const int fel = fidx / NF, fslotN = fidx % NF;
  // SPLIT: slots take the sides in the order south, north, east, west, so that the lanes of a
  // 16-lane LDS group read opposite faces, whose face lines are the same nodes (no bank
  // conflicts between them)
  const int side = SPLIT  ? ((fslotN / NP) >> 1 | ((fslotN / NP) & 1) << 1)
                   : FPAD ? ((tid >> 4) & 1) + (fslot >= NP ? 2 : 0)
                          : fslotN / NP;
  const int fi = fslotN % NP, fnode = side * NP + fi;
  // element-to-wavefront locality of the node and face lanes (see phase_sync): at N = 7 a
  // wavefront owns one element (face lanes split in pairs); at N = 3 an element's 16 nodes and 16
  // face nodes share the same 16 lanes
  constexpr bool WAVE = L::WAVE;
  const int pf = blk * EPB + fel;
  const bool faceActive = ONE_FACE_PASS && (!SPLIT || fslot < NF) && (!FPAD || fslot < 2 * NP) && fel < EPB &&
                          pf < a.nEl;
  // one element per wavefront (N = 7, face lanes split in pairs): the face lanes' element is the
  // node lanes' one, already in a scalar register
  int fe;
  if constexpr (SPLIT && NF == 32 && NN == 64) fe = e;
  else fe = a.elems ? (faceActive ? a.elems[pf] : 0) : pf;
  // SPLIT: LDS slots of the half face line this lane interpolates (phases 1 and 3).  South and
  // north faces run along j through column fi, east and west along i through row fi.
  constexpr int FH0 = NP / 2, FHN = NP - FH0;
  int foff[SPLIT ? FHN : 1];
  if constexpr (SPLIT) {
    const bool alongJ = (side == 0 || side == 2);
    const int lo = fhalf ? FH0 : 0;
#pragma unroll
    for (int t = 0; t < FHN; t++) foff[t] = alongJ ? slot(fi, lo + t) : slot(lo + t, fi);
  }
  // ---- loads ----
  // The vector-memory counter completes in issue order, and a wait after a branch must cover
  // every path into it.  So every load is issued before any of them is waited for: the
  // neighbour codes (first, since the trace loads depend on them), the matrices (unless
  // preloaded), the node lanes' streams (on every path: inactive lanes and absent fields read
  // element 0 or the state, and the values are discarded), then the loads addressed through
  // the codes, which wait for the codes alone; the matrices go to LDS last.
  // one element per wavefront (N = 7, face lanes split in pairs): the element's four codes are
  // one scalar load further down; otherwise each face lane's code is the first vector load
  constexpr bool SCALAR_CODES = SPLIT && NF == 32 && NN == 64;
  int code = 0;
  if constexpr (!SCALAR_CODES) {
    const bool own = faceActive && fhalf == 0;
    const int c = a.nbr[4 * (size_t)(own ? fe : 0) + side];
    code = own ? c : 0;
  }
  constexpr int MCNT = 2 * NN + 4 * NP, MPL = PRELOADED ? 1 : (MCNT + THREADS - 1) / THREADS;
  double mreg[MPL];
  if constexpr (!PRELOADED) {
#pragma unroll
    for (int q = 0; q < MPL; q++) mreg[q] = a.mats[min(tid + q * THREADS, MCNT - 1)];
  }
  const size_t sBase = (size_t)e * NV * NN;  // S(0,0,0,e)
  const size_t nodeAt = nodeActive ? (size_t)e * NN + node : (size_t)node;  // inactive: element 0
  double s[NV], gReg[NV];
  double auxVal = a.p.auxConst, statVal = a.p.stat2Const;
  double dsdx[4] = {0, 0, 0, 0}, invJ = 1.0;
  {
    const size_t sAt = nodeActive ? sBase + node : (size_t)node;
    // state and RK register: non-temporal loads.  This stage overwrites both, so caching the
    // lines read here buys nothing; measured 5 % faster per launch than cached loads at N = 7
    // (non-temporal loads of the static field or of the neighbour traces, which later stages
    // or neighbours reuse, were slower).
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = __builtin_nontemporal_load(&a.S[sAt + (size_t)v * NN]);
    if constexpr (MODE == 0) {
      // the first stage (rkA = 0) reads element 0's, a few cached lines, and discards them
      const bool useG = a.rkA != 0.0;
      const size_t gAt = useG ? sAt : (size_t)node;
      double gl[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) gl[v] = __builtin_nontemporal_load(&a.G[gAt + (size_t)v * NN]);
#pragma unroll
      for (int v = 0; v < NV; v++) gReg[v] = useG ? gl[v] : 0.0;
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) gReg[v] = 0.0;
    }
    if constexpr (USE_AUX) {
      const double av = a.aux ? a.aux[nodeAt] : a.S[node];
      if (a.aux) auxVal = av;
    }
    const double sv = a.stat2 ? a.stat2[nodeAt] : a.S[node];
    if (a.stat2) statVal = sv;
    if constexpr (NODAL) {
      const double *gn = a.ndsdx + 4 * nodeAt;
      dsdx[0] = gn[0]; dsdx[1] = gn[1]; dsdx[2] = gn[2]; dsdx[3] = gn[3];
      invJ = 1.0 / a.nJ[nodeAt];
    } else {
      const double *gr = a.geom + (nodeActive ? (size_t)e : 0) * kGeomStride;
      dsdx[0] = gr[0]; dsdx[1] = gr[1]; dsdx[2] = gr[2]; dsdx[3] = gr[3];
      invJ = gr[6];  // 1/J, precomputed in the element record
    }
  }
  if constexpr (SCALAR_CODES) {
    // (a wavefront past the last element of a ragged last tile reads element 0's codes; the
    // wavefront's node and face lanes are the same element, so nodeActive, a scalar, decides)
    const int4 c4 = *reinterpret_cast<const int4 *>(a.nbr + 4 * (size_t)(nodeActive ? fe : 0));
    // all four in scalar registers at once (one load, one wait), then a per-lane select
    const int cx = __builtin_amdgcn_readfirstlane(c4.x), cy = __builtin_amdgcn_readfirstlane(c4.y);
    const int cz = __builtin_amdgcn_readfirstlane(c4.z), cw = __builtin_amdgcn_readfirstlane(c4.w);
    code = (side & 2) ? ((side & 1) ? cw : cz) : ((side & 1) ? cy : cx);
    if (!(faceActive && fhalf == 0)) code = 0;
  }

  double qR[NT], unL = 0.0, unR = 0.0, nx = 0.0, ny = 0.0, nmag = 0.0;
  if (faceActive && fhalf == 0) {
    // normals: first in nodal mode, where the lagged normal velocities need them; after the
    // neighbour loads in affine mode (measured slightly faster there)
    if constexpr (NODAL) {
      const size_t fb = (size_t)fi + NP * (side + 4 * (size_t)fe);
      nx = a.nnHat[2 * fb]; ny = a.nnHat[2 * fb + 1]; nmag = a.nnScale[fb];
    }
    if (code >= 0) {
      const int flip = code & 1, s2 = (code >> 1) & 3;
      const size_t e2 = (size_t)(code >> 3);
      const size_t nb = (size_t)(flip ? N - fi : fi) + NP * (s2 + 4 * e2);
      const double *tr = a.traceIn + (size_t)(flip ? N - fi : fi) + NP * (size_t)NT * (s2 + 4 * e2);
#pragma unroll
      for (int k = 0; k < NT; k++) qR[k] = tr[(size_t)k * NP];
      if (Phys::kLag) unR = lag_outer_raw<NODAL>(a.unIn, nb, nx, ny);
    }
    if (Phys::kLag) unL = lag_own<NODAL>(a.unIn, (size_t)fi + NP * (side + 4 * (size_t)fe), nx, ny);
    if constexpr (!NODAL) {
      const double *gr = a.geom + (size_t)fe * kGeomStride;
      nx = gr[8 + side]; ny = gr[12 + side]; nmag = gr[16 + side];
    }
  }

  // the matrices into LDS once every load is in flight
  if constexpr (!PRELOADED) {
#pragma unroll
    for (int q = 0; q < MPL; q++) {
      const int k = tid + q * THREADS;
      if (k < MCNT) sM[tile_matrix_slot<Phys, N, NODAL, THREADS>(k)] = mreg[q];
    }
  }
  const double *dgM = sM, *Dm = sM + NP * MRS, *bM0 = sM + 2 * NP * MRS, *bM1 = bM0 + NP, *qW = bM1 + NP;

  // ---- phase 0: volume fluxes, nodal trace quantities ----
  if (nodeActive) {
    double fx[NV], fy[NV], q[NT];
    Phys::volume(s, auxVal, a.p, fx, fy, q);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      // ContravariantProjection (src/SELF_MappedData.f90:2220-2269)
      sF1[v][el][nl] = dsdx[0] * fx[v] + dsdx[1] * fy[v];
      sF2[v][el][nl] = dsdx[2] * fx[v] + dsdx[3] * fy[v];
      if (MODE == 1 && a.fluxOut) {
        double *fo = a.fluxOut + 2 * (sBase + (size_t)v * NN + node);
        fo[0] = sF1[v][el][nl];
        fo[1] = sF2[v][el][nl];
      }
    }
#pragma unroll
    for (int k = 0; k < NT; k++) sQ[k][el][nl] = q[k];
    if (USE_AUX) sAux[el][nl] = auxVal;
    if constexpr (GRADN) {
#pragma unroll
      for (int k = 0; k < 4; k++) sGb[k][el][nl] = auxVal * dsdx[k];
    }
  }
  // the matrices (written by the whole block) and the element's node values
  if constexpr (PRELOADED) phase_sync<WAVE>();
  else __syncthreads();

  // ---- phase 1: faces ----
  auto face = [&](int fel_, int side_, int fi_, int fe_, int fnode_, int code_, double *qR_, double unL_, double unR_,
                  double nx_, double ny_, double nmag_) {
    // own traces, ScalarBoundaryInterp_2D order (src/SELF_Lagrange.f90:2241-2270)
    const double *bm = (side_ == 1 || side_ == 2) ? bM1 : bM0;
    // nodes of the face line: column i = fi for south / north, row j = fi for east / west
    const bool alongJ = (side_ == 0 || side_ == 2);
    double qL[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) qL[k] = 0.0;
    // a static field's trace that enters only boundary conditions (SW topography) is skipped on
    // meshes without boundary sides (the condition is uniform over the launch, so the lane pairs
    // of SPLIT stay together)
    const bool auxTrace = Phys::kAuxTrace && a.aux && (Phys::kAuxTraceInFlux || a.anyBoundary);
    double auxB = a.p.auxConst;
    if (auxTrace) auxB = 0.0;
    if constexpr (SPLIT) {
      const int lo = fhalf ? FH0 : 0;
#pragma unroll
      for (int t = 0; t < FHN; t++) {
        const int ii = lo + t;
        if (fhalf == 0 && t >= FH0) break;
        const int nd = foff[t];
#pragma unroll
        for (int k = 0; k < NT; k++) qL[k] += bm[ii] * sQ[k][fel_][nd];
        if (auxTrace) auxB += bm[ii] * sAux[fel_][nd];
      }
#pragma unroll
      for (int k = 0; k < NT; k++) qL[k] = pair_sum32(qL[k]);
      if (auxTrace) auxB = pair_sum32(auxB);
      if (fhalf) return;  // the pair's first lane does the boundary and Riemann work
    } else {
#pragma unroll
      for (int ii = 0; ii < NP; ii++) {
        const int nd = alongJ ? slot(fi_, ii) : slot(ii, fi_);
#pragma unroll
        for (int k = 0; k < NT; k++) qL[k] += bm[ii] * sQ[k][fel_][nd];
        if (auxTrace) auxB += bm[ii] * sAux[fel_][nd];
      }
    }
    const bool interior = code_ >= 0;
    if (!interior) {
      if (code_ == kNbrOrphan) {
#pragma unroll
        for (int k = 0; k < NT; k++) qR_[k] = 0.0;
      } else {
        const double *pres =
            a.presTrace ? a.presTrace + (size_t)fi_ + NP * (size_t)NT * (side_ + 4 * (size_t)fe_) : nullptr;
        // the prescribed values go by value with a flag (a pointer that may be null would keep
        // the array in scratch memory)
        double pv[NT];
#pragma unroll
        for (int k = 0; k < NT; k++) pv[k] = pres ? pres[(size_t)k * NP] : 0.0;
        Phys::bc(qL, auxB, nx_, ny_, -code_ - 1, pv, pres != nullptr, qR_);
      }
    }
    if (Phys::kLag) {
      // velocity traces lag one stage (src/models/SELF_ShallowWater.f90:222-237): the stored
      // velocities are used now and this stage's are stored for the next one
      const double invH = 1.0 / qL[2];
      const size_t fb = (size_t)fi_ + NP * (side_ + 4 * (size_t)fe_);
      if constexpr (NODAL) {
        a.unOut[2 * fb] = qL[0] * invH;
        a.unOut[2 * fb + 1] = qL[1] * invH;
      } else {
        a.unOut[fb] = (qL[0] * invH) * nx_ + (qL[1] * invH) * ny_;
      }
    }
    double fstar[NV];
    Phys::riemann(qL, qR_, nx_, ny_, nmag_, auxB, unL_, outer_un<NODAL>(unR_), a.p, fstar);
#pragma unroll
    for (int v = 0; v < NV; v++) sBF[v][fel_][fnode_] = fstar[v];
  };
  if constexpr (ONE_FACE_PASS) {
    if (faceActive) face(fel, side, fi, fe, fnode, code, qR, unL, unR, nx, ny, nmag);
  } else {
    // N = 1, 2: more face nodes than lanes; load and compute face by face
    for (int ft = tid; ft < EPB * NF; ft += THREADS) {
      const int fel2 = ft / NF, fnode2 = ft % NF, side2 = fnode2 / NP, fi2 = fnode2 % NP;
      const int pf2 = blk * EPB + fel2;
      if (pf2 >= a.nEl) continue;
      const int fe2 = a.elems ? a.elems[pf2] : pf2;
      double qR2[NT], unL2 = 0.0, unR2 = 0.0, nx2, ny2, nmag2;
      const int code2 = a.nbr[4 * (size_t)fe2 + side2];
      if constexpr (NODAL) {
        const size_t fb = (size_t)fi2 + NP * (side2 + 4 * (size_t)fe2);
        nx2 = a.nnHat[2 * fb]; ny2 = a.nnHat[2 * fb + 1]; nmag2 = a.nnScale[fb];
      } else {
        const double *gr = a.geom + (size_t)fe2 * kGeomStride;
        nx2 = gr[8 + side2]; ny2 = gr[12 + side2]; nmag2 = gr[16 + side2];
      }
      if (code2 >= 0) {
        const int flip = code2 & 1, s2 = (code2 >> 1) & 3;
        const size_t e2 = (size_t)(code2 >> 3);
        const double *tr = a.traceIn + (size_t)(flip ? N - fi2 : fi2) + NP * (size_t)NT * (s2 + 4 * e2);
        for (int k = 0; k < NT; k++) qR2[k] = tr[(size_t)k * NP];
        if (Phys::kLag) unR2 = lag_outer_raw<NODAL>(a.unIn, (size_t)(flip ? N - fi2 : fi2) + NP * (s2 + 4 * e2), nx2, ny2);
      }
      if (Phys::kLag) unL2 = lag_own<NODAL>(a.unIn, (size_t)fi2 + NP * (side2 + 4 * (size_t)fe2), nx2, ny2);
      face(fel2, side2, fi2, fe2, fnode2, code2, qR2, unL2, unR2, nx2, ny2, nmag2);
    }
  }
  phase_sync<WAVE>();

  // ---- phase 2: divergence, source, update ----
  if (nodeActive) {
    // One pass over the line index ii for every sum of the phase: the volume terms of the weak
    // divergence of each variable (VectorDGDivergence_2D, src/SELF_Lagrange.f90:1589-1625; xi
    // and eta sums in separate FMA chains) and GradientSF of the static field
    // (src/SELF_MappedData.f90:1080-1149).  Each matrix entry is read from LDS once for all
    // variables; every sum keeps its own accumulation order.
    constexpr bool GRAD = Phys::kAuxGrad;
    const bool grad = GRAD && a.aux;
    double dxs[NV], des[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) { dxs[v] = 0.0; des[v] = 0.0; }
    double gx = 0.0, gy = 0.0, dxi = 0.0, deta = 0.0;
    // one step of the sums at line index ii: matrix entries and row-wise (xi) operands come in
    // as pairs (ii, ii + 1) when NP is even (16-byte LDS reads: the pairs are adjacent and
    // aligned in every node and matrix layout); the loop is not unrolled further, so the
    // compiler does not hoist every LDS read of the phase (that took 192 VGPRs)
    auto line = [&](int ii, double ai, double aj, const double *f1, double di, double dj, double bx) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        dxs[v] = dxs[v] + ai * f1[v];
        des[v] = des[v] + aj * sF2[v][el][slot(i, ii)];
      }
      if (GRAD && grad) {
        if constexpr (GRADN) {
          gx = gx + di * sGb[0][el][slot(ii, j)] + dj * sGb[2][el][slot(i, ii)];
          gy = gy + di * sGb[1][el][slot(ii, j)] + dj * sGb[3][el][slot(i, ii)];
        } else {
          dxi = dxi + di * bx;
          deta = deta + dj * sAux[el][slot(i, ii)];
        }
      }
    };
    if constexpr (NP % 2 == 0) {
#pragma unroll 1
      for (int ii = 0; ii < NP; ii += 2) {
        const double2 ai = *reinterpret_cast<const double2 *>(&dgM[ii + MRS * i]);
        const double2 aj = *reinterpret_cast<const double2 *>(&dgM[ii + MRS * j]);
        double2 di = {0.0, 0.0}, dj = {0.0, 0.0}, bx = {0.0, 0.0};
        if (GRAD && grad) {
          di = *reinterpret_cast<const double2 *>(&Dm[ii + MRS * i]);
          dj = *reinterpret_cast<const double2 *>(&Dm[ii + MRS * j]);
          if constexpr (!GRADN) bx = *reinterpret_cast<const double2 *>(&sAux[el][slot(ii, j)]);
        }
        double f1a[NV], f1b[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double2 f = *reinterpret_cast<const double2 *>(&sF1[v][el][slot(ii, j)]);
          f1a[v] = f.x;
          f1b[v] = f.y;
        }
        line(ii, ai.x, aj.x, f1a, di.x, dj.x, bx.x);
        line(ii + 1, ai.y, aj.y, f1b, di.y, dj.y, bx.y);
      }
    } else {
#pragma unroll 2
      for (int ii = 0; ii < NP; ii++) {
        double f1[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) f1[v] = sF1[v][el][slot(ii, j)];
        line(ii, dgM[ii + MRS * i], dgM[ii + MRS * j], f1, Dm[ii + MRS * i], Dm[ii + MRS * j],
             USE_AUX ? sAux[el][slot(ii, j)] : 0.0);
      }
    }
    if (GRAD && grad) {
      if constexpr (!GRADN) {
        // affine element: the metric terms are constant, so they factor out of the sums
        gx = dsdx[0] * dxi + dsdx[2] * deta;
        gy = dsdx[1] * dxi + dsdx[3] * deta;
      }
      gx = gx * invJ;
      gy = gy * invJ;
    }
    double src[NV];
    Phys::source(s, gx, gy, statVal, a.p, src);
    // per-run step size of a batched launch: the same IEEE product g_m*dt as a single run
    const double gdt = (MODE == 0 && a.perElemDt) ? a.rkG * a.geom[(size_t)e * kGeomStride + 5] : a.rkGdt;
    const double iwi = qW[NP + i], iwj = qW[NP + j];  // 1/w_i, 1/w_j
    const double bi1 = bM1[i], bi0 = bM0[i], bj1 = bM1[j], bj0 = bM0[j];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      // surface terms of the weak divergence, then JacobianWeight
      double d = dxs[v] + des[v];
      d = d + (bi1 * sBF[v][el][NP + j] + bi0 * sBF[v][el][3 * NP + j]) * iwi +
          (bj1 * sBF[v][el][2 * NP + i] + bj0 * sBF[v][el][i]) * iwj;
      d = d * invJ;
      if (MODE == 1 && a.fluxDivOut) a.fluxDivOut[sBase + (size_t)v * NN + node] = d;
      double dsdt = src[v] - d;
      if (a.corr) {
        // closure: mass correction made mean-free over the mesh (variable 2 is H / eta / rho)
        double c = a.corr[sBase + (size_t)v * NN + node];
        if (v == 2) c -= a.corrShift[0];
        dsdt += c;
      }
      if (MODE == 1) {
        a.dSdt[sBase + (size_t)v * NN + node] = dsdt;
      } else {
        const double gn = a.rkA * gReg[v] + dsdt;
        s[v] = s[v] + gdt * gn;
        if (a.writeG) a.G[sBase + (size_t)v * NN + node] = gn;
        a.S[sBase + (size_t)v * NN + node] = s[v];
      }
    }
    if (MODE == 0) {
      double q[NT];
      phys_nodal<Phys>(s, a.p, q);
#pragma unroll
      for (int k = 0; k < NT; k++) sQ[k][el][nl] = q[k];
    }
  }
  if (MODE == 1) return;
  phase_sync<WAVE>();

  // ---- phase 3: traces of the updated state ----
  if constexpr (SPLIT) {
    if (faceActive) {
      const double *bm = (side == 1 || side == 2) ? bM1 : bM0;
      const int lo = fhalf ? FH0 : 0;
      double qn[NT];
#pragma unroll
      for (int k = 0; k < NT; k++) qn[k] = 0.0;
#pragma unroll
      for (int t = 0; t < FHN; t++) {
        const int ii = lo + t;
        if (fhalf == 0 && t >= FH0) break;
        const int nd = foff[t];
#pragma unroll
        for (int k = 0; k < NT; k++) qn[k] += bm[ii] * sQ[k][fel][nd];
      }
#pragma unroll
      for (int k = 0; k < NT; k++) qn[k] = pair_sum32(qn[k]);
      if (fhalf == 0) {
        double *tr = a.traceOut + (size_t)fi + NP * (size_t)NT * (side + 4 * (size_t)fe);
#pragma unroll
        for (int k = 0; k < NT; k++) tr[(size_t)k * NP] = qn[k];
      }
    }
    return;
  }
  if constexpr (FPAD) {
    // the face lanes of phase 1, in their padded order
    if (faceActive) {
      const double *bm = (side == 1 || side == 2) ? bM1 : bM0;
      const bool alongJ = (side == 0 || side == 2);
      double qn[NT];
#pragma unroll
      for (int k = 0; k < NT; k++) qn[k] = 0.0;
#pragma unroll
      for (int ii = 0; ii < NP; ii++) {
        const int nd = alongJ ? slot(fi, ii) : slot(ii, fi);
#pragma unroll
        for (int k = 0; k < NT; k++) qn[k] += bm[ii] * sQ[k][fel][nd];
      }
      double *tr = a.traceOut + (size_t)fi + NP * (size_t)NT * (side + 4 * (size_t)fe);
#pragma unroll
      for (int k = 0; k < NT; k++) tr[(size_t)k * NP] = qn[k];
    }
    return;
  }
  for (int ft = tid; ft < EPB * NF; ft += THREADS) {
    const int fel3 = ft / NF, fnode3 = ft % NF;
    const int side3 = fnode3 / NP, fi3 = fnode3 % NP;
    const int pf3 = blk * EPB + fel3;
    if (pf3 >= a.nEl) continue;
    const int fe3 = a.elems ? a.elems[pf3] : pf3;
    const double *bm = (side3 == 1 || side3 == 2) ? bM1 : bM0;
    const bool alongJ = (side3 == 0 || side3 == 2);
    double qn[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) qn[k] = 0.0;
#pragma unroll
    for (int ii = 0; ii < NP; ii++) {
      const int nd = alongJ ? slot(fi3, ii) : slot(ii, fi3);
#pragma unroll
      for (int k = 0; k < NT; k++) qn[k] += bm[ii] * sQ[k][fel3][nd];
    }
    double *tr = a.traceOut + (size_t)fi3 + NP * (size_t)NT * (side3 + 4 * (size_t)fe3);
#pragma unroll
    for (int k = 0; k < NT; k++) tr[(size_t)k * NP] = qn[k];
  }
}

// The stage kernel: one tile per block.  XCD orders the tiles so that the blocks one XCD
// receives (the dispatcher deals blocks to the 8 XCDs in turn) cover a contiguous run of
// elements, whose neighbours then share that XCD's L2; a.reverse visits the tiles last to
// first (successive stages alternate direction).
template <class Phys, int N, int MODE, bool NODAL, int THREADS, int WPE = 1, bool XCD = false>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WPE, 8))) k_fused(const FusedArgs a) {
  __shared__ __align__(16) double lds[TileLds<Phys, N, NODAL, THREADS>::SIZE];
  int blk = blockIdx.x;
  if constexpr (XCD) {
    const int nb = gridDim.x, q = nb >> 3, rem = nb & 7, x = blk & 7, k = blk >> 3;
    blk = x < rem ? x * (q + 1) + k : rem * (q + 1) + (x - rem) * q + k;
  }
  if (a.reverse) blk = (int)gridDim.x - 1 - blk;
  fused_tile<Phys, N, MODE, NODAL, THREADS>(a, blk, lds, threadIdx.x);
}

// Batched runs of several degrees in ONE launch per RK stage: a persistent grid (as many blocks
// as the device holds at once) over the tile lists of the degree groups N = 3, 5, 7, in that
// order (a tile of 256 work-items: 16, 7 or 4 elements).  Blocks are dispatched to the eight
// XCDs in turn (block b to XCD b mod 8), so each group's tile list is cut into eight contiguous
// parts and the blocks of XCD x stride through part x: every block gets the same mix of
// degrees and all blocks sweep the groups together.  A block loads a group's matrices into its
// LDS once; where each element lives in one wavefront (N = 3 and 7) the tiles then need no
// block barrier at all.  Per tile the arithmetic is that of the degree's own stage kernel, so
// every run stays bit-identical to its separate solve; each group's element order alternates
// between stages.
constexpr int kPersistentDegrees[3] = {3, 5, 7};
struct PersistentArgs {
  FusedArgs g[3];  // the groups of degree 3, 5, 7 (nEl = 0 when a degree is absent)
};
static_assert(sizeof(PersistentArgs) <= 4096, "kernel arguments exceed the 4 KiB kernarg segment");

template <class Phys>
__global__ void __launch_bounds__(256) k_batch_persistent(const PersistentArgs b) {
  constexpr int S3 = TileLds<Phys, 3, false, 256>::SIZE, S5 = TileLds<Phys, 5, false, 256>::SIZE,
                S7 = TileLds<Phys, 7, false, 256>::SIZE;
  constexpr int SMAX = S3 > S5 ? (S3 > S7 ? S3 : S7) : (S5 > S7 ? S5 : S7);
  __shared__ __align__(16) double lds[SMAX];
  const int G = gridDim.x, x = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int cx = (G >> 3) + (x < (G & 7) ? 1 : 0);  // blocks on this XCD
  // one loop per degree, so that what the compiler keeps in registers across one degree's
  // tiles is not live during another's
  auto group = [&](auto deg, int k) {
    constexpr int N = decltype(deg)::value;
    using L = TileLds<Phys, N, false, 256>;
    const FusedArgs &a = b.g[k];
    const int nt = (a.nEl + L::EPB - 1) / L::EPB;
    const int l0 = (int)(((long long)nt * x) >> 3) + r, hi = (int)(((long long)nt * (x + 1)) >> 3);
    if (l0 >= hi) return;  // uniform over the block
    // the previous group's tiles are done with the LDS
    __syncthreads();
    load_tile_matrices<Phys, N, false, 256>(a.mats, lds, threadIdx.x);
    __syncthreads();
#pragma unroll 1
    for (int l = l0; l < hi; l += cx) {
      const int blk = a.reverse ? nt - 1 - l : l;
      // the work-item index, opaque to the compiler in each iteration: the tile's lane-derived
      // offsets are then formed per tile as in the one-tile kernel, not hoisted out of the loop
      // and held in registers across it
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      __builtin_assume(tid >= 0 && tid < 256);
      fused_tile<Phys, N, 0, false, 256, true>(a, blk, lds, tid);
      // the next tile reuses the LDS: where each element lives in one wavefront, that
      // wavefront's own accesses are in program order
      phase_sync<L::WAVE>();
    }
  };
  group(std::integral_constant<int, 3>{}, 0);
  group(std::integral_constant<int, 5>{}, 1);
  group(std::integral_constant<int, 7>{}, 2);
}

// ------------------------------------------------------------------------------
// Integrals of the state: CalculateEntropy and mass-type integrals
// ------------------------------------------------------------------------------
// One work-item per node forms w_i w_j J f(S) (the per-node J as SEMQuad stores it); blocks
// sum in a fixed tree order into partial sums, and one block adds those in a fixed order, so
// the result does not depend on scheduling.
constexpr int kIntBlock = 256;
enum IntegralKind { kIntSWEntropy = 0, kIntLSWEntropy = 1, kIntEulerEntropy = 2, kIntVariable = 3 };

__device__ inline void block_sum_store(double v, double *out) {
  __shared__ double red[kIntBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = kIntBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <int WHAT>
__global__ void __launch_bounds__(kIntBlock) k_integral_partial(const double *S, const double *aux, const double *J,
                                                                const double *qW, Params p, int N, int nVar, int var,
                                                                size_t nNodes, double *partial) {
  const size_t n = blockIdx.x * (size_t)kIntBlock + threadIdx.x;
  double v = 0.0;
  if (n < nNodes) {
    const int P = N + 1, PP = P * P;
    const size_t e = n / PP;
    const int node = (int)(n - e * PP), i = node % P, j = node / P;
    const double jac = J[n], wi = qW[i], wj = qW[j];
    const double *s = S + e * (size_t)nVar * PP + node;  // variable k at s[k * PP]
    if (WHAT == kIntSWEntropy) {
      // CalculateEntropy_ShallowWater (src/models/SELF_ShallowWater.f90:149-188)
      const double Hu = s[0], Hv = s[PP], H = s[2 * PP], b = aux ? aux[n] : p.auxConst;
      v = (0.5 * (Hu * Hu / H + Hv * Hv / H) + 0.5 * p.g * H * H - p.g * H * b) * jac * wi * wj;
    } else if (WHAT == kIntLSWEntropy) {
      // CalculateEntropy_LinearShallowWater (src/models/SELF_LinearShallowWater.f90:171-216)
      const double u = s[0], vv = s[PP], eta = s[2 * PP], H = aux ? aux[n] : p.auxConst;
      v = 0.5 * (H * (u * u + vv * vv) + p.g * eta * eta) * jac * wi * wj;
    } else if (WHAT == kIntEulerEntropy) {
      // CalculateEntropy_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:782-823)
      // with the pressure of the current state (ConservativeToPrimitive, :1129-1170)
      const double rhoU = s[0], rhoV = s[PP], rho = s[2 * PP], rhoE = s[3 * PP];
      const double u = rhoU / rho, vv = rhoV / rho;
      const double pr = (p.gamma - 1.0) * (rhoE - 0.5 * (rhoU * u + rhoV * vv));
      v = rho * (log(pr) - p.gamma * log(rho)) / (p.gamma - 1.0) * wi * wj * jac;
    } else {
      v = s[(size_t)var * PP] * wi * wj * jac;
    }
  }
  block_sum_store(v, partial + blockIdx.x);
}

// one block: work-item t adds partial[t], partial[t + kIntBlock], ... in order, then the
// block sums the work-items' totals in a fixed tree order
__global__ void __launch_bounds__(kIntBlock) k_sum_partials(const double *partial, int n, double *out) {
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += kIntBlock) v += partial[k];
  block_sum_store(v, out);
}

// Traces of the nodal quantities of the current state (SetSolution's BoundaryInterp).
template <class Phys, int N>
__global__ void __launch_bounds__(kThreads) k_traces(const double *S, double *traceOut, const double *mats, Params p,
                                                     int nEl) {
  constexpr int NP = N + 1, NN = NP * NP, EPB = kThreads / NN, NF = 4 * NP;
  constexpr int NV = Phys::NV, NT = Phys::NT;
  __shared__ double sQ[NT][EPB][NN];
  const int tid = threadIdx.x;
  const int el = tid / NN, node = tid % NN;
  const int e = blockIdx.x * EPB + el;
  if (el < EPB && e < nEl) {
    double s[NV], q[NT];
    for (int v = 0; v < NV; v++) s[v] = S[(size_t)e * NV * NN + (size_t)v * NN + node];
    phys_nodal<Phys>(s, p, q);
    for (int k = 0; k < NT; k++) sQ[k][el][node] = q[k];
  }
  __syncthreads();
  for (int ft = tid; ft < EPB * NF; ft += kThreads) {
    const int fel = ft / NF, fnode = ft % NF, side = fnode / NP, fi = fnode % NP;
    const int fe = blockIdx.x * EPB + fel;
    if (fe >= nEl) continue;
    const double *bm = mats + 2 * NN + ((side == 1 || side == 2) ? NP : 0);
    const bool alongJ = (side == 0 || side == 2);
    double qn[NT];
    for (int k = 0; k < NT; k++) qn[k] = 0.0;
    for (int ii = 0; ii < NP; ii++) {
      const int nd = alongJ ? fi + NP * ii : ii + NP * fi;
      for (int k = 0; k < NT; k++) qn[k] += bm[ii] * sQ[k][fel][nd];
    }
    double *tr = traceOut + (size_t)fi + NP * (size_t)NT * (side + 4 * (size_t)fe);
    for (int k = 0; k < NT; k++) tr[(size_t)k * NP] = qn[k];
  }
}

// Initial SW lagged normal velocities (own side of each face node) from the current traces:
// at the first stage after SetSolution the reference's velocity traces come from
// SetSolution's BoundaryInterp.
template <int N>
__global__ void k_sw_un_init(const double *trace, double *un, const double *geom, const double *nnHat, int nEl) {
  constexpr int NP = N + 1;
  const size_t total = (size_t)NP * 4 * nEl;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int fi = t % NP, side = (t / NP) % 4;
    const size_t e = t / (4 * NP);
    double nx, ny;
    if (nnHat) {
      nx = nnHat[2 * t]; ny = nnHat[2 * t + 1];
    } else {
      const double *gr = geom + e * kGeomStride;
      nx = gr[8 + side]; ny = gr[12 + side];
    }
    const double *tl = trace + fi + NP * 3 * (side + 4 * e);
    if (nnHat) {
      un[2 * t] = tl[0] / tl[2 * NP];
      un[2 * t + 1] = tl[NP] / tl[2 * NP];
    } else {
      un[t] = (tl[0] / tl[2 * NP]) * nx + (tl[NP] / tl[2 * NP]) * ny;
    }
  }
}

// Block size (tile) per polynomial degree: one wavefront per element at N = 7, larger
// blocks where an element covers only part of a wavefront.
template <int N>
constexpr int tile_threads() {
  return N == 7 ? 64 : N == 6 ? 64 : N == 5 ? 128 : N == 4 ? 128 : 256;
}

// Tuned tiles: the stage kernel (affine metrics, RK update) of SW at N = 3, 5, 7 and of LSW
// and Euler at N = 7 comes in 18 variants.  v < 12: v = 4 t + 2 r + x with THREADS = 64 << t,
// element order alternating between stages when r = 1, XCD-contiguous element order when
// x = 1.  v = 12 + 2 t + w: THREADS = 64 << t, alternating order, and an occupancy target of
// 5 + w wavefronts per SIMD for the register allocator (a few registers spill to scratch).  The table below (per N; -1 =
// the default tile) is set by selfamd_tile_set, normally from the autotuner's measurements.
// Variants whose blocks leave room for two lanes per face node split the face-line sums
// between the pair (round-off level differences); otherwise the arithmetic is identical.
constexpr int kTileVariants = 18;
// g_tile[kind][N]: SW at N = 3, 5, 7; LSW and Euler at N = 7
int g_tile[4][8] = {{-1, -1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1, -1},
                    {-1, -1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1, -1}};
inline bool tunable(int kind, int N) {
  return (kind == SELFAMD_MODEL_SW && (N == 3 || N == 5 || N == 7)) ||
         ((kind == SELFAMD_MODEL_LSW || kind == SELFAMD_MODEL_EULER) && N == 7);
}
inline int tile_of(int kind, int N) { return tunable(kind, N) ? g_tile[kind][N] : -1; }
// Launches of at most kSmallLaunch elements (a few per CU) take small tiles in plain element
// order, whatever the table (tuned on 256^2 meshes) says.  Measured per launch: SW at N = 7,
// 64 wide, 4.75 against 5.54 us at 10^2 and 10.8 against 14.7 us at 64^2 (the tuned 256-wide
// alternating tile wins from 96^2 on); SW at N = 5, 128 wide in XCD order, 8.6 against 12.1 us
// at 64^2 (equal at 128^2); LSW and Euler at N = 7, 64 wide, 4.4 against 5.1 us (10^2) and
// 7.2 against 9.6 us (32^2).  The face sums split as the table's tile does (N = 5 keeps a
// table entry of 64 wide, which splits), so a run's results do not depend on launch sizes.
constexpr int kSmallLaunch = 8192;
inline int tile_for(int kind, int N, int nEl) {
  const int v = tile_of(kind, N);
  if (tunable(kind, N) && nEl <= kSmallLaunch) {
    if (N == 7) return 0;
    const bool split64 = v >= 0 && ((v < 12 && (v >> 2) == 0) || v == 12 || v == 13);
    if (N == 5 && !split64) return 5;
  }
  return v;
}
// element order alternates between stages (the default tiles alternate too)
inline bool tile_alternates(int kind, int N, int nEl) {
  const int v = tile_for(kind, N, nEl);
  return v < 0 || v >= 12 || (v & 2);
}
int g_tileEpoch = 0;


template <class Phys, int N>
hipError_t launch_tile(int v, const FusedArgs &a, hipStream_t st) {
  constexpr int NN = (N + 1) * (N + 1);
  switch (v) {
#define V(IDX, T, W, X)                                                                        \
  case IDX: {                                                                                  \
    constexpr int EPB = T / NN;                                                                \
    hipLaunchKernelGGL((k_fused<Phys, N, 0, false, T, W, X>), dim3((a.nEl + EPB - 1) / EPB), T, 0, st, a); \
    return hipGetLastError();                                                                  \
  }
    V(0, 64, 1, false) V(1, 64, 1, true) V(2, 64, 1, false) V(3, 64, 1, true)
    V(4, 128, 1, false) V(5, 128, 1, true) V(6, 128, 1, false) V(7, 128, 1, true)
    V(8, 256, 1, false) V(9, 256, 1, true) V(10, 256, 1, false) V(11, 256, 1, true)
    V(12, 64, 5, false) V(13, 64, 6, false) V(14, 128, 5, false) V(15, 128, 6, false)
    V(16, 256, 5, false) V(17, 256, 6, false)
#undef V
    default: return hipErrorInvalidValue;
  }
}

template <class Phys>
struct Launcher {
  static hipError_t stage(int N, int mode, bool nodal, const FusedArgs &a, hipStream_t st) {
    const int v = (mode == 0 && !nodal) ? tile_for(Phys::kKind, N, a.nEl) : -1;
    if (v >= 0) {
      if constexpr (std::is_same<Phys, PhysSW>::value) {
        if (N == 3) return launch_tile<Phys, 3>(v, a, st);
        if (N == 5) return launch_tile<Phys, 5>(v, a, st);
      }
      if (N == 7) return launch_tile<Phys, 7>(v, a, st);
    }
    switch (N) {
#define CASE(NN_)                                                                                     \
  case NN_: {                                                                                         \
    constexpr int T = tile_threads<NN_>();                                                            \
    constexpr int EPB = T / ((NN_ + 1) * (NN_ + 1));                                                  \
    dim3 grid((a.nEl + EPB - 1) / EPB);                                                               \
    if (nodal) {                                                                                      \
      if (mode == 0) hipLaunchKernelGGL((k_fused<Phys, NN_, 0, true, T>), grid, T, 0, st, a);         \
      else hipLaunchKernelGGL((k_fused<Phys, NN_, 1, true, T>), grid, T, 0, st, a);                   \
    } else {                                                                                          \
      if (mode == 0) hipLaunchKernelGGL((k_fused<Phys, NN_, 0, false, T>), grid, T, 0, st, a);        \
      else hipLaunchKernelGGL((k_fused<Phys, NN_, 1, false, T>), grid, T, 0, st, a);                  \
    }                                                                                                 \
    return hipGetLastError();                                                                         \
  }
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
      default: return hipErrorInvalidValue;
    }
  }
  static hipError_t traces(int N, const double *S, double *tr, const double *mats, const Params &p, int nEl,
                           hipStream_t st) {
    switch (N) {
#define CASE(NN_)                                                                                   \
  case NN_: {                                                                                       \
    constexpr int EPB = kThreads / ((NN_ + 1) * (NN_ + 1));                                         \
    hipLaunchKernelGGL((k_traces<Phys, NN_>), dim3((nEl + EPB - 1) / EPB), kThreads, 0, st, S, tr, mats, p, nEl); \
    return hipGetLastError();                                                                       \
  }
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
      default: return hipErrorInvalidValue;
    }
  }
};

hipError_t launch_un_init(int N, const double *tr, double *un, const double *geom, const double *nnHat, int nEl,
                          hipStream_t st) {
  const size_t n = (size_t)(N + 1) * 4 * nEl;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  switch (N) {
#define CASE(NN_) \
  case NN_: hipLaunchKernelGGL((k_sw_un_init<NN_>), dim3(blocks), 256, 0, st, tr, un, geom, nnHat, nEl); return hipGetLastError();
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// ------------------------------------------------------------------------------
// Model2D engine
// ------------------------------------------------------------------------------
struct selfamd_model2d {
  int kind = 0, N = 0, nEl = 0, nVar = 0, nT = 0;
  int device = 0;  // the GPU the model's buffers live on (current device at creation)
  hipStream_t stream = nullptr;
  bool ownStream = false;
  double *S = nullptr, *G = nullptr, *trace[2] = {nullptr, nullptr}, *un[2] = {nullptr, nullptr};
  double *geom = nullptr, *mats = nullptr, *aux = nullptr, *stat2 = nullptr, *presTrace = nullptr;
  // per-node metrics (copied from SEMQuad at creation); used when `nodal` is set
  double *ndsdx = nullptr, *nJ = nullptr, *nnHat = nullptr, *nnScale = nullptr;
  bool affine = true, nodal = false;
  bool anyBoundary = false;  // some element side is a physical boundary (neighbour code < 0)
  int32_t *nbr = nullptr;
  int cur = 0;  // trace buffer holding the traces of the current S
  Params p{};
  double t = 0.0;
  // captured one-step graphs, one per starting trace buffer
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // graphs of 2^k steps (k = 1 .. top_level()), per starting trace buffer, captured (for both
  // buffers) when a call first needs them; 2^k steps flip the trace buffer an even number of
  // times, so the start parity is unchanged after each.  A call replays the largest as often as it fits, then each
  // smaller one at most once and, for an odd count, the one-step graph.  Between two replays
  // the GPU idles ~9 us, a tenth of a stage launch on small meshes.
  static constexpr int kGraphLevels = 7;        // up to 64 steps per graph
  static constexpr int kGraphLevelsBanded = 4;  // banded stepping: up to 8 (many more nodes per step)
  hipGraphExec_t graphP[kGraphLevels][2] = {};
  int top_level() const { return (nBands > 0 ? kGraphLevelsBanded : kGraphLevels) - 1; }
  double graphDt = -1.0;
  int graphEpoch = -1;  // tile table version the graphs were captured with
  ClosureBuffers *closure = nullptr;
  double *intScratch = nullptr;  // partial sums of the state integrals (allocated on first use)
  size_t intScratchLen = 0;
  double *adjTrace = nullptr, *adjChi = nullptr;  // tendency adjoint: traces of its state, neighbour-trace cotangents

  // Banded stepping: the elements split into bands of whole BFS levels of the element graph
  // (neighbours differ by at most one band); one RK3 step runs stage s of band b in round
  // b + s, rounds in order and, within a round, stages in order.  A band's data are then
  // touched three times within ~3 bands' worth of traffic, i.e. while they are still in the
  // Infinity Cache.  Dependencies (neighbour traces of the previous stage in the buffer of
  // that stage's parity, none overwritten before its last reader) hold by this order.
  int nBands = 0;
  int32_t *bandElems = nullptr;   // elements sorted by band
  std::vector<int> bandStart;     // nBands + 1 offsets into bandElems
  std::vector<int32_t> hnbr;      // host copy of the neighbour codes

  hipError_t launch_stage(int mode, int stage, double dt, int traceIn, double *dSdt, int band = -1,
                          double *fluxOut = nullptr, double *fluxDivOut = nullptr) {
    FusedArgs a{};
    a.fluxOut = fluxOut;
    a.fluxDivOut = fluxDivOut;
    if (closure) {
      hipError_t e = closure_evaluate(*closure, S, geom, mats + 2 * (N + 1) * (N + 1) + 2 * (N + 1), N, nVar, 2, nEl,
                                      stream);
      if (e != hipSuccess) return e;
      a.corr = closure->corr;
      a.corrShift = closure->shift;
    }
    a.S = S; a.G = G; a.dSdt = dSdt;
    a.traceIn = trace[traceIn];
    a.traceOut = trace[traceIn ^ 1];
    a.unIn = un[traceIn]; a.unOut = un[traceIn ^ 1];
    a.aux = aux; a.stat2 = stat2; a.presTrace = presTrace;
    a.geom = geom; a.nbr = nbr; a.mats = mats; a.nEl = nEl;
    if (band >= 0) {
      a.elems = bandElems + bandStart[band];
      a.nEl = bandStart[band + 1] - bandStart[band];
    }
    a.ndsdx = ndsdx; a.nJ = nJ; a.nnHat = nnHat; a.nnScale = nnScale;
    a.rkA = kRk3A[stage];
    a.rkGdt = kRk3G[stage] * dt;
    a.rkG = kRk3G[stage];
    a.writeG = stage != 2;
    // element order alternates with the trace buffer parity, i.e. from stage to stage
    a.reverse = tile_alternates(kind, N, a.nEl) ? traceIn : 0;
    a.perElemDt = 0;
    a.anyBoundary = anyBoundary;
    a.p = p;
    switch (kind) {
      case SELFAMD_MODEL_SW: return Launcher<PhysSW>::stage(N, mode, nodal, a, stream);
      case SELFAMD_MODEL_LSW: return Launcher<PhysLSW>::stage(N, mode, nodal, a, stream);
      default: return Launcher<PhysEuler>::stage(N, mode, nodal, a, stream);
    }
  }
  hipError_t launch_traces() {
    switch (kind) {
      case SELFAMD_MODEL_SW: return Launcher<PhysSW>::traces(N, S, trace[cur], mats, p, nEl, stream);
      case SELFAMD_MODEL_LSW: return Launcher<PhysLSW>::traces(N, S, trace[cur], mats, p, nEl, stream);
      default: return Launcher<PhysEuler>::traces(N, S, trace[cur], mats, p, nEl, stream);
    }
  }
  void drop_graphs() {
    for (auto &g : graph)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto &lv : graphP)
      for (auto &g : lv)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    graphDt = -1.0;
  }
  // captures `steps` RK3 steps starting from trace buffer `start`
  hipError_t capture(int start, int steps, double dt, hipGraphExec_t *out) {
    hipGraph_t g;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    int tin = start;
    hipError_t le = hipSuccess;
    for (int n = 0; n < steps; n++) {
      if (nBands > 0) {
        for (int r = 0; r < nBands + 2 && le == hipSuccess; r++)
          for (int st = 0; st < 3 && le == hipSuccess; st++) {
            const int b = r - st;
            if (b >= 0 && b < nBands) le = launch_stage(0, st, dt, tin ^ (st & 1), nullptr, b);
          }
        tin ^= 1;  // three stages
      } else {
        for (int st = 0; st < 3 && le == hipSuccess; st++) {
          le = launch_stage(0, st, dt, tin, nullptr);
          tin ^= 1;
        }
      }
    }
    hipError_t ce = hipStreamEndCapture(stream, &g);
    if (le != hipSuccess) return le;
    if (ce != hipSuccess) return ce;
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    // the graph's first replay would otherwise carry its upload to the device (~25 us at 8
    // steps), inside whatever timed region first needs it
    if (ie == hipSuccess) ie = hipGraphUpload(*out, stream);
    return ie;
  }
  // `steps` RK3 steps launched stage by stage on the stream, without a graph
  hipError_t launch_steps(int steps, double dt) {
    for (int n = 0; n < steps; n++) {
      if (nBands > 0) {
        for (int r = 0; r < nBands + 2; r++)
          for (int st = 0; st < 3; st++) {
            const int b = r - st;
            if (b >= 0 && b < nBands) {
              hipError_t e = launch_stage(0, st, dt, cur ^ (st & 1), nullptr, b);
              if (e != hipSuccess) return e;
            }
          }
        cur ^= 1;
      } else {
        for (int st = 0; st < 3; st++) {
          hipError_t e = launch_stage(0, st, dt, cur, nullptr);
          if (e != hipSuccess) return e;
          cur ^= 1;
        }
      }
      t += dt;
    }
    return hipSuccess;
  }
};

using Model = selfamd_model2d;

namespace {
void free_closure(ClosureBuffers *c) {
  if (!c) return;
  void *bufs[] = {c->X,    c->Y,     c->inMean, c->inScale, c->outMean,   c->outScale,
                  c->corr, c->elemMass, c->shift, c->dY,      c->dX,        c->vjpPartial};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete c;
}
}  // namespace

extern "C" {

int selfamd_model2d_nvar(int kind) { return kind == SELFAMD_MODEL_EULER ? 4 : 3; }

int selfamd_model2d_create(selfamd_model2d **out, int kind, int N, int nEl, const int32_t *sideInfo,
                           const double *dsdx, const double *J, const double *nHatb, const double *nScaleb,
                           void *stream) {
  SELFAMD_REQUIRE(out, "model2d_create: null output handle");
  *out = nullptr;
  SELFAMD_REQUIRE(kind == SELFAMD_MODEL_LSW || kind == SELFAMD_MODEL_SW || kind == SELFAMD_MODEL_EULER,
                  "model2d_create: unknown model kind %d", kind);
  SELFAMD_REQUIRE(N >= 1 && N <= 7, "model2d_create: fused kernels cover 1 <= N <= 7 (got %d)", N);
  SELFAMD_REQUIRE(nEl >= 1, "model2d_create: nEl=%d", nEl);
  SELFAMD_REQUIRE(sideInfo && dsdx && J && nHatb && nScaleb, "model2d_create: null geometry/connectivity");
  const int P = N + 1, PP = P * P;
  // neighbour codes from sideInfo (validated: bit-exact copy of the connectivity)
  std::vector<int32_t> nbr((size_t)4 * nEl);
  for (int e = 0; e < nEl; e++)
    for (int s = 0; s < 4; s++) {
      const int32_t *si = sideInfo + 5 * (s + 4 * (size_t)e);
      const int e2 = si[2], bcid = si[4];
      int code;
      if (bcid == 0) {
        const int s2 = si[3] / 10, flip = si[3] - 10 * s2;
        SELFAMD_REQUIRE(e2 >= 1 && e2 <= nEl && s2 >= 1 && s2 <= 4 && (flip == 0 || flip == 1),
                        "model2d_create: bad interior side (elem %d side %d): e2=%d s2=%d flip=%d", e + 1, s + 1,
                        e2, s2, flip);
        SELFAMD_REQUIRE(e2 - 1 < (1 << 27), "model2d_create: too many elements");
        code = ((e2 - 1) << 3) | ((s2 - 1) << 1) | flip;
      } else if (e2 == 0) {
        code = -bcid - 1;
      } else {
        code = kNbrOrphan;  // periodic-tagged HOPr side: never exchanged, ext state stays 0
      }
      nbr[4 * (size_t)e + s] = code;
    }
  // per-element affine geometry records
  const size_t nn = (size_t)PP * nEl, nb = (size_t)P * 4 * nEl;
  std::vector<double> hdsdx(4 * nn), hJ(nn), hn(2 * nb), hs(nb);
  SELFAMD_HIP(hipMemcpy(hdsdx.data(), dsdx, sizeof(double) * 4 * nn, hipMemcpyDeviceToHost));
  SELFAMD_HIP(hipMemcpy(hJ.data(), J, sizeof(double) * nn, hipMemcpyDeviceToHost));
  SELFAMD_HIP(hipMemcpy(hn.data(), nHatb, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost));
  SELFAMD_HIP(hipMemcpy(hs.data(), nScaleb, sizeof(double) * nb, hipMemcpyDeviceToHost));
  // Per-element affine records: metric terms constant over each element up to round-off use
  // one record; otherwise the nodal metrics are used.  The round-off of a derivative of
  // interpolated coordinates grows like (coordinate magnitude / element size) x eps: on the
  // unit square it is ~3e-12 relative at 64 x 64 elements, ~1e-11 at 256 x 256 and ~2e-11 at
  // 512 x 512, so the tolerance is 1e-10 (meshes up to ~2048 elements across).
  constexpr double kAffineTol = 1e-10;
  std::vector<double> rec((size_t)kGeomStride * nEl, 0.0);
  bool affine = true;
  for (int e = 0; e < nEl && affine; e++) {
    double *r = rec.data() + (size_t)kGeomStride * e;
    const double *d0 = hdsdx.data() + 4 * (size_t)PP * e;
    double scale = 0.0;
    for (int k = 0; k < 4; k++) { r[k] = d0[k]; scale = std::fmax(scale, std::fabs(d0[k])); }
    r[4] = hJ[(size_t)PP * e];
    r[6] = 1.0 / r[4];
    for (int n = 0; n < PP && affine; n++) {
      for (int k = 0; k < 4; k++)
        if (std::fabs(d0[4 * n + k] - r[k]) > kAffineTol * scale) affine = false;
      if (std::fabs(hJ[(size_t)PP * e + n] - r[4]) > kAffineTol * std::fabs(r[4])) affine = false;
    }
    for (int s = 0; s < 4 && affine; s++) {
      const size_t b0 = (size_t)P * (s + 4 * (size_t)e);
      r[8 + s] = hn[2 * b0];
      r[12 + s] = hn[2 * b0 + 1];
      r[16 + s] = hs[b0];
      for (int i = 1; i < P; i++)
        if (std::fabs(hn[2 * (b0 + i)] - r[8 + s]) > kAffineTol || std::fabs(hn[2 * (b0 + i) + 1] - r[12 + s]) > kAffineTol ||
            std::fabs(hs[b0 + i] - r[16 + s]) > kAffineTol * std::fabs(r[16 + s]))
          affine = false;
    }
  }
  // Lagrange matrices on Gauss points
  double cp[8], tp[8], bw[8], qw[8], iM[64], dM[64], dgM[64], bM[16];
  int rc = selfamd_lagrange_init(N, SELFAMD_GAUSS, N, SELFAMD_GAUSS, cp, tp, bw, qw, iM, dM, dgM, bM);
  if (rc) return rc;
  std::vector<double> mats(2 * PP + 4 * P);
  for (int k = 0; k < PP; k++) { mats[k] = dgM[k]; mats[PP + k] = dM[k]; }
  for (int k = 0; k < 2 * P; k++) mats[2 * PP + k] = bM[k];
  for (int k = 0; k < P; k++) mats[2 * PP + 2 * P + k] = qw[k];
  for (int k = 0; k < P; k++) mats[2 * PP + 3 * P + k] = 1.0 / qw[k];

  Model *m = new Model();
  m->kind = kind; m->N = N; m->nEl = nEl;
  m->hnbr = nbr;
  m->anyBoundary = std::any_of(nbr.begin(), nbr.end(), [](int32_t c) { return c < 0; });
  m->affine = affine;
  m->nodal = !affine;
  m->nVar = selfamd_model2d_nvar(kind);
  m->nT = kind == SELFAMD_MODEL_EULER ? 9 : 3;
  m->p.g = 1.0; m->p.fCori = 0.0; m->p.gamma = 1.4; m->p.Rgas = 1.0;
  m->p.auxConst = 0.0;  // static fields default to zero until a Set* call provides them
  m->p.stat2Const = 0.0;
  auto fail = [&](int code) { selfamd_model2d_destroy(m); return code; };
  if (hipGetDevice(&m->device) != hipSuccess) {
    set_error("model2d_create: hipGetDevice failed");
    return fail(kHipError);
  }
  if (stream) {
    m->stream = reinterpret_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("model2d_create: hipStreamCreate failed");
      return fail(kHipError);
    }
    m->ownStream = true;
  }
  const size_t nS = nn * m->nVar, nT = (size_t)P * m->nT * 4 * nEl;
  if (hipMalloc(&m->S, sizeof(double) * nS) != hipSuccess || hipMalloc(&m->G, sizeof(double) * nS) != hipSuccess ||
      hipMalloc(&m->trace[0], sizeof(double) * nT) != hipSuccess ||
      hipMalloc(&m->trace[1], sizeof(double) * nT) != hipSuccess ||
      hipMalloc(&m->geom, sizeof(double) * rec.size()) != hipSuccess ||
      hipMalloc(&m->mats, sizeof(double) * mats.size()) != hipSuccess ||
      hipMalloc(&m->nbr, sizeof(int32_t) * nbr.size()) != hipSuccess ||
      hipMalloc(&m->ndsdx, sizeof(double) * 4 * nn) != hipSuccess || hipMalloc(&m->nJ, sizeof(double) * nn) != hipSuccess ||
      hipMalloc(&m->nnHat, sizeof(double) * 2 * nb) != hipSuccess ||
      hipMalloc(&m->nnScale, sizeof(double) * nb) != hipSuccess) {
    set_error("model2d_create: device allocation failed (%zu doubles of state)", nS);
    return fail(kHipError);
  }
  // lagged velocities: room for the nodal mode's vectors (2 doubles per face node)
  if (kind == SELFAMD_MODEL_SW && (hipMalloc(&m->un[0], sizeof(double) * 2 * (size_t)P * 4 * nEl) != hipSuccess ||
                                   hipMalloc(&m->un[1], sizeof(double) * 2 * (size_t)P * 4 * nEl) != hipSuccess)) {
    set_error("model2d_create: device allocation failed (un)");
    return fail(kHipError);
  }
  if (hipMemcpy(m->geom, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->mats, mats.data(), sizeof(double) * mats.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->nbr, nbr.data(), sizeof(int32_t) * nbr.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->ndsdx, dsdx, sizeof(double) * 4 * nn, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(m->nJ, J, sizeof(double) * nn, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(m->nnHat, nHatb, sizeof(double) * 2 * nb, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(m->nnScale, nScaleb, sizeof(double) * nb, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemsetAsync(m->S, 0, sizeof(double) * nS, m->stream) != hipSuccess ||
      hipMemsetAsync(m->G, 0, sizeof(double) * nS, m->stream) != hipSuccess ||
      hipMemsetAsync(m->trace[0], 0, sizeof(double) * nT, m->stream) != hipSuccess ||
      hipMemsetAsync(m->trace[1], 0, sizeof(double) * nT, m->stream) != hipSuccess ||
      (m->un[0] && hipMemsetAsync(m->un[0], 0, sizeof(double) * 2 * (size_t)P * 4 * nEl, m->stream) != hipSuccess) ||
      (m->un[1] && hipMemsetAsync(m->un[1], 0, sizeof(double) * 2 * (size_t)P * 4 * nEl, m->stream) != hipSuccess) ||
      hipStreamSynchronize(m->stream) != hipSuccess) {
    set_error("model2d_create: device initialisation failed");
    return fail(kHipError);
  }
  *out = m;
  return kOk;
}

int selfamd_model2d_destroy(selfamd_model2d *m) {
  if (!m) return kOk;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  m->drop_graphs();
  free_closure(m->closure);
  double *bufs[] = {m->S,    m->G,    m->trace[0], m->trace[1], m->un[0], m->un[1], m->geom,  m->mats,
                    m->aux,  m->stat2, m->presTrace, m->ndsdx, m->nJ,    m->nnHat, m->nnScale};
  if (m->bandElems) (void)hipFree(m->bandElems);
  if (m->intScratch) (void)hipFree(m->intScratch);
  if (m->adjTrace) (void)hipFree(m->adjTrace);
  if (m->adjChi) (void)hipFree(m->adjChi);
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (m->nbr) (void)hipFree(m->nbr);
  if (m->ownStream && m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return kOk;
}

int selfamd_model2d_set_param(selfamd_model2d *m, int which, double value) {
  SELFAMD_REQUIRE(m, "model2d_set_param: null model");
  switch (which) {
    case SELFAMD_PARAM_G: m->p.g = value; break;
    case SELFAMD_PARAM_FCORI: m->p.fCori = value; break;
    case SELFAMD_PARAM_GAMMA: m->p.gamma = value; break;
    case SELFAMD_PARAM_RGAS: m->p.Rgas = value; break;
    case SELFAMD_PARAM_NODAL_METRICS:
      SELFAMD_REQUIRE(value != 0.0 || m->affine, "model2d_set_param: the mesh is not affine; nodal metrics required");
      if (m->nodal != (value != 0.0) && m->kind == SELFAMD_MODEL_SW) {
        // the two modes keep the SW velocity lag in different layouts: restart it from the
        // current traces, as SetSolution does
        m->nodal = value != 0.0;
        SELFAMD_HIP(launch_un_init(m->N, m->trace[m->cur], m->un[m->cur], m->geom, m->nodal ? m->nnHat : nullptr,
                                   m->nEl, m->stream));
        SELFAMD_HIP(hipStreamSynchronize(m->stream));
      }
      m->nodal = value != 0.0;
      break;
    default: SELFAMD_REQUIRE(false, "model2d_set_param: unknown parameter %d", which);
  }
  m->drop_graphs();
  return kOk;
}

int selfamd_model2d_set_field(selfamd_model2d *m, int which, const double *field) {
  SELFAMD_REQUIRE(m && field, "model2d_set_field: null argument");
  const size_t n = (size_t)(m->N + 1) * (m->N + 1) * m->nEl;
  double **slot = nullptr;
  if ((which == SELFAMD_FIELD_TOPOGRAPHY && m->kind == SELFAMD_MODEL_SW) ||
      (which == SELFAMD_FIELD_BATHYMETRY && m->kind == SELFAMD_MODEL_LSW) ||
      (which == SELFAMD_FIELD_POTENTIAL && m->kind == SELFAMD_MODEL_EULER))
    slot = &m->aux;
  else if ((which == SELFAMD_FIELD_CORIOLIS && m->kind == SELFAMD_MODEL_LSW) ||
           (which == SELFAMD_FIELD_DRAG && m->kind == SELFAMD_MODEL_EULER))
    slot = &m->stat2;
  SELFAMD_REQUIRE(slot, "model2d_set_field: field %d does not apply to model kind %d", which, m->kind);
  if (!*slot) SELFAMD_HIP(hipMalloc(slot, sizeof(double) * n));
  SELFAMD_HIP(hipMemcpyAsync(*slot, field, sizeof(double) * n, hipMemcpyDeviceToDevice, m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  m->drop_graphs();
  return kOk;
}

int selfamd_model2d_set_solution(selfamd_model2d *m, const double *interior) {
  SELFAMD_REQUIRE(m && interior, "model2d_set_solution: null argument");
  const size_t n = (size_t)(m->N + 1) * (m->N + 1) * m->nVar * m->nEl;
  SELFAMD_HIP(hipMemcpyAsync(m->S, interior, sizeof(double) * n, hipMemcpyDeviceToDevice, m->stream));
  SELFAMD_HIP(m->launch_traces());
  if (m->kind == SELFAMD_MODEL_SW)
    SELFAMD_HIP(launch_un_init(m->N, m->trace[m->cur], m->un[m->cur], m->geom, m->nodal ? m->nnHat : nullptr, m->nEl,
                               m->stream));
  if (m->kind == SELFAMD_MODEL_EULER) {
    // prescribed boundary state := traces of this state (SetPrescribedSolutionFromSolution)
    const size_t nT = (size_t)(m->N + 1) * 9 * 4 * m->nEl;
    if (!m->presTrace) SELFAMD_HIP(hipMalloc(&m->presTrace, sizeof(double) * nT));
    SELFAMD_HIP(hipMemcpyAsync(m->presTrace, m->trace[m->cur], sizeof(double) * nT, hipMemcpyDeviceToDevice, m->stream));
    m->drop_graphs();
  }
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_get_solution(selfamd_model2d *m, double *interior) {
  SELFAMD_REQUIRE(m && interior, "model2d_get_solution: null argument");
  const size_t n = (size_t)(m->N + 1) * (m->N + 1) * m->nVar * m->nEl;
  SELFAMD_HIP(hipMemcpyAsync(interior, m->S, sizeof(double) * n, hipMemcpyDeviceToDevice, m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

// integral of the current state over the mesh (the integrand kind, and the variable for
// kIntVariable); synchronises the model's stream
static int model_integral(selfamd_model2d *m, int what, int var, double *value, const char *who) {
  const size_t PP = (size_t)(m->N + 1) * (m->N + 1), nNodes = PP * m->nEl;
  const int nBlocks = (int)((nNodes + kIntBlock - 1) / kIntBlock);
  if (m->intScratchLen < (size_t)nBlocks + 1) {
    if (m->intScratch) SELFAMD_HIP(hipFree(m->intScratch));
    m->intScratch = nullptr;
    m->intScratchLen = 0;
    if (hipMalloc(&m->intScratch, sizeof(double) * (nBlocks + 1)) != hipSuccess) {
      set_error("%s: device allocation failed", who);
      return kHipError;
    }
    m->intScratchLen = (size_t)nBlocks + 1;
  }
  const double *qW = m->mats + 2 * PP + 2 * (m->N + 1);
  switch (what) {
#define INT_CASE(W)                                                                                        \
  case W:                                                                                                 \
    hipLaunchKernelGGL((k_integral_partial<W>), dim3(nBlocks), dim3(kIntBlock), 0, m->stream, m->S, m->aux, \
                       m->nJ, qW, m->p, m->N, m->nVar, var, nNodes, m->intScratch);                         \
    break;
    INT_CASE(kIntSWEntropy) INT_CASE(kIntLSWEntropy) INT_CASE(kIntEulerEntropy) INT_CASE(kIntVariable)
#undef INT_CASE
    default: set_error("%s: unknown integral", who); return kBadArgument;
  }
  SELFAMD_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kIntBlock), 0, m->stream, m->intScratch, nBlocks,
                     m->intScratch + nBlocks);
  SELFAMD_CHECK_LAUNCH();
  SELFAMD_HIP(hipMemcpyAsync(value, m->intScratch + nBlocks, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

namespace {
// Partial sums of selfamd_model2d_integral_noalloc: one per block of kIntBlock nodes, then the
// total.  A file-scope device array, so that entry point allocates nothing; every GPU that runs
// it holds its own copy (512 KiB) until the process ends.  Host code must not pass `scratch`
// itself to a kernel (in host code the name is a host-side stand-in): it uses the device
// address that scratch_address() looks up once per device.
constexpr int kNoAllocPartials = 65536;  // meshes of up to 65536 * 256 = 16,777,216 nodes
__device__ __align__(256) double scratch[kNoAllocPartials + 1];
constexpr int kMaxDevices = 64;
double *g_scratchDev[kMaxDevices] = {};
std::mutex g_scratchMutex;

// device address of `scratch` on device `dev`, which must be the current device
hipError_t scratch_address(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchDev[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchDev[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratchDev[dev];
  return hipSuccess;
}
}  // namespace

int selfamd_model2d_integral_noalloc(selfamd_model2d *m, int var, double *value) {
  SELFAMD_REQUIRE(m && value, "model2d_integral_noalloc: null argument");
  SELFAMD_REQUIRE(var >= 0 && var < m->nVar, "model2d_integral_noalloc: variable %d out of range (nVar = %d)", var,
                  m->nVar);
  const size_t PP = (size_t)(m->N + 1) * (m->N + 1), nNodes = PP * m->nEl;
  const size_t nBlocks = (nNodes + kIntBlock - 1) / kIntBlock;
  SELFAMD_REQUIRE(nBlocks <= (size_t)kNoAllocPartials,
                  "model2d_integral_noalloc: the mesh has %zu nodes; this entry point takes at most %zu "
                  "(use selfamd_model2d_integral)", nNodes, (size_t)kNoAllocPartials * kIntBlock);
  int dev = -1;
  SELFAMD_HIP(hipGetDevice(&dev));
  SELFAMD_REQUIRE(dev == m->device, "model2d_integral_noalloc: the model lives on device %d, the current device is %d",
                  m->device, dev);
  SELFAMD_REQUIRE(dev >= 0 && dev < kMaxDevices, "model2d_integral_noalloc: device %d out of range", dev);
  double *part = nullptr;
  SELFAMD_HIP(scratch_address(dev, &part));
  // every block writes its own partial sum, so nothing in the array needs clearing first
  const double *qW = m->mats + 2 * PP + 2 * (m->N + 1);
  hipLaunchKernelGGL((k_integral_partial<kIntVariable>), dim3((unsigned)nBlocks), dim3(kIntBlock), 0, m->stream, m->S,
                     m->aux, m->nJ, qW, m->p, m->N, m->nVar, var, nNodes, part);
  SELFAMD_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kIntBlock), 0, m->stream, part, (int)nBlocks, value);
  SELFAMD_CHECK_LAUNCH();
  return kOk;
}

int selfamd_model2d_entropy(selfamd_model2d *m, double *entropy) {
  SELFAMD_REQUIRE(m && entropy, "model2d_entropy: null argument");
  const int what = m->kind == SELFAMD_MODEL_SW ? kIntSWEntropy
                   : m->kind == SELFAMD_MODEL_LSW ? kIntLSWEntropy : kIntEulerEntropy;
  return model_integral(m, what, 0, entropy, "model2d_entropy");
}

int selfamd_model2d_integral(selfamd_model2d *m, int var, double *value) {
  SELFAMD_REQUIRE(m && value, "model2d_integral: null argument");
  SELFAMD_REQUIRE(var >= 0 && var < m->nVar, "model2d_integral: variable %d out of range (nVar = %d)", var, m->nVar);
  return model_integral(m, kIntVariable, var, value, "model2d_integral");
}

int selfamd_model2d_tendency(selfamd_model2d *m, double *dSdt) {
  SELFAMD_REQUIRE(m && dSdt, "model2d_tendency: null argument");
  SELFAMD_HIP(m->launch_stage(1, 0, 0.0, m->cur, dSdt));
  if (m->un[0]) {
    // the evaluation's velocities become the lag state of the next stage
    SELFAMD_HIP(hipMemcpyAsync(m->un[m->cur], m->un[m->cur ^ 1],
                               sizeof(double) * (m->nodal ? 2 : 1) * (m->N + 1) * 4 * (size_t)m->nEl,
                               hipMemcpyDeviceToDevice, m->stream));
  }
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_flux_registers(selfamd_model2d *m, double *fluxDivergence, double *flux) {
  SELFAMD_REQUIRE(m && fluxDivergence && flux, "model2d_flux_registers: null argument");
  // a tendency evaluation of the current state that keeps the lag state and writes no dSdt
  // register of the caller's (the model's own buffer takes it)
  double *scratch = nullptr;
  const size_t n = (size_t)(m->N + 1) * (m->N + 1) * m->nVar * m->nEl;
  SELFAMD_HIP(hipMalloc(&scratch, sizeof(double) * n));
  hipError_t e = m->launch_stage(1, 0, 0.0, m->cur, scratch, -1, flux, fluxDivergence);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  (void)hipFree(scratch);
  SELFAMD_HIP(e);
  return kOk;
}

namespace {
// a few steps of another size than the captured graphs' (e.g. ForwardStep's shorter last step
// up to tn) are launched directly, so the graphs of the running step size survive
bool direct_steps(const selfamd_model2d *m, int nSteps, double dt) {
  return m->graphDt > 0.0 && m->graphDt != dt && m->graphEpoch == g_tileEpoch && nSteps <= kDirectSteps;
}
// captures the graphs that nSteps steps of size dt from the current trace buffer replay
hipError_t prepare_graphs(selfamd_model2d *m, int nSteps, double dt) {
  if (nSteps == 0 || direct_steps(m, nSteps, dt)) return hipSuccess;
  if (m->graphDt != dt || m->graphEpoch != g_tileEpoch) {
    m->drop_graphs();
    for (int start = 0; start < 2; start++) {
      hipError_t e = m->capture(start, 1, dt, &m->graph[start]);
      if (e != hipSuccess) return e;
    }
    m->graphDt = dt;
    m->graphEpoch = g_tileEpoch;
  }
  // the 2^k-step graphs of nSteps' decomposition (they keep the current buffer), for both
  // starting buffers, so that a later call of the same size finds them whichever is current
  const int top = m->top_level();
  int rem = nSteps;
  for (int k = top; k >= 1; k--) {
    const int s = 1 << k;
    if (rem < s) continue;
    for (int start = 0; start < 2; start++)
      if (!m->graphP[k][start]) {
        const hipError_t e = m->capture(start, s, dt, &m->graphP[k][start]);
        if (e != hipSuccess) return e;
      }
    rem = k == top ? rem % s : rem - s;
  }
  return hipSuccess;
}
}  // namespace

int selfamd_model2d_forward_step(selfamd_model2d *m, int nSteps, double dt) {
  SELFAMD_REQUIRE(m && nSteps >= 0 && dt > 0.0, "model2d_forward_step: bad arguments (nSteps=%d dt=%g)", nSteps, dt);
  if (nSteps == 0) return kOk;
  if (direct_steps(m, nSteps, dt)) {
    SELFAMD_HIP(m->launch_steps(nSteps, dt));
    return kOk;
  }
  SELFAMD_HIP(prepare_graphs(m, nSteps, dt));
  const int top = m->top_level();
  int n = nSteps;
  for (int k = top; k >= 1; k--) {
    const int s = 1 << k;
    while (n >= s) {
      SELFAMD_HIP(hipGraphLaunch(m->graphP[k][m->cur], m->stream));  // 3s stage flips: parity unchanged
      for (int q = 0; q < s; q++) m->t += dt;
      n -= s;
      if (k != top) break;
    }
  }
  if (n == 1) {
    SELFAMD_HIP(hipGraphLaunch(m->graph[m->cur], m->stream));
    m->cur ^= 1;  // three stages flip the trace buffer three times
    m->t += dt;
  }
  return kOk;
}

int selfamd_model2d_synchronize(selfamd_model2d *m) {
  SELFAMD_REQUIRE(m, "model2d_synchronize: null model");
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_stream(selfamd_model2d *m, void **stream) {
  SELFAMD_REQUIRE(m && stream, "model2d_stream: null argument");
  *stream = (void *)m->stream;
  return kOk;
}

int selfamd_model2d_time_steps(selfamd_model2d *m, int nSteps, double dt, double *elapsedMs) {
  SELFAMD_REQUIRE(m && elapsedMs && nSteps >= 0 && dt > 0.0, "model2d_time_steps: bad arguments");
  // graph capture is host work: done before the first event, so the GPU's idle time during it
  // is not timed
  SELFAMD_HIP(prepare_graphs(m, nSteps, dt));
  hipEvent_t e0, e1;
  SELFAMD_HIP(hipEventCreate(&e0));
  SELFAMD_HIP(hipEventCreate(&e1));
  SELFAMD_HIP(hipEventRecord(e0, m->stream));
  int rc = selfamd_model2d_forward_step(m, nSteps, dt);
  if (rc) return rc;
  SELFAMD_HIP(hipEventRecord(e1, m->stream));
  SELFAMD_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SELFAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  *elapsedMs = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return kOk;
}

int selfamd_model2d_set_closure(selfamd_model2d *m, selfamd_mlp *mlp, const float *inMean, const float *inScale,
                                const float *outMean, const float *outScale) {
  SELFAMD_REQUIRE(m, "model2d_set_closure: null model");
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  m->drop_graphs();
  free_closure(m->closure);
  m->closure = nullptr;
  if (!mlp) return kOk;
  SELFAMD_REQUIRE(inMean && inScale && outMean && outScale, "model2d_set_closure: null normalisation array");
  SELFAMD_REQUIRE(m->affine, "model2d_set_closure: the closure reads the affine element records; mesh is not affine");
  SELFAMD_REQUIRE(m->nBands == 0, "model2d_set_closure: banded stepping is on (selfamd_model2d_set_bands(model, 0))");
  const int P = m->N + 1, NN = P * P;
  int nIn = 0, nOut = 0, maxB = 0;
  mlp_shape(mlp, &nIn, &nOut, &maxB);
  SELFAMD_REQUIRE(nIn == m->nVar * NN + 3 && nOut == m->nVar * NN,
                  "model2d_set_closure: network shape %d -> %d does not match %d nodal values + 3 metric scalars", nIn,
                  nOut, m->nVar * NN);
  SELFAMD_REQUIRE(maxB >= m->nEl, "model2d_set_closure: network batch %d < %d elements", maxB, m->nEl);
  // Sum of w_i w_j J over the mesh, for the mean-free mass correction
  std::vector<double> rec((size_t)kGeomStride * m->nEl), mats(2 * NN + 3 * P);
  SELFAMD_HIP(hipMemcpy(rec.data(), m->geom, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
  SELFAMD_HIP(hipMemcpy(mats.data(), m->mats, sizeof(double) * mats.size(), hipMemcpyDeviceToHost));
  double sw = 0.0, area = 0.0;
  for (int i = 0; i < P; i++) sw += mats[2 * NN + 2 * P + i];
  for (int e = 0; e < m->nEl; e++) area += rec[(size_t)kGeomStride * e + 4] * sw * sw;
  ClosureBuffers *c = new ClosureBuffers();
  c->mlp = mlp; c->nIn = nIn; c->nOut = nOut; c->invArea = 1.0 / area;
  const size_t nEl = m->nEl;
  bool ok = hipMalloc(&c->X, 4 * nEl * nIn) == hipSuccess && hipMalloc(&c->Y, 4 * nEl * nOut) == hipSuccess &&
            hipMalloc(&c->inMean, 4 * nIn) == hipSuccess && hipMalloc(&c->inScale, 4 * nIn) == hipSuccess &&
            hipMalloc(&c->outMean, 4 * nOut) == hipSuccess && hipMalloc(&c->outScale, 4 * nOut) == hipSuccess &&
            hipMalloc(&c->corr, 8 * nEl * nOut) == hipSuccess && hipMalloc(&c->elemMass, 8 * nEl) == hipSuccess &&
            hipMalloc(&c->shift, 8) == hipSuccess && hipMalloc(&c->dY, 4 * nEl * nOut) == hipSuccess &&
            hipMalloc(&c->dX, 4 * nEl * nIn) == hipSuccess &&
            hipMalloc(&c->vjpPartial, 8 * closure_vjp_partials(m->N, m->nEl)) == hipSuccess;
  ok = ok && hipMemcpy(c->inMean, inMean, 4 * nIn, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->inScale, inScale, 4 * nIn, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->outMean, outMean, 4 * nOut, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->outScale, outScale, 4 * nOut, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    free_closure(c);
    set_error("model2d_set_closure: device allocation or copy failed");
    return kHipError;
  }
  m->closure = c;
  return kOk;
}

int selfamd_model2d_stage(selfamd_model2d *m, int stage, double dt, double *Ssave, double *unSave) {
  SELFAMD_REQUIRE(m && stage >= 0 && stage < 3 && dt > 0.0, "model2d_stage: bad arguments (stage=%d dt=%g)", stage, dt);
  SELFAMD_REQUIRE(m->nBands == 0, "model2d_stage: banded stepping is on");
  const size_t nS = (size_t)(m->N + 1) * (m->N + 1) * m->nVar * m->nEl;
  if (Ssave) SELFAMD_HIP(hipMemcpyAsync(Ssave, m->S, sizeof(double) * nS, hipMemcpyDeviceToDevice, m->stream));
  if (unSave) {
    SELFAMD_REQUIRE(m->un[0], "model2d_stage: this model has no lagged velocities to save");
    const size_t nU = (size_t)(m->nodal ? 2 : 1) * (m->N + 1) * 4 * m->nEl;
    SELFAMD_HIP(hipMemcpyAsync(unSave, m->un[m->cur], sizeof(double) * nU, hipMemcpyDeviceToDevice, m->stream));
  }
  SELFAMD_HIP(m->launch_stage(0, stage, dt, m->cur, nullptr));
  m->cur ^= 1;
  if (stage == 2) m->t += dt;
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_tendency_vjp(selfamd_model2d *m, const double *S, const double *un, const double *lambda,
                                 double *mu, const double *Sprev, double *muPrev) {
  SELFAMD_REQUIRE(m && S && un && lambda && mu, "model2d_tendency_vjp: null argument");
  SELFAMD_REQUIRE(!Sprev == !muPrev, "model2d_tendency_vjp: Sprev and muPrev go together");
  SELFAMD_REQUIRE(m->kind == SELFAMD_MODEL_SW && !m->nodal && !m->anyBoundary,
                  "model2d_tendency_vjp: only nonlinear shallow water on affine meshes without boundary sides");
  const size_t nTr = (size_t)(m->N + 1) * 3 * 4 * m->nEl;
  // scratch: traces of S and of Sprev, neighbour-trace cotangents, lagged-velocity cotangents
  if (!m->adjTrace) SELFAMD_HIP(hipMalloc(&m->adjTrace, sizeof(double) * 2 * nTr));
  if (!m->adjChi) SELFAMD_HIP(hipMalloc(&m->adjChi, sizeof(double) * nTr * 5 / 3));
  SELFAMD_HIP(Launcher<PhysSW>::traces(m->N, S, m->adjTrace, m->mats, m->p, m->nEl, m->stream));
  if (Sprev) SELFAMD_HIP(Launcher<PhysSW>::traces(m->N, Sprev, m->adjTrace + nTr, m->mats, m->p, m->nEl, m->stream));
  SELFAMD_HIP(sw_tendency_vjp(m->N, m->nEl, S, m->adjTrace, un, m->aux, m->geom, m->nbr, m->mats, m->p.g, m->p.fCori,
                              m->p.auxConst, lambda, mu, m->adjChi, m->adjTrace + nTr, muPrev, m->adjChi + nTr,
                              m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_closure_vjp(selfamd_model2d *m, const double *S, const double *lambda, double *mu) {
  SELFAMD_REQUIRE(m && S && lambda && mu, "model2d_closure_vjp: null argument");
  SELFAMD_REQUIRE(m->closure, "model2d_closure_vjp: no closure attached");
  const int P = m->N + 1;
  SELFAMD_HIP(closure_vjp(*m->closure, S, m->geom, m->mats + 2 * P * P + 2 * P, lambda, mu, m->N, m->nVar, 2, m->nEl,
                          m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_model2d_set_bands(selfamd_model2d *m, int nBands) {
  SELFAMD_REQUIRE(m && nBands >= 0, "model2d_set_bands: bad arguments");
  SELFAMD_REQUIRE(nBands == 0 || !m->closure, "model2d_set_bands: banded stepping does not take a closure");
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  m->drop_graphs();
  if (m->bandElems) (void)hipFree(m->bandElems);
  m->bandElems = nullptr;
  m->bandStart.clear();
  m->nBands = 0;
  if (nBands == 0) return kOk;
  const int nEl = m->nEl;
  SELFAMD_REQUIRE(nBands >= 2 && nBands <= nEl, "model2d_set_bands: need 2 <= nBands <= nEl (got %d)", nBands);
  // BFS levels of the element graph (interior neighbours), component after component
  std::vector<int> level(nEl, -1), queue;
  queue.reserve(nEl);
  int base = 0;
  for (int seed = 0; seed < nEl; seed++) {
    if (level[seed] >= 0) continue;
    size_t head = queue.size();
    level[seed] = base;
    queue.push_back(seed);
    int top = base;
    while (head < queue.size()) {
      const int e = queue[head++];
      for (int sd = 0; sd < 4; sd++) {
        const int c = m->hnbr[4 * (size_t)e + sd];
        if (c < 0) continue;
        const int e2 = c >> 3;
        if (level[e2] < 0) {
          level[e2] = level[e] + 1;
          top = std::max(top, level[e2]);
          queue.push_back(e2);
        }
      }
    }
    base = top + 1;
  }
  const int nLevels = base;
  std::vector<int> perLevel(nLevels, 0);
  for (int e = 0; e < nEl; e++) perLevel[level[e]]++;
  // consecutive whole levels per band, about nEl / nBands elements each
  std::vector<int> bandOf(nLevels, 0);
  int b = 0;
  long acc = 0;
  for (int l = 0; l < nLevels; l++) {
    bandOf[l] = b;
    acc += perLevel[l];
    if (b < nBands - 1 && acc >= (long)nEl * (b + 1) / nBands) b++;
  }
  const int used = bandOf[nLevels - 1] + 1;
  SELFAMD_REQUIRE(used >= 2, "model2d_set_bands: the element graph has too few BFS levels (%d) for %d bands",
                  nLevels, nBands);
  nBands = used;  // a band boundary can only fall between levels
  for (int e = 0; e < nEl; e++)
    for (int sd = 0; sd < 4; sd++) {
      const int c = m->hnbr[4 * (size_t)e + sd];
      if (c < 0) continue;
      const int d = bandOf[level[e]] - bandOf[level[c >> 3]];
      SELFAMD_REQUIRE(d >= -1 && d <= 1, "model2d_set_bands: neighbours %d and %d are more than one band apart", e,
                      c >> 3);
    }
  std::vector<int32_t> order;
  order.reserve(nEl);
  std::vector<int> start(nBands + 1, 0);
  for (int e = 0; e < nEl; e++) start[bandOf[level[e]] + 1]++;
  for (int k = 0; k < nBands; k++) start[k + 1] += start[k];
  order.assign(nEl, 0);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (int e = 0; e < nEl; e++) order[fill[bandOf[level[e]]]++] = e;  // by band, then element index
  SELFAMD_HIP(hipMalloc(&m->bandElems, sizeof(int32_t) * nEl));
  SELFAMD_HIP(hipMemcpy(m->bandElems, order.data(), sizeof(int32_t) * nEl, hipMemcpyHostToDevice));
  m->bandStart = start;
  m->nBands = nBands;
  return kOk;
}

int selfamd_model2d_closure_features(selfamd_model2d *m, float *X) {
  SELFAMD_REQUIRE(m && X, "model2d_closure_features: null argument");
  SELFAMD_REQUIRE(m->closure, "model2d_closure_features: no closure attached");
  SELFAMD_HIP(closure_features(m->S, m->geom, m->closure->inMean, m->closure->inScale, X, m->N, m->nVar, m->nEl,
                               m->stream));
  SELFAMD_HIP(hipStreamSynchronize(m->stream));
  return kOk;
}

int selfamd_tile_count(int kind, int N, int *count) {
  SELFAMD_REQUIRE(count, "tile_count: null argument");
  *count = tunable(kind, N) ? kTileVariants : 0;
  return kOk;
}

int selfamd_tile_info(int kind, int N, int variant, int *threads, int *alternate, int *xcdOrder, int *minWaves) {
  SELFAMD_REQUIRE(threads && alternate && xcdOrder && minWaves, "tile_info: null argument");
  SELFAMD_REQUIRE(tunable(kind, N) && variant >= 0 && variant < kTileVariants,
                  "tile_info: no variant %d for kind %d, N=%d", variant, kind, N);
  if (variant < 12) {
    *threads = 64 << (variant >> 2);
    *alternate = (variant >> 1) & 1;
    *xcdOrder = variant & 1;
    *minWaves = 1;
  } else {
    *threads = 64 << ((variant - 12) >> 1);
    *alternate = 1;
    *xcdOrder = 0;
    *minWaves = 5 + ((variant - 12) & 1);
  }
  return kOk;
}

int selfamd_tile_set(int kind, int N, int variant) {
  SELFAMD_REQUIRE(tunable(kind, N), "tile_set: kind %d, N=%d has no tile variants", kind, N);
  SELFAMD_REQUIRE(variant >= -1 && variant < kTileVariants, "tile_set: variant %d out of range", variant);
  g_tile[kind][N] = variant;
  g_tileEpoch++;
  return kOk;
}

int selfamd_tile_get(int kind, int N, int *variant) {
  SELFAMD_REQUIRE(variant, "tile_get: null argument");
  *variant = tile_of(kind, N);
  return kOk;
}

int selfamd_model2d_get_time(selfamd_model2d *m, double *t) {
  SELFAMD_REQUIRE(m && t, "model2d_get_time: null argument");
  *t = m->t;
  return kOk;
}
int selfamd_model2d_set_time(selfamd_model2d *m, double t) {
  SELFAMD_REQUIRE(m, "model2d_set_time: null model");
  m->t = t;
  return kOk;
}

}  // extern "C"

// ------------------------------------------------------------------------------
// Batched runs: many independent models advanced together
// ------------------------------------------------------------------------------
// Models of one polynomial degree are concatenated into one virtual mesh (state, traces,
// static fields, neighbour codes offset by each run's first element); each element record
// carries its run's dt.  One fused launch per degree group and RK stage advances every run
// of the group, and one hipGraph holds all groups' three stages of a step.  Per element the
// arithmetic is that of the single-run engine, so a batched run is bit-identical to the same
// run advanced alone.
namespace {
struct BatchGroup {
  int N = 0, nEl = 0;
  double *S = nullptr, *G = nullptr, *trace[2] = {nullptr, nullptr}, *un[2] = {nullptr, nullptr}, *geom = nullptr;
  double *mats = nullptr, *aux = nullptr, *stat2 = nullptr, *presTrace = nullptr;
  int32_t *nbr = nullptr;
  std::vector<int> runs;
  std::vector<size_t> first;  // first element of each run in the group
};
}  // namespace

struct selfamd_batch {
  int kind = 0;
  std::vector<Model *> models;
  std::vector<double> dt;
  std::vector<BatchGroup> groups;
  Params p{};
  bool anyBoundary = false;  // some run's mesh has a physical boundary side
  hipStream_t stream = nullptr;
  int cur = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // one step, per starting trace buffer
  // graphs of 2^k steps (k = 1 .. kLevels - 1), per starting buffer, captured when a call first
  // needs them (2^k steps keep the starting buffer); a call replays the largest as often as it
  // fits and each smaller one at most once, as the single-run engine does
  static constexpr int kLevels = 5;  // up to 16 steps per graph
  hipGraphExec_t graphP[kLevels][2] = {};

  // SELFAMD_BATCH_PER_GROUP: one launch per degree group and stage; SELFAMD_BATCH_PERSISTENT:
  // one persistent launch per stage for all groups (SW, degrees 3, 5, 7)
  int mode = SELFAMD_BATCH_PER_GROUP;
  int persistentBlocks = 0;

  FusedArgs group_args(const BatchGroup &g, int stage, int tin) const {
    FusedArgs a{};
    a.S = g.S; a.G = g.G; a.dSdt = nullptr;
    a.traceIn = g.trace[tin]; a.traceOut = g.trace[tin ^ 1];
    a.unIn = g.un[tin]; a.unOut = g.un[tin ^ 1];
    a.aux = g.aux; a.stat2 = g.stat2; a.presTrace = g.presTrace;
    a.geom = g.geom; a.nbr = g.nbr; a.mats = g.mats; a.nEl = g.nEl;
    a.rkA = kRk3A[stage]; a.rkG = kRk3G[stage]; a.rkGdt = 0.0; a.perElemDt = 1;
    a.writeG = stage != 2;
    a.reverse = tile_alternates(kind, g.N, g.nEl) ? tin : 0;
    a.anyBoundary = anyBoundary;
    a.p = p;
    return a;
  }
  hipError_t launch_group(BatchGroup &g, int stage, int tin) {
    const FusedArgs a = group_args(g, stage, tin);
    switch (kind) {
      case SELFAMD_MODEL_SW: return Launcher<PhysSW>::stage(g.N, 0, false, a, stream);
      case SELFAMD_MODEL_LSW: return Launcher<PhysLSW>::stage(g.N, 0, false, a, stream);
      default: return Launcher<PhysEuler>::stage(g.N, 0, false, a, stream);
    }
  }
  // whether the persistent kernel covers this batch: nonlinear SW, every degree 3, 5 or 7, and
  // the same face-sum split as the per-group tiles (its 256-wide tiles split at N = 7 and not
  // at N = 3 or 5; among the per-group tiles only the 64-wide ones at N = 5 split), so that
  // the results stay bit-identical to the separate solves
  bool persistent_ok() const {
    if (kind != SELFAMD_MODEL_SW) return false;
    for (const auto &g : groups) {
      if (g.N != 3 && g.N != 5 && g.N != 7) return false;
      if (g.N == 5) {
        const int v = tile_for(kind, 5, g.nEl);
        const int threads = v < 0 ? tile_threads<5>() : v < 12 ? 64 << (v >> 2) : 64 << ((v - 12) >> 1);
        if (threads == 64) return false;
      }
    }
    return true;
  }
  hipError_t launch_persistent(int stage, int tin) {
    PersistentArgs pa{};
    for (int k = 0; k < 3; k++)
      for (const auto &g : groups)
        if (g.N == kPersistentDegrees[k]) {
          pa.g[k] = group_args(g, stage, tin);
          pa.g[k].reverse = tin;  // per-element arithmetic does not depend on the order
        }
    hipLaunchKernelGGL(k_batch_persistent<PhysSW>, dim3(persistentBlocks), 256, 0, stream, pa);
    return hipGetLastError();
  }
  // graphs of one RK3 step for both starting trace buffers
  // captures `steps` steps starting from trace buffer `start` (each step flips the buffer)
  hipError_t capture_steps(int start, int steps, hipGraphExec_t *out) {
    hipGraph_t gr;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = hipSuccess;
    for (int n = 0; n < steps && le == hipSuccess; n++) {
      const int tin = start ^ (n & 1);
      if (mode == SELFAMD_BATCH_PERSISTENT) {
        for (int st = 0; st < 3 && le == hipSuccess; st++) le = launch_persistent(st, tin ^ (st & 1));
      } else {
        // The groups are independent, so each group's three stages run back to back (a stage
        // then starts on the data its predecessor left in the Infinity Cache), and the group
        // order reverses from step to step (with the buffer parity), so that a group's last
        // stage of one step is followed by its first stage of the next.
        const int ng = (int)groups.size();
        for (int k = 0; k < ng; k++) {
          BatchGroup &g = groups[tin == 0 ? k : ng - 1 - k];
          for (int st = 0; st < 3 && le == hipSuccess; st++) le = launch_group(g, st, tin ^ (st & 1));
        }
      }
    }
    hipError_t ce = hipStreamEndCapture(stream, &gr);
    if (le != hipSuccess) return le;
    if (ce != hipSuccess) return ce;
    hipError_t ie = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ie == hipSuccess) ie = hipGraphUpload(*out, stream);  // not on the first replay
    return ie;
  }
  void drop_graphs() {
    for (auto &gx : graph)
      if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
    for (auto &lv : graphP)
      for (auto &gx : lv)
        if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
  }
  // the one-step graphs of the current mode (the longer ones follow when first needed)
  hipError_t capture() {
    drop_graphs();
    for (int start = 0; start < 2; start++) {
      const hipError_t e = capture_steps(start, 1, &graph[start]);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  // the 2^k-step graphs of nSteps' decomposition, for both starting buffers
  hipError_t prepare(int nSteps) {
    const int top = kLevels - 1;
    int rem = nSteps;
    for (int k = top; k >= 1; k--) {
      const int s = 1 << k;
      if (rem < s) continue;
      for (int start = 0; start < 2; start++)
        if (!graphP[k][start]) {
          const hipError_t e = capture_steps(start, s, &graphP[k][start]);
          if (e != hipSuccess) return e;
        }
      rem = k == top ? rem % s : rem - s;
    }
    return hipSuccess;
  }
};

extern "C" {

int selfamd_batch_destroy(selfamd_batch *b) {
  if (!b) return kOk;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graphs();
  for (auto &g : b->groups) {
    double *bufs[] = {g.S, g.G, g.trace[0], g.trace[1], g.un[0], g.un[1], g.geom, g.mats, g.aux, g.stat2, g.presTrace};
    for (double *p : bufs)
      if (p) (void)hipFree(p);
    if (g.nbr) (void)hipFree(g.nbr);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return kOk;
}

int selfamd_batch_create(selfamd_batch **out, int nModels, selfamd_model2d *const *models, const double *dt) {
  SELFAMD_REQUIRE(out && models && dt && nModels >= 1, "batch_create: bad arguments");
  *out = nullptr;
  const Model *m0 = models[0];
  SELFAMD_REQUIRE(m0, "batch_create: null model");
  for (int r = 0; r < nModels; r++) {
    const Model *m = models[r];
    SELFAMD_REQUIRE(m && m->kind == m0->kind, "batch_create: run %d: all runs must be the same model kind", r);
    SELFAMD_REQUIRE(!m->nodal, "batch_create: run %d uses nodal metrics; batched runs take affine meshes", r);
    SELFAMD_REQUIRE(!m->closure, "batch_create: run %d has a closure attached; batched launches do not apply it", r);
    SELFAMD_REQUIRE(dt[r] > 0.0, "batch_create: run %d: dt must be positive", r);
    SELFAMD_REQUIRE(memcmp(&m->p, &m0->p, sizeof(Params)) == 0,
                    "batch_create: run %d: physical parameters differ from run 0", r);
    SELFAMD_REQUIRE((m->aux != nullptr) == (m0->aux != nullptr) && (m->stat2 != nullptr) == (m0->stat2 != nullptr) &&
                        (m->presTrace != nullptr) == (m0->presTrace != nullptr),
                    "batch_create: run %d: static fields must be set on all runs or on none", r);
  }
  selfamd_batch *b = new selfamd_batch();
  b->kind = m0->kind;
  b->p = m0->p;
  for (int r = 0; r < nModels; r++) b->anyBoundary = b->anyBoundary || models[r]->anyBoundary;
  b->models.assign(models, models + nModels);
  b->dt.assign(dt, dt + nModels);
  auto fail = [&](const char *what, int code = kHipError) {
    set_error("batch_create: %s", what);
    selfamd_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream creation failed");
  for (int N = 1; N <= 7; N++) {
    BatchGroup g;
    g.N = N;
    size_t total = 0;
    for (int r = 0; r < nModels; r++)
      if (models[r]->N == N) { g.runs.push_back(r); g.first.push_back(total); total += models[r]->nEl; }
    if (g.runs.empty()) continue;
    if (total >= (size_t)1 << 27) return fail("a degree group has too many elements (at most 2^27)", kBadArgument);
    g.nEl = (int)total;
    const size_t P = N + 1, PP = P * P, nV = m0->nVar, nT = m0->nT;
    const size_t nS = total * PP * nV, nTr = total * P * nT * 4;
    if (hipMalloc(&g.S, 8 * nS) || hipMalloc(&g.G, 8 * nS) || hipMalloc(&g.trace[0], 8 * nTr) ||
        hipMalloc(&g.trace[1], 8 * nTr) || hipMalloc(&g.geom, 8 * kGeomStride * total) ||
        hipMalloc(&g.mats, 8 * (2 * PP + 4 * P)) || hipMalloc(&g.nbr, 4 * 4 * total) ||
        (m0->un[0] && (hipMalloc(&g.un[0], 8 * total * P * 4) || hipMalloc(&g.un[1], 8 * total * P * 4))) || (m0->aux && hipMalloc(&g.aux, 8 * total * PP)) ||
        (m0->stat2 && hipMalloc(&g.stat2, 8 * total * PP)) || (m0->presTrace && hipMalloc(&g.presTrace, 8 * nTr))) {
      b->groups.push_back(g);
      return fail("device allocation failed");
    }
    b->groups.push_back(g);
    BatchGroup &G = b->groups.back();
    if (hipMemcpy(G.mats, models[G.runs[0]]->mats, 8 * (2 * PP + 4 * P), hipMemcpyDeviceToDevice))
      return fail("copy failed");
    std::vector<double> rec;
    std::vector<int32_t> nb;
    for (size_t k = 0; k < G.runs.size(); k++) {
      const Model *m = models[G.runs[k]];
      const size_t off = G.first[k], ne = m->nEl;
      hipError_t e = hipSuccess;
      e = e ? e : hipMemcpy(G.S + off * PP * nV, m->S, 8 * ne * PP * nV, hipMemcpyDeviceToDevice);
      e = e ? e : hipMemcpy(G.G + off * PP * nV, m->G, 8 * ne * PP * nV, hipMemcpyDeviceToDevice);
      e = e ? e : hipMemcpy(G.trace[0] + off * P * nT * 4, m->trace[m->cur], 8 * ne * P * nT * 4, hipMemcpyDeviceToDevice);
      if (G.un[0]) e = e ? e : hipMemcpy(G.un[0] + off * P * 4, m->un[m->cur], 8 * ne * P * 4, hipMemcpyDeviceToDevice);
      if (G.aux) e = e ? e : hipMemcpy(G.aux + off * PP, m->aux, 8 * ne * PP, hipMemcpyDeviceToDevice);
      if (G.stat2) e = e ? e : hipMemcpy(G.stat2 + off * PP, m->stat2, 8 * ne * PP, hipMemcpyDeviceToDevice);
      if (G.presTrace)
        e = e ? e : hipMemcpy(G.presTrace + off * P * nT * 4, m->presTrace, 8 * ne * P * nT * 4, hipMemcpyDeviceToDevice);
      rec.resize(kGeomStride * ne);
      nb.resize(4 * ne);
      e = e ? e : hipMemcpy(rec.data(), m->geom, 8 * kGeomStride * ne, hipMemcpyDeviceToHost);
      e = e ? e : hipMemcpy(nb.data(), m->nbr, 4 * 4 * ne, hipMemcpyDeviceToHost);
      if (e) return fail("copy failed");
      for (size_t q = 0; q < ne; q++) rec[kGeomStride * q + 5] = b->dt[G.runs[k]];
      for (auto &c : nb)
        if (c >= 0) c += (int32_t)(off << 3);  // neighbour element index shifts by the run's offset
      e = hipMemcpy(G.geom + kGeomStride * off, rec.data(), 8 * kGeomStride * ne, hipMemcpyHostToDevice);
      e = e ? e : hipMemcpy(G.nbr + 4 * off, nb.data(), 4 * 4 * ne, hipMemcpyHostToDevice);
      if (e) return fail("copy failed");
    }
  }
  if (b->capture() != hipSuccess) return fail("stage graph capture failed");
  *out = b;
  return kOk;
}

int selfamd_batch_set_mode(selfamd_batch *b, int mode) {
  SELFAMD_REQUIRE(b && (mode == SELFAMD_BATCH_PER_GROUP || mode == SELFAMD_BATCH_PERSISTENT),
                  "batch_set_mode: bad arguments (mode=%d)", mode);
  SELFAMD_REQUIRE(mode == SELFAMD_BATCH_PER_GROUP || b->persistent_ok(),
                  "batch_set_mode: the persistent launch takes nonlinear SW runs of degrees 3, 5 and 7 "
                  "whose per-group tile at N = 5 is not 64 wide (that tile splits the face sums)");
  SELFAMD_HIP(hipStreamSynchronize(b->stream));
  if (mode == SELFAMD_BATCH_PERSISTENT && b->persistentBlocks == 0) {
    int dev = 0, cus = 0, perCU = 0;
    SELFAMD_HIP(hipGetDevice(&dev));
    SELFAMD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SELFAMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_batch_persistent<PhysSW>, 256, 0));
    SELFAMD_REQUIRE(cus > 0 && perCU > 0, "batch_set_mode: no resident block for the persistent kernel");
    b->persistentBlocks = cus * perCU;
  }
  const int previous = b->mode;
  b->mode = mode;
  const hipError_t e = b->capture();
  if (e != hipSuccess) {
    // keep the batch usable: back to the previous mode's graphs
    b->mode = previous;
    const hipError_t e2 = b->capture();
    set_error("batch_set_mode: graph capture failed: %s%s", hipGetErrorString(e),
              e2 == hipSuccess ? "" : " (the previous mode's graphs could not be restored either)");
    return kHipError;
  }
  return kOk;
}

int selfamd_batch_forward_step(selfamd_batch *b, int nSteps) {
  SELFAMD_REQUIRE(b && nSteps >= 0, "batch_forward_step: bad arguments");
  SELFAMD_HIP(b->prepare(nSteps));
  const int top = selfamd_batch::kLevels - 1;
  int n = nSteps;
  for (int k = top; k >= 1; k--) {
    const int s = 1 << k;
    while (n >= s) {
      SELFAMD_HIP(hipGraphLaunch(b->graphP[k][b->cur], b->stream));  // an even step count: same buffer
      n -= s;
      if (k != top) break;
    }
  }
  if (n == 1) {
    SELFAMD_HIP(hipGraphLaunch(b->graph[b->cur], b->stream));
    b->cur ^= 1;
  }
  for (size_t r = 0; r < b->models.size(); r++) b->models[r]->t += nSteps * b->dt[r];
  return kOk;
}

int selfamd_batch_time_steps(selfamd_batch *b, int nSteps, double *elapsedMs) {
  SELFAMD_REQUIRE(b && elapsedMs && nSteps >= 0, "batch_time_steps: bad arguments");
  // graph capture is host work: done before the first event
  SELFAMD_HIP(b->prepare(nSteps));
  hipEvent_t e0, e1;
  SELFAMD_HIP(hipEventCreate(&e0));
  SELFAMD_HIP(hipEventCreate(&e1));
  SELFAMD_HIP(hipEventRecord(e0, b->stream));
  int rc = selfamd_batch_forward_step(b, nSteps);
  if (rc) return rc;
  SELFAMD_HIP(hipEventRecord(e1, b->stream));
  SELFAMD_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SELFAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  *elapsedMs = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return kOk;
}

// Copies every run's state (solution, RK register, traces, lagged velocities) back into its
// model handle, which then continues on its own exactly where the batch left it.
int selfamd_batch_scatter(selfamd_batch *b) {
  SELFAMD_REQUIRE(b, "batch_scatter: null batch");
  SELFAMD_HIP(hipStreamSynchronize(b->stream));
  for (auto &G : b->groups) {
    const size_t P = G.N + 1, PP = P * P;
    for (size_t k = 0; k < G.runs.size(); k++) {
      Model *m = b->models[G.runs[k]];
      const size_t off = G.first[k], ne = m->nEl, nV = m->nVar, nT = m->nT;
      SELFAMD_HIP(hipMemcpy(m->S, G.S + off * PP * nV, 8 * ne * PP * nV, hipMemcpyDeviceToDevice));
      SELFAMD_HIP(hipMemcpy(m->G, G.G + off * PP * nV, 8 * ne * PP * nV, hipMemcpyDeviceToDevice));
      SELFAMD_HIP(hipMemcpy(m->trace[m->cur], G.trace[b->cur] + off * P * nT * 4, 8 * ne * P * nT * 4,
                            hipMemcpyDeviceToDevice));
      if (m->un[0])
        SELFAMD_HIP(hipMemcpy(m->un[m->cur], G.un[b->cur] + off * P * 4, 8 * ne * P * 4, hipMemcpyDeviceToDevice));
    }
  }
  return kOk;
}

int selfamd_batch_info(selfamd_batch *b, int *nGroups, double *totalDof) {
  SELFAMD_REQUIRE(b && nGroups && totalDof, "batch_info: null argument");
  *nGroups = (int)b->groups.size();
  double dof = 0.0;
  for (auto &g : b->groups) dof += (double)g.nEl * (g.N + 1) * (g.N + 1) * b->models[0]->nVar;
  *totalDof = dof;
  return kOk;
}

}  // extern "C"
