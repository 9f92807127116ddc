// Internal interface between the closure kernels (closure.hip) and the model engine
// (dg2d_fused.hip): a model with a closure attached evaluates, before each stage's fused
// launch, features -> MLP -> correction, and the fused kernel adds the correction to dS/dt.
#pragma once
#include "selfamd_common.hpp"

struct selfamd_mlp;

namespace selfamd {

struct ClosureBuffers {
  selfamd_mlp *mlp = nullptr;
  int nIn = 0, nOut = 0;
  float *X = nullptr, *Y = nullptr;          // MLP input / output, one row per element
  float *inMean = nullptr, *inScale = nullptr, *outMean = nullptr, *outScale = nullptr;
  double *corr = nullptr;                    // correction in the interior layout (0:N,0:N,nVar,nEl)
  double *elemMass = nullptr;                // per-element integral of the mass-variable correction
  double *shift = nullptr;                   // one double: global mean of the mass correction
  double invArea = 0.0;                      // 1 / (sum over elements of the element area in xi-space times J)
  // adjoint (closure_vjp) work space
  float *dY = nullptr, *dX = nullptr;        // output / input cotangents, one row per element
  double *vjpPartial = nullptr;              // block partial sums of the mass-variable cotangent, then the total
};

// Normalised MLP inputs of state S: per element the nVar*(N+1)^2 nodal values, then the
// element metric scalars J, dsdx(1,1), dsdx(2,2) of the affine element record.
hipError_t closure_features(const double *S, const double *geom, const float *inMean, const float *inScale, float *X,
                            int N, int nVar, int nEl, hipStream_t st);

// features -> MLP -> correction (+ per-element mass integrals and the global mean shift)
hipError_t closure_evaluate(const ClosureBuffers &c, const double *S, const double *geom, const double *qW, int N,
                            int nVar, int massVar, int nEl, hipStream_t st);

// Adjoint of the correction that closure_evaluate adds to dS/dt at state S: for the cotangent
// lam of that correction (interior layout), the network's parameter gradients are added to its
// gradient buffer and the state cotangent through the network inputs is added to mu.
hipError_t closure_vjp(const ClosureBuffers &c, const double *S, const double *geom, const double *qW,
                       const double *lam, double *mu, int N, int nVar, int massVar, int nEl, hipStream_t st);
// number of block partial sums closure_vjp needs (+1 for the total)
size_t closure_vjp_partials(int N, int nEl);

hipError_t mlp_forward(selfamd_mlp *m, const float *X, float *Y, int B, hipStream_t st);
int mlp_shape(const selfamd_mlp *m, int *nIn, int *nOut, int *maxB);

}  // namespace selfamd
