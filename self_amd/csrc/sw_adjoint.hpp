// Internal interface of the shallow-water tendency adjoint (sw_adjoint.hip), used by the model
// engine (dg2d_fused.hip) for closure training through the solver.
#pragma once
#include "selfamd_common.hpp"

namespace selfamd {

// mu = (d dSdt / dS)^T lam for the SW tendency at state S (affine elements, every side an
// interior side), the lagged normal velocities un held fixed.  traces: the face traces of S
// (0:N,3,4,nEl); chiR: scratch of the same size.  With muPrev, also the lag path:
// muPrev = (d dSdt / d Sp)^T lam, where un = (Hu/H, Hv/H).n of the face traces tracesPrev of
// the previous stage's input state Sp; chiUn: scratch of 2 (N+1) 4 nEl doubles.
// Asynchronous on st.
hipError_t sw_tendency_vjp(int N, int nEl, const double *S, const double *traces, const double *un,
                           const double *aux, const double *geom, const int32_t *nbr, const double *mats, double g,
                           double fCori, double auxConst, const double *lam, double *mu, double *chiR,
                           const double *tracesPrev, double *muPrev, double *chiUn, hipStream_t st);

}  // namespace selfamd
