/*
 * self_oracle.h -- CPU restatement of the SELF 2D DGSEM tendency + low-storage RK3 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  The product
 * library (self_amd) never links, loads or calls anything here.
 *
 * Every routine restates one reference routine (cited as file:line relative to the
 * reference checkout) with the same loop order and the same quirks.  Arrays use the
 * reference's Fortran column-major shapes, zero-based:
 *   interior(0:N,0:N,nVar,nEl)          SC(i,j,v,e)      = i + P*(j + P*(v + nVar*e))
 *   vector interior(2,0:N,0:N,nVar,nEl) VE(d,i,j,v,e)    = d + 2*SC(i,j,v,e)
 *   tensor interior(2,2,0:N,0:N,..)     TE(r,c,i,j,v,e)  = r + 2*c + 4*SC(i,j,v,e)
 *   boundary(0:N,nVar,4,nEl)            SCB(i,v,s,e)     = i + P*(v + nVar*(s + 4*e))
 *   vector boundary(2,0:N,nVar,4,nEl)   VEB(d,i,v,s,e)   = d + 2*SCB(i,v,s,e)
 *   tensor boundary(2,2,0:N,...)        TEB(r,c,i,v,s,e) = r + 2*c + 4*SCB(i,v,s,e)
 *   sideInfo(5,4,nEl) int32             SI(k,s,e)        = k + 5*(s + 4*e)
 * with P = N+1.  Element ids stored inside sideInfo are 1-based, as in the reference.
 */
#ifndef SELF_ORACLE_H
#define SELF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature codes (src/SELF_Constants.f90:78-84) */
enum { OR_GAUSS = 1, OR_GAUSS_LOBATTO = 2, OR_UNIFORM = 5 };

/* Boundary condition ids (src/SELF_Model.f90:70-72) */
enum { OR_BC_PRESCRIBED = 100, OR_BC_RADIATION = 101, OR_BC_NONORMALFLOW = 102 };

/* ---- a1: quadrature and Lagrange matrices ---- */
void or_legendre_polynomial(int N, double x, double *l, double *dl);
void or_legendre_gauss(int N, double *nodes, double *weights);
void or_legendre_gauss_lobatto(int N, double *nodes, double *weights);
/* Fills controlPoints(0:N), targetPoints(0:M), bWeights, qWeights(0:N),
 * iMatrix(0:N,0:M), dMatrix(0:N,0:N), dgMatrix(0:N,0:N), bMatrix(0:N,0:1). */
void or_lagrange_init(int N, int ctrlType, int M, int tgtType,
                      double *controlPoints, double *targetPoints,
                      double *bWeights, double *qWeights, double *iMatrix,
                      double *dMatrix, double *dgMatrix, double *bMatrix);

/* ---- a2, a3, a12, a16: Lagrange 2D operators on raw arrays ---- */
void or_scalar_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl);
void or_vector_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl);
void or_tensor_boundary_interp_2d(int N, const double *bM, const double *f, double *fb, int nVar, int nEl);
void or_scalar_grid_interp_2d(int N, int M, const double *iM, const double *f, double *fNew, int nVar, int nEl);
void or_vector_grid_interp_2d(int N, int M, const double *iM, const double *f, double *fNew, int nVar, int nEl);
void or_vector_gradient_2d(int N, const double *dM, const double *f, double *gradF, int nVar, int nEl);
void or_vector_divergence_2d(int N, const double *dM, const double *f, double *dF, int nVar, int nEl);
void or_vector_dg_divergence_2d(int N, const double *dgM, const double *bM, const double *qW,
                                const double *f, const double *bF, double *dF, int nVar, int nEl);
void or_determinant_tensor_2d(int N, const double *t, double *det, int nVar, int nEl);

/* ---- a18: mesh connectivity ---- */
void or_recalculate_flip(int nElem, int nGeo, int32_t *sideInfo, const int32_t *globalNodeIDs);
void or_reset_boundary_condition_type(int nElem, int32_t *sideInfo, int bcid);
/* 3-D HOPr extruded mesh -> 2-D mesh (src/SELF_Mesh.f90:852-1016, without the HDF5 reads).
 * sideInfo3D(5,nSides3D), nodeCoords3D(3,nNodes3D), globalNodeIDs3D(nNodes3D) as read.
 * Outputs sideInfo(5,4,nElem), nodeCoords(2,0:nGeo,0:nGeo,nElem), globalNodeIDs(0:nGeo,0:nGeo,nElem). */
void or_hopr_to_mesh2d(int nElem, int nGeo, const int32_t *sideInfo3D, const double *nodeCoords3D,
                       const int32_t *globalNodeIDs3D, int32_t *sideInfo, double *nodeCoords,
                       int32_t *globalNodeIDs);

/* ---- a16: geometry (SEMQuad) ---- */
typedef struct {
  int N, nElem;
  double *x, *xb;          /* Vector2D x interior / boundary */
  double *dxds, *dxdsb;    /* Tensor2D */
  double *dsdx, *dsdxb;    /* Tensor2D */
  double *nHatb;           /* Vector2D boundary */
  double *nScaleb;         /* Scalar2D boundary */
  double *J, *Jb;          /* Scalar2D */
} or_geometry;
/* nodeCoords(2,0:nGeo,0:nGeo,nElem) with uniform control points (HOPr / structured mesh) */
or_geometry *or_geometry_from_mesh(int N, int ctrlType, int nGeo, int nElem, const double *nodeCoords);
void or_geometry_free(or_geometry *g);
double or_covariant_arc_min(const or_geometry *g);

/* ---- a4, a5, a11, a12, a17: mapped operators ---- */
void or_side_exchange_scalar_2d(int N, int nVar, int nEl, const int32_t *sideInfo,
                                const double *boundary, double *extBoundary);
void or_side_exchange_vector_2d(int N, int nVar, int nEl, const int32_t *sideInfo,
                                const double *boundary, double *extBoundary);
void or_gradient_sf_2d(int N, const double *dM, const double *dsdx, const double *J,
                       const double *f, double *gradF, int nVar, int nEl);
void or_jacobian_weight_2d(int N, const double *J, double *f, int nVar, int nEl);
void or_contravariant_projection_2d(int N, const double *dsdx, double *f, int nVar, int nEl);

/* ---- a6-a15, a19: models ---- */
enum { OR_MODEL_LSW = 1, OR_MODEL_SW = 2, OR_MODEL_EULER = 3 };

typedef struct {
  int kind, N, nVar, nElem;
  /* Lagrange */
  double *qWeights, *dMatrix, *dgMatrix, *bMatrix;
  /* mesh + geometry (borrowed) */
  const int32_t *sideInfo;
  const or_geometry *geom;
  /* registers (Model2D, src/SELF_Model2D.f90:18-31) */
  double *sol, *solB, *solE;     /* solution interior / boundary / extBoundary */
  double *workSol;               /* RK register G */
  double *dSdt, *source, *fluxDiv;
  double *flux, *fluxBN;         /* flux interior (vector) / boundaryNormal (scalar boundary) */
  double *vel, *velB, *velE;     /* SW velocity (vector, nVar=1) */
  /* physics fields */
  double g, fCoriConst;          /* SW: g, fCori (scalar) ; LSW: g */
  double *topo, *topoB, *gradTopo;           /* SW this%H (topography) + gradH ; LSW bathymetry H */
  double *fCori;                             /* LSW fCori field */
  /* Euler */
  double gamma, Rgas;
  double *prim, *primB, *primE;
  double *diag, *diagB, *diagE;
  double *ent, *entB, *entE;
  double *env, *envGrad;                     /* environmentals(2 vars): gravity potential, drag */
  double *presSolB, *presPrimB, *presDiagB;  /* prescribed boundary states */
  double t, dt;
} or_model;

or_model *or_model_create(int kind, int N, int nElem, const int32_t *sideInfo, const or_geometry *geom);
void or_model_free(or_model *m);
/* Setup helpers (mirror Set* routines with arrays in place of equation strings) */
void or_model_set_solution(or_model *m, const double *interior);            /* SetSolution */
void or_sw_set_topography(or_model *m, const double *topo);                  /* SetTopography */
void or_lsw_set_bathymetry(or_model *m, const double *H);                    /* SetBathymetry */
void or_lsw_set_coriolis(or_model *m, const double *f);                      /* SetCoriolis */
void or_euler_set_gravity(or_model *m, const double *phi);                   /* SetGravity */
void or_euler_set_drag(or_model *m, const double *cd);                       /* SetDrag */
void or_euler_set_prescribed(or_model *m);                                   /* prescribed := solution */
/* Stage pieces */
void or_pre_tendency(or_model *m);
void or_set_boundary_condition(or_model *m);
void or_source(or_model *m);
void or_riemann(or_model *m);
void or_flux(or_model *m);
void or_calculate_tendency(or_model *m);
void or_update_grk3(or_model *m, int stage);
/* LowStorageRK3 over nSteps full steps of size dt (src/SELF_Model.f90:704-732) */
void or_rk3_steps(or_model *m, int nSteps, double dt);
double or_entropy(const or_model *m);
double or_mass(const or_model *m, int var);
/* Euler helpers used by PreTendency (exposed for operator-level parity) */
void or_euler_diagnostics(or_model *m);
void or_euler_primitive(or_model *m);
void or_euler_entropy_vars(or_model *m);

/* Named access to model registers / geometry arrays and scalars (test harness) */
double *or_model_field(or_model *m, const char *name);
void or_model_set_scalar(or_model *m, const char *name, double v);
double or_model_get_scalar(or_model *m, const char *name);
double *or_geometry_field(or_geometry *g, const char *name);

/* Threads for the element loops (OpenMP); 1 = the reference's serial order */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
