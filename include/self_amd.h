/*
 * self_amd.h -- C ABI of the MI355X-native DGSEM tendency engine.
 *
 * Conventions
 *   - Every function returns 0 on success, else a positive status (see SELFAMD_E*);
 *     selfamd_last_error() returns the message of the last failure on this thread.
 *   - Pointers named *_dev are device pointers (hipMalloc'd) passed BY VALUE; every other
 *     pointer is host memory.  The library never frees caller memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).  Calls only
 *     enqueue work; no call synchronises the device except where stated.
 *   - Arrays use the reference's Fortran column-major shapes (zero-based):
 *       interior(0:N,0:N,nVar,nEl), vector interior(2,0:N,0:N,nVar,nEl),
 *       tensor interior(2,2,0:N,0:N,nVar,nEl), boundary(0:N,nVar,4,nEl),
 *       vector boundary(2,0:N,nVar,4,nEl), tensor boundary(2,2,0:N,nVar,4,nEl),
 *       sideInfo(5,4,nEl) int32 with 1-based element ids.
 *   - Reference interfaces are cited file:line relative to the SELF checkout.  The
 *     reference wrappers take Fortran c_ptr arguments by reference (real**) and return
 *     nothing; these entry points take device pointers by value, an explicit stream and
 *     return a status.
 */
#ifndef SELF_AMD_H
#define SELF_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SELFAMD_VERSION 1

/* status codes */
#define SELFAMD_OK 0
#define SELFAMD_EBADARG 1
#define SELFAMD_EUNSUPPORTED 2
#define SELFAMD_EHIP 3
#define SELFAMD_ENOTAFFINE 4

/* quadrature codes (src/SELF_Constants.f90:78-84) */
#define SELFAMD_GAUSS 1
#define SELFAMD_GAUSS_LOBATTO 2
#define SELFAMD_UNIFORM 5

/* boundary condition ids (src/SELF_Model.f90:70-72) */
#define SELFAMD_BC_PRESCRIBED 100
#define SELFAMD_BC_RADIATION 101
#define SELFAMD_BC_NONORMALFLOW 102

/* physics models (src/models/SELF_LinearShallowWater.f90, SELF_ShallowWater.f90,
 * SELF_CompressibleIdealGas2D.f90) */
#define SELFAMD_MODEL_LSW 1
#define SELFAMD_MODEL_SW 2
#define SELFAMD_MODEL_EULER 3

const char *selfamd_last_error(void);
int selfamd_version(void);

/* ------------------------------------------------------------------------------------ */
/* Device plumbing: hosts without a HIP runtime of their own (ctypes, iso_c_binding)      */
/* allocate and copy through these, so a process loads only the runtime this library     */
/* links.  No reference counterpart (the reference uses hipfort, src/SELF_Memory.f90).    */
/* ------------------------------------------------------------------------------------ */
int selfamd_device_count(int *count);
int selfamd_device_name(int device, char *name, int len);
int selfamd_malloc(void **ptr, size_t bytes);
int selfamd_free(void *ptr);
/* kind: 1 host->device, 2 device->host, 3 device->device (synchronous) */
int selfamd_memcpy(void *dst, const void *src, size_t bytes, int kind);
int selfamd_memset(void *ptr, int value, size_t bytes);
int selfamd_device_synchronize(void);
/* dst = src for n doubles (n even, 16-byte aligned pointers): a streaming kernel of known
 * traffic (16 n bytes) for bandwidth and counter calibration.  No reference counterpart. */
int selfamd_copy_f64(const double *src_dev, double *dst_dev, size_t n, void *stream);
/* hipEvent-timed repetitions of selfamd_copy_f64 on `stream`; returns total milliseconds. */
int selfamd_time_copy_f64(const double *src_dev, double *dst_dev, size_t n, int reps, double *elapsedMs);

/* ------------------------------------------------------------------------------------ */
/* Host setup                                                                            */
/* ------------------------------------------------------------------------------------ */

/* Replaces Init_Lagrange (src/SELF_Lagrange.f90:521-615) with LegendreGauss
 * (src/SELF_Quadrature.f90:201-262).  Host arrays: controlPoints/bWeights/qWeights(0:N),
 * targetPoints(0:M), iMatrix(0:N,0:M), dMatrix/dgMatrix(0:N,0:N), bMatrix(0:N,0:1). */
int selfamd_lagrange_init(int N, int controlNodeType, int M, int targetNodeType,
                          double *controlPoints, double *targetPoints, double *bWeights,
                          double *qWeights, double *iMatrix, double *dMatrix, double *dgMatrix,
                          double *bMatrix);

/* Structured nx*ny quad mesh (nGeo = 1, raster element order).  No reference counterpart:
 * the reference only reads HOPr files (src/SELF_Mesh.f90:852-1016); this emits the same
 * sideInfo convention (src/SELF_Mesh.f90:84-91).  Periodic directions become interior
 * sides.  Host outputs: sideInfo(5,4,nx*ny), nodeCoords(2,0:1,0:1,nx*ny),
 * globalNodeIDs(0:1,0:1,nx*ny). */
int selfamd_mesh2d_structured(int nx, int ny, double x0, double x1, double y0, double y1,
                              int periodicX, int periodicY, int boundaryBcid, int32_t *sideInfo,
                              double *nodeCoords, int32_t *globalNodeIDs);
/* Replaces RecalculateFlip_Mesh2D (src/SELF_Mesh.f90:1018-1174), single rank. */
int selfamd_mesh2d_recalculate_flip(int nElem, int nGeo, int32_t *sideInfo, const int32_t *globalNodeIDs);
/* Replaces ResetBoundaryConditionType_Mesh2D (src/SELF_Mesh.f90:695-719). */
int selfamd_mesh2d_reset_boundary_condition_type(int nElem, int32_t *sideInfo, int bcid);
/* Replaces the 3-D -> 2-D part of Read_HOPr_Mesh2D (src/SELF_Mesh.f90:962-1007) on arrays
 * already read from the file: sideInfo3D(5,6*nElem), nodeCoords3D(3,(nGeo+1)^3*nElem),
 * globalNodeIDs3D((nGeo+1)^3*nElem). */
int selfamd_mesh2d_from_hopr(int nElem, int nGeo, const int32_t *sideInfo3D, const double *nodeCoords3D,
                             const int32_t *globalNodeIDs3D, int32_t *sideInfo, double *nodeCoords,
                             int32_t *globalNodeIDs);

/* ------------------------------------------------------------------------------------ */
/* Lagrange / data operators (device)                                                    */
/* ------------------------------------------------------------------------------------ */

/* ScalarBoundaryInterp_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:393-399) */
int selfamd_scalar_boundary_interp_2d(const double *bMatrix_dev, const double *f_dev, double *fBound_dev,
                                      int N, int nVar, int nEl, void *stream);
/* VectorBoundaryInterp_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:422-428) */
int selfamd_vector_boundary_interp_2d(const double *bMatrix_dev, const double *f_dev, double *fBound_dev,
                                      int N, int nVar, int nEl, void *stream);
/* TensorBoundaryInterp_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp, TensorBoundaryInterp_2D) */
int selfamd_tensor_boundary_interp_2d(const double *bMatrix_dev, const double *f_dev, double *fBound_dev,
                                      int N, int nVar, int nEl, void *stream);
/* ScalarGridInterp_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:85); iMatrix(0:N,0:M) */
int selfamd_scalar_grid_interp_2d(const double *iMatrix_dev, const double *f_dev, double *fNew_dev,
                                  int N, int M, int nVar, int nEl, void *stream);
/* VectorGridInterp_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:114) */
int selfamd_vector_grid_interp_2d(const double *iMatrix_dev, const double *f_dev, double *fNew_dev,
                                  int N, int M, int nVar, int nEl, void *stream);
/* VectorGradient_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:707) */
int selfamd_vector_gradient_2d(const double *dMatrix_dev, const double *f_dev, double *gradF_dev,
                               int N, int nVar, int nEl, void *stream);
/* VectorDivergence_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:777), strong form */
int selfamd_vector_divergence_2d(const double *dMatrix_dev, const double *f_dev, double *dF_dev,
                                 int N, int nVar, int nEl, void *stream);
/* VectorDGDivergence_2D_gpu_wrapper (src/hip/SELF_Lagrange.cpp:825-831), weak form */
int selfamd_vector_dg_divergence_2d(const double *dgMatrix_dev, const double *bMatrix_dev,
                                    const double *qWeights_dev, const double *f_dev, const double *bF_dev,
                                    double *dF_dev, int N, int nVar, int nEl, void *stream);
/* Determinant_Tensor2D_gpu_wrapper (src/hip/SELF_Data.cpp, Determinant_Tensor2D) */
int selfamd_determinant_tensor_2d(const double *t_dev, double *det_dev, int N, int nVar, int nEl, void *stream);

/* ------------------------------------------------------------------------------------ */
/* Mapped operators and model register updates (device)                                  */
/* ------------------------------------------------------------------------------------ */

/* SideExchange_MappedScalar2D_gpu_wrapper (src/hip/SELF_MappedData.cpp:941-947); sideInfo
 * is a device copy.  Only sides with bcid == 0 are written (src/SELF_MappedData.f90:878). */
int selfamd_side_exchange_scalar_2d(const double *boundary_dev, double *extBoundary_dev,
                                    const int32_t *sideInfo_dev, int N, int nVar, int nEl, void *stream);
/* SideExchange_MappedVector2D_gpu_wrapper (src/hip/SELF_MappedData.cpp:950) */
int selfamd_side_exchange_vector_2d(const double *boundary_dev, double *extBoundary_dev,
                                    const int32_t *sideInfo_dev, int N, int nVar, int nEl, void *stream);
/* GradientSF_MappedScalar2D_gpu_wrapper (src/hip/SELF_MappedData.cpp:171); dsdx and J are
 * the geometry arrays (nVar = 1). */
int selfamd_gradient_sf_2d(const double *dMatrix_dev, const double *dsdx_dev, const double *J_dev,
                           const double *f_dev, double *gradF_dev, int N, int nVar, int nEl, void *stream);
/* JacobianWeight_MappedScalar2D_gpu_wrapper (src/hip/SELF_MappedData.cpp:210), in place */
int selfamd_jacobian_weight_2d(const double *J_dev, double *f_dev, int N, int nVar, int nEl, void *stream);
/* ContravariantProjection_MappedVector2D_gpu_wrapper (src/hip/SELF_MappedData.cpp:401), in place */
int selfamd_contravariant_projection_2d(const double *dsdx_dev, double *f_dev, int N, int nVar, int nEl,
                                        void *stream);
/* UpdateGRK_Model2D_gpu_wrapper (src/hip/SELF_Model2D.cpp:153-157):
 * G = a*G + dSdt ; S = S + (g*dt)*G */
int selfamd_update_grk(double *G_dev, double *S_dev, const double *dSdt_dev, double a, double g, double dt,
                       int N, int nVar, int nEl, void *stream);
/* CalculateDSDt_Model2D_gpu_wrapper (src/hip/SELF_Model2D.cpp:192-196): dSdt = source - fluxDiv */
int selfamd_calculate_dsdt(const double *fluxDiv_dev, const double *source_dev, double *dSdt_dev, int N,
                           int nVar, int nEl, void *stream);

/* GenerateFromMesh_SEMQuad + CalculateMetricTerms_SEMQuad (src/SELF_Geometry.f90:393-552),
 * computed on the device (the reference runs these on the host; its GPU branch is
 * commented out at :419-429,532-543).  nodeCoords(2,0:nGeo,0:nGeo,nEl) on uniform control
 * points; outputs are SEMQuad's x, dxds, dsdx, nHat, nScale, J in the shapes above
 * (interior and boundary).  Synchronises `stream` before returning. */
int selfamd_semquad_generate(int N, int controlNodeType, int nGeo, int nEl, const double *nodeCoords_dev,
                             double *x_dev, double *xb_dev, double *dxds_dev, double *dxdsb_dev,
                             double *dsdx_dev, double *dsdxb_dev, double *nHatb_dev, double *nScaleb_dev,
                             double *J_dev, double *Jb_dev, void *stream);

/* ------------------------------------------------------------------------------------ */
/* Physics operators (device), one per reference plug-in method                          */
/* ------------------------------------------------------------------------------------ */

/* SW PreTendency (src/models/SELF_ShallowWater.f90:190-235): velocity interior from the
 * current interior, velocity boundary from the solution boundary as it stands (the
 * previous stage's traces).  Follow with selfamd_side_exchange_vector_2d. */
int selfamd_sw2d_velocity(const double *sol_dev, const double *solB_dev, double *vel_dev, double *velB_dev,
                          int N, int nEl, void *stream);
/* SetBoundaryCondition_ShallowWater (src/models/SELF_ShallowWater.f90:284-343) */
int selfamd_sw2d_set_boundary_condition(const double *solB_dev, double *solE_dev, const double *topoB_dev,
                                        const double *nHatb_dev, const int32_t *sideInfo_dev, int N, int nEl,
                                        void *stream);
/* Source_ShallowWater (src/models/SELF_ShallowWater.f90:345-384) */
int selfamd_sw2d_source(double *source_dev, const double *sol_dev, const double *gradTopo_dev, double g,
                        double fCori, int N, int nEl, void *stream);
/* Flux_ShallowWater (src/models/SELF_ShallowWater.f90:386-449) */
int selfamd_sw2d_flux(double *flux_dev, const double *sol_dev, const double *vel_dev, double g, int N, int nEl,
                      void *stream);
/* RiemannSolver_ShallowWater (src/models/SELF_ShallowWater.f90:451-529), Rusanov */
int selfamd_sw2d_riemann(double *fluxBN_dev, const double *solB_dev, const double *solE_dev,
                         const double *velB_dev, const double *velE_dev, const double *nHatb_dev,
                         const double *nScaleb_dev, double g, int N, int nEl, void *stream);

/* SetBoundaryCondition_LinearShallowWater_gpu_wrapper (src/models/hip/SELF_LinearShallowWater.cpp:6),
 * restated from the CPU branch (src/models/SELF_LinearShallowWater.f90:298-355) */
int selfamd_lsw2d_set_boundary_condition(const double *solB_dev, double *solE_dev, const double *nHatb_dev,
                                         const int32_t *sideInfo_dev, int N, int nEl, void *stream);
/* Source_LinearShallowWater_gpu_wrapper (src/models/hip/SELF_LinearShallowWater.cpp:54) */
int selfamd_lsw2d_source(double *source_dev, const double *sol_dev, const double *fCori_dev, int N, int nEl,
                         void *stream);
/* Flux_LinearShallowWater_gpu_wrapper (src/models/hip/SELF_LinearShallowWater.cpp:80) */
int selfamd_lsw2d_flux(double *flux_dev, const double *sol_dev, const double *H_dev, double g, int N, int nEl,
                       void *stream);
/* RiemannSolver_LinearShallowWater_gpu_wrapper (src/models/hip/SELF_LinearShallowWater.cpp:108) */
int selfamd_lsw2d_riemann(double *fluxBN_dev, const double *solB_dev, const double *solE_dev,
                          const double *Hb_dev, const double *nHatb_dev, const double *nScaleb_dev, double g,
                          int N, int nEl, void *stream);

/* CalculateDiagnostics_CompressibleIdealGas2D_gpu_wrapper (src/models/hip/SELF_CompressibleIdealGas2D.cpp:324) */
int selfamd_euler2d_diagnostics(const double *sol_dev, double *diag_dev, double gamma, double R, int N, int nEl,
                                void *stream);
/* ConservativeToPrimitive_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:376) */
int selfamd_euler2d_primitive(const double *sol_dev, double *prim_dev, double gamma, int N, int nEl, void *stream);
/* ConservativeToEntropy_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:413) */
int selfamd_euler2d_entropy_vars(const double *sol_dev, double *ent_dev, double gamma, int N, int nEl,
                                 void *stream);
/* SetBoundaryCondition_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:208),
 * restated from the CPU branch (src/models/SELF_CompressibleIdealGas2D.f90:1217-1323) */
int selfamd_euler2d_set_boundary_condition(const double *solB_dev, double *solE_dev, const double *presSolB_dev,
                                           const double *primB_dev, double *primE_dev,
                                           const double *presPrimB_dev, const double *diagB_dev,
                                           double *diagE_dev, const double *presDiagB_dev,
                                           const double *nHatb_dev, const int32_t *sideInfo_dev, int N, int nEl,
                                           void *stream);
/* Source_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:5), restated from
 * the CPU branch, drag term included (src/models/SELF_CompressibleIdealGas2D.f90:1325-1366).
 * env(0:N,0:N,2,nEl) = (potential, drag); envGrad(2,0:N,0:N,2,nEl). */
int selfamd_euler2d_source(double *source_dev, const double *sol_dev, const double *env_dev,
                           const double *envGrad_dev, int N, int nEl, void *stream);
/* Flux_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:34) */
int selfamd_euler2d_flux(double *flux_dev, const double *sol_dev, const double *prim_dev, const double *diag_dev,
                         int N, int nEl, void *stream);
/* NaiveLLF_CompressibleIdealGas2D_gpu_wrapper (...CompressibleIdealGas2D.cpp:101) */
int selfamd_euler2d_riemann(double *fluxBN_dev, const double *solB_dev, const double *solE_dev,
                            const double *primB_dev, const double *primE_dev, const double *diagB_dev,
                            const double *diagE_dev, const double *nHatb_dev, const double *nScaleb_dev, int N,
                            int nEl, void *stream);

/* ------------------------------------------------------------------------------------ */
/* Model2D engine: the fused tendency + low-storage RK3 path                             */
/* ------------------------------------------------------------------------------------ */
/* One handle replaces a Model2D extension (src/SELF_Model2D.f90:18-71) on one GPU:
 * CalculateTendency (:816-859) and ForwardStep/LowStorageRK3 (src/SELF_Model.f90:474-523,
 * 704-732) run as one fused kernel per RK stage.  The state stays resident on the device. */
typedef struct selfamd_model2d selfamd_model2d;

/* scalar parameters (selfamd_model2d_set_param) */
#define SELFAMD_PARAM_G 1      /* gravity g (LSW, SW) */
#define SELFAMD_PARAM_FCORI 2  /* SW constant Coriolis parameter */
#define SELFAMD_PARAM_GAMMA 3  /* Euler expansion factor Cp/Cv */
#define SELFAMD_PARAM_RGAS 4   /* Euler gas constant */
#define SELFAMD_PARAM_NODAL_METRICS 5  /* 1: per-node metric terms (as SEMQuad stores them);
                                          0: one affine record per element (affine meshes).
                                          The affine record takes node (0,0) and side-node-0
                                          values, so it departs from the reference's per-node
                                          metrics by their round-off (~1e-12 relative on
                                          affine meshes), which moves the tendency by up to
                                          ~1e-10 relative (7e-11 at 256^2 on the unit square):
                                          set 1 for the reference's arithmetic to 1e-12 */
/* static per-node fields (selfamd_model2d_set_field), interior(0:N,0:N,1,nEl) device arrays */
#define SELFAMD_FIELD_TOPOGRAPHY 1  /* SW this%H (SetTopography) */
#define SELFAMD_FIELD_BATHYMETRY 2  /* LSW this%H (SetBathymetry) */
#define SELFAMD_FIELD_CORIOLIS 3    /* LSW this%fCori (SetCoriolis) */
#define SELFAMD_FIELD_POTENTIAL 4   /* Euler gravity potential (SetGravity) */
#define SELFAMD_FIELD_DRAG 5        /* Euler drag coefficient (SetDrag) */

/* Init_Model2D + the model's Init (e.g. src/models/SELF_ShallowWater.f90:96-147).  sideInfo is
 * host memory; geometry arrays are SEMQuad device arrays (dsdx interior, J interior, nHat and
 * nScale boundary) from selfamd_semquad_generate, copied into the engine.  Affine meshes
 * (metric terms constant per element to 1e-10 relative) default to one metric record per
 * element, a representation that departs from the reference's per-node metric values at the
 * 1e-10 level (see SELFAMD_PARAM_NODAL_METRICS); other meshes use the per-node metrics. */
int selfamd_model2d_create(selfamd_model2d **model, int kind, int N, int nEl, const int32_t *sideInfo,
                           const double *dsdx_dev, const double *J_dev, const double *nHatb_dev,
                           const double *nScaleb_dev, void *stream);
int selfamd_model2d_destroy(selfamd_model2d *model);
int selfamd_model2d_set_param(selfamd_model2d *model, int which, double value);
int selfamd_model2d_set_field(selfamd_model2d *model, int which, const double *field_dev);
/* SetSolution (src/SELF_Model2D.f90:313-332): copies interior(0:N,0:N,nVar,nEl) and
 * interpolates it to the element boundaries. */
int selfamd_model2d_set_solution(selfamd_model2d *model, const double *interior_dev);
int selfamd_model2d_get_solution(selfamd_model2d *model, double *interior_dev);
/* CalculateTendency (src/SELF_Model2D.f90:816-859): dSdt(0:N,0:N,nVar,nEl) */
int selfamd_model2d_tendency(selfamd_model2d *model, double *dSdt_dev);
/* The flux and fluxDivergence registers that CalculateTendency_Model2D
 * (src/SELF_Model2D.f90:816-859) leaves behind, for the current state: flux(2,0:N,0:N,nVar,nEl)
 * after ContravariantProjection and fluxDivergence(0:N,0:N,nVar,nEl) after JacobianWeight, as
 * Write_Model2D (:861-1059) stores them.  Neither the state nor the SW lag state changes.
 * Synchronises the model's stream. */
int selfamd_model2d_flux_registers(selfamd_model2d *model, double *fluxDivergence_dev, double *flux_dev);
/* CalculateEntropy of the current state into *entropy (host): SW
 * src/models/SELF_ShallowWater.f90:149-188, LSW src/models/SELF_LinearShallowWater.f90:171-216,
 * Euler src/models/SELF_CompressibleIdealGas2D.f90:782-823.  Euler takes the pressure of the
 * current state; the reference takes it from the primitive variables of its last tendency
 * evaluation.  Deterministic reduction; synchronises the model's stream. */
int selfamd_model2d_entropy(selfamd_model2d *model, double *entropy);
/* sum over elements and nodes of w_i w_j J S(i,j,var,e) into *value (host): the mass of
 * variable var (H for SW, eta for LSW, rho for Euler with var = 2).  No reference routine:
 * the pattern of Integral_MappedScalar2D (src/SELF_MappedData.f90:1270-1311) with the
 * variable as an argument. */
int selfamd_model2d_integral(selfamd_model2d *model, int var, double *value);
/* selfamd_model2d_integral without any allocation: the same kernels and the same result, bit
 * for bit, written to *value_dev (device memory) on the model's stream, without
 * synchronising.  No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial sums live in a 512 KiB device array of the library, one per GPU,
 * kept until the process ends; meshes of more than 16,777,216 nodes ((N+1)^2 nEl) are
 * rejected with an error.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (other streams, other host threads) race on it.  The current device must be the
 * model's. */
int selfamd_model2d_integral_noalloc(selfamd_model2d *model, int var, double *value_dev);
/* nSteps full low-storage RK3 steps of size dt (ForwardStep with tn = t + nSteps*dt). */
int selfamd_model2d_forward_step(selfamd_model2d *model, int nSteps, double dt);
/* Blocks until the model's stream has drained. */
int selfamd_model2d_synchronize(selfamd_model2d *model);
/* The HIP stream (hipStream_t, as void*) that the model's kernels and graphs run on, so that
 * a caller can order its own work with the model's.  No reference counterpart (the reference
 * launches everything on the default stream). */
int selfamd_model2d_stream(selfamd_model2d *model, void **stream);
/* forward_step bracketed by two hipEvents on the model's stream; returns the elapsed
 * milliseconds between them (synchronises).  Graphs that the steps need are captured before
 * the first event. */
int selfamd_model2d_time_steps(selfamd_model2d *model, int nSteps, double dt, double *elapsedMs);
/* Tile variants of the stage kernel (affine metrics) of nonlinear SW at N = 3, 5, 7 and of
 * LSW and Euler at N = 7: block size (64, 128, 256 work-items), element order alternating
 * between stages (so a stage starts on the data the previous one left in the Infinity Cache),
 * XCD-contiguous element order, and an occupancy target (minWaves wavefronts per SIMD; 1 =
 * none) for the register allocator.
 * No reference counterpart; the autotuner (self_amd/autotune.py) times them and records the
 * fastest per N.  A set variant (-1 = default tile) applies to every model's next graph
 * capture and to batches created afterwards.  Variants agree to round-off (blocks with room
 * for two lanes per face node split the face-line sums between them). */
int selfamd_tile_count(int kind, int N, int *count);
int selfamd_tile_info(int kind, int N, int variant, int *threads, int *alternate, int *xcdOrder, int *minWaves);
int selfamd_tile_set(int kind, int N, int variant);
int selfamd_tile_get(int kind, int N, int *variant);
/* Banded stepping (no reference counterpart): nBands > 0 splits the elements into bands of
 * whole BFS levels of the element graph and runs each RK3 step as a wavefront over the bands
 * (stage s of band b after stage s-1 of band b+1), so a band's data are reused from the
 * Infinity Cache across its three stages.  Results are bit-identical to nBands = 0 (the
 * default, one launch per stage over all elements).  The number of bands used can be lower
 * (band limits fall between BFS levels).  Not with a closure attached. */
int selfamd_model2d_set_bands(selfamd_model2d *model, int nBands);
/* Model time t */
int selfamd_model2d_get_time(selfamd_model2d *model, double *t);
int selfamd_model2d_set_time(selfamd_model2d *model, double t);
/* Number of conserved variables (3 for LSW/SW, 4 for Euler). */
int selfamd_model2d_nvar(int kind);

/* ------------------------------------------------------------------------------------ */
/* Batched runs: many independent models advanced together                               */
/* ------------------------------------------------------------------------------------ */
/* No reference counterpart (the reference runs one model per process).  Models of the same
 * kind, affine metrics and identical physical parameters; runs of one polynomial degree are
 * concatenated into one mesh and advanced by one fused launch per RK stage, all degree groups
 * of a step in one hipGraph.  dt[r] is run r's step.  Creation copies the runs' current
 * state; selfamd_batch_scatter copies it back, after which each model continues alone
 * (bit-identical to having been advanced alone). */
typedef struct selfamd_batch selfamd_batch;
int selfamd_batch_create(selfamd_batch **batch, int nModels, selfamd_model2d *const *models, const double *dt);
int selfamd_batch_forward_step(selfamd_batch *batch, int nSteps);
int selfamd_batch_time_steps(selfamd_batch *batch, int nSteps, double *elapsedMs);
int selfamd_batch_scatter(selfamd_batch *batch);
int selfamd_batch_info(selfamd_batch *batch, int *nGroups, double *totalDof);
int selfamd_batch_destroy(selfamd_batch *batch);
/* Launch structure of the batch's RK3 stages (no reference counterpart):
 * SELFAMD_BATCH_PER_GROUP (default) one fused launch per degree group and stage, each with its
 * degree's tuned tile; SELFAMD_BATCH_PERSISTENT one launch per stage for all groups, a
 * persistent grid (the device's CUs times the resident blocks per CU) whose blocks stride
 * through every group's 256-wide tiles, group after group (nonlinear SW, degrees 3, 5 and 7;
 * refused when the tuned N = 5 tile is 64 wide).  Either way each run is bit-identical to its
 * separate solve.  Synchronizes the batch's stream and recaptures the step graphs. */
#define SELFAMD_BATCH_PER_GROUP 0
#define SELFAMD_BATCH_PERSISTENT 1
int selfamd_batch_set_mode(selfamd_batch *batch, int mode);

/* ------------------------------------------------------------------------------------ */
/* Learned subgrid closure: FP32 MLP, fine -> coarse projection, model hook               */
/* ------------------------------------------------------------------------------------ */
/* No reference counterpart.  An MLP n_in -> H -> H -> n_out with GELU hidden layers, FP32
 * weights and FP32 accumulation, trained by the library's own forward/backward kernels
 * (mean-squared-error loss, Adam).  Parameters are packed on the host as
 * W1[n_in][H] b1[H] W2[H][H] b2[H] W3[H][n_out] b3[n_out] (row-major W[in][out], so a layer
 * is Y = X W + b).  Rows of X and Y are samples, device memory, row-major. */
typedef struct selfamd_mlp selfamd_mlp;
int selfamd_mlp_create(selfamd_mlp **mlp, int nIn, int nHidden, int nOut, int maxBatch);
int selfamd_mlp_destroy(selfamd_mlp *mlp);
int selfamd_mlp_num_params(selfamd_mlp *mlp, size_t *nParams);
/* Loads parameters (host) and resets the Adam moments. */
int selfamd_mlp_set_params(selfamd_mlp *mlp, const float *params);
int selfamd_mlp_get_params(selfamd_mlp *mlp, float *params);
/* Gradients of the last selfamd_mlp_loss_grad call, packed like the parameters (host). */
int selfamd_mlp_get_grads(selfamd_mlp *mlp, float *grads);
int selfamd_mlp_forward(selfamd_mlp *mlp, const float *X_dev, float *Y_dev, int batch, void *stream);
/* Forward + backward of L = mean((Y - T)^2) over batch rows; *loss (host, may be null) is the
 * loss of this batch (synchronises the stream when requested). */
int selfamd_mlp_loss_grad(selfamd_mlp *mlp, const float *X_dev, const float *T_dev, int batch, float *loss,
                          void *stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected; step counts from 1). */
int selfamd_mlp_adam_step(selfamd_mlp *mlp, float lr, int step, void *stream);
/* Backward pass from a given output cotangent dY[batch][nOut] (device): forward of X, then the
 * parameter gradients written to the gradient buffer (accumulate = 0) or added to it
 * (accumulate = 1), and the input cotangent dX[batch][nIn] (device, may be NULL).  For
 * training through a solver (the closure's a-posteriori training).  No reference counterpart. */
int selfamd_mlp_backward(selfamd_mlp *mlp, const float *X_dev, const float *dY_dev, int batch, int accumulate,
                         float *dX_dev, void *stream);
/* Sets the gradient buffer to zero (before accumulating backward passes). */
int selfamd_mlp_zero_grads(selfamd_mlp *mlp, void *stream);

/* L2 projection of interior(0:Nf,0:Nf,nVar,(r nxc)(r nyc)) on a structured raster-ordered
 * mesh onto interior(0:Nc,0:Nc,nVar,nxc nyc): each coarse element covers r x r fine elements.
 * Exact for coarse-space polynomials and mass-conserving. */
int selfamd_project_l2_structured_2d(const double *fine_dev, double *coarse_dev, int Nf, int Nc, int nxc, int nyc,
                                     int r, int nVar, void *stream);

/* Attaches an MLP as a tendency correction: before each stage (and in CalculateTendency) the
 * element features x = ((S nodal values, J, dsdx(1,1), dsdx(2,2)) - inMean) / inScale go
 * through the network, and dS/dt += y * outScale + outMean, with the mass variable's
 * correction made mean-free over the mesh (mass is conserved to round-off).  Needs an affine
 * mesh and n_in = nVar (N+1)^2 + 3, n_out = nVar (N+1)^2, maxBatch >= nEl.  Normalisation
 * arrays are host memory; mlp = NULL detaches.  The model keeps a pointer to the network. */
int selfamd_model2d_set_closure(selfamd_model2d *model, selfamd_mlp *mlp, const float *inMean, const float *inScale,
                                const float *outMean, const float *outScale);
/* Closure training through the solver (no reference counterpart).
 * selfamd_model2d_stage: one low-storage RK3 stage (0, 1, 2 in turn; a step is three calls)
 * of ForwardStep's arithmetic, closure included; before it the stage's input state is copied to
 * Ssave_dev (interior layout) and the lagged SW face normal velocities it reads to unSave_dev
 * ((0:N,4,nEl), affine mode), either may be NULL.  Synchronises the model's stream. */
int selfamd_model2d_stage(selfamd_model2d *model, int stage, double dt, double *Ssave_dev, double *unSave_dev);
/* mu = (d dSdt / dS)^T lambda: the adjoint of the nonlinear shallow-water tendency
 * (CalculateTendency_Model2D, src/SELF_Model2D.f90:816-859, with the plug-ins of
 * src/models/SELF_ShallowWater.f90) at state S with the lagged normal velocities un (as saved
 * by selfamd_model2d_stage); affine meshes whose sides are all interior (periodic).  With
 * Sprev (the previous stage's input state, from which un was formed) also the lag path,
 * muPrev = (d dSdt / d Sprev)^T lambda; both NULL holds un fixed.  The closure's correction is
 * not included (selfamd_model2d_closure_vjp).  Synchronises the model's stream. */
int selfamd_model2d_tendency_vjp(selfamd_model2d *model, const double *S_dev, const double *un_dev,
                                 const double *lambda_dev, double *mu_dev, const double *Sprev_dev, double *muPrev_dev);
/* Adjoint of the attached closure's correction at state S for its cotangent lambda: the
 * network's parameter gradients are ADDED to its gradient buffer (selfamd_mlp_zero_grads
 * clears it) and the state cotangent through the network's inputs is ADDED to mu.
 * Synchronises the model's stream. */
int selfamd_model2d_closure_vjp(selfamd_model2d *model, const double *S_dev, const double *lambda_dev, double *mu_dev);
/* The closure's normalised features of the current state: X(nEl, n_in) FP32 device. */
int selfamd_model2d_closure_features(selfamd_model2d *model, float *X_dev);

#ifdef __cplusplus
}
#endif
#endif
