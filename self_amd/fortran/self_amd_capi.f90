! self_amd_capi.f90 -- iso_c_binding interfaces to libself_amd.so (include/self_amd.h),
! including its device allocation and copy entry points.
!
! Device pointers are passed BY VALUE (type(c_ptr), value); every entry point returns an
! integer status (0 = success).  selfamd_check stops with the library's message otherwise.
MODULE SELF_AMD_CAPI
  USE ISO_C_BINDING
  IMPLICIT NONE

  ! selfamd_memcpy kinds
  INTEGER(C_INT), PARAMETER :: hipMemcpyHostToDevice = 1
  INTEGER(C_INT), PARAMETER :: hipMemcpyDeviceToHost = 2
  INTEGER(C_INT), PARAMETER :: hipMemcpyDeviceToDevice = 3

  INTERFACE
    ! ---- device plumbing (libself_amd; the process loads only the HIP runtime it links) ----
    INTEGER(C_INT) FUNCTION selfamd_malloc(ptr, bytes) BIND(C, name="selfamd_malloc")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR) :: ptr
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_free(ptr) BIND(C, name="selfamd_free")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ptr
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_memcpy(dst, src, bytes, kind) BIND(C, name="selfamd_memcpy")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: dst, src
      INTEGER(C_SIZE_T), VALUE :: bytes
      INTEGER(C_INT), VALUE :: kind
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_memset(ptr, val, bytes) BIND(C, name="selfamd_memset")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: ptr
      INTEGER(C_INT), VALUE :: val
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_device_synchronize() BIND(C, name="selfamd_device_synchronize")
      IMPORT :: C_INT
    END FUNCTION

    ! ---- libself_amd: errors ----
    TYPE(C_PTR) FUNCTION selfamd_last_error() BIND(C, name="selfamd_last_error")
      IMPORT :: C_PTR
    END FUNCTION

    ! ---- host setup ----
    INTEGER(C_INT) FUNCTION selfamd_lagrange_init(N, ctrl, M, tgt, cp, tp, bw, qw, iM, dM, dgM, bM) &
        BIND(C, name="selfamd_lagrange_init")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, ctrl, M, tgt
      REAL(C_DOUBLE) :: cp(*), tp(*), bw(*), qw(*), iM(*), dM(*), dgM(*), bM(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_mesh2d_structured(nx, ny, x0, x1, y0, y1, px, py, bcid, sideInfo, &
        nodeCoords, gnid) BIND(C, name="selfamd_mesh2d_structured")
      IMPORT :: C_INT, C_DOUBLE, C_INT32_T
      INTEGER(C_INT), VALUE :: nx, ny, px, py, bcid
      REAL(C_DOUBLE), VALUE :: x0, x1, y0, y1
      INTEGER(C_INT32_T) :: sideInfo(*), gnid(*)
      REAL(C_DOUBLE) :: nodeCoords(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_mesh2d_recalculate_flip(nElem, nGeo, sideInfo, gnid) &
        BIND(C, name="selfamd_mesh2d_recalculate_flip")
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT), VALUE :: nElem, nGeo
      INTEGER(C_INT32_T) :: sideInfo(*), gnid(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_mesh2d_reset_boundary_condition_type(nElem, sideInfo, bcid) &
        BIND(C, name="selfamd_mesh2d_reset_boundary_condition_type")
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT), VALUE :: nElem, bcid
      INTEGER(C_INT32_T) :: sideInfo(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_semquad_generate(N, ctrl, nGeo, nEl, nodeCoords, x, xb, dxds, dxdsb, &
        dsdx, dsdxb, nHatb, nScaleb, J, Jb, stream) BIND(C, name="selfamd_semquad_generate")
      IMPORT :: C_INT, C_PTR
      INTEGER(C_INT), VALUE :: N, ctrl, nGeo, nEl
      TYPE(C_PTR), VALUE :: nodeCoords, x, xb, dxds, dxdsb, dsdx, dsdxb, nHatb, nScaleb, J, Jb, stream
    END FUNCTION

    ! ---- operators (device) ----
    INTEGER(C_INT) FUNCTION selfamd_scalar_boundary_interp_2d(bM, f, fb, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_scalar_boundary_interp_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: bM, f, fb, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_vector_boundary_interp_2d(bM, f, fb, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_vector_boundary_interp_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: bM, f, fb, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_scalar_grid_interp_2d(iM, f, fNew, N, M, nVar, nEl, stream) &
        BIND(C, name="selfamd_scalar_grid_interp_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: iM, f, fNew, stream
      INTEGER(C_INT), VALUE :: N, M, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_vector_grid_interp_2d(iM, f, fNew, N, M, nVar, nEl, stream) &
        BIND(C, name="selfamd_vector_grid_interp_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: iM, f, fNew, stream
      INTEGER(C_INT), VALUE :: N, M, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_vector_dg_divergence_2d(dgM, bM, qW, f, bF, dF, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_vector_dg_divergence_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: dgM, bM, qW, f, bF, dF, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_side_exchange_scalar_2d(b, ext, sideInfo, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_side_exchange_scalar_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: b, ext, sideInfo, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_side_exchange_vector_2d(b, ext, sideInfo, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_side_exchange_vector_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: b, ext, sideInfo, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_gradient_sf_2d(dM, dsdx, J, f, gradF, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_gradient_sf_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: dM, dsdx, J, f, gradF, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_jacobian_weight_2d(J, f, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_jacobian_weight_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: J, f, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_contravariant_projection_2d(dsdx, f, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_contravariant_projection_2d")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: dsdx, f, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION

    ! ---- Model2D engine ----
    INTEGER(C_INT) FUNCTION selfamd_model2d_create(model, kind, N, nEl, sideInfo, dsdx, J, nHatb, nScaleb, &
        stream) BIND(C, name="selfamd_model2d_create")
      IMPORT :: C_INT, C_PTR, C_INT32_T
      TYPE(C_PTR) :: model
      INTEGER(C_INT), VALUE :: kind, N, nEl
      INTEGER(C_INT32_T) :: sideInfo(*)
      TYPE(C_PTR), VALUE :: dsdx, J, nHatb, nScaleb, stream
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_destroy(model) BIND(C, name="selfamd_model2d_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_set_param(model, which, value) BIND(C, name="selfamd_model2d_set_param")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: model
      INTEGER(C_INT), VALUE :: which
      REAL(C_DOUBLE), VALUE :: value
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_set_field(model, which, field) BIND(C, name="selfamd_model2d_set_field")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model, field
      INTEGER(C_INT), VALUE :: which
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_set_solution(model, interior) &
        BIND(C, name="selfamd_model2d_set_solution")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model, interior
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_get_solution(model, interior) &
        BIND(C, name="selfamd_model2d_get_solution")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model, interior
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_tendency(model, dSdt) BIND(C, name="selfamd_model2d_tendency")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model, dSdt
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_entropy(model, entropy) BIND(C, name="selfamd_model2d_entropy")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: model
      REAL(C_DOUBLE) :: entropy
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_integral(model, var, val) BIND(C, name="selfamd_model2d_integral")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: model
      INTEGER(C_INT), VALUE :: var
      REAL(C_DOUBLE) :: val
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_forward_step(model, nSteps, dt) &
        BIND(C, name="selfamd_model2d_forward_step")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: model
      INTEGER(C_INT), VALUE :: nSteps
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_flux_registers(model, fluxDiv, flux) &
        BIND(C, name="selfamd_model2d_flux_registers")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model, fluxDiv, flux
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_synchronize(model) BIND(C, name="selfamd_model2d_synchronize")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: model
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_model2d_set_time(model, t) BIND(C, name="selfamd_model2d_set_time")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: model
      REAL(C_DOUBLE), VALUE :: t
    END FUNCTION

    ! ---- register updates (the per-operator path of CalculateTendency / LowStorageRK3) ----
    INTEGER(C_INT) FUNCTION selfamd_update_grk(workSol, S, dSdt, a, g, dt, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_update_grk")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: workSol, S, dSdt, stream
      REAL(C_DOUBLE), VALUE :: a, g, dt
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_calculate_dsdt(fluxDiv, source, dSdt, N, nVar, nEl, stream) &
        BIND(C, name="selfamd_calculate_dsdt")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: fluxDiv, source, dSdt, stream
      INTEGER(C_INT), VALUE :: N, nVar, nEl
    END FUNCTION

    ! ---- physics operators, one per plug-in method ----
    INTEGER(C_INT) FUNCTION selfamd_sw2d_velocity(sol, solB, vel, velB, N, nEl, stream) &
        BIND(C, name="selfamd_sw2d_velocity")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: sol, solB, vel, velB, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_sw2d_set_boundary_condition(solB, solE, topoB, nHatb, sideInfo, N, nEl, stream) &
        BIND(C, name="selfamd_sw2d_set_boundary_condition")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: solB, solE, topoB, nHatb, sideInfo, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_sw2d_source(source, sol, gradTopo, g, fCori, N, nEl, stream) &
        BIND(C, name="selfamd_sw2d_source")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: source, sol, gradTopo, stream
      REAL(C_DOUBLE), VALUE :: g, fCori
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_sw2d_flux(flux, sol, vel, g, N, nEl, stream) BIND(C, name="selfamd_sw2d_flux")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: flux, sol, vel, stream
      REAL(C_DOUBLE), VALUE :: g
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_sw2d_riemann(fluxBN, solB, solE, velB, velE, nHatb, nScaleb, g, N, nEl, stream) &
        BIND(C, name="selfamd_sw2d_riemann")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: fluxBN, solB, solE, velB, velE, nHatb, nScaleb, stream
      REAL(C_DOUBLE), VALUE :: g
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_lsw2d_set_boundary_condition(solB, solE, nHatb, sideInfo, N, nEl, stream) &
        BIND(C, name="selfamd_lsw2d_set_boundary_condition")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: solB, solE, nHatb, sideInfo, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_lsw2d_source(source, sol, fCori, N, nEl, stream) BIND(C, name="selfamd_lsw2d_source")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: source, sol, fCori, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_lsw2d_flux(flux, sol, H, g, N, nEl, stream) BIND(C, name="selfamd_lsw2d_flux")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: flux, sol, H, stream
      REAL(C_DOUBLE), VALUE :: g
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_lsw2d_riemann(fluxBN, solB, solE, Hb, nHatb, nScaleb, g, N, nEl, stream) &
        BIND(C, name="selfamd_lsw2d_riemann")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: fluxBN, solB, solE, Hb, nHatb, nScaleb, stream
      REAL(C_DOUBLE), VALUE :: g
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_diagnostics(sol, diag, gamma, R, N, nEl, stream) &
        BIND(C, name="selfamd_euler2d_diagnostics")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: sol, diag, stream
      REAL(C_DOUBLE), VALUE :: gamma, R
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_primitive(sol, prim, gamma, N, nEl, stream) &
        BIND(C, name="selfamd_euler2d_primitive")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: sol, prim, stream
      REAL(C_DOUBLE), VALUE :: gamma
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_entropy_vars(sol, ent, gamma, N, nEl, stream) &
        BIND(C, name="selfamd_euler2d_entropy_vars")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: sol, ent, stream
      REAL(C_DOUBLE), VALUE :: gamma
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_set_boundary_condition(solB, solE, presSolB, primB, primE, presPrimB, &
        diagB, diagE, presDiagB, nHatb, sideInfo, N, nEl, stream) BIND(C, name="selfamd_euler2d_set_boundary_condition")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: solB, solE, presSolB, primB, primE, presPrimB, diagB, diagE, presDiagB, nHatb, sideInfo
      TYPE(C_PTR), VALUE :: stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_source(source, sol, env, envGrad, N, nEl, stream) &
        BIND(C, name="selfamd_euler2d_source")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: source, sol, env, envGrad, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_flux(flux, sol, prim, diag, N, nEl, stream) &
        BIND(C, name="selfamd_euler2d_flux")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: flux, sol, prim, diag, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_euler2d_riemann(fluxBN, solB, solE, primB, primE, diagB, diagE, nHatb, nScaleb, &
        N, nEl, stream) BIND(C, name="selfamd_euler2d_riemann")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: fluxBN, solB, solE, primB, primE, diagB, diagE, nHatb, nScaleb, stream
      INTEGER(C_INT), VALUE :: N, nEl
    END FUNCTION
  END INTERFACE

CONTAINS

  SUBROUTINE selfamd_check(rc, what)
    !! Stop with the library's error message when a call failed
    INTEGER(C_INT), INTENT(in) :: rc
    CHARACTER(*), INTENT(in) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    CHARACTER(LEN=512) :: text
    INTEGER :: k
    IF (rc == 0) RETURN
    CALL C_F_POINTER(selfamd_last_error(), msg, [512])
    text = ''
    DO k = 1, 512
      IF (msg(k) == C_NULL_CHAR) EXIT
      text(k:k) = msg(k)
    END DO
    PRINT '(A,A,A,I0,A,A)', 'self_amd: ', what, ' failed (status ', rc, '): ', TRIM(text)
    ERROR STOP 3
  END SUBROUTINE selfamd_check

  SUBROUTINE hip_check(rc, what)
    !! Device plumbing calls report through the same channel as the other entry points
    INTEGER(C_INT), INTENT(in) :: rc
    CHARACTER(*), INTENT(in) :: what
    CALL selfamd_check(rc, what)
  END SUBROUTINE hip_check

  FUNCTION device_alloc(nbytes) RESULT(p)
    !! Device allocation, zero filled
    INTEGER(C_SIZE_T), INTENT(in) :: nbytes
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    CALL selfamd_check(selfamd_malloc(p, MAX(nbytes, 8_C_SIZE_T)), 'selfamd_malloc')
    CALL selfamd_check(selfamd_memset(p, 0_C_INT, MAX(nbytes, 8_C_SIZE_T)), 'selfamd_memset')
  END FUNCTION device_alloc

  SUBROUTINE device_free(p)
    TYPE(C_PTR), INTENT(inout) :: p
    IF (C_ASSOCIATED(p)) CALL selfamd_check(selfamd_free(p), 'selfamd_free')
    p = C_NULL_PTR
  END SUBROUTINE device_free

END MODULE SELF_AMD_CAPI
