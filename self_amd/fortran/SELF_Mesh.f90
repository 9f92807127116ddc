! SELF_Mesh.f90 -- Mesh2D (src/SELF_Mesh.f90:228-254) and SEMQuad (src/SELF_Geometry.f90:38-61)
! for the self_amd engine.
!
! Mesh2D keeps the reference connectivity convention sideInfo(1:5,1:4,nElem) (side type,
! global side id, neighbour element, 10*neighbour side + flip, boundary condition id).
! StructuredMesh has no reference counterpart (the reference reads HOPr files only); it emits
! the same convention.  SEMQuad%GenerateFromMesh computes the metric terms on the device.
MODULE SELF_Mesh
  USE SELF_Constants
  USE SELF_AMD_CAPI
  USE SELF_Lagrange
  IMPLICIT NONE

  TYPE, PUBLIC :: Mesh2D
    INTEGER :: nGeo = 1, nElem = 0, quadrature = UNIFORM
    INTEGER(C_INT32_T), ALLOCATABLE :: sideInfo(:,:,:)        ! (5,4,nElem)
    REAL(prec), ALLOCATABLE :: nodeCoords(:,:,:,:)            ! (2,0:nGeo,0:nGeo,nElem)
    INTEGER(C_INT32_T), ALLOCATABLE :: globalNodeIDs(:,:,:)   ! (0:nGeo,0:nGeo,nElem)
    TYPE(C_PTR) :: sideInfo_dev = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: StructuredMesh => StructuredMesh_Mesh2D
    PROCEDURE, PUBLIC :: ResetBoundaryConditionType => ResetBoundaryConditionType_Mesh2D
    PROCEDURE, PUBLIC :: RecalculateFlip => RecalculateFlip_Mesh2D
    PROCEDURE, PUBLIC :: UpdateDevice => UpdateDevice_Mesh2D
    PROCEDURE, PUBLIC :: Free => Free_Mesh2D
  END TYPE Mesh2D

  TYPE, PUBLIC :: SEMQuad
    INTEGER :: nElem = 0
    TYPE(Lagrange), POINTER :: interp => NULL()
    ! device arrays in the reference shapes (interior / boundary)
    TYPE(C_PTR) :: x = C_NULL_PTR, xb = C_NULL_PTR, dxds = C_NULL_PTR, dxdsb = C_NULL_PTR
    TYPE(C_PTR) :: dsdx = C_NULL_PTR, dsdxb = C_NULL_PTR, nHatb = C_NULL_PTR, nScaleb = C_NULL_PTR
    TYPE(C_PTR) :: J = C_NULL_PTR, Jb = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_SEMQuad
    PROCEDURE, PUBLIC :: GenerateFromMesh => GenerateFromMesh_SEMQuad
    PROCEDURE, PUBLIC :: CovariantArcMin => CovariantArcMin_SEMQuad
    PROCEDURE, PUBLIC :: CopyX => CopyX_SEMQuad
    PROCEDURE, PUBLIC :: Free => Free_SEMQuad
  END TYPE SEMQuad

CONTAINS

  SUBROUTINE StructuredMesh_Mesh2D(myMesh, nx, ny, x0, x1, y0, y1, periodicX, periodicY, bcid)
    CLASS(Mesh2D), INTENT(out) :: myMesh
    INTEGER, INTENT(in) :: nx, ny, bcid
    REAL(prec), INTENT(in) :: x0, x1, y0, y1
    LOGICAL, INTENT(in) :: periodicX, periodicY
    INTEGER(C_INT) :: px, py
    myMesh % nGeo = 1
    myMesh % nElem = nx*ny
    myMesh % quadrature = UNIFORM
    ALLOCATE (myMesh % sideInfo(5, 4, nx*ny), myMesh % nodeCoords(2, 0:1, 0:1, nx*ny), &
              myMesh % globalNodeIDs(0:1, 0:1, nx*ny))
    px = MERGE(1, 0, periodicX)
    py = MERGE(1, 0, periodicY)
    CALL selfamd_check(selfamd_mesh2d_structured(INT(nx, C_INT), INT(ny, C_INT), x0, x1, y0, y1, px, py, &
                       INT(bcid, C_INT), myMesh % sideInfo, myMesh % nodeCoords, myMesh % globalNodeIDs), &
                       'Mesh2D%StructuredMesh')
    CALL myMesh % UpdateDevice()
  END SUBROUTINE StructuredMesh_Mesh2D

  ! ResetBoundaryConditionType_Mesh2D (src/SELF_Mesh.f90:695-719)
  SUBROUTINE ResetBoundaryConditionType_Mesh2D(myMesh, bcid)
    CLASS(Mesh2D), INTENT(inout) :: myMesh
    INTEGER, INTENT(in) :: bcid
    CALL selfamd_check(selfamd_mesh2d_reset_boundary_condition_type(INT(myMesh % nElem, C_INT), myMesh % sideInfo, &
                       INT(bcid, C_INT)), 'Mesh2D%ResetBoundaryConditionType')
    CALL myMesh % UpdateDevice()
  END SUBROUTINE ResetBoundaryConditionType_Mesh2D

  ! RecalculateFlip_Mesh2D (src/SELF_Mesh.f90:1018-1174)
  SUBROUTINE RecalculateFlip_Mesh2D(myMesh)
    CLASS(Mesh2D), INTENT(inout) :: myMesh
    CALL selfamd_check(selfamd_mesh2d_recalculate_flip(INT(myMesh % nElem, C_INT), INT(myMesh % nGeo, C_INT), &
                       myMesh % sideInfo, myMesh % globalNodeIDs), 'Mesh2D%RecalculateFlip')
    CALL myMesh % UpdateDevice()
  END SUBROUTINE RecalculateFlip_Mesh2D

  SUBROUTINE UpdateDevice_Mesh2D(myMesh)
    CLASS(Mesh2D), INTENT(inout), TARGET :: myMesh
    INTEGER(C_SIZE_T) :: nb
    nb = INT(SIZE(myMesh % sideInfo), C_SIZE_T)*4_C_SIZE_T
    IF (.NOT. C_ASSOCIATED(myMesh % sideInfo_dev)) myMesh % sideInfo_dev = device_alloc(nb)
    CALL hip_check(selfamd_memcpy(myMesh % sideInfo_dev, C_LOC(myMesh % sideInfo), nb, hipMemcpyHostToDevice), 'selfamd_memcpy')
  END SUBROUTINE UpdateDevice_Mesh2D

  SUBROUTINE Free_Mesh2D(myMesh)
    CLASS(Mesh2D), INTENT(inout) :: myMesh
    CALL device_free(myMesh % sideInfo_dev)
    IF (ALLOCATED(myMesh % sideInfo)) DEALLOCATE (myMesh % sideInfo, myMesh % nodeCoords, myMesh % globalNodeIDs)
  END SUBROUTINE Free_Mesh2D

  ! Init_SEMQuad (src/SELF_Geometry.f90:320-352): device storage for the metric terms
  SUBROUTINE Init_SEMQuad(myGeom, interp, nElem)
    CLASS(SEMQuad), INTENT(out) :: myGeom
    TYPE(Lagrange), POINTER, INTENT(in) :: interp
    INTEGER, INTENT(in) :: nElem
    INTEGER(C_SIZE_T) :: nn, nb
    myGeom % interp => interp
    myGeom % nElem = nElem
    nn = INT(interp % N + 1, C_SIZE_T)**2*INT(nElem, C_SIZE_T)*8_C_SIZE_T
    nb = INT(interp % N + 1, C_SIZE_T)*4_C_SIZE_T*INT(nElem, C_SIZE_T)*8_C_SIZE_T
    myGeom % x = device_alloc(2*nn)
    myGeom % xb = device_alloc(2*nb)
    myGeom % dxds = device_alloc(4*nn)
    myGeom % dxdsb = device_alloc(4*nb)
    myGeom % dsdx = device_alloc(4*nn)
    myGeom % dsdxb = device_alloc(4*nb)
    myGeom % nHatb = device_alloc(2*nb)
    myGeom % nScaleb = device_alloc(nb)
    myGeom % J = device_alloc(nn)
    myGeom % Jb = device_alloc(nb)
  END SUBROUTINE Init_SEMQuad

  ! GenerateFromMesh_SEMQuad + CalculateMetricTerms_SEMQuad (src/SELF_Geometry.f90:393-552), on the device
  SUBROUTINE GenerateFromMesh_SEMQuad(myGeom, mesh)
    CLASS(SEMQuad), INTENT(inout) :: myGeom
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(C_PTR) :: nc
    INTEGER(C_SIZE_T) :: nb
    nb = INT(SIZE(mesh % nodeCoords), C_SIZE_T)*8_C_SIZE_T
    nc = device_alloc(nb)
    CALL hip_check(selfamd_memcpy(nc, C_LOC(mesh % nodeCoords), nb, hipMemcpyHostToDevice), 'selfamd_memcpy')
    CALL selfamd_check(selfamd_semquad_generate(INT(myGeom % interp % N, C_INT), &
                       INT(myGeom % interp % controlNodeType, C_INT), INT(mesh % nGeo, C_INT), &
                       INT(mesh % nElem, C_INT), nc, myGeom % x, myGeom % xb, myGeom % dxds, myGeom % dxdsb, &
                       myGeom % dsdx, myGeom % dsdxb, myGeom % nHatb, myGeom % nScaleb, myGeom % J, myGeom % Jb, &
                       C_NULL_PTR), 'SEMQuad%GenerateFromMesh')
    CALL device_free(nc)
  END SUBROUTINE GenerateFromMesh_SEMQuad

  ! Physical node positions x(2,0:N,0:N,nElem) on the host
  SUBROUTINE CopyX_SEMQuad(myGeom, x)
    CLASS(SEMQuad), INTENT(in) :: myGeom
    REAL(prec), INTENT(out), TARGET :: x(:,:,:,:)
    CALL hip_check(selfamd_memcpy(C_LOC(x), myGeom % x, INT(SIZE(x), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), &
                   'selfamd_memcpy')
  END SUBROUTINE CopyX_SEMQuad

  ! CovariantArcMin_SEMQuad (src/SELF_Geometry.f90:554-582)
  FUNCTION CovariantArcMin_SEMQuad(myGeom) RESULT(dxMin)
    CLASS(SEMQuad), INTENT(in) :: myGeom
    REAL(prec) :: dxMin
    REAL(prec), ALLOCATABLE, TARGET :: dxds(:,:,:,:,:)
    REAL(prec) :: dx, dy
    INTEGER :: i, j, e, N
    N = myGeom % interp % N
    ALLOCATE (dxds(2, 2, 0:N, 0:N, myGeom % nElem))
    CALL hip_check(selfamd_memcpy(C_LOC(dxds), myGeom % dxds, INT(SIZE(dxds), C_SIZE_T)*8_C_SIZE_T, &
                             hipMemcpyDeviceToHost), 'selfamd_memcpy')
    dxMin = HUGE(1.0_prec)
    DO e = 1, myGeom % nElem
      DO j = 0, N
        DO i = 0, N
          dx = SQRT(dxds(1, 1, i, j, e)**2 + dxds(1, 2, i, j, e)**2)
          dy = SQRT(dxds(2, 1, i, j, e)**2 + dxds(2, 2, i, j, e)**2)
          dxMin = MIN(dxMin, 2.0_prec*MIN(dx, dy)/(REAL(N, prec)**2))
        END DO
      END DO
    END DO
  END FUNCTION CovariantArcMin_SEMQuad

  SUBROUTINE Free_SEMQuad(myGeom)
    CLASS(SEMQuad), INTENT(inout) :: myGeom
    CALL device_free(myGeom % x); CALL device_free(myGeom % xb)
    CALL device_free(myGeom % dxds); CALL device_free(myGeom % dxdsb)
    CALL device_free(myGeom % dsdx); CALL device_free(myGeom % dsdxb)
    CALL device_free(myGeom % nHatb); CALL device_free(myGeom % nScaleb)
    CALL device_free(myGeom % J); CALL device_free(myGeom % Jb)
  END SUBROUTINE Free_SEMQuad

END MODULE SELF_Mesh
