! SELF_MappedData.f90 -- MappedScalar2D / MappedVector2D (src/SELF_MappedData.f90) for the
! self_amd engine: device-resident interior/boundary/extBoundary arrays in the reference
! shapes with the hot-path methods running the library's HIP kernels.  Host mirrors are
! filled only by UpdateHost (IO / comparison), never implicitly.
MODULE SELF_MappedData
  USE SELF_Constants
  USE SELF_AMD_CAPI
  USE SELF_Lagrange
  USE SELF_Mesh
  IMPLICIT NONE

  TYPE, PUBLIC :: MappedScalar2D
    INTEGER :: nVar = 0, nElem = 0
    TYPE(Lagrange), POINTER :: interp => NULL()
    REAL(prec), ALLOCATABLE :: interior(:,:,:,:)    ! host mirror (0:N,0:N,nVar,nElem)
    REAL(prec), ALLOCATABLE :: boundary(:,:,:,:)    ! host mirror (0:N,nVar,4,nElem)
    TYPE(C_PTR) :: interior_dev = C_NULL_PTR, boundary_dev = C_NULL_PTR, extBoundary_dev = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_MappedScalar2D
    PROCEDURE, PUBLIC :: Free => Free_MappedScalar2D
    PROCEDURE, PUBLIC :: UpdateHost => UpdateHost_MappedScalar2D
    PROCEDURE, PUBLIC :: UpdateDevice => UpdateDevice_MappedScalar2D
    PROCEDURE, PUBLIC :: BoundaryInterp => BoundaryInterp_MappedScalar2D
    PROCEDURE, PUBLIC :: SideExchange => SideExchange_MappedScalar2D
    PROCEDURE, PUBLIC :: JacobianWeight => JacobianWeight_MappedScalar2D
  END TYPE MappedScalar2D

  TYPE, PUBLIC :: MappedVector2D
    INTEGER :: nVar = 0, nElem = 0
    TYPE(Lagrange), POINTER :: interp => NULL()
    REAL(prec), ALLOCATABLE :: interior(:,:,:,:,:)  ! host mirror (2,0:N,0:N,nVar,nElem)
    TYPE(C_PTR) :: interior_dev = C_NULL_PTR, boundary_dev = C_NULL_PTR, extBoundary_dev = C_NULL_PTR
    TYPE(C_PTR) :: boundaryNormal_dev = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_MappedVector2D
    PROCEDURE, PUBLIC :: Free => Free_MappedVector2D
    PROCEDURE, PUBLIC :: BoundaryInterp => BoundaryInterp_MappedVector2D
    PROCEDURE, PUBLIC :: SideExchange => SideExchange_MappedVector2D
    PROCEDURE, PUBLIC :: ContravariantProjection => ContravariantProjection_MappedVector2D
    PROCEDURE, PUBLIC :: Divergence => Divergence_MappedVector2D
  END TYPE MappedVector2D

CONTAINS

  SUBROUTINE Init_MappedScalar2D(this, interp, nVar, nElem)
    CLASS(MappedScalar2D), INTENT(out) :: this
    TYPE(Lagrange), TARGET, INTENT(in) :: interp
    INTEGER, INTENT(in) :: nVar, nElem
    INTEGER(C_SIZE_T) :: P
    this % interp => interp
    this % nVar = nVar
    this % nElem = nElem
    P = INT(interp % N + 1, C_SIZE_T)
    ALLOCATE (this % interior(0:interp % N, 0:interp % N, nVar, nElem), this % boundary(0:interp % N, nVar, 4, nElem))
    this % interior = 0.0_prec
    this % boundary = 0.0_prec
    this % interior_dev = device_alloc(P*P*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
    this % boundary_dev = device_alloc(P*4_C_SIZE_T*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
    this % extBoundary_dev = device_alloc(P*4_C_SIZE_T*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
  END SUBROUTINE Init_MappedScalar2D

  SUBROUTINE Free_MappedScalar2D(this)
    CLASS(MappedScalar2D), INTENT(inout) :: this
    CALL device_free(this % interior_dev)
    CALL device_free(this % boundary_dev)
    CALL device_free(this % extBoundary_dev)
    IF (ALLOCATED(this % interior)) DEALLOCATE (this % interior, this % boundary)
  END SUBROUTINE Free_MappedScalar2D

  SUBROUTINE UpdateHost_MappedScalar2D(this)
    CLASS(MappedScalar2D), INTENT(inout), TARGET :: this
    CALL hip_check(selfamd_memcpy(C_LOC(this % interior), this % interior_dev, &
                   INT(SIZE(this % interior), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), 'selfamd_memcpy')
    CALL hip_check(selfamd_memcpy(C_LOC(this % boundary), this % boundary_dev, &
                   INT(SIZE(this % boundary), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), 'selfamd_memcpy')
  END SUBROUTINE UpdateHost_MappedScalar2D

  SUBROUTINE UpdateDevice_MappedScalar2D(this)
    CLASS(MappedScalar2D), INTENT(inout), TARGET :: this
    CALL hip_check(selfamd_memcpy(this % interior_dev, C_LOC(this % interior), &
                   INT(SIZE(this % interior), C_SIZE_T)*8_C_SIZE_T, hipMemcpyHostToDevice), 'selfamd_memcpy')
  END SUBROUTINE UpdateDevice_MappedScalar2D

  ! BoundaryInterp_Scalar2D (src/SELF_Data.f90:630-647), device branch
  SUBROUTINE BoundaryInterp_MappedScalar2D(this)
    CLASS(MappedScalar2D), INTENT(inout) :: this
    CALL this % interp % ScalarBoundaryInterp_2D(this % interior_dev, this % boundary_dev, this % nVar, this % nElem)
  END SUBROUTINE BoundaryInterp_MappedScalar2D

  ! SideExchange_MappedScalar2D (src/SELF_MappedData.f90:834-917), single GPU: no host round trip
  SUBROUTINE SideExchange_MappedScalar2D(this, mesh)
    CLASS(MappedScalar2D), INTENT(inout) :: this
    TYPE(Mesh2D), INTENT(in) :: mesh
    CALL selfamd_check(selfamd_side_exchange_scalar_2d(this % boundary_dev, this % extBoundary_dev, &
                       mesh % sideInfo_dev, INT(this % interp % N, C_INT), INT(this % nVar, C_INT), &
                       INT(this % nElem, C_INT), C_NULL_PTR), 'MappedScalar2D%SideExchange')
  END SUBROUTINE SideExchange_MappedScalar2D

  ! JacobianWeight_MappedScalar2D (src/SELF_MappedData.f90:1235-1268)
  SUBROUTINE JacobianWeight_MappedScalar2D(this, geometry)
    CLASS(MappedScalar2D), INTENT(inout) :: this
    TYPE(SEMQuad), INTENT(in) :: geometry
    CALL selfamd_check(selfamd_jacobian_weight_2d(geometry % J, this % interior_dev, INT(this % interp % N, C_INT), &
                       INT(this % nVar, C_INT), INT(this % nElem, C_INT), C_NULL_PTR), 'JacobianWeight')
  END SUBROUTINE JacobianWeight_MappedScalar2D

  SUBROUTINE Init_MappedVector2D(this, interp, nVar, nElem)
    CLASS(MappedVector2D), INTENT(out) :: this
    TYPE(Lagrange), TARGET, INTENT(in) :: interp
    INTEGER, INTENT(in) :: nVar, nElem
    INTEGER(C_SIZE_T) :: P
    this % interp => interp
    this % nVar = nVar
    this % nElem = nElem
    P = INT(interp % N + 1, C_SIZE_T)
    ALLOCATE (this % interior(2, 0:interp % N, 0:interp % N, nVar, nElem))
    this % interior = 0.0_prec
    this % interior_dev = device_alloc(2_C_SIZE_T*P*P*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
    this % boundary_dev = device_alloc(2_C_SIZE_T*P*4_C_SIZE_T*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
    this % extBoundary_dev = device_alloc(2_C_SIZE_T*P*4_C_SIZE_T*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
    this % boundaryNormal_dev = device_alloc(P*4_C_SIZE_T*INT(nVar*nElem, C_SIZE_T)*8_C_SIZE_T)
  END SUBROUTINE Init_MappedVector2D

  SUBROUTINE Free_MappedVector2D(this)
    CLASS(MappedVector2D), INTENT(inout) :: this
    CALL device_free(this % interior_dev)
    CALL device_free(this % boundary_dev)
    CALL device_free(this % extBoundary_dev)
    CALL device_free(this % boundaryNormal_dev)
    IF (ALLOCATED(this % interior)) DEALLOCATE (this % interior)
  END SUBROUTINE Free_MappedVector2D

  SUBROUTINE BoundaryInterp_MappedVector2D(this)
    CLASS(MappedVector2D), INTENT(inout) :: this
    CALL this % interp % VectorBoundaryInterp_2D(this % interior_dev, this % boundary_dev, this % nVar, this % nElem)
  END SUBROUTINE BoundaryInterp_MappedVector2D

  ! SideExchange_MappedVector2D (src/SELF_MappedData.f90:1870-1958)
  SUBROUTINE SideExchange_MappedVector2D(this, mesh)
    CLASS(MappedVector2D), INTENT(inout) :: this
    TYPE(Mesh2D), INTENT(in) :: mesh
    CALL selfamd_check(selfamd_side_exchange_vector_2d(this % boundary_dev, this % extBoundary_dev, &
                       mesh % sideInfo_dev, INT(this % interp % N, C_INT), INT(this % nVar, C_INT), &
                       INT(this % nElem, C_INT), C_NULL_PTR), 'MappedVector2D%SideExchange')
  END SUBROUTINE SideExchange_MappedVector2D

  ! ContravariantProjection_MappedVector2D (src/SELF_MappedData.f90:2220-2269)
  SUBROUTINE ContravariantProjection_MappedVector2D(this, geometry)
    CLASS(MappedVector2D), INTENT(inout) :: this
    TYPE(SEMQuad), INTENT(in) :: geometry
    CALL selfamd_check(selfamd_contravariant_projection_2d(geometry % dsdx, this % interior_dev, &
                       INT(this % interp % N, C_INT), INT(this % nVar, C_INT), INT(this % nElem, C_INT), &
                       C_NULL_PTR), 'ContravariantProjection')
  END SUBROUTINE ContravariantProjection_MappedVector2D

  ! Divergence_MappedVector2D, weak DG form (src/SELF_MappedData.f90:1996-2044)
  SUBROUTINE Divergence_MappedVector2D(this, geometry, divVector)
    CLASS(MappedVector2D), INTENT(in) :: this
    TYPE(SEMQuad), INTENT(in) :: geometry
    TYPE(MappedScalar2D), INTENT(inout) :: divVector
    CALL this % interp % VectorDGDivergence_2D(this % interior_dev, this % boundaryNormal_dev, &
                                               divVector % interior_dev, this % nVar, this % nElem)
    CALL divVector % JacobianWeight(geometry)
  END SUBROUTINE Divergence_MappedVector2D

END MODULE SELF_MappedData
