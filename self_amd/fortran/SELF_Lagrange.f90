! SELF_Lagrange.f90 -- Fortran host mirror of the reference's SELF_Constants / SELF_Lagrange
! modules (src/SELF_Constants.f90, src/SELF_Lagrange.f90:25-192) over libself_amd.
!
! Lagrange%Init (src/SELF_Lagrange.f90:521-615) fills the host matrices through the C ABI
! (selfamd_lagrange_init) and keeps device copies; the 2D interpolation/derivative methods
! run the library's HIP kernels on device arrays (TYPE(c_ptr)).
MODULE SELF_Constants
  USE ISO_C_BINDING
  USE ISO_FORTRAN_ENV
  IMPLICIT NONE
  INTEGER, PARAMETER :: prec = real64
  INTEGER, PARAMETER :: c_prec = C_DOUBLE
  REAL(prec), PARAMETER :: pi = 4.0_prec*ATAN(1.0_prec)
  REAL(prec), PARAMETER :: TOL = EPSILON(1.0_prec)
  INTEGER, PARAMETER :: GAUSS = 1
  INTEGER, PARAMETER :: GAUSS_LOBATTO = 2
  INTEGER, PARAMETER :: UNIFORM = 5
  ! boundary condition ids (src/SELF_Model.f90:70-72)
  INTEGER, PARAMETER :: SELF_BC_PRESCRIBED = 100
  INTEGER, PARAMETER :: SELF_BC_RADIATION = 101
  INTEGER, PARAMETER :: SELF_BC_NONORMALFLOW = 102
END MODULE SELF_Constants

MODULE SELF_Lagrange
  USE SELF_Constants
  USE SELF_AMD_CAPI
  IMPLICIT NONE

  TYPE, PUBLIC :: Lagrange
    INTEGER :: N = 0, M = 0
    INTEGER :: controlNodeType = GAUSS, targetNodeType = UNIFORM
    REAL(prec), ALLOCATABLE :: controlPoints(:), targetPoints(:), bWeights(:), qWeights(:)
    REAL(prec), ALLOCATABLE :: iMatrix(:,:), dMatrix(:,:), dgMatrix(:,:), bMatrix(:,:)
    TYPE(C_PTR) :: iMatrix_dev = C_NULL_PTR, dMatrix_dev = C_NULL_PTR, dgMatrix_dev = C_NULL_PTR
    TYPE(C_PTR) :: bMatrix_dev = C_NULL_PTR, qWeights_dev = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_Lagrange
    PROCEDURE, PUBLIC :: Free => Free_Lagrange
    PROCEDURE, PUBLIC :: ScalarBoundaryInterp_2D
    PROCEDURE, PUBLIC :: VectorBoundaryInterp_2D
    PROCEDURE, PUBLIC :: ScalarGridInterp_2D
    PROCEDURE, PUBLIC :: VectorDGDivergence_2D
  END TYPE Lagrange

CONTAINS

  SUBROUTINE Init_Lagrange(myPoly, N, controlNodeType, M, targetNodeType)
    CLASS(Lagrange), INTENT(out) :: myPoly
    INTEGER, INTENT(in) :: N, controlNodeType, M, targetNodeType
    myPoly % N = N
    myPoly % M = M
    myPoly % controlNodeType = controlNodeType
    myPoly % targetNodeType = targetNodeType
    ALLOCATE (myPoly % controlPoints(0:N), myPoly % targetPoints(0:M), myPoly % bWeights(0:N), &
              myPoly % qWeights(0:N), myPoly % iMatrix(0:N, 0:M), myPoly % dMatrix(0:N, 0:N), &
              myPoly % dgMatrix(0:N, 0:N), myPoly % bMatrix(0:N, 0:1))
    CALL selfamd_check(selfamd_lagrange_init(INT(N, C_INT), INT(controlNodeType, C_INT), INT(M, C_INT), &
                                             INT(targetNodeType, C_INT), myPoly % controlPoints, &
                                             myPoly % targetPoints, myPoly % bWeights, myPoly % qWeights, &
                                             myPoly % iMatrix, myPoly % dMatrix, myPoly % dgMatrix, &
                                             myPoly % bMatrix), 'Lagrange%Init')
    CALL to_device(myPoly % iMatrix, myPoly % iMatrix_dev)
    CALL to_device(myPoly % dMatrix, myPoly % dMatrix_dev)
    CALL to_device(myPoly % dgMatrix, myPoly % dgMatrix_dev)
    CALL to_device(myPoly % bMatrix, myPoly % bMatrix_dev)
    CALL to_device1(myPoly % qWeights, myPoly % qWeights_dev)
  CONTAINS
    SUBROUTINE to_device(a, p)
      REAL(prec), INTENT(in), TARGET :: a(:,:)
      TYPE(C_PTR), INTENT(out) :: p
      INTEGER(C_SIZE_T) :: nb
      nb = INT(SIZE(a), C_SIZE_T)*8_C_SIZE_T
      p = device_alloc(nb)
      CALL hip_check(selfamd_memcpy(p, C_LOC(a), nb, hipMemcpyHostToDevice), 'selfamd_memcpy')
    END SUBROUTINE to_device
    SUBROUTINE to_device1(a, p)
      REAL(prec), INTENT(in), TARGET :: a(:)
      TYPE(C_PTR), INTENT(out) :: p
      INTEGER(C_SIZE_T) :: nb
      nb = INT(SIZE(a), C_SIZE_T)*8_C_SIZE_T
      p = device_alloc(nb)
      CALL hip_check(selfamd_memcpy(p, C_LOC(a), nb, hipMemcpyHostToDevice), 'selfamd_memcpy')
    END SUBROUTINE to_device1
  END SUBROUTINE Init_Lagrange

  SUBROUTINE Free_Lagrange(myPoly)
    CLASS(Lagrange), INTENT(inout) :: myPoly
    CALL device_free(myPoly % iMatrix_dev)
    CALL device_free(myPoly % dMatrix_dev)
    CALL device_free(myPoly % dgMatrix_dev)
    CALL device_free(myPoly % bMatrix_dev)
    CALL device_free(myPoly % qWeights_dev)
    IF (ALLOCATED(myPoly % controlPoints)) DEALLOCATE (myPoly % controlPoints, myPoly % targetPoints, &
        myPoly % bWeights, myPoly % qWeights, myPoly % iMatrix, myPoly % dMatrix, myPoly % dgMatrix, myPoly % bMatrix)
  END SUBROUTINE Free_Lagrange

  ! ScalarBoundaryInterp_2D_gpu (src/SELF_Lagrange.f90:2272-2282)
  SUBROUTINE ScalarBoundaryInterp_2D(myPoly, f, fBound, nVariables, nElements)
    CLASS(Lagrange), INTENT(in) :: myPoly
    TYPE(C_PTR), INTENT(in) :: f, fBound
    INTEGER, INTENT(in) :: nVariables, nElements
    CALL selfamd_check(selfamd_scalar_boundary_interp_2d(myPoly % bMatrix_dev, f, fBound, INT(myPoly % N, C_INT), &
                       INT(nVariables, C_INT), INT(nElements, C_INT), C_NULL_PTR), 'ScalarBoundaryInterp_2D')
  END SUBROUTINE ScalarBoundaryInterp_2D

  ! VectorBoundaryInterp_2D_gpu (src/SELF_Lagrange.f90:2318-2328)
  SUBROUTINE VectorBoundaryInterp_2D(myPoly, f, fBound, nVariables, nElements)
    CLASS(Lagrange), INTENT(in) :: myPoly
    TYPE(C_PTR), INTENT(in) :: f, fBound
    INTEGER, INTENT(in) :: nVariables, nElements
    CALL selfamd_check(selfamd_vector_boundary_interp_2d(myPoly % bMatrix_dev, f, fBound, INT(myPoly % N, C_INT), &
                       INT(nVariables, C_INT), INT(nElements, C_INT), C_NULL_PTR), 'VectorBoundaryInterp_2D')
  END SUBROUTINE VectorBoundaryInterp_2D

  ! ScalarGridInterp_2D_gpu (src/SELF_Lagrange.f90:791-822)
  SUBROUTINE ScalarGridInterp_2D(myPoly, f, fNew, nVariables, nElements)
    CLASS(Lagrange), INTENT(in) :: myPoly
    TYPE(C_PTR), INTENT(in) :: f, fNew
    INTEGER, INTENT(in) :: nVariables, nElements
    CALL selfamd_check(selfamd_scalar_grid_interp_2d(myPoly % iMatrix_dev, f, fNew, INT(myPoly % N, C_INT), &
                       INT(myPoly % M, C_INT), INT(nVariables, C_INT), INT(nElements, C_INT), C_NULL_PTR), &
                       'ScalarGridInterp_2D')
  END SUBROUTINE ScalarGridInterp_2D

  ! VectorDGDivergence_2D_gpu (src/SELF_Lagrange.f90:1627-1640)
  SUBROUTINE VectorDGDivergence_2D(myPoly, f, bF, dF, nVariables, nElements)
    CLASS(Lagrange), INTENT(in) :: myPoly
    TYPE(C_PTR), INTENT(in) :: f, bF, dF
    INTEGER, INTENT(in) :: nVariables, nElements
    CALL selfamd_check(selfamd_vector_dg_divergence_2d(myPoly % dgMatrix_dev, myPoly % bMatrix_dev, &
                       myPoly % qWeights_dev, f, bF, dF, INT(myPoly % N, C_INT), INT(nVariables, C_INT), &
                       INT(nElements, C_INT), C_NULL_PTR), 'VectorDGDivergence_2D')
  END SUBROUTINE VectorDGDivergence_2D

END MODULE SELF_Lagrange
