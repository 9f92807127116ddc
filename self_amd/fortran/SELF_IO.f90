! SELF_IO.f90 -- HDF5 file access for the Fortran layer, over libself_amd_io.so:
!   Read_HOPr   Mesh2D%Read_HOPr (Read_HOPr_Mesh2D, src/SELF_Mesh.f90:852-1016)
!   WriteModel  Model%WriteModel (Write_Model2D, src/SELF_Model2D.f90:861-1059), with the
!               attribute N that ReadModel needs
!   ReadModel   Model%ReadModel (Read_Model2D, src/SELF_Model2D.f90:1061-1099)
! Free subroutines in a module of their own (CALL WriteModel(model, file) for the reference's
! CALL model % WriteModel(file)), so that programs that do no file IO need no HDF5 library.
MODULE SELF_IO
  USE SELF_Constants
  USE SELF_AMD_CAPI
  USE SELF_Lagrange
  USE SELF_Mesh
  USE SELF_Model2D
  IMPLICIT NONE

  INTERFACE
    TYPE(C_PTR) FUNCTION selfamd_io_last_error() BIND(C, name="selfamd_io_last_error")
      IMPORT :: C_PTR
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_hopr_attribute_int(path, name, val) BIND(C, name="selfamd_hopr_attribute_int")
      IMPORT :: C_INT, C_CHAR, C_INT32_T
      CHARACTER(KIND=C_CHAR) :: path(*), name(*)
      INTEGER(C_INT32_T) :: val
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_hopr_dataset_dims(path, name, rank, dims) BIND(C, name="selfamd_hopr_dataset_dims")
      IMPORT :: C_INT, C_CHAR, C_INT32_T, C_INT64_T
      CHARACTER(KIND=C_CHAR) :: path(*), name(*)
      INTEGER(C_INT32_T) :: rank
      INTEGER(C_INT64_T) :: dims(5)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_hopr_read_int(path, name, out, count) BIND(C, name="selfamd_hopr_read_int")
      IMPORT :: C_INT, C_CHAR, C_INT32_T, C_INT64_T
      CHARACTER(KIND=C_CHAR) :: path(*), name(*)
      INTEGER(C_INT32_T) :: out(*)
      INTEGER(C_INT64_T), VALUE :: count
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_hopr_read_double(path, name, out, count) BIND(C, name="selfamd_hopr_read_double")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE, C_INT64_T
      CHARACTER(KIND=C_CHAR) :: path(*), name(*)
      REAL(C_DOUBLE) :: out(*)
      INTEGER(C_INT64_T), VALUE :: count
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_pickup_write(path, N, nVar, nEl, xi, w, dgM, dM, bM, sol, solB, x, xb) &
        BIND(C, name="selfamd_pickup_write")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE
      CHARACTER(KIND=C_CHAR) :: path(*)
      INTEGER(C_INT), VALUE :: N, nVar, nEl
      REAL(C_DOUBLE) :: xi(*), w(*), dgM(*), dM(*), bM(*), sol(*), solB(*), x(*), xb(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_pickup_write_registers(path, N, nVar, nEl, fluxDiv, flux, fluxB, grad, gradB) &
        BIND(C, name="selfamd_pickup_write_registers")
      IMPORT :: C_INT, C_CHAR, C_PTR
      CHARACTER(KIND=C_CHAR) :: path(*)
      INTEGER(C_INT), VALUE :: N, nVar, nEl
      TYPE(C_PTR), VALUE :: fluxDiv, flux, fluxB, grad, gradB  ! host arrays; C_NULL_PTR: zeros
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_pickup_read(path, N, nVar, nEl, sol) BIND(C, name="selfamd_pickup_read")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE
      CHARACTER(KIND=C_CHAR) :: path(*)
      INTEGER(C_INT), VALUE :: N, nVar, nEl
      REAL(C_DOUBLE) :: sol(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION selfamd_mesh2d_from_hopr(nElem, nGeo, si3, nc3, gn3, sideInfo, nodeCoords, gnid) &
        BIND(C, name="selfamd_mesh2d_from_hopr")
      IMPORT :: C_INT, C_DOUBLE, C_INT32_T
      INTEGER(C_INT), VALUE :: nElem, nGeo
      INTEGER(C_INT32_T) :: si3(*), gn3(*), sideInfo(*), gnid(*)
      REAL(C_DOUBLE) :: nc3(*), nodeCoords(*)
    END FUNCTION
  END INTERFACE

CONTAINS

  SUBROUTINE io_check(rc, what)
    !! Stop with the IO library's message when a call failed
    INTEGER(C_INT), INTENT(in) :: rc
    CHARACTER(*), INTENT(in) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    CHARACTER(LEN=512) :: text
    INTEGER :: k
    IF (rc == 0) RETURN
    CALL C_F_POINTER(selfamd_io_last_error(), msg, [512])
    text = ''
    DO k = 1, 512
      IF (msg(k) == C_NULL_CHAR) EXIT
      text(k:k) = msg(k)
    END DO
    PRINT '(A,A,A,I0,A,A)', 'self_amd: ', what, ' failed (status ', rc, '): ', TRIM(text)
    ERROR STOP 3
  END SUBROUTINE io_check

  FUNCTION cstr(s) RESULT(c)
    CHARACTER(*), INTENT(in) :: s
    CHARACTER(KIND=C_CHAR, LEN=LEN_TRIM(s) + 1) :: c
    c = TRIM(s)//C_NULL_CHAR
  END FUNCTION cstr

  ! Read_HOPr_Mesh2D (src/SELF_Mesh.f90:852-1016): the file's Ngeo, SideInfo, NodeCoords and
  ! GlobalNodeIDs, converted from the 3-D HOPr description to Mesh2D
  SUBROUTINE Read_HOPr(mesh, meshFile)
    TYPE(Mesh2D), INTENT(out) :: mesh
    CHARACTER(*), INTENT(in) :: meshFile
    INTEGER(C_INT32_T) :: nGeo, rank
    INTEGER(C_INT64_T) :: dims(5)
    INTEGER(C_INT32_T), ALLOCATABLE :: si3(:,:), gn3(:)
    REAL(C_DOUBLE), ALLOCATABLE :: nc3(:,:)
    INTEGER :: nElem
    CALL io_check(selfamd_hopr_attribute_int(cstr(meshFile), cstr('Ngeo'), nGeo), 'Read_HOPr')
    ! row-major (nSides, 5) in the file is (5, nSides) here
    CALL io_check(selfamd_hopr_dataset_dims(cstr(meshFile), cstr('SideInfo'), rank, dims), 'Read_HOPr')
    ALLOCATE (si3(dims(2), dims(1)))
    CALL io_check(selfamd_hopr_read_int(cstr(meshFile), cstr('SideInfo'), si3, SIZE(si3, KIND=C_INT64_T)), 'Read_HOPr')
    CALL io_check(selfamd_hopr_dataset_dims(cstr(meshFile), cstr('NodeCoords'), rank, dims), 'Read_HOPr')
    ALLOCATE (nc3(dims(2), dims(1)))
    CALL io_check(selfamd_hopr_read_double(cstr(meshFile), cstr('NodeCoords'), nc3, SIZE(nc3, KIND=C_INT64_T)), &
                  'Read_HOPr')
    CALL io_check(selfamd_hopr_dataset_dims(cstr(meshFile), cstr('GlobalNodeIDs'), rank, dims), 'Read_HOPr')
    ALLOCATE (gn3(PRODUCT(dims(1:rank))))
    CALL io_check(selfamd_hopr_read_int(cstr(meshFile), cstr('GlobalNodeIDs'), gn3, SIZE(gn3, KIND=C_INT64_T)), &
                  'Read_HOPr')
    nElem = SIZE(si3, 2)/6  ! six sides per HOPr (3-D) element
    mesh % nGeo = nGeo
    mesh % nElem = nElem
    mesh % quadrature = UNIFORM
    ALLOCATE (mesh % sideInfo(5, 4, nElem), mesh % nodeCoords(2, 0:nGeo, 0:nGeo, nElem), &
              mesh % globalNodeIDs(0:nGeo, 0:nGeo, nElem))
    CALL selfamd_check(selfamd_mesh2d_from_hopr(INT(nElem, C_INT), INT(nGeo, C_INT), si3, nc3, gn3, mesh % sideInfo, &
                       mesh % nodeCoords, mesh % globalNodeIDs), 'Read_HOPr')
    CALL mesh % UpdateDevice()
  END SUBROUTINE Read_HOPr

  ! Write_Model2D (src/SELF_Model2D.f90:861-1059) plus the attribute N
  SUBROUTINE WriteModel(model, fileName)
    CLASS(Model2D), INTENT(inout) :: model
    CHARACTER(*), INTENT(in) :: fileName
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: x(:), xb(:), fdiv(:), fl(:), flb(:)
    TYPE(C_PTR) :: fdiv_dev, fl_dev
    INTEGER :: P, nEl, nVar
    TYPE(Lagrange), POINTER :: interp
    interp => model % solution % interp
    P = interp % N + 1
    nEl = model % solution % nElem
    CALL model % UpdateHost()
    CALL model % solution % BoundaryInterp()
    CALL model % solution % UpdateHost()
    ALLOCATE (x(2*P*P*nEl), xb(2*P*4*nEl))
    CALL hip_check(selfamd_memcpy(C_LOC(x), model % geometry % x, INT(SIZE(x), C_SIZE_T)*8_C_SIZE_T, &
                   hipMemcpyDeviceToHost), 'WriteModel')
    CALL hip_check(selfamd_memcpy(C_LOC(xb), model % geometry % xb, INT(SIZE(xb), C_SIZE_T)*8_C_SIZE_T, &
                   hipMemcpyDeviceToHost), 'WriteModel')
    CALL io_check(selfamd_pickup_write(cstr(fileName), INT(interp % N, C_INT), INT(model % solution % nVar, C_INT), &
                  INT(nEl, C_INT), interp % controlPoints, interp % qWeights, interp % dgMatrix, interp % dMatrix, &
                  interp % bMatrix, model % solution % interior, model % solution % boundary, x, xb), 'WriteModel')
    ! the flux and fluxDivergence registers: on the per-operator path those of the last
    ! tendency evaluation (as the reference's), on the fused engine those of the current state
    nVar = model % solution % nVar
    ALLOCATE (fdiv(P*P*nVar*nEl), fl(2*P*P*nVar*nEl), flb(2*P*nVar*4*nEl))
    IF (model % plugins) THEN
      fdiv_dev = model % fluxDivergence % interior_dev
      fl_dev = model % flux % interior_dev
      CALL hip_check(selfamd_memcpy(C_LOC(flb), model % flux % boundary_dev, INT(SIZE(flb), C_SIZE_T)*8_C_SIZE_T, &
                     hipMemcpyDeviceToHost), 'WriteModel')
    ELSE
      fdiv_dev = device_alloc(INT(SIZE(fdiv), C_SIZE_T)*8_C_SIZE_T)
      fl_dev = device_alloc(INT(SIZE(fl), C_SIZE_T)*8_C_SIZE_T)
      CALL selfamd_check(selfamd_model2d_flux_registers(model % engine, fdiv_dev, fl_dev), 'WriteModel')
      flb = 0.0_prec
    END IF
    CALL hip_check(selfamd_memcpy(C_LOC(fdiv), fdiv_dev, INT(SIZE(fdiv), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), &
                   'WriteModel')
    CALL hip_check(selfamd_memcpy(C_LOC(fl), fl_dev, INT(SIZE(fl), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), &
                   'WriteModel')
    IF (.NOT. model % plugins) THEN
      CALL device_free(fdiv_dev)
      CALL device_free(fl_dev)
    END IF
    CALL io_check(selfamd_pickup_write_registers(cstr(fileName), INT(interp % N, C_INT), INT(nVar, C_INT), &
                  INT(nEl, C_INT), C_LOC(fdiv), C_LOC(fl), C_LOC(flb), C_NULL_PTR, C_NULL_PTR), 'WriteModel')
  END SUBROUTINE WriteModel

  ! Read_Model2D (src/SELF_Model2D.f90:1061-1099): the solution of a pickup file (its N must
  ! match), then SetSolution
  SUBROUTINE ReadModel(model, fileName)
    CLASS(Model2D), INTENT(inout) :: model
    CHARACTER(*), INTENT(in) :: fileName
    REAL(prec), ALLOCATABLE :: sol(:,:,:,:)
    ALLOCATE (sol, MOLD=model % solution % interior)
    CALL io_check(selfamd_pickup_read(cstr(fileName), INT(model % solution % interp % N, C_INT), &
                  INT(model % solution % nVar, C_INT), INT(model % solution % nElem, C_INT), sol), 'ReadModel')
    CALL model % SetSolution(sol)
  END SUBROUTINE ReadModel

END MODULE SELF_IO
