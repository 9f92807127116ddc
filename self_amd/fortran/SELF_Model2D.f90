! SELF_Model2D.f90 -- Model2D (src/SELF_Model2D.f90:18-71) and the three physics plug-ins
! (src/models/SELF_LinearShallowWater.f90, SELF_ShallowWater.f90, SELF_CompressibleIdealGas2D.f90)
! for the self_amd engine.
!
! The state lives on the GPU inside one engine handle; CalculateTendency and ForwardStep run
! one fused HIP kernel per RK3 stage (hipGraph replay).  Host copies of the solution are
! made only by UpdateHost.  Initial conditions, topography, bathymetry, Coriolis and gravity
! fields are given as arrays (the reference's FEQParse equation strings are not used).
MODULE SELF_Model2D
  USE SELF_Constants
  USE SELF_AMD_CAPI
  USE SELF_Lagrange
  USE SELF_Mesh
  USE SELF_MappedData
  IMPLICIT NONE

  INTEGER, PARAMETER :: SELF_MODEL_LSW = 1, SELF_MODEL_SW = 2, SELF_MODEL_EULER = 3
  INTEGER(C_INT), PARAMETER :: PARAM_G = 1, PARAM_FCORI = 2, PARAM_GAMMA = 3, PARAM_RGAS = 4
  INTEGER(C_INT), PARAMETER :: PARAM_NODAL_METRICS = 5
  INTEGER(C_INT), PARAMETER :: FIELD_TOPOGRAPHY = 1, FIELD_BATHYMETRY = 2, FIELD_CORIOLIS = 3
  INTEGER(C_INT), PARAMETER :: FIELD_POTENTIAL = 4, FIELD_DRAG = 5

  TYPE, PUBLIC :: Model2D
    INTEGER :: kind = 0
    REAL(prec) :: t = 0.0_prec
    REAL(prec) :: dt = 0.0_prec
    REAL(prec) :: entropy = 0.0_prec
    TYPE(MappedScalar2D) :: solution
    TYPE(MappedScalar2D) :: dSdt
    TYPE(Mesh2D), POINTER :: mesh => NULL()
    TYPE(SEMQuad), POINTER :: geometry => NULL()
    TYPE(C_PTR) :: engine = C_NULL_PTR
  CONTAINS
    PROCEDURE, PUBLIC :: Init_Model2D
    PROCEDURE, PUBLIC :: Free => Free_Model2D
    PROCEDURE, PUBLIC :: SetSolution => SetSolution_Model2D
    PROCEDURE, PUBLIC :: SetTimeIntegrator => SetTimeIntegrator_Model2D
    PROCEDURE, PUBLIC :: CalculateTendency => CalculateTendency_Model2D
    PROCEDURE, PUBLIC :: ForwardStep => ForwardStep_Model2D
    PROCEDURE, PUBLIC :: UpdateHost => UpdateHost_Model2D
    PROCEDURE, PUBLIC :: CalculateEntropy => CalculateEntropy_Model2D
    PROCEDURE, PUBLIC :: ReportEntropy => ReportEntropy_Model2D
    PROCEDURE, PUBLIC :: Integral => Integral_Model2D
    PROCEDURE, PUBLIC :: SetNodalMetrics => SetNodalMetrics_Model2D
    PROCEDURE, PUBLIC :: WriteTecplot => WriteTecplot_Model2D
    PROCEDURE, PRIVATE :: SetField => SetField_Model2D
  END TYPE Model2D

  TYPE, EXTENDS(Model2D), PUBLIC :: ShallowWater
    REAL(prec) :: g = 1.0_prec
    REAL(prec) :: fCori = 0.0_prec
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_ShallowWater
    PROCEDURE, PUBLIC :: SetTopography => SetTopography_ShallowWater
  END TYPE ShallowWater

  TYPE, EXTENDS(Model2D), PUBLIC :: LinearShallowWater
    REAL(prec) :: g = 1.0_prec
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_LinearShallowWater
    PROCEDURE, PUBLIC :: SetBathymetry => SetBathymetry_LinearShallowWater
    PROCEDURE, PUBLIC :: SetCoriolis => SetCoriolis_LinearShallowWater
  END TYPE LinearShallowWater

  TYPE, EXTENDS(Model2D), PUBLIC :: CompressibleIdealGas2D
    REAL(prec) :: expansionFactor = 1.4_prec
    REAL(prec) :: R = 1.0_prec
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SetGravity => SetGravity_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SetDrag => SetDrag_CompressibleIdealGas2D
  END TYPE CompressibleIdealGas2D

CONTAINS

  ! Init_Model2D (src/SELF_Model2D.f90:142-178)
  SUBROUTINE Init_Model2D(this, kind, nvar, mesh, geometry)
    CLASS(Model2D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: kind, nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    this % kind = kind
    this % mesh => mesh
    this % geometry => geometry
    this % t = 0.0_prec
    CALL this % solution % Init(geometry % interp, nvar, mesh % nElem)
    CALL this % dSdt % Init(geometry % interp, nvar, mesh % nElem)
    CALL selfamd_check(selfamd_model2d_create(this % engine, INT(kind, C_INT), INT(geometry % interp % N, C_INT), &
                       INT(mesh % nElem, C_INT), this % mesh % sideInfo, geometry % dsdx, geometry % J, &
                       geometry % nHatb, geometry % nScaleb, C_NULL_PTR), 'Model2D%Init')
  END SUBROUTINE Init_Model2D

  SUBROUTINE Free_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    IF (C_ASSOCIATED(this % engine)) CALL selfamd_check(selfamd_model2d_destroy(this % engine), 'Model2D%Free')
    this % engine = C_NULL_PTR
    CALL this % solution % Free()
    CALL this % dSdt % Free()
  END SUBROUTINE Free_Model2D

  ! Per-node metric terms (the reference's representation, needed on curved meshes) or one
  ! affine record per element (the default on affine meshes)
  SUBROUTINE SetNodalMetrics_Model2D(this, nodal)
    CLASS(Model2D), INTENT(inout) :: this
    LOGICAL, INTENT(in) :: nodal
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_NODAL_METRICS, &
                       MERGE(1.0_prec, 0.0_prec, nodal)), 'Model2D%SetNodalMetrics')
  END SUBROUTINE SetNodalMetrics_Model2D

  ! SetSolution (src/SELF_Model2D.f90:313-332) from an array interior(0:N,0:N,nVar,nElem)
  SUBROUTINE SetSolution_Model2D(this, interior)
    CLASS(Model2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: interior(:,:,:,:)
    this % solution % interior = interior
    CALL this % solution % UpdateDevice()
    CALL selfamd_check(selfamd_model2d_set_solution(this % engine, this % solution % interior_dev), &
                       'Model2D%SetSolution')
  END SUBROUTINE SetSolution_Model2D

  ! SetTimeIntegrator (src/SELF_Model.f90:292-382): the engine implements Williamson's
  ! low-storage RK3 (the integrator of the tendency path)
  SUBROUTINE SetTimeIntegrator_Model2D(this, integrator)
    CLASS(Model2D), INTENT(inout) :: this
    CHARACTER(*), INTENT(in) :: integrator
    IF (TRIM(integrator) /= 'rk3' .AND. TRIM(integrator) /= 'RK3') THEN
      PRINT '(A,A)', 'self_amd: unsupported time integrator ', TRIM(integrator)
      ERROR STOP 2
    END IF
  END SUBROUTINE SetTimeIntegrator_Model2D

  ! CalculateTendency_Model2D (src/SELF_Model2D.f90:816-859): dSdt on the device
  SUBROUTINE CalculateTendency_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_model2d_tendency(this % engine, this % dSdt % interior_dev), &
                       'Model2D%CalculateTendency')
  END SUBROUTINE CalculateTendency_Model2D

  ! ForwardStep_Model + LowStorageRK3_timeIntegrator (src/SELF_Model.f90:474-523, 704-732):
  ! the reference's step loop dt = MIN(dtLim, tn - t) is replayed on the host, full steps are
  ! batched into one engine call and a shorter final step gets its own call.
  SUBROUTINE ForwardStep_Model2D(this, tn, dt)
    CLASS(Model2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: tn
    REAL(prec), OPTIONAL, INTENT(in) :: dt
    REAL(prec) :: t, dtStep
    INTEGER :: nFull
    IF (PRESENT(dt)) this % dt = dt
    t = this % t
    nFull = 0
    DO WHILE (t < tn)
      dtStep = MIN(this % dt, tn - t)
      IF (dtStep == this % dt) THEN
        nFull = nFull + 1
      ELSE
        IF (nFull > 0) CALL selfamd_check(selfamd_model2d_forward_step(this % engine, INT(nFull, C_INT), this % dt), &
                                          'Model2D%ForwardStep')
        nFull = 0
        CALL selfamd_check(selfamd_model2d_forward_step(this % engine, 1_C_INT, dtStep), 'Model2D%ForwardStep')
      END IF
      t = t + dtStep
    END DO
    IF (nFull > 0) CALL selfamd_check(selfamd_model2d_forward_step(this % engine, INT(nFull, C_INT), this % dt), &
                                      'Model2D%ForwardStep')
    CALL selfamd_check(selfamd_model2d_synchronize(this % engine), 'Model2D%ForwardStep')
    this % t = t
    CALL selfamd_check(selfamd_model2d_set_time(this % engine, t), 'Model2D%ForwardStep')
  END SUBROUTINE ForwardStep_Model2D

  ! Solution (and the last tendency) back to the host mirrors
  SUBROUTINE UpdateHost_Model2D(this)
    CLASS(Model2D), INTENT(inout), TARGET :: this
    CALL selfamd_check(selfamd_model2d_get_solution(this % engine, this % solution % interior_dev), &
                       'Model2D%UpdateHost')
    CALL hip_check(selfamd_memcpy(C_LOC(this % solution % interior), this % solution % interior_dev, &
                   INT(SIZE(this % solution % interior), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), 'selfamd_memcpy')
    CALL hip_check(selfamd_memcpy(C_LOC(this % dSdt % interior), this % dSdt % interior_dev, &
                   INT(SIZE(this % dSdt % interior), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), 'selfamd_memcpy')
  END SUBROUTINE UpdateHost_Model2D

  ! CalculateEntropy (src/models/SELF_ShallowWater.f90:149-188, SELF_LinearShallowWater.f90:171-216,
  ! SELF_CompressibleIdealGas2D.f90:782-823), reduced on the device
  SUBROUTINE CalculateEntropy_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    REAL(C_DOUBLE) :: e
    CALL selfamd_check(selfamd_model2d_entropy(this % engine, e), 'Model2D%CalculateEntropy')
    this % entropy = e
  END SUBROUTINE CalculateEntropy_Model2D

  ! ReportEntropy_Model (src/SELF_Model.f90:440-469): time and entropy on standard output, in
  ! the reference's ES16.7E3 format (ASCII labels)
  SUBROUTINE ReportEntropy_Model2D(this)
    CLASS(Model2D), INTENT(in) :: this
    CHARACTER(LEN=20) :: modelTime, entropy
    WRITE (modelTime, "(ES16.7E3)") this % t
    WRITE (entropy, "(ES16.7E3)") this % entropy
    WRITE (*, '(A)') 't_i =' // TRIM(modelTime) // '  |  e_i =' // TRIM(entropy)
  END SUBROUTINE ReportEntropy_Model2D

  ! WriteTecplot_Model2D (src/SELF_Model2D.f90:1101-1207): element positions and solution
  ! interpolated on the device to the interpolant's target grid ((M+1)^2 points per element,
  ! GridInterp), ASCII Tecplot, one zone per element: X, Y, the variables, then d/dx and d/dy
  ! of each variable.  The gradient columns are zero: no tendency on this path forms the
  ! solution gradient (the reference's 2-D models allocate that register and never fill it).
  ! Default file name solution.<TimeStamp(t)>.tec.
  SUBROUTINE WriteTecplot_Model2D(this, filename)
    CLASS(Model2D), INTENT(inout) :: this
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    TYPE(Lagrange), POINTER :: interp
    REAL(prec), ALLOCATABLE, TARGET :: xT(:,:,:,:), sT(:,:,:,:)
    REAL(prec), ALLOCATABLE :: zeros(:)
    TYPE(C_PTR) :: xT_dev, sT_dev
    CHARACTER(LEN=512) :: tecFile
    CHARACTER(LEN=1024) :: header
    CHARACTER(LEN=32) :: fmat
    INTEGER :: N, M, nVar, nEl, iEl, i, j, k, u
    interp => this % solution % interp
    N = interp % N
    M = interp % M
    nVar = this % solution % nVar
    nEl = this % solution % nElem
    IF (PRESENT(filename)) THEN
      tecFile = filename
    ELSE
      tecFile = 'solution.'//TimeStamp(this % t)//'.tec'
    END IF
    CALL selfamd_check(selfamd_model2d_get_solution(this % engine, this % solution % interior_dev), &
                       'Model2D%WriteTecplot')
    xT_dev = device_alloc(INT(2*(M + 1)**2*nEl, C_SIZE_T)*8_C_SIZE_T)
    sT_dev = device_alloc(INT(nVar*(M + 1)**2*nEl, C_SIZE_T)*8_C_SIZE_T)
    CALL selfamd_check(selfamd_vector_grid_interp_2d(interp % iMatrix_dev, this % geometry % x, xT_dev, INT(N, C_INT), &
                       INT(M, C_INT), 1_C_INT, INT(nEl, C_INT), C_NULL_PTR), 'Model2D%WriteTecplot')
    CALL selfamd_check(selfamd_scalar_grid_interp_2d(interp % iMatrix_dev, this % solution % interior_dev, sT_dev, &
                       INT(N, C_INT), INT(M, C_INT), INT(nVar, C_INT), INT(nEl, C_INT), C_NULL_PTR), &
                       'Model2D%WriteTecplot')
    ALLOCATE (xT(2, 0:M, 0:M, nEl), sT(0:M, 0:M, nVar, nEl), zeros(2*nVar))
    zeros = 0.0_prec
    CALL hip_check(selfamd_memcpy(C_LOC(xT), xT_dev, INT(SIZE(xT), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), &
                   'selfamd_memcpy')
    CALL hip_check(selfamd_memcpy(C_LOC(sT), sT_dev, INT(SIZE(sT), C_SIZE_T)*8_C_SIZE_T, hipMemcpyDeviceToHost), &
                   'selfamd_memcpy')
    CALL device_free(xT_dev)
    CALL device_free(sT_dev)
    header = 'VARIABLES = "X", "Y"'
    DO k = 1, nVar
      header = TRIM(header)//', "'//TRIM(VarName(this % kind, k))//'"'
    END DO
    DO k = 1, nVar
      header = TRIM(header)//', "d/dx('//TRIM(VarName(this % kind, k))//')"'
    END DO
    DO k = 1, nVar
      header = TRIM(header)//', "d/dy('//TRIM(VarName(this % kind, k))//')"'
    END DO
    WRITE (fmat, '(A,I0,A)') '(', 3*nVar + 2, '(ES16.7E3,1X))'
    OPEN (NEWUNIT=u, FILE=TRIM(tecFile), FORM='formatted', STATUS='replace', ACTION='write')
    WRITE (u, '(A)') ' '//TRIM(header)
    DO iEl = 1, nEl
      WRITE (u, '(A,I8.8,A,I0,A,I0)') ' ZONE T="el', iEl, '", I=', M + 1, ', J=', M + 1
      DO j = 0, M
        DO i = 0, M
          WRITE (u, fmat) xT(1, i, j, iEl), xT(2, i, j, iEl), sT(i, j, 1:nVar, iEl), zeros
        END DO
      END DO
    END DO
    CLOSE (u)
  END SUBROUTINE WriteTecplot_Model2D

  ! solution variable names (SetName in the models' Init: src/models/SELF_ShallowWater.f90:135-143,
  ! SELF_LinearShallowWater.f90:139-147, SELF_CompressibleIdealGas2D.f90:306-318)
  FUNCTION VarName(kind, k) RESULT(name)
    INTEGER, INTENT(in) :: kind, k
    CHARACTER(LEN=8) :: name
    CHARACTER(LEN=8), PARAMETER :: sw(3) = [CHARACTER(LEN=8) :: 'Hu', 'Hv', 'H']
    CHARACTER(LEN=8), PARAMETER :: lsw(3) = [CHARACTER(LEN=8) :: 'u', 'v', 'eta']
    CHARACTER(LEN=8), PARAMETER :: eu(4) = [CHARACTER(LEN=8) :: 'rho*u', 'rho*v', 'rho', 'rho*E']
    SELECT CASE (kind)
    CASE (SELF_MODEL_SW)
      name = sw(k)
    CASE (SELF_MODEL_LSW)
      name = lsw(k)
    CASE DEFAULT
      name = eu(k)
    END SELECT
  END FUNCTION VarName

  ! TimeStamp with units 's' (src/SELF_SupportRoutines.f90:856-924): dddd hh mm ss mmm
  FUNCTION TimeStamp(t) RESULT(stamp)
    REAL(prec), INTENT(in) :: t
    CHARACTER(LEN=13) :: stamp
    INTEGER :: day, hour, minute, second, ms
    REAL(prec) :: rest
    day = INT(t/86400.0_prec)
    hour = INT((t - 86400.0_prec*day)/3600.0_prec)
    minute = INT((t - 3600.0_prec*hour - 86400.0_prec*day)/60.0_prec)
    rest = t - 60.0_prec*minute - 3600.0_prec*hour - 86400.0_prec*day
    second = INT(rest)
    ms = NINT((rest - second)*1000.0_prec)
    IF (ms >= 1000) THEN
      ms = ms - 1000
      second = second + 1
    END IF
    IF (second >= 60) THEN
      second = second - 60
      minute = minute + 1
    END IF
    IF (minute >= 60) THEN
      minute = minute - 60
      hour = hour + 1
    END IF
    IF (hour >= 24) THEN
      hour = hour - 24
      day = day + 1
    END IF
    WRITE (stamp, '(I4.4,I2.2,I2.2,I2.2,I3.3)') day, hour, minute, second, ms
  END FUNCTION TimeStamp

  ! sum of w_i w_j J S(i,j,var) over the mesh (var counted from 1, as in the reference)
  REAL(prec) FUNCTION Integral_Model2D(this, var) RESULT(v)
    CLASS(Model2D), INTENT(in) :: this
    INTEGER, INTENT(in) :: var
    REAL(C_DOUBLE) :: r
    CALL selfamd_check(selfamd_model2d_integral(this % engine, INT(var - 1, C_INT), r), 'Model2D%Integral')
    v = r
  END FUNCTION Integral_Model2D

  SUBROUTINE SetField_Model2D(this, which, values)
    CLASS(Model2D), INTENT(inout) :: this
    INTEGER(C_INT), INTENT(in) :: which
    REAL(prec), INTENT(in), TARGET :: values(:,:,:)
    TYPE(C_PTR) :: d
    INTEGER(C_SIZE_T) :: nb
    nb = INT(SIZE(values), C_SIZE_T)*8_C_SIZE_T
    d = device_alloc(nb)
    CALL hip_check(selfamd_memcpy(d, C_LOC(values), nb, hipMemcpyHostToDevice), 'selfamd_memcpy')
    CALL selfamd_check(selfamd_model2d_set_field(this % engine, which, d), 'Model2D%SetField')
    CALL device_free(d)
  END SUBROUTINE SetField_Model2D

  ! ---- ShallowWater (src/models/SELF_ShallowWater.f90:96-147) ----
  SUBROUTINE Init_ShallowWater(this, nvar, mesh, geometry)
    CLASS(ShallowWater), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    CALL this % Init_Model2D(SELF_MODEL_SW, 3, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_G, this % g), 'ShallowWater%Init')
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_FCORI, this % fCori), 'ShallowWater%Init')
  END SUBROUTINE Init_ShallowWater

  ! SetTopography (src/models/SELF_ShallowWater.f90:241-282), topography(0:N,0:N,nElem)
  SUBROUTINE SetTopography_ShallowWater(this, topography)
    CLASS(ShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: topography(:,:,:)
    CALL this % SetField(FIELD_TOPOGRAPHY, topography)
  END SUBROUTINE SetTopography_ShallowWater

  ! ---- LinearShallowWater (src/models/SELF_LinearShallowWater.f90:101-151) ----
  SUBROUTINE Init_LinearShallowWater(this, nvar, mesh, geometry)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    REAL(prec), ALLOCATABLE :: zero(:,:,:)
    CALL this % Init_Model2D(SELF_MODEL_LSW, 3, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_G, this % g), 'LinearShallowWater%Init')
    ALLOCATE (zero(0:geometry % interp % N, 0:geometry % interp % N, mesh % nElem))
    zero = 0.0_prec
    CALL this % SetCoriolis(zero)  ! fCori starts at zero (hfReal allocation zero-fills)
  END SUBROUTINE Init_LinearShallowWater

  SUBROUTINE SetBathymetry_LinearShallowWater(this, H)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: H(:,:,:)
    CALL this % SetField(FIELD_BATHYMETRY, H)
  END SUBROUTINE SetBathymetry_LinearShallowWater

  SUBROUTINE SetCoriolis_LinearShallowWater(this, f)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: f(:,:,:)
    CALL this % SetField(FIELD_CORIOLIS, f)
  END SUBROUTINE SetCoriolis_LinearShallowWater

  ! ---- CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:261-358) ----
  SUBROUTINE Init_CompressibleIdealGas2D(this, nvar, mesh, geometry)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    CALL this % Init_Model2D(SELF_MODEL_EULER, 4, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_GAMMA, this % expansionFactor), &
                       'CompressibleIdealGas2D%Init')
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_RGAS, this % R), 'CompressibleIdealGas2D%Init')
  END SUBROUTINE Init_CompressibleIdealGas2D

  SUBROUTINE SetGravity_CompressibleIdealGas2D(this, potential)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: potential(:,:,:)
    CALL this % SetField(FIELD_POTENTIAL, potential)
  END SUBROUTINE SetGravity_CompressibleIdealGas2D

  SUBROUTINE SetDrag_CompressibleIdealGas2D(this, Cd)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: Cd(:,:,:)
    CALL this % SetField(FIELD_DRAG, Cd)
  END SUBROUTINE SetDrag_CompressibleIdealGas2D

END MODULE SELF_Model2D
