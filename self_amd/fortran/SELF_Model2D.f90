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

  ! Williamson low-storage RK3 (src/SELF_Model.f90:25-27)
  REAL(prec), PARAMETER :: rk3_a(1:3) = [0.0_prec, -5.0_prec/9.0_prec, -153.0_prec/128.0_prec]
  REAL(prec), PARAMETER :: rk3_b(1:3) = [0.0_prec, 1.0_prec/3.0_prec, 3.0_prec/4.0_prec]
  REAL(prec), PARAMETER :: rk3_g(1:3) = [1.0_prec/3.0_prec, 15.0_prec/16.0_prec, 8.0_prec/15.0_prec]

  ! Two tendency paths share this type:
  ! * the fused engine (default for the three built-in physics): one HIP kernel per RK3
  !   stage inside the engine handle;
  ! * the per-operator path (plugins = .TRUE.): CalculateTendency_Model2D's sequence
  !   (src/SELF_Model2D.f90:816-859) over the registers below, calling the overridable
  !   plug-in methods PreTendency, SetBoundaryCondition, SourceMethod, RiemannSolver and
  !   FluxMethod (src/SELF_Model.f90:128-132) between the library's device operators.  A
  !   type extension of one of the physics is switched to it by that physics' Init, so a user
  !   extension that overrides a plug-in method runs its own method on the GPU path; a direct
  !   extension of Model2D (kind 0) has no engine and always uses it.
  TYPE, PUBLIC :: Model2D
    INTEGER :: kind = 0
    REAL(prec) :: t = 0.0_prec
    REAL(prec) :: dt = 0.0_prec
    REAL(prec) :: entropy = 0.0_prec
    TYPE(MappedScalar2D) :: solution
    TYPE(MappedScalar2D) :: dSdt
    ! per-operator path registers (src/SELF_Model2D.f90:18-31), allocated by EnablePlugins
    TYPE(MappedScalar2D) :: source
    TYPE(MappedVector2D) :: flux
    TYPE(MappedScalar2D) :: fluxDivergence
    TYPE(MappedScalar2D) :: workSol
    LOGICAL :: plugins = .FALSE.
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
    PROCEDURE, PUBLIC :: UpdateGRK3 => UpdateGRK3_Model2D
    PROCEDURE, PUBLIC :: UpdateHost => UpdateHost_Model2D
    PROCEDURE, PUBLIC :: CalculateEntropy => CalculateEntropy_Model2D
    PROCEDURE, PUBLIC :: ReportEntropy => ReportEntropy_Model2D
    PROCEDURE, PUBLIC :: Integral => Integral_Model2D
    PROCEDURE, PUBLIC :: SetNodalMetrics => SetNodalMetrics_Model2D
    PROCEDURE, PUBLIC :: WriteTecplot => WriteTecplot_Model2D
    ! plug-in methods (src/SELF_Model.f90:128-132); the defaults do nothing, as the reference's
    PROCEDURE, PUBLIC :: PreTendency => PreTendency_Model2D
    PROCEDURE, PUBLIC :: SetBoundaryCondition => SetBoundaryCondition_Model2D
    PROCEDURE, PUBLIC :: SourceMethod => Source_Model2D
    PROCEDURE, PUBLIC :: FluxMethod => Flux_Model2D
    PROCEDURE, PUBLIC :: RiemannSolver => RiemannSolver_Model2D
    PROCEDURE, PUBLIC :: EnablePlugins => EnablePlugins_Model2D
    ! hooks of the physics for the per-operator path: their own registers, and what a new
    ! solution sets (e.g. the prescribed boundary state)
    PROCEDURE, PUBLIC :: InitPlugins => InitPlugins_Model2D
    PROCEDURE, PUBLIC :: PluginSetSolution => PluginSetSolution_Model2D
    PROCEDURE, PUBLIC :: StateToDevice => StateToDevice_Model2D
    PROCEDURE, PRIVATE :: SyncEngine => SyncEngine_Model2D
    PROCEDURE, PRIVATE :: SetField => SetField_Model2D
  END TYPE Model2D

  TYPE, EXTENDS(Model2D), PUBLIC :: ShallowWater
    REAL(prec) :: g = 1.0_prec
    REAL(prec) :: fCori = 0.0_prec
    TYPE(MappedScalar2D) :: H               ! topography (the reference's this%H)
    TYPE(MappedVector2D) :: velocity        ! per-operator path: u, v and their traces
    TYPE(C_PTR) :: gradH_dev = C_NULL_PTR   ! GradientSF of H, (2,0:N,0:N,1,nElem)
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_ShallowWater
    PROCEDURE, PUBLIC :: Free => Free_ShallowWater
    PROCEDURE, PUBLIC :: SetTopography => SetTopography_ShallowWater
    PROCEDURE, PUBLIC :: InitPlugins => InitPlugins_ShallowWater
    PROCEDURE, PUBLIC :: PreTendency => PreTendency_ShallowWater
    PROCEDURE, PUBLIC :: SetBoundaryCondition => SetBoundaryCondition_ShallowWater
    PROCEDURE, PUBLIC :: SourceMethod => Source_ShallowWater
    PROCEDURE, PUBLIC :: FluxMethod => Flux_ShallowWater
    PROCEDURE, PUBLIC :: RiemannSolver => RiemannSolver_ShallowWater
  END TYPE ShallowWater

  TYPE, EXTENDS(Model2D), PUBLIC :: LinearShallowWater
    REAL(prec) :: g = 1.0_prec
    TYPE(MappedScalar2D) :: H       ! bathymetry
    TYPE(MappedScalar2D) :: fCori   ! Coriolis parameter
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_LinearShallowWater
    PROCEDURE, PUBLIC :: Free => Free_LinearShallowWater
    PROCEDURE, PUBLIC :: SetBathymetry => SetBathymetry_LinearShallowWater
    PROCEDURE, PUBLIC :: SetCoriolis => SetCoriolis_LinearShallowWater
    PROCEDURE, PUBLIC :: InitPlugins => InitPlugins_LinearShallowWater
    PROCEDURE, PUBLIC :: SetBoundaryCondition => SetBoundaryCondition_LinearShallowWater
    PROCEDURE, PUBLIC :: SourceMethod => Source_LinearShallowWater
    PROCEDURE, PUBLIC :: FluxMethod => Flux_LinearShallowWater
    PROCEDURE, PUBLIC :: RiemannSolver => RiemannSolver_LinearShallowWater
  END TYPE LinearShallowWater

  TYPE, EXTENDS(Model2D), PUBLIC :: CompressibleIdealGas2D
    REAL(prec) :: expansionFactor = 1.4_prec
    REAL(prec) :: R = 1.0_prec
    TYPE(MappedScalar2D) :: environmentals    ! (gravity potential, drag coefficient)
    TYPE(C_PTR) :: environmentalsGradient_dev = C_NULL_PTR
    TYPE(MappedScalar2D) :: primitive, diagnostics, entropyVars
    TYPE(C_PTR) :: prescribedSolution_dev = C_NULL_PTR, prescribedPrimitive_dev = C_NULL_PTR
    TYPE(C_PTR) :: prescribedDiagnostics_dev = C_NULL_PTR   ! boundary traces of the prescribed state
  CONTAINS
    PROCEDURE, PUBLIC :: Init => Init_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: Free => Free_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SetGravity => SetGravity_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SetDrag => SetDrag_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: InitPlugins => InitPlugins_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: PluginSetSolution => PluginSetSolution_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: PreTendency => PreTendency_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SetBoundaryCondition => SetBoundaryCondition_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: SourceMethod => Source_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: FluxMethod => Flux_CompressibleIdealGas2D
    PROCEDURE, PUBLIC :: RiemannSolver => RiemannSolver_CompressibleIdealGas2D
  END TYPE CompressibleIdealGas2D

CONTAINS

  ! Init_Model2D (src/SELF_Model2D.f90:142-178).  kind 0 (a direct extension of Model2D) has no
  ! engine: its tendency is the per-operator path over its own plug-in methods.
  SUBROUTINE Init_Model2D(this, kind, nvar, mesh, geometry)
    CLASS(Model2D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: kind, nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    this % kind = kind
    this % mesh => mesh
    this % geometry => geometry
    this % t = 0.0_prec
    this % plugins = .FALSE.
    CALL this % solution % Init(geometry % interp, nvar, mesh % nElem)
    CALL this % dSdt % Init(geometry % interp, nvar, mesh % nElem)
    IF (kind /= 0) THEN
      CALL selfamd_check(selfamd_model2d_create(this % engine, INT(kind, C_INT), INT(geometry % interp % N, C_INT), &
                         INT(mesh % nElem, C_INT), this % mesh % sideInfo, geometry % dsdx, geometry % J, &
                         geometry % nHatb, geometry % nScaleb, C_NULL_PTR), 'Model2D%Init')
    ELSE
      CALL this % EnablePlugins()
    END IF
  END SUBROUTINE Init_Model2D

  SUBROUTINE Free_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    IF (C_ASSOCIATED(this % engine)) CALL selfamd_check(selfamd_model2d_destroy(this % engine), 'Model2D%Free')
    this % engine = C_NULL_PTR
    CALL this % solution % Free()
    CALL this % dSdt % Free()
    IF (this % plugins) THEN
      CALL this % source % Free()
      CALL this % flux % Free()
      CALL this % fluxDivergence % Free()
      CALL this % workSol % Free()
    END IF
    this % plugins = .FALSE.
  END SUBROUTINE Free_Model2D

  ! Per-node metric terms (the reference's representation, needed on curved meshes) or one
  ! affine record per element (affine meshes); fused engine only
  SUBROUTINE SetNodalMetrics_Model2D(this, nodal)
    CLASS(Model2D), INTENT(inout) :: this
    LOGICAL, INTENT(in) :: nodal
    IF (.NOT. C_ASSOCIATED(this % engine)) RETURN
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_NODAL_METRICS, &
                       MERGE(1.0_prec, 0.0_prec, nodal)), 'Model2D%SetNodalMetrics')
  END SUBROUTINE SetNodalMetrics_Model2D

  ! Switches the model to the per-operator path: allocates its registers (and the physics'
  ! through InitPlugins) and takes the engine's current state
  SUBROUTINE EnablePlugins_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    INTEGER :: nVar, nEl
    IF (this % plugins) RETURN
    nVar = this % solution % nVar
    nEl = this % solution % nElem
    CALL this % source % Init(this % geometry % interp, nVar, nEl)
    CALL this % flux % Init(this % geometry % interp, nVar, nEl)
    CALL this % fluxDivergence % Init(this % geometry % interp, nVar, nEl)
    CALL this % workSol % Init(this % geometry % interp, nVar, nEl)
    this % plugins = .TRUE.
    IF (C_ASSOCIATED(this % engine)) &
      CALL selfamd_check(selfamd_model2d_get_solution(this % engine, this % solution % interior_dev), &
                         'Model2D%EnablePlugins')
    CALL this % solution % BoundaryInterp()
    CALL this % InitPlugins()
    CALL this % PluginSetSolution()
  END SUBROUTINE EnablePlugins_Model2D

  SUBROUTINE InitPlugins_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE InitPlugins_Model2D

  SUBROUTINE PluginSetSolution_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE PluginSetSolution_Model2D

  ! The default plug-in methods do nothing (src/SELF_Model.f90:242-290)
  SUBROUTINE PreTendency_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE PreTendency_Model2D

  SUBROUTINE SetBoundaryCondition_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE SetBoundaryCondition_Model2D

  SUBROUTINE Source_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE Source_Model2D

  SUBROUTINE Flux_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE Flux_Model2D

  SUBROUTINE RiemannSolver_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
  END SUBROUTINE RiemannSolver_Model2D

  ! The current state in solution % interior_dev (the engine keeps its own copy)
  SUBROUTINE StateToDevice_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    IF (this % plugins .OR. .NOT. C_ASSOCIATED(this % engine)) RETURN
    CALL selfamd_check(selfamd_model2d_get_solution(this % engine, this % solution % interior_dev), &
                       'Model2D%StateToDevice')
  END SUBROUTINE StateToDevice_Model2D

  ! After per-operator steps: the engine takes the state and time, for its diagnostics
  SUBROUTINE SyncEngine_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    IF (.NOT. C_ASSOCIATED(this % engine)) RETURN
    CALL hip_check(selfamd_device_synchronize(), 'Model2D%SyncEngine')
    CALL selfamd_check(selfamd_model2d_set_solution(this % engine, this % solution % interior_dev), 'Model2D%SyncEngine')
    CALL selfamd_check(selfamd_model2d_set_time(this % engine, this % t), 'Model2D%SyncEngine')
  END SUBROUTINE SyncEngine_Model2D

  ! SetSolution (src/SELF_Model2D.f90:313-332) from an array interior(0:N,0:N,nVar,nElem)
  SUBROUTINE SetSolution_Model2D(this, interior)
    CLASS(Model2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: interior(:,:,:,:)
    this % solution % interior = interior
    CALL this % solution % UpdateDevice()
    IF (C_ASSOCIATED(this % engine)) &
      CALL selfamd_check(selfamd_model2d_set_solution(this % engine, this % solution % interior_dev), &
                         'Model2D%SetSolution')
    IF (this % plugins) THEN
      CALL this % solution % BoundaryInterp()
      CALL this % PluginSetSolution()
    END IF
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

  ! CalculateTendency_Model2D (src/SELF_Model2D.f90:816-859): dSdt on the device, by the fused
  ! engine or, on the per-operator path, by the reference's sequence of plug-in methods and
  ! device operators
  SUBROUTINE CalculateTendency_Model2D(this)
    CLASS(Model2D), INTENT(inout) :: this
    IF (.NOT. this % plugins) THEN
      CALL selfamd_check(selfamd_model2d_tendency(this % engine, this % dSdt % interior_dev), &
                         'Model2D%CalculateTendency')
      RETURN
    END IF
    CALL this % PreTendency()
    CALL this % solution % BoundaryInterp()
    CALL this % solution % SideExchange(this % mesh)
    CALL this % SetBoundaryCondition()
    CALL this % SourceMethod()
    CALL this % RiemannSolver()
    CALL this % FluxMethod()
    CALL this % flux % ContravariantProjection(this % geometry)
    CALL this % flux % Divergence(this % geometry, this % fluxDivergence)
    CALL selfamd_check(selfamd_calculate_dsdt(this % fluxDivergence % interior_dev, this % source % interior_dev, &
                       this % dSdt % interior_dev, INT(this % solution % interp % N, C_INT), &
                       INT(this % solution % nVar, C_INT), INT(this % solution % nElem, C_INT), C_NULL_PTR), &
                       'Model2D%CalculateTendency')
  END SUBROUTINE CalculateTendency_Model2D

  ! UpdateGRK3_Model2D (src/SELF_Model2D.f90:710-752): G = a_m G + dSdt, S = S + g_m dt G
  SUBROUTINE UpdateGRK3_Model2D(this, m)
    CLASS(Model2D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: m
    CALL selfamd_check(selfamd_update_grk(this % workSol % interior_dev, this % solution % interior_dev, &
                       this % dSdt % interior_dev, rk3_a(m), rk3_g(m), this % dt, &
                       INT(this % solution % interp % N, C_INT), INT(this % solution % nVar, C_INT), &
                       INT(this % solution % nElem, C_INT), C_NULL_PTR), 'Model2D%UpdateGRK3')
  END SUBROUTINE UpdateGRK3_Model2D

  ! ForwardStep_Model + LowStorageRK3_timeIntegrator (src/SELF_Model.f90:474-523, 704-732).
  ! Fused engine: the reference's step loop dt = MIN(dtLim, tn - t) is replayed on the host,
  ! full steps are batched into one engine call and a shorter final step gets its own call.
  ! Per-operator path: the reference's loop itself, three CalculateTendency / UpdateGRK3 per
  ! step, with the stage times of the reference.
  SUBROUTINE ForwardStep_Model2D(this, tn, dt)
    CLASS(Model2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: tn
    REAL(prec), OPTIONAL, INTENT(in) :: dt
    REAL(prec) :: t, dtStep, dtLim, t0
    INTEGER :: nFull, m
    IF (PRESENT(dt)) this % dt = dt
    IF (this % plugins) THEN
      dtLim = this % dt
      DO WHILE (this % t < tn)
        t0 = this % t
        this % dt = MIN(dtLim, tn - this % t)
        DO m = 1, 3
          CALL this % CalculateTendency()
          CALL this % UpdateGRK3(m)
          this % t = t0 + rk3_b(m)*this % dt
        END DO
        this % t = t0 + this % dt
      END DO
      this % dt = dtLim
      CALL this % SyncEngine()
      RETURN
    END IF
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
    CALL this % StateToDevice()
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
    IF (.NOT. C_ASSOCIATED(this % engine)) RETURN  ! a kind-0 model defines its own
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
    CALL this % StateToDevice()
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
    IF (.NOT. C_ASSOCIATED(this % engine)) THEN
      PRINT '(A)', 'self_amd: Model2D%Integral needs one of the built-in physics (kind 1 to 3)'
      ERROR STOP 2
    END IF
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

  ! ---- shared helpers of the physics' per-operator methods ----
  INTEGER(C_INT) FUNCTION nN(this)
    CLASS(Model2D), INTENT(in) :: this
    nN = INT(this % solution % interp % N, C_INT)
  END FUNCTION nN

  INTEGER(C_INT) FUNCTION nE(this)
    CLASS(Model2D), INTENT(in) :: this
    nE = INT(this % solution % nElem, C_INT)
  END FUNCTION nE

  FUNCTION nodes_bytes(this, k) RESULT(nb)
    CLASS(Model2D), INTENT(in) :: this
    INTEGER, INTENT(in) :: k   ! doubles per node
    INTEGER(C_SIZE_T) :: nb
    nb = INT(k, C_SIZE_T)*INT(this % solution % interp % N + 1, C_SIZE_T)**2*INT(this % solution % nElem, C_SIZE_T) &
         *8_C_SIZE_T
  END FUNCTION nodes_bytes

  FUNCTION trace_bytes(this, k) RESULT(nb)
    CLASS(Model2D), INTENT(in) :: this
    INTEGER, INTENT(in) :: k   ! doubles per face node
    INTEGER(C_SIZE_T) :: nb
    nb = INT(k, C_SIZE_T)*INT(this % solution % interp % N + 1, C_SIZE_T)*4_C_SIZE_T* &
         INT(this % solution % nElem, C_SIZE_T)*8_C_SIZE_T
  END FUNCTION trace_bytes

  ! ---- ShallowWater (src/models/SELF_ShallowWater.f90:96-147) ----
  SUBROUTINE Init_ShallowWater(this, nvar, mesh, geometry)
    CLASS(ShallowWater), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    TYPE(ShallowWater) :: base
    CALL this % Init_Model2D(SELF_MODEL_SW, 3, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_G, this % g), 'ShallowWater%Init')
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_FCORI, this % fCori), 'ShallowWater%Init')
    CALL this % H % Init(geometry % interp, 1, mesh % nElem)
    ! a type extension may override plug-in methods: it runs the per-operator path
    IF (.NOT. SAME_TYPE_AS(this, base)) CALL this % EnablePlugins()
  END SUBROUTINE Init_ShallowWater

  SUBROUTINE Free_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    IF (this % plugins) THEN
      CALL this % velocity % Free()
      CALL device_free(this % gradH_dev)
    END IF
    CALL this % H % Free()
    CALL Free_Model2D(this)
  END SUBROUTINE Free_ShallowWater

  ! SetTopography (src/models/SELF_ShallowWater.f90:241-282), topography(0:N,0:N,nElem):
  ! H, its traces and its gradient (GradientSF)
  SUBROUTINE SetTopography_ShallowWater(this, topography)
    CLASS(ShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: topography(:,:,:)
    this % H % interior = RESHAPE(topography, SHAPE(this % H % interior))
    CALL this % H % UpdateDevice()
    CALL this % SetField(FIELD_TOPOGRAPHY, topography)
    IF (this % plugins) CALL topography_derived(this)
  END SUBROUTINE SetTopography_ShallowWater

  SUBROUTINE topography_derived(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL this % H % BoundaryInterp()
    CALL selfamd_check(selfamd_gradient_sf_2d(this % geometry % interp % dMatrix_dev, this % geometry % dsdx, &
                       this % geometry % J, this % H % interior_dev, this % gradH_dev, nN(this), 1_C_INT, nE(this), &
                       C_NULL_PTR), 'ShallowWater%SetTopography')
  END SUBROUTINE topography_derived

  SUBROUTINE InitPlugins_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL this % velocity % Init(this % geometry % interp, 1, this % solution % nElem)
    this % gradH_dev = device_alloc(nodes_bytes(this, 2))
    CALL topography_derived(this)
  END SUBROUTINE InitPlugins_ShallowWater

  ! PreTendency_ShallowWater (src/models/SELF_ShallowWater.f90:190-239): velocity from the
  ! interior and from the solution traces as they stand (the previous stage's: the velocity
  ! traces lag one stage), then exchanged across sides
  SUBROUTINE PreTendency_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_sw2d_velocity(this % solution % interior_dev, this % solution % boundary_dev, &
                       this % velocity % interior_dev, this % velocity % boundary_dev, nN(this), nE(this), &
                       C_NULL_PTR), 'ShallowWater%PreTendency')
    CALL this % velocity % SideExchange(this % mesh)
  END SUBROUTINE PreTendency_ShallowWater

  ! SetBoundaryCondition_ShallowWater (src/models/SELF_ShallowWater.f90:284-343)
  SUBROUTINE SetBoundaryCondition_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_sw2d_set_boundary_condition(this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % H % boundary_dev, this % geometry % nHatb, &
                       this % mesh % sideInfo_dev, nN(this), nE(this), C_NULL_PTR), 'ShallowWater%SetBoundaryCondition')
  END SUBROUTINE SetBoundaryCondition_ShallowWater

  ! Source_ShallowWater (src/models/SELF_ShallowWater.f90:345-384)
  SUBROUTINE Source_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_sw2d_source(this % source % interior_dev, this % solution % interior_dev, &
                       this % gradH_dev, this % g, this % fCori, nN(this), nE(this), C_NULL_PTR), &
                       'ShallowWater%SourceMethod')
  END SUBROUTINE Source_ShallowWater

  ! Flux_ShallowWater (src/models/SELF_ShallowWater.f90:386-449)
  SUBROUTINE Flux_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_sw2d_flux(this % flux % interior_dev, this % solution % interior_dev, &
                       this % velocity % interior_dev, this % g, nN(this), nE(this), C_NULL_PTR), &
                       'ShallowWater%FluxMethod')
  END SUBROUTINE Flux_ShallowWater

  ! RiemannSolver_ShallowWater (src/models/SELF_ShallowWater.f90:451-529), Rusanov
  SUBROUTINE RiemannSolver_ShallowWater(this)
    CLASS(ShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_sw2d_riemann(this % flux % boundaryNormal_dev, this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % velocity % boundary_dev, &
                       this % velocity % extBoundary_dev, this % geometry % nHatb, this % geometry % nScaleb, &
                       this % g, nN(this), nE(this), C_NULL_PTR), 'ShallowWater%RiemannSolver')
  END SUBROUTINE RiemannSolver_ShallowWater

  ! ---- LinearShallowWater (src/models/SELF_LinearShallowWater.f90:101-151) ----
  SUBROUTINE Init_LinearShallowWater(this, nvar, mesh, geometry)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    TYPE(LinearShallowWater) :: base
    REAL(prec), ALLOCATABLE :: zero(:,:,:)
    CALL this % Init_Model2D(SELF_MODEL_LSW, 3, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_G, this % g), 'LinearShallowWater%Init')
    CALL this % H % Init(geometry % interp, 1, mesh % nElem)
    CALL this % fCori % Init(geometry % interp, 1, mesh % nElem)
    ALLOCATE (zero(0:geometry % interp % N, 0:geometry % interp % N, mesh % nElem))
    zero = 0.0_prec
    CALL this % SetCoriolis(zero)  ! fCori starts at zero (hfReal allocation zero-fills)
    IF (.NOT. SAME_TYPE_AS(this, base)) CALL this % EnablePlugins()
  END SUBROUTINE Init_LinearShallowWater

  SUBROUTINE Free_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL this % H % Free()
    CALL this % fCori % Free()
    CALL Free_Model2D(this)
  END SUBROUTINE Free_LinearShallowWater

  ! SetBathymetry (src/models/SELF_LinearShallowWater.f90:251-296): H and its traces
  SUBROUTINE SetBathymetry_LinearShallowWater(this, H)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: H(:,:,:)
    this % H % interior = RESHAPE(H, SHAPE(this % H % interior))
    CALL this % H % UpdateDevice()
    CALL this % SetField(FIELD_BATHYMETRY, H)
    IF (this % plugins) CALL this % H % BoundaryInterp()
  END SUBROUTINE SetBathymetry_LinearShallowWater

  ! SetCoriolis (src/models/SELF_LinearShallowWater.f90:218-249)
  SUBROUTINE SetCoriolis_LinearShallowWater(this, f)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: f(:,:,:)
    this % fCori % interior = RESHAPE(f, SHAPE(this % fCori % interior))
    CALL this % fCori % UpdateDevice()
    CALL this % SetField(FIELD_CORIOLIS, f)
  END SUBROUTINE SetCoriolis_LinearShallowWater

  SUBROUTINE InitPlugins_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL this % H % BoundaryInterp()
  END SUBROUTINE InitPlugins_LinearShallowWater

  ! SetBoundaryCondition_LinearShallowWater (src/models/SELF_LinearShallowWater.f90:298-355)
  SUBROUTINE SetBoundaryCondition_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_lsw2d_set_boundary_condition(this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % geometry % nHatb, this % mesh % sideInfo_dev, &
                       nN(this), nE(this), C_NULL_PTR), 'LinearShallowWater%SetBoundaryCondition')
  END SUBROUTINE SetBoundaryCondition_LinearShallowWater

  ! Source_LinearShallowWater (src/models/SELF_LinearShallowWater.f90:413-446)
  SUBROUTINE Source_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_lsw2d_source(this % source % interior_dev, this % solution % interior_dev, &
                       this % fCori % interior_dev, nN(this), nE(this), C_NULL_PTR), 'LinearShallowWater%SourceMethod')
  END SUBROUTINE Source_LinearShallowWater

  ! Flux_LinearShallowWater (src/models/SELF_LinearShallowWater.f90:448-498)
  SUBROUTINE Flux_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_lsw2d_flux(this % flux % interior_dev, this % solution % interior_dev, &
                       this % H % interior_dev, this % g, nN(this), nE(this), C_NULL_PTR), &
                       'LinearShallowWater%FluxMethod')
  END SUBROUTINE Flux_LinearShallowWater

  ! RiemannSolver_LinearShallowWater (src/models/SELF_LinearShallowWater.f90:500-557)
  SUBROUTINE RiemannSolver_LinearShallowWater(this)
    CLASS(LinearShallowWater), INTENT(inout) :: this
    CALL selfamd_check(selfamd_lsw2d_riemann(this % flux % boundaryNormal_dev, this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % H % boundary_dev, this % geometry % nHatb, &
                       this % geometry % nScaleb, this % g, nN(this), nE(this), C_NULL_PTR), &
                       'LinearShallowWater%RiemannSolver')
  END SUBROUTINE RiemannSolver_LinearShallowWater

  ! ---- CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:261-358) ----
  SUBROUTINE Init_CompressibleIdealGas2D(this, nvar, mesh, geometry)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: nvar
    TYPE(Mesh2D), INTENT(in), TARGET :: mesh
    TYPE(SEMQuad), INTENT(in), TARGET :: geometry
    TYPE(CompressibleIdealGas2D) :: base
    CALL this % Init_Model2D(SELF_MODEL_EULER, 4, mesh, geometry)
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_GAMMA, this % expansionFactor), &
                       'CompressibleIdealGas2D%Init')
    CALL selfamd_check(selfamd_model2d_set_param(this % engine, PARAM_RGAS, this % R), 'CompressibleIdealGas2D%Init')
    CALL this % environmentals % Init(geometry % interp, 2, mesh % nElem)
    IF (.NOT. SAME_TYPE_AS(this, base)) CALL this % EnablePlugins()
  END SUBROUTINE Init_CompressibleIdealGas2D

  SUBROUTINE Free_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    IF (this % plugins) THEN
      CALL this % primitive % Free()
      CALL this % diagnostics % Free()
      CALL this % entropyVars % Free()
      CALL device_free(this % environmentalsGradient_dev)
      CALL device_free(this % prescribedSolution_dev)
      CALL device_free(this % prescribedPrimitive_dev)
      CALL device_free(this % prescribedDiagnostics_dev)
    END IF
    CALL this % environmentals % Free()
    CALL Free_Model2D(this)
  END SUBROUTINE Free_CompressibleIdealGas2D

  ! SetGravity (src/models/SELF_CompressibleIdealGas2D.f90:608-630): the potential is
  ! environmentals variable 1; GradientSF of both environmentals variables
  SUBROUTINE SetGravity_CompressibleIdealGas2D(this, potential)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: potential(:,:,:)
    this % environmentals % interior(:,:,1,:) = potential
    CALL this % environmentals % UpdateDevice()
    CALL this % SetField(FIELD_POTENTIAL, potential)
    IF (this % plugins) CALL environmentals_gradient(this)
  END SUBROUTINE SetGravity_CompressibleIdealGas2D

  ! SetDrag (src/models/SELF_CompressibleIdealGas2D.f90:632-672): environmentals variable 2
  SUBROUTINE SetDrag_CompressibleIdealGas2D(this, Cd)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: Cd(:,:,:)
    this % environmentals % interior(:,:,2,:) = Cd
    CALL this % environmentals % UpdateDevice()
    CALL this % SetField(FIELD_DRAG, Cd)
  END SUBROUTINE SetDrag_CompressibleIdealGas2D

  SUBROUTINE environmentals_gradient(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_gradient_sf_2d(this % geometry % interp % dMatrix_dev, this % geometry % dsdx, &
                       this % geometry % J, this % environmentals % interior_dev, this % environmentalsGradient_dev, &
                       nN(this), 2_C_INT, nE(this), C_NULL_PTR), 'CompressibleIdealGas2D%SetGravity')
  END SUBROUTINE environmentals_gradient

  SUBROUTINE InitPlugins_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL this % primitive % Init(this % geometry % interp, 4, this % solution % nElem)
    CALL this % diagnostics % Init(this % geometry % interp, 4, this % solution % nElem)
    CALL this % entropyVars % Init(this % geometry % interp, 4, this % solution % nElem)
    this % environmentalsGradient_dev = device_alloc(nodes_bytes(this, 4))
    this % prescribedSolution_dev = device_alloc(trace_bytes(this, 4))
    this % prescribedPrimitive_dev = device_alloc(trace_bytes(this, 4))
    this % prescribedDiagnostics_dev = device_alloc(trace_bytes(this, 4))
    CALL environmentals_gradient(this)
  END SUBROUTINE InitPlugins_CompressibleIdealGas2D

  ! the prescribed boundary state := traces of the new solution, its primitive variables and
  ! diagnostics (SetPrescribedSolutionFromSolution, src/models/SELF_CompressibleIdealGas2D.f90:694-740)
  SUBROUTINE PluginSetSolution_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL euler_nodal(this, .FALSE.)
    CALL this % solution % interp % ScalarBoundaryInterp_2D(this % solution % interior_dev, &
                                                             this % prescribedSolution_dev, 4, this % solution % nElem)
    CALL this % solution % interp % ScalarBoundaryInterp_2D(this % primitive % interior_dev, &
                                                             this % prescribedPrimitive_dev, 4, this % solution % nElem)
    CALL this % solution % interp % ScalarBoundaryInterp_2D(this % diagnostics % interior_dev, &
                                                             this % prescribedDiagnostics_dev, 4, this % solution % nElem)
  END SUBROUTINE PluginSetSolution_CompressibleIdealGas2D

  ! CalculateDiagnostics, ConservativeToPrimitive and (optionally) ConservativeToEntropy
  ! (src/models/SELF_CompressibleIdealGas2D.f90:984-1215) of the interior state
  SUBROUTINE euler_nodal(this, entropy)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    LOGICAL, INTENT(in) :: entropy
    CALL selfamd_check(selfamd_euler2d_diagnostics(this % solution % interior_dev, this % diagnostics % interior_dev, &
                       this % expansionFactor, this % R, nN(this), nE(this), C_NULL_PTR), 'CalculateDiagnostics')
    CALL selfamd_check(selfamd_euler2d_primitive(this % solution % interior_dev, this % primitive % interior_dev, &
                       this % expansionFactor, nN(this), nE(this), C_NULL_PTR), 'ConservativeToPrimitive')
    IF (entropy) CALL selfamd_check(selfamd_euler2d_entropy_vars(this % solution % interior_dev, &
                                    this % entropyVars % interior_dev, this % expansionFactor, nN(this), nE(this), &
                                    C_NULL_PTR), 'ConservativeToEntropy')
  END SUBROUTINE euler_nodal

  ! PreTendency_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:955-982)
  SUBROUTINE PreTendency_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL euler_nodal(this, .TRUE.)
    CALL this % primitive % BoundaryInterp()
    CALL this % diagnostics % BoundaryInterp()
    CALL this % entropyVars % BoundaryInterp()
    CALL this % primitive % SideExchange(this % mesh)
    CALL this % diagnostics % SideExchange(this % mesh)
    CALL this % entropyVars % SideExchange(this % mesh)
  END SUBROUTINE PreTendency_CompressibleIdealGas2D

  ! SetBoundaryCondition_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:1217-1323)
  SUBROUTINE SetBoundaryCondition_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_euler2d_set_boundary_condition(this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % prescribedSolution_dev, this % primitive % boundary_dev, &
                       this % primitive % extBoundary_dev, this % prescribedPrimitive_dev, &
                       this % diagnostics % boundary_dev, this % diagnostics % extBoundary_dev, &
                       this % prescribedDiagnostics_dev, this % geometry % nHatb, this % mesh % sideInfo_dev, nN(this), &
                       nE(this), C_NULL_PTR), 'CompressibleIdealGas2D%SetBoundaryCondition')
  END SUBROUTINE SetBoundaryCondition_CompressibleIdealGas2D

  ! Source_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:1325-1366)
  SUBROUTINE Source_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_euler2d_source(this % source % interior_dev, this % solution % interior_dev, &
                       this % environmentals % interior_dev, this % environmentalsGradient_dev, nN(this), nE(this), &
                       C_NULL_PTR), 'CompressibleIdealGas2D%SourceMethod')
  END SUBROUTINE Source_CompressibleIdealGas2D

  ! Flux_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:1368-1440)
  SUBROUTINE Flux_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_euler2d_flux(this % flux % interior_dev, this % solution % interior_dev, &
                       this % primitive % interior_dev, this % diagnostics % interior_dev, nN(this), nE(this), &
                       C_NULL_PTR), 'CompressibleIdealGas2D%FluxMethod')
  END SUBROUTINE Flux_CompressibleIdealGas2D

  ! NaiveLLF_CompressibleIdealGas2D (src/models/SELF_CompressibleIdealGas2D.f90:1455-1542)
  SUBROUTINE RiemannSolver_CompressibleIdealGas2D(this)
    CLASS(CompressibleIdealGas2D), INTENT(inout) :: this
    CALL selfamd_check(selfamd_euler2d_riemann(this % flux % boundaryNormal_dev, this % solution % boundary_dev, &
                       this % solution % extBoundary_dev, this % primitive % boundary_dev, &
                       this % primitive % extBoundary_dev, this % diagnostics % boundary_dev, &
                       this % diagnostics % extBoundary_dev, this % geometry % nHatb, this % geometry % nScaleb, &
                       nN(this), nE(this), C_NULL_PTR), 'CompressibleIdealGas2D%RiemannSolver')
  END SUBROUTINE RiemannSolver_CompressibleIdealGas2D

END MODULE SELF_Model2D
