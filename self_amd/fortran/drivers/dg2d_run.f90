! dg2d_run -- runs one 2D DGSEM model on a structured mesh through the Fortran API, in the
! manner of the reference's example programs (test/ShallowWater2D_*.f90): Lagrange%Init,
! Mesh2D, SEMQuad%GenerateFromMesh, <Model>%Init, Set<Field>, SetSolution,
! CalculateTendency, ForwardStep.
!
! Usage: dg2d_run <parameter file>.  The namelist &run gives the model (1 LSW, 2 SW,
! 3 Euler), N, nx, ny, periodic, bcid, nSteps, dt, nodal (metric representation) and
! prefix, and optionally meshFile, a HOPr mesh file read by Read_HOPr in place of the
! structured nx x ny mesh (bcid > 0 then resets its boundary conditions, as the reference's
! programs do with ResetBoundaryConditionType); <prefix>_sol.bin holds the initial interior(0:N,0:N,nVar,nElem) and
! <prefix>_field1.bin / _field2.bin the static fields (0:N,0:N,nElem) in native doubles.
! Outputs: <prefix>_tend.bin (tendency of the initial state) and <prefix>_out.bin
! (solution after nSteps RK3 steps).  With Mtarget >= 1 in the namelist the interpolant's target
! grid is Mtarget+1 uniform points and the final state is also written by WriteTecplot to <prefix>.tec.
! With plugins = .true. the model is a user type extension of the physics (module
! user_models below), so CalculateTendency and ForwardStep take the per-operator path through
! the plug-in methods; massSource /= 0 selects the extension of ShallowWater whose own
! SourceMethod adds that constant to the thickness tendency.
MODULE user_models
  USE SELF_Constants
  USE SELF_Model2D
  IMPLICIT NONE

  ! extensions that override nothing: the physics' own plug-in methods, per operator
  TYPE, EXTENDS(ShallowWater) :: UserShallowWater
  END TYPE UserShallowWater
  TYPE, EXTENDS(LinearShallowWater) :: UserLinearShallowWater
  END TYPE UserLinearShallowWater
  TYPE, EXTENDS(CompressibleIdealGas2D) :: UserCompressibleIdealGas2D
  END TYPE UserCompressibleIdealGas2D

  ! a user's own source method, written on host arrays: the parent's source plus a constant
  ! mass source
  TYPE, EXTENDS(ShallowWater) :: ForcedShallowWater
    REAL(prec) :: massSource = 0.0_prec
  CONTAINS
    PROCEDURE :: SourceMethod => Source_ForcedShallowWater
  END TYPE ForcedShallowWater

CONTAINS

  SUBROUTINE Source_ForcedShallowWater(this)
    CLASS(ForcedShallowWater), INTENT(inout) :: this
    CALL this % ShallowWater % SourceMethod()
    CALL this % source % UpdateHost()
    this % source % interior(:,:,3,:) = this % source % interior(:,:,3,:) + this % massSource
    CALL this % source % UpdateDevice()
  END SUBROUTINE Source_ForcedShallowWater

END MODULE user_models

PROGRAM dg2d_run
  USE SELF_Constants
  USE SELF_Lagrange
  USE SELF_Mesh
  USE SELF_Model2D
  USE SELF_IO
  USE user_models
  IMPLICIT NONE
  INTEGER :: model = 2, N = 7, Mtarget = -1, nx = 4, ny = 4, bcid = SELF_BC_NONORMALFLOW, nSteps = 10
  LOGICAL :: periodic = .FALSE., nodal = .FALSE., plugins = .FALSE.
  REAL(prec) :: dt = 1.0e-4_prec, massSource = 0.0_prec
  CHARACTER(LEN=400) :: prefix = 'dg2d', meshFile = '', parfile
  NAMELIST /run/ model, N, Mtarget, nx, ny, periodic, bcid, nSteps, dt, nodal, prefix, meshFile, plugins, &
    massSource
  TYPE(Lagrange), TARGET :: interp
  TYPE(Mesh2D), TARGET :: mesh
  TYPE(SEMQuad), TARGET :: geometry
  TYPE(ShallowWater) :: sw
  TYPE(LinearShallowWater) :: lsw
  TYPE(CompressibleIdealGas2D) :: euler
  TYPE(UserShallowWater) :: usw
  TYPE(UserLinearShallowWater) :: ulsw
  TYPE(UserCompressibleIdealGas2D) :: ueuler
  TYPE(ForcedShallowWater) :: fsw
  REAL(prec), ALLOCATABLE :: sol(:,:,:,:), f1(:,:,:), f2(:,:,:)
  INTEGER :: u, nVar

  CALL GET_COMMAND_ARGUMENT(1, parfile)
  OPEN (NEWUNIT=u, FILE=TRIM(parfile), STATUS='old', ACTION='read')
  READ (u, NML=run)
  CLOSE (u)
  nVar = MERGE(4, 3, model == 3)

  IF (Mtarget >= 1) THEN
    CALL interp % Init(N, GAUSS, Mtarget, UNIFORM)
  ELSE
    CALL interp % Init(N, GAUSS, N, GAUSS)
  END IF
  IF (LEN_TRIM(meshFile) > 0) THEN
    CALL Read_HOPr(mesh, TRIM(meshFile))
    IF (bcid > 0) CALL mesh % ResetBoundaryConditionType(bcid)
  ELSE
    CALL mesh % StructuredMesh(nx, ny, 0.0_prec, 1.0_prec, 0.0_prec, 1.0_prec, periodic, periodic, bcid)
  END IF
  CALL geometry % Init(interp, mesh % nElem)
  CALL geometry % GenerateFromMesh(mesh)

  ALLOCATE (sol(0:N, 0:N, nVar, mesh % nElem), f1(0:N, 0:N, mesh % nElem), f2(0:N, 0:N, mesh % nElem))
  CALL read_array(TRIM(prefix)//'_sol.bin', sol, SIZE(sol))
  CALL read_array(TRIM(prefix)//'_field1.bin', f1, SIZE(f1))
  CALL read_array(TRIM(prefix)//'_field2.bin', f2, SIZE(f2))

  IF (plugins) THEN
    SELECT CASE (model)
    CASE (1)
      CALL ulsw % Init(nVar, mesh, geometry)
      CALL ulsw % SetBathymetry(f1)
      CALL ulsw % SetCoriolis(f2)
      CALL run_model(ulsw)
    CASE (2)
      IF (massSource /= 0.0_prec) THEN
        fsw % massSource = massSource
        CALL fsw % Init(nVar, mesh, geometry)
        CALL fsw % SetTopography(f1)
        CALL run_model(fsw)
      ELSE
        CALL usw % Init(nVar, mesh, geometry)
        CALL usw % SetTopography(f1)
        CALL run_model(usw)
      END IF
    CASE (3)
      CALL ueuler % Init(nVar, mesh, geometry)
      CALL ueuler % SetGravity(f1)
      CALL ueuler % SetDrag(f2)
      CALL run_model(ueuler)
    CASE DEFAULT
      PRINT *, 'dg2d_run: unknown model ', model
      ERROR STOP 2
    END SELECT
  ELSE
    SELECT CASE (model)
    CASE (1)
      CALL lsw % Init(nVar, mesh, geometry)
      CALL lsw % SetNodalMetrics(nodal)
      CALL lsw % SetBathymetry(f1)
      CALL lsw % SetCoriolis(f2)
      CALL run_model(lsw)
    CASE (2)
      CALL sw % Init(nVar, mesh, geometry)
      CALL sw % SetNodalMetrics(nodal)
      CALL sw % SetTopography(f1)
      CALL run_model(sw)
    CASE (3)
      CALL euler % Init(nVar, mesh, geometry)
      CALL euler % SetNodalMetrics(nodal)
      CALL euler % SetGravity(f1)
      CALL euler % SetDrag(f2)
      CALL run_model(euler)
    CASE DEFAULT
      PRINT *, 'dg2d_run: unknown model ', model
      ERROR STOP 2
    END SELECT
  END IF
  PRINT '(A,I0,A,I0,A)', 'dg2d_run: model ', model, ' done (', nSteps, ' RK3 steps)'

CONTAINS

  SUBROUTINE run_model(m)
    CLASS(Model2D), INTENT(inout) :: m
    INTEGER :: v
    CALL m % SetTimeIntegrator('rk3')
    CALL m % SetSolution(sol)
    CALL m % CalculateTendency()
    CALL m % UpdateHost()
    CALL write_array(TRIM(prefix)//'_tend.bin', m % dSdt % interior, SIZE(m % dSdt % interior))
    ! entropy and mass (H, eta or rho: variable 3) of the initial state
    CALL m % CalculateEntropy()
    CALL m % ReportEntropy()
    OPEN (NEWUNIT=v, FILE=TRIM(prefix)//'_diag.txt', STATUS='replace', ACTION='write')
    WRITE (v, '(2ES26.17)') m % entropy, m % Integral(3)
    CLOSE (v)
    m % t = 0.0_prec
    CALL m % ForwardStep(tn=REAL(nSteps, prec)*dt, dt=dt)
    CALL m % UpdateHost()
    CALL write_array(TRIM(prefix)//'_out.bin', m % solution % interior, SIZE(m % solution % interior))
    IF (Mtarget >= 1) CALL m % WriteTecplot(TRIM(prefix)//'.tec')
    ! pickup file of the final state, then read back into the model
    CALL WriteModel(m, TRIM(prefix)//'_pickup.h5')
    CALL ReadModel(m, TRIM(prefix)//'_pickup.h5')
    CALL m % UpdateHost()
    CALL write_array(TRIM(prefix)//'_reread.bin', m % solution % interior, SIZE(m % solution % interior))
    IF (m % plugins) PRINT '(A)', 'dg2d_run: per-operator path (plug-in methods)'
    CALL m % Free()
  END SUBROUTINE run_model

  SUBROUTINE read_array(path, a, n)
    CHARACTER(*), INTENT(in) :: path
    INTEGER, INTENT(in) :: n
    REAL(prec), INTENT(out) :: a(n)
    INTEGER :: v
    OPEN (NEWUNIT=v, FILE=path, ACCESS='stream', FORM='unformatted', STATUS='old', ACTION='read')
    READ (v) a
    CLOSE (v)
  END SUBROUTINE read_array

  SUBROUTINE write_array(path, a, n)
    CHARACTER(*), INTENT(in) :: path
    INTEGER, INTENT(in) :: n
    REAL(prec), INTENT(in) :: a(n)
    INTEGER :: v
    OPEN (NEWUNIT=v, FILE=path, ACCESS='stream', FORM='unformatted', STATUS='replace', ACTION='write')
    WRITE (v) a
    CLOSE (v)
  END SUBROUTINE write_array

END PROGRAM dg2d_run
