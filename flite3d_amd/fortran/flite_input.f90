! flite_input -- <case>.inp reading for the Fortran host.
!
! The control file is the reference's NAMELIST /InputVariables/ ivd
! (SolverSource/mgnsg3d_def_par.F90:201-422), so InputVariablesData keeps the
! component names and kinds of the reference type (DEF:64-196): a namelist
! read accepts exactly the names its type declares.  REAL is 8 bytes as in
! the reference build (-r8).  readInputVariables sets the reference's
! defaults (DEF:217-379), reads the namelist and applies the same
! post-processing (DEF:386-415); to_flite_params hands the hot path's
! scalars to the C-ABI.
module flite_input
  use, intrinsic :: iso_c_binding
  use flite_gpu, only: flite_params
  implicit none
  private
  public :: InputVariablesData, readInputVariables, check_supported, to_flite_params

  integer, parameter :: rk = c_double

  type InputVariablesData
     integer :: numberOfProcesses
     character :: dataDirectory*120
     character :: solverCommand*200
     integer :: numberOfMGIterations
     integer :: numberOfSGIterations
     integer :: numberOfCFLIncrements
     integer :: writeToFileInterval
     real(rk) :: ReynoldsNumber, MachNumber, PrandtlNumber, CFLNumber
     real(rk) :: turbulentPrandtlNumber, turbulentCFLNumber
     real(rk) :: inflowTemperature
     real(rk) :: gamma
     real(rk) :: alpha
     real(rk) :: beta
     real(rk) :: wallTemperature
     real(rk) :: HartensCorrectionFactor
     logical :: useMatrixDissipation
     real(rk) :: secondOrderDissipationFactor, fourthOrderDissipationFactor
     real(rk) :: coarseGridDissipationFactor
     integer :: numberOfRelaxationSteps
     integer :: numberOfRSSteps
     real(rk) :: inflowField(6, 2)
     logical :: patchInitialization
     integer :: patchType
     real(rk) :: patchBox(3, 2)
     real(rk) :: patchSphere(4)
     integer :: HLLCFluxOrder
     integer :: HLLCICLS
     integer :: Automaticlipping
     integer :: GradientMethod
     real(rk) :: HLLCClipping
     logical :: explicit
     real(rk) :: residualSmoothingFactor
     real(rk) :: prolongationSmoothingFactor
     integer :: numberOfPSSteps
     integer :: prolongationMappingScheme
     integer :: viscosityScheme
     integer :: numberOfTurbulenceSteps
     real(rk) :: turbulenceSmoothingFactor
     real(rk) :: maxIncrementFactor
     real(rk) :: prolongationRelaxation
     integer :: engineFlowType
     real(rk) :: enginesFrontMassFlow(2)
     real(rk) :: enginesRearParameters(7, 2)
     real(rk) :: engineBCRelaxation
     real(rk) :: datum(3)
     real(rk) :: nwheels
     real(rk) :: wheelCentre(4, 3)
     real(rk) :: wheelDimensions(4, 2)
     real(rk) :: groundplane
     real(rk) :: tripFactor
     real(rk) :: maxTurbulenceValue
     real(rk) :: firstoffwallpoint
     real(rk) :: kinfoverUinf2
     real(rk) :: maxTurbulenceValueK, maxTurbulenceValueW, minTurbulenceValueW, maxTurbulenceValueMU
     real(rk) :: winf
     real(rk) :: nuwall
     logical :: SAS
     real(rk) :: XPER(2), YPER(2), ZPER(2)
     logical :: GUSTMODEL
     real(rk) :: UDS
     integer :: typegust
     integer :: TTS
     integer :: nspan
     integer :: spanCor
     real(rk) :: tipspan, rootspan
     logical :: wallsAreIsentropic
     logical :: movingwall
     real(rk) :: turbProlongationRelaxation
     logical :: addTimeStep
     integer :: flowType
     integer :: timeScheme
     integer :: numberOfStepsPerCycle
     integer :: cycleNumber
     integer :: numberofphysicaltimesteps
     integer :: maximumnumberofcycles
     integer :: restartNumber
     real(rk) :: physicalTimeStep
     real(rk) :: stopTime
     real(rk) :: currentTime
     integer :: fluxType
     real(rk) :: referenceArea
     integer :: forceCalcType
     integer :: boundaryTerm
     real(rk) :: multigridBCRelaxation
     real(rk) :: minimumDensity
     logical :: useDissipationWeighting
     integer :: numberOfTurbulenceGrids
     integer :: numberOfEORelaxationSteps
     integer :: numberOfTriggerSteps
     real(rk) :: turbulenceTriggerValue
     real(rk) :: triggerRadius
     real(rk) :: turbK
     integer :: multigridScheme
     integer :: turbulenceModel
     integer :: dissipationScheme
     integer :: coarseGridDissipationScheme
     integer :: numberOfGridsToUse
     logical :: useTimeResidual
     real(rk) :: initialResidual
  end type InputVariablesData

contains

  ! defaults (DEF:217-379), namelist read from `unit` (none when unit <= 0)
  ! and post-processing (DEF:386-415)
  subroutine readInputVariables(ivd, unit)
    type(InputVariablesData), intent(inout) :: ivd
    integer, intent(in) :: unit
    real(rk) :: PI, alphaRad, betaRad
    namelist /InputVariables/ ivd

    ivd%numberOfProcesses = 1
    ivd%dataDirectory = ""
    ivd%solverCommand = ""
    ivd%numberOfMGIterations = 100
    ivd%numberOfSGIterations = 0
    ivd%numberOfCFLIncrements = 0
    ivd%writeToFileInterval = 25
    ivd%gamma = 1.4_rk
    ivd%MachNumber = 0.5_rk
    ivd%alpha = 0.0_rk
    ivd%beta = 0.0_rk
    ivd%CFLNumber = 1.0_rk
    ivd%ReynoldsNumber = 1.0e6_rk
    ivd%viscosityScheme = 1
    ivd%inflowField(1:5, 1) = [1.0_rk, 1.0_rk, 0.0_rk, 0.0_rk, 1.0_rk]
    ivd%inflowField(6, 1) = ivd%inflowField(1, 1)*(ivd%MachNumber**2)/ivd%gamma
    ivd%inflowField(:, 2) = ivd%inflowField(:, 1)
    ivd%inflowTemperature = 528.0_rk
    ivd%wallsAreIsentropic = .false.
    ivd%movingwall = .false.
    ivd%wallTemperature = 528.0_rk
    ivd%patchInitialization = .false.
    ivd%patchType = 1
    ivd%patchSphere = [0.0_rk, 0.0_rk, 0.0_rk, 1.0_rk]
    ivd%patchBox = 0.0_rk
    ivd%engineFlowType = 1
    ivd%enginesFrontMassFlow = 0.0_rk
    ivd%enginesRearParameters = 0.0_rk
    ivd%numberOfEORelaxationSteps = 0
    ivd%engineBCRelaxation = 1.0_rk
    ivd%dissipationScheme = 2
    ivd%coarseGridDissipationScheme = 2
    ivd%secondOrderDissipationFactor = 0.4_rk
    ivd%fourthOrderDissipationFactor = 0.2_rk
    ivd%coarseGridDissipationFactor = 0.75_rk
    ivd%HartensCorrectionFactor = 0.0_rk
    ivd%useMatrixDissipation = .false.
    ivd%residualSmoothingFactor = 0.0_rk
    ivd%multigridScheme = 3
    ivd%numberOfRelaxationSteps = 1
    ivd%numberOfGridsToUse = 1
    ivd%numberOfRSSteps = 0
    ivd%numberOfPSSteps = 0
    ivd%multigridBCRelaxation = 1.0_rk
    ivd%turbulenceModel = 0
    ivd%numberOfTurbulenceGrids = 1
    ivd%turbulentCFLNumber = 1.0_rk
    ivd%tripFactor = 1.0_rk
    ivd%triggerRadius = 9999999.9_rk
    ivd%turbulenceTriggerValue = 0.0_rk
    ivd%referenceArea = 1.0_rk
    ivd%forceCalcType = 1
    ivd%useTimeResidual = .false.
    ivd%flowType = 0
    ivd%addTimeStep = .true.
    ivd%timeScheme = 2
    ivd%physicalTimeStep = 0.1_rk
    ivd%numberofphysicaltimesteps = 1
    ivd%maximumnumberofcycles = 10000
    ivd%restartNumber = 1
    ivd%initialResidual = 10000.0_rk
    ivd%fluxType = 1
    ivd%HLLCFluxOrder = 2
    ivd%HLLCClipping = 1.0_rk
    ivd%Automaticlipping = 0
    ivd%GradientMethod = 1
    ivd%HLLCICLS = 1
    ivd%explicit = .false.
    ivd%stopTime = 1.0_rk
    ivd%currentTime = 0.0_rk
    ivd%turbProlongationRelaxation = 1.0_rk
    ivd%prolongationSmoothingFactor = 0.0_rk
    ivd%prolongationMappingScheme = 1
    ivd%prolongationRelaxation = 1.0_rk
    ivd%PrandtlNumber = 0.72_rk
    ivd%turbulenceSmoothingFactor = 0.002_rk
    ivd%turbulentPrandtlNumber = 0.9_rk
    ivd%numberOfTurbulenceSteps = 9999999
    ivd%useDissipationWeighting = .false.
    ivd%maxTurbulenceValue = 3000.0_rk
    ivd%numberOfTriggerSteps = 0
    ivd%maxIncrementFactor = 0.25_rk
    ivd%minimumDensity = 0.1_rk
    ivd%turbK = 0.1_rk
    ivd%boundaryTerm = 1
    ivd%maxTurbulenceValueK = 1.5_rk
    ivd%maxTurbulenceValueW = 1.0e7_rk
    ivd%maxTurbulenceValueMU = 3000.0_rk
    ivd%minTurbulenceValueW = 5.0_rk
    ivd%firstoffwallpoint = 1.0e-4_rk
    ivd%kinfoverUinf2 = 1.0e-6_rk
    ivd%winf = 5.0_rk
    ivd%nuwall = 1.0_rk
    ivd%SAS = .false.
    ivd%XPER = 0.0_rk
    ivd%YPER = 0.0_rk
    ivd%ZPER = 0.0_rk
    ivd%GUSTMODEL = .false.
    ivd%UDS = 0.0_rk
    ivd%typegust = 0
    ivd%TTS = 0
    ivd%nspan = 0
    ivd%spanCor = 2
    ivd%rootspan = 0.0_rk
    ivd%tipspan = 15.0_rk
    ivd%groundplane = -9999999.9_rk
    ivd%datum = 0.0_rk
    ivd%nwheels = 0.0_rk
    ivd%wheelCentre = 0.0_rk
    ivd%wheelDimensions = 0.0_rk
    ivd%numberOfStepsPerCycle = 1
    ivd%cycleNumber = 0

    if (unit > 0) read (unit, nml=InputVariables)

    if (ivd%flowType == 0) then
       ivd%maximumnumberofcycles = 1000000
       ivd%addTimeStep = .false.
    end if
    if (ivd%explicit) then
       ivd%numberofphysicaltimesteps = 1
       ivd%addTimeStep = .false.
    end if
    if (ivd%fluxType == 4) then
       ivd%secondOrderDissipationFactor = 0.0_rk
       ivd%fourthOrderDissipationFactor = 0.0_rk
    end if
    PI = 4.0_rk*atan(1.0_rk)
    if (.not. ivd%patchInitialization) then
       ivd%wallTemperature = (1.0_rk/((ivd%gamma - 1.0_rk)*(ivd%MachNumber**2)))* &
            (ivd%wallTemperature/ivd%inflowTemperature)
       alphaRad = ivd%alpha*PI/180.0_rk
       betaRad = ivd%beta*PI/180.0_rk
       ivd%inflowField(1, 1) = 1.0_rk
       ivd%inflowField(2, 1) = cos(alphaRad)*cos(betaRad)
       ivd%inflowField(3, 1) = sin(betaRad)
       ivd%inflowField(4, 1) = sin(alphaRad)
       ivd%inflowField(5, 1) = (1.0_rk/((ivd%gamma - 1.0_rk)*ivd%gamma*ivd%MachNumber**2)) + 0.5_rk
       ivd%inflowField(6, 1) = 1.0_rk/(ivd%gamma*ivd%MachNumber**2)
    end if
  end subroutine readInputVariables

  ! options outside the implemented hot path stop the run instead of
  ! silently running something else
  subroutine check_supported(ivd)
    type(InputVariablesData), intent(in) :: ivd
    if (ivd%flowType < 0 .or. ivd%flowType > 2) &
         error stop "ERROR: flowType must be 0 (steady), 1 (fixed mesh) or 2 (moving mesh)"
    if (ivd%flowType /= 0 .and. ivd%explicit) error stop "ERROR: explicit time stepping is not on the GPU path"
    if (ivd%turbulenceModel /= 0 .and. ivd%turbulenceModel /= 1) &
         error stop "ERROR: only turbulenceModel 0 and 1 (Spalart-Allmaras) run on the GPU path"
    if (ivd%fluxType < 1 .or. ivd%fluxType > 4) error stop "ERROR: only fluxType 1-4 run on the GPU path"
    if (ivd%fluxType /= 4 .and. (ivd%dissipationScheme /= 2 .or. ivd%coarseGridDissipationScheme /= 2)) &
         error stop "ERROR: only dissipationScheme 2 (JST) runs on the GPU path"
    if (ivd%fluxType /= 4 .and. ivd%useDissipationWeighting) &
         error stop "ERROR: useDissipationWeighting is not on the GPU path"
    if (ivd%HLLCFluxOrder > 1 .and. ivd%GradientMethod /= 1) &
         error stop "ERROR: only Green-Gauss gradients (GradientMethod=1) run on the GPU path"
    if (ivd%movingwall) error stop "ERROR: moving walls are not on the GPU path"
    if (ivd%patchInitialization) error stop "ERROR: patch initialisation is not on the GPU path"
    if (ivd%residualSmoothingFactor > 0.0_rk .or. ivd%prolongationSmoothingFactor > 0.0_rk) &
         error stop "ERROR: residual/prolongation smoothing is not on the GPU path"
    if (ivd%multigridScheme /= 1 .and. ivd%multigridScheme /= 3 .and. ivd%multigridScheme > 0) &
         error stop "ERROR: only multigridScheme 0, 1 and 3 run on the GPU path"
  end subroutine check_supported

  function to_flite_params(ivd, coarse_fp32) result(p)
    type(InputVariablesData), intent(in) :: ivd
    logical, intent(in) :: coarse_fp32
    type(flite_params) :: p
    p%gamma = ivd%gamma
    p%mach_number = ivd%MachNumber
    p%reynolds_number = ivd%ReynoldsNumber
    p%prandtl_number = ivd%PrandtlNumber
    p%turbulent_prandtl_number = ivd%turbulentPrandtlNumber
    p%alpha = ivd%alpha
    p%beta = ivd%beta
    p%inflow_field = ivd%inflowField
    p%inflow_temperature = ivd%inflowTemperature
    p%walls_are_isentropic = merge(1, 0, ivd%wallsAreIsentropic)
    p%turbulence_model = ivd%turbulenceModel
    p%flux_type = ivd%fluxType
    p%hllc_flux_order = ivd%HLLCFluxOrder
    p%automatic_clipping = ivd%Automaticlipping
    p%gradient_method = ivd%GradientMethod
    p%hllc_clipping = ivd%HLLCClipping
    p%boundary_term = ivd%boundaryTerm
    p%cfl_number = ivd%CFLNumber
    p%turbulent_cfl_number = ivd%turbulentCFLNumber
    p%rk_coefficients = [0.6_rk, 0.6_rk, 1.0_rk, 0.0_rk, 0.0_rk]
    p%rk_stages = 3
    p%max_increment_factor = ivd%maxIncrementFactor
    p%minimum_density = ivd%minimumDensity
    p%max_turbulence_value = ivd%maxTurbulenceValue
    p%turb_k = ivd%turbK
    p%second_order_dissipation_factor = ivd%secondOrderDissipationFactor
    p%fourth_order_dissipation_factor = ivd%fourthOrderDissipationFactor
    p%multigrid_scheme = ivd%multigridScheme
    p%number_of_relaxation_steps = ivd%numberOfRelaxationSteps
    p%number_of_turbulence_grids = ivd%numberOfTurbulenceGrids
    p%number_of_turbulence_steps = ivd%numberOfTurbulenceSteps
    p%prolongation_relaxation = ivd%prolongationRelaxation
    p%turb_prolongation_relaxation = ivd%turbProlongationRelaxation
    p%multigrid_bc_relaxation = ivd%multigridBCRelaxation
    p%reference_area = ivd%referenceArea
    p%groundplane = ivd%groundplane
    p%datum = ivd%datum
    p%coarse_fp32 = merge(1, 0, coarse_fp32)
    p%coarse_grid_dissipation_factor = ivd%coarseGridDissipationFactor
    p%dissipation_scheme = ivd%dissipationScheme
    p%coarse_grid_dissipation_scheme = ivd%coarseGridDissipationScheme
    p%use_matrix_dissipation = merge(1, 0, ivd%useMatrixDissipation)
    p%hartens_correction_factor = ivd%HartensCorrectionFactor
    p%flow_type = ivd%flowType
    p%time_scheme = ivd%timeScheme
    p%add_time_step = merge(1, 0, ivd%addTimeStep)
    p%physical_time_step = ivd%physicalTimeStep
  end function to_flite_params

end module flite_input
