! mgnsg3d_gpu -- steady multigrid driver whose grid work runs on the GPU.
!
! The control flow, inputs and outputs are those of the reference driver
! (SolverSource/mgnsg3d_mg_par.F90): run.inp names the case (MG:107-110),
! <case>.inp holds the &inputVariables namelist (DEF:201-422), <case>.sol_1
! the preprocessed grids (MG:1198-1268), <case>.rst_1 an optional start-up
! field (restartNumber > 0, MG:247-260); multiGridCycles' loop (MG:606-800)
! runs oneCycle either numberOfMGIterations times or, when that is negative,
! until the monitor residual has fallen by that many decades; every cycle
! appends a line to <case>.rsd (writeResiduals, MG:1478-1514), every
! writeToFileInterval cycles and at the end the solution goes to <case>.res_1
! (writeResultsToFile, MG:1518-1531), liftdrag.res gets the forces of each
! physical step (MG:847-884), and with a plot register (plotreg.reg) present
! the result is merged into <case>.unk (makeplot, MG:1599-1707).  Unsteady
! runs on a fixed mesh (flowType 1) loop over the physical time steps
! restartNumber .. numberOfPhysicalTimesteps around multiGridCycles' loop
! (MG:486-890); a moving mesh (flowType 2) reads <case>_<k>.sol_1 with k the
! step's place in numberOfStepsPerCycle, first at the start step (MG:164-193)
! and then before every later step (MG:505-548), and hands each level's new
! geometry to flite_level_geometry.  Only single-domain runs exist here: the MPI partitioning
! collapses to one domain on one GPU.
!
! Usage: mgnsg3d_gpu [--coarse-fp32]
!   --coarse-fp32  smooth and transfer on the coarse levels in FP32
!                  (the fine level, the solution update and the forces stay FP64);
!                  FP32 coarse corrections stall near a 1e-5 residual drop, so
!                  a run to a deeper drop (numberOfMGIterations < -4) switches
!                  the coarse levels to FP64 once the residual has fallen 1e-4
program mgnsg3d_gpu
  use, intrinsic :: iso_c_binding
  use flite_gpu
  use flite_input
  use flite_gridio
  implicit none

  type(InputVariablesData) :: ivd
  type(GridLevelData), allocatable, target :: grp(:), grpNext(:)
  type(flite_params) :: p
  type(flite_level_view) :: v
  character(len=256) :: fileName, arg, solName
  integer :: numberOfGrids, i, l, nvar, nres, k, physicalTimestepNumber, firstStep, lastStep
  logical :: coarseFp32, switched, havePlotReg
  real(c_double), allocatable :: u(:, :)
  real(c_double) :: res(7), maxres(7), forces(9)
  integer(c_int) :: maxind(7)
  real(c_double) :: initialResidual, currentResidual, stopAtResidualReduction, currentRes
  real(c_double) :: averageTimePerCycle
  integer(c_int64_t) :: t0, t1, rate

  coarseFp32 = .false.
  do k = 1, command_argument_count()
     call get_command_argument(k, arg)
     if (trim(arg) == "--coarse-fp32") then
        coarseFp32 = .true.
     else
        write (*, '(A)') "usage: mgnsg3d_gpu [--coarse-fp32]"
        error stop "ERROR: unknown argument"
     end if
  end do

  open (1, file='run.inp', status='old')
  read (1, "(a)") fileName
  close (1)
  fileName = adjustl(fileName)

  open (21, file=trim(fileName)//".inp", form='formatted', status='old')
  call readInputVariables(ivd, 21)
  close (21)
  call check_supported(ivd)
  nvar = 5
  if (ivd%turbulenceModel == 1) nvar = 6
  nres = nvar

  solName = trim(fileName)//".sol_1"
  if (ivd%flowType == 2) solName = step_sol_name(ivd%restartNumber)
  write (*, '(A)') "reading "//trim(solName)//" ..."
  open (20, file=trim(solName), form='unformatted', status='old', access='sequential')
  call readComputationData(20, ivd%numberOfGridsToUse, grp, numberOfGrids, ivd%timeScheme)
  close (20)
  write (*, '(A,I3,A,I10,A,I11,A)') "Number of grids:", numberOfGrids, "  fine grid:", &
       grp(1)%numberOfNodes, " nodes,", grp(1)%numberOfSides, " sides"

  p = to_flite_params(ivd, coarseFp32)
  call flite_check(flite_ctx_create(p, numberOfGrids, 1, gpu_ctx), "flite_ctx_create")
  do l = 1, numberOfGrids
     v = level_view(grp(l), l == 1, ivd%flowType == 2)
     call flite_check(flite_level_upload(gpu_ctx, l, v), "readGridComputationData")
  end do

  allocate (u(grp(1)%numberOfNodes, nvar))
  if (ivd%restartNumber > 0) then
     open (22, file=trim(fileName)//".rst_1", form='unformatted', status='old')
     call readStartupField(22, u, grp(1)%numberOfNodes, nvar)
     close (22)
     call flite_check(flite_state_set(gpu_ctx, 1, u, grp(1)%numberOfNodes, nvar, 0), "setUpInitialField")
     call flite_check(flite_init_from_state(gpu_ctx), "setUpInitialField")
  else
     call flite_check(flite_init_freestream(gpu_ctx), "setUpInitialField")
  end if

  initialResidual = ivd%initialResidual
  averageTimePerCycle = 0.0_c_double
  call system_clock(count_rate=rate)
  switched = .false.
  ! physical time steps (MG:477-486): one for steady flow
  firstStep = ivd%restartNumber
  lastStep = ivd%numberofphysicaltimesteps
  if (ivd%flowType == 0) then
     firstStep = 1
     lastStep = 1
  end if

  do physicalTimestepNumber = firstStep, lastStep
  if (ivd%flowType /= 0) write (*, '(A,I7)') " >>>>>>>>>----> number of timestep = ", physicalTimestepNumber
  if (ivd%flowType == 2 .and. physicalTimestepNumber > firstStep) then
     ! the step's mesh (MG:505-548): new geometry, the solution kept
     solName = step_sol_name(physicalTimestepNumber)
     write (*, '(A)') " new data File "//trim(solName)//" opened ..."
     open (20, file=trim(solName), form='unformatted', status='old', access='sequential')
     call readComputationData(20, ivd%numberOfGridsToUse, grpNext, l, ivd%timeScheme)
     close (20)
     if (l /= numberOfGrids) error stop "ERROR: the step's grid file has another number of grids"
     do l = 1, numberOfGrids
        v = level_view(grpNext(l), l == 1, .true.)
        call flite_check(flite_level_geometry(gpu_ctx, l, v), "readComputationData")
     end do
     call move_alloc(grpNext, grp)
  end if
  currentResidual = 10000.0_c_double
  stopAtResidualReduction = -1.0_c_double
  if (ivd%numberOfMGIterations < 0) stopAtResidualReduction = abs(dble(ivd%numberOfMGIterations))

  write (*, '(A)') "Starting timestepping..."
  i = 1
  do
     if (stopAtResidualReduction < 0.0_c_double) then
        if (i > ivd%numberOfMGIterations) exit
     else
        if (.not. (log(initialResidual) - log(currentResidual) < log(10.0_c_double)*stopAtResidualReduction)) exit
     end if
     call system_clock(t0)
     call gpu_oneCycle(i)
     call flite_check(flite_sync(gpu_ctx), "oneCycle")
     call system_clock(t1)
     averageTimePerCycle = averageTimePerCycle + dble(t1 - t0)/dble(rate)

     call writeResiduals(i)
     if (coarseFp32 .and. .not. switched .and. stopAtResidualReduction > 4.0_c_double) then
        if (log(initialResidual) - log(currentResidual) >= 4.0_c_double*log(10.0_c_double)) then
           call flite_check(flite_set_coarse_precision(gpu_ctx, 0_c_int), "coarse precision")
           switched = .true.
           write (*, '(A,I7)') " coarse levels FP32 -> FP64 from cycle ", i + 1
        end if
     end if
     if (mod(i, ivd%writeToFileInterval) == 0) call writeResultsToFile()
     i = i + 1
     if (stopAtResidualReduction >= 0.0_c_double .and. i > ivd%maximumnumberofcycles) exit
  end do
  call writeResultsToFile()

  ! lift file of the step (MG:855-884): a new file on the first step
  call gpu_monitor(res, maxres, maxind, forces)
  if (physicalTimestepNumber > 1) then
     open (28, file='liftdrag.res', form='formatted', status='unknown', position='append')
  else
     open (28, file='liftdrag.res', form='formatted', status='unknown')
  end if
  write (28, '(7E14.5)') physicalTimestepNumber*ivd%physicalTimeStep, forces(1:6)
  close (28)

  ! end of the physical step: uold2 <- uold <- u (MG:889-890)
  if (ivd%flowType /= 0) call flite_check(flite_advance_physical_time(gpu_ctx), "physical time step")
  end do

  ! merging results to the .unk file (MG:776-781), when the preprocessor
  ! left its plot register here
  inquire (file='plotreg.reg', exist=havePlotReg)
  if (havePlotReg) then
     write (*, '(A)') " Merging results to .unk file ..."
     call makePlot(fileName, nvar)
  end if
  if (i > 1) write (*, *) "Average time per cycle: ", averageTimePerCycle/dble(i - 1)

  call flite_ctx_destroy(gpu_ctx)

contains

  ! <case>_<k>.sol_1 of a moving mesh, k = mod(step, numberOfStepsPerCycle)
  ! with 0 read as numberOfStepsPerCycle (MG:164-193, 505-548)
  function step_sol_name(step) result(name)
    integer, intent(in) :: step
    character(len=256) :: name
    character(len=16) :: ext
    integer :: k
    k = mod(step, ivd%numberOfStepsPerCycle)
    if (k == 0) k = ivd%numberOfStepsPerCycle
    write (ext, '(I0)') k
    name = trim(fileName)//"_"//trim(ext)//".sol_1"
  end function step_sol_name

  ! writeResiduals (MG:1271-1516): screen line, NaN check, .rsd line
  subroutine writeResiduals(iterationNumber)
    integer, intent(in) :: iterationNumber
    integer :: j
    call gpu_monitor(res, maxres, maxind, forces)
    currentResidual = res(1)
    if (initialResidual == 10000.0_c_double) then
       initialResidual = res(1)
       write (*, *) ' Initial Residual is : ', initialResidual
    end if
    write (*, 501) iterationNumber, res(1), " (", maxres(1), ") ", res(2), " (", maxres(2), ") ", &
         res(3), " (", maxres(3), ") ", res(4), " (", maxres(4), ") ", res(5), " (", maxres(5), ") ", &
         forces(1), forces(2)
    do j = 1, nres
       if (isnan(res(j))) then
          write (*, *) "---> NAN in residual, exiting now ...."
          write (*, *) "---> stop in sub writeResiduals"
          stop
       end if
    end do
    if (res(1)/initialResidual > 1.0e-15_c_double) then
       currentRes = log(res(1)/initialResidual)
    else
       currentRes = -100.0_c_double
    end if
    if (iterationNumber == 1) then
       open (25, file=trim(fileName)//".rsd", form='formatted', status='unknown')
    else
       open (25, file=trim(fileName)//".rsd", form='formatted', status='unknown', position='append')
    end if
    write (25, 504) iterationNumber, currentRes/log(10.0_c_double), forces(1:6), &
         (log(res(j))/log(10.0_c_double), j=2, nres)
    close (25)
501 format(I7, E14.5, A, E8.2, A, E14.5, A, E8.2, A, E14.5, A, E8.2, A, E14.5, A, E8.2, A, E14.5, A, E8.2, A, &
         E12.4, E12.4)
504 format(I7, E14.5, E14.5, E14.5, 12(f13.6, 1x))
  end subroutine writeResiduals

  subroutine writeResultsToFile()
    call flite_check(flite_state_get(gpu_ctx, 1, u, grp(1)%numberOfNodes, nvar, 0), "writeResultsToFile")
    open (26, file=trim(fileName)//".res_1", form='unformatted', status='replace')
    call writeGridResults(26, u, grp(1)%numberOfNodes, nvar)
    close (26)
  end subroutine writeResultsToFile

end program mgnsg3d_gpu
