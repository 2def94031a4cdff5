! flite_gridio -- <case>.sol_1 reader and .res_1/.rst_1 solution files for
! the Fortran host.
!
! readComputationData follows the record order of the reference reader
! (SolverSource/mgnsg3d_mg_par.F90:1198-1268 and readGridComputationData,
! SolverSource/mgnsg3d_def_par.F90:13812-14255, with readBoundarySolverData
! DEF:791-865, readGridParallelizationData DEF:594-690, the prolongation and
! restriction readers DEF:452-478 and 516-542 and readComParallelizationData
! DEF:700-745) for a single-domain file, keeping the arrays the hot path
! needs in the reference's layout so that level_view can hand them to the
! C-ABI without copies.  Trip lines, engine inlets and multi-domain
! communication lists stop the run: they are not on the GPU path.
module flite_gridio
  use, intrinsic :: iso_c_binding
  use flite_gpu, only: flite_level_view
  implicit none
  private
  public :: GridLevelData, readComputationData, level_view, writeGridResults, readStartupField, makePlot

  type GridLevelData
     integer(c_int) :: numberOfNodes = 0, numberOfSides = 0
     integer(c_int) :: numberOfBoundarySides = 0, numberOfBoundaryNodes = 0, numberOfNormals = 0
     integer(c_int), allocatable :: sideIndexArray(:, :)        ! (E,2)
     real(c_double), allocatable :: sideWeightsArray(:, :)      ! (E,3)
     integer(c_int), allocatable :: bsideIndexArray(:, :)       ! brp%sideIndexArray(Eb,3)
     real(c_double), allocatable :: bsideWeightsArray(:, :)     ! brp%sideWeightsArray(Eb,3)
     real(c_double), allocatable :: nodeVolume(:), wallDistance(:)
     real(c_double), allocatable :: coordinates(:, :)           ! (N,3), level 1 only
     integer(c_int), allocatable :: wallDistanceBoundaryNodeArray(:), seedNodeRegister(:)
     integer(c_int) :: nodeIndicatorRegister(-20:0) = 0
     integer(c_int), allocatable :: nodeIndicatorArray(:)
     integer(c_int) :: numberOfFineNodes = 0
     integer(c_int), allocatable :: prolongationArray(:)        ! pod%prolongationArray(Nfine)
     real(c_double), allocatable :: restrictionArray(:)         ! rod%restrictionArray(Nfine)
     ! moving meshes, as readGridComputationData leaves them (DEF:816-819,
     ! 13855-13861, 13891-13900)
     real(c_double), allocatable :: GCLWeightsArray(:)          ! (E)
     real(c_double), allocatable :: bGCLWeightsArray(:, :)      ! brp%GCLWeightsArray(Eb,2)
     real(c_double), allocatable :: nodeVolumeP(:), nodeVolumeP2(:)  ! level 1
     real(c_double), allocatable :: coordinateMovement(:, :)    ! (N,3), level 1
  end type GridLevelData

contains

  subroutine readComputationData(unit, numberOfGridsToUse, grp, numberOfGrids, timeScheme)
    integer, intent(in) :: unit, numberOfGridsToUse
    integer, intent(in) :: timeScheme  ! ivd%timeScheme: 2 combines the GCL weights of two steps
    type(GridLevelData), allocatable, intent(out) :: grp(:)
    integer, intent(out) :: numberOfGrids
    integer :: i, domain, numberOfGridDomains, nInFile
    read (unit) nInFile
    numberOfGrids = nInFile
    if (numberOfGridsToUse > 0) numberOfGrids = min(nInFile, numberOfGridsToUse)
    read (unit) domain, numberOfGridDomains
    if (numberOfGridDomains /= 1) error stop "ERROR: only single-domain grid files run on the GPU path"
    allocate (grp(numberOfGrids))
    do i = 1, numberOfGrids
       if (i == 1) then
          call readGridComputationData(unit, grp(i), .true., 0, timeScheme)
       else
          call readGridComputationData(unit, grp(i), .false., grp(i - 1)%numberOfNodes, timeScheme)
       end if
    end do
  end subroutine readComputationData

  subroutine readBoundarySolverData(unit, g, isFine, timeScheme)
    integer, intent(in) :: unit, timeScheme
    type(GridLevelData), intent(inout) :: g
    logical, intent(in) :: isFine
    integer :: i, k, nEngine
    real(c_double) :: gcl(2), gclPrev(2)
    read (unit) g%numberOfBoundarySides
    allocate (g%bsideIndexArray(g%numberOfBoundarySides, 3), g%bsideWeightsArray(g%numberOfBoundarySides, 3))
    allocate (g%bGCLWeightsArray(g%numberOfBoundarySides, 2))
    do i = 1, g%numberOfBoundarySides
       read (unit) g%bsideIndexArray(i, 1:3), g%bsideWeightsArray(i, 1:3), gcl, gclPrev
       g%bGCLWeightsArray(i, :) = gcl
       if (timeScheme == 2 .and. isFine) g%bGCLWeightsArray(i, :) = 1.5_c_double * gcl - 0.5_c_double * gclPrev
    end do
    read (unit) g%numberOfBoundaryNodes, g%numberOfNormals
    do k = -20, 0
       read (unit) g%nodeIndicatorRegister(k)
    end do
    allocate (g%nodeIndicatorArray(max(g%numberOfBoundaryNodes, g%numberOfNormals)))
    g%nodeIndicatorArray = 0
    do i = 1, g%numberOfNormals
       read (unit) g%nodeIndicatorArray(i)
    end do
    read (unit) nEngine
    if (nEngine /= 0) error stop "ERROR: engine inlet sides are not on the GPU path"
  end subroutine readBoundarySolverData

  subroutine readGridComputationData(unit, g, isFine, numberOfFineNodes, timeScheme)
    integer, intent(in) :: unit, numberOfFineNodes, timeScheme
    type(GridLevelData), intent(inout) :: g
    logical, intent(in) :: isFine
    integer :: i, idx, wdn, seed, ntrip, ntripField, nComSides, nBComSides, nSend, nRecv
    integer :: nIntFine, nIntCoarse, nCom, f, c
    real(c_double) :: vol, volP, volP2, wd, gcl, gclPrev, length, x(3), xP(3), xP2(3), rw

    call readBoundarySolverData(unit, g, isFine, timeScheme)
    read (unit) g%numberOfSides
    allocate (g%sideIndexArray(g%numberOfSides, 2), g%sideWeightsArray(g%numberOfSides, 3))
    allocate (g%GCLWeightsArray(g%numberOfSides))
    do i = 1, g%numberOfSides
       read (unit) g%sideIndexArray(i, 1:2), g%sideWeightsArray(i, 1:3), gcl, gclPrev, length
       g%GCLWeightsArray(i) = gcl
       if (timeScheme == 2 .and. isFine) g%GCLWeightsArray(i) = 1.5_c_double * gcl - 0.5_c_double * gclPrev
    end do
    read (unit) g%numberOfNodes
    allocate (g%nodeVolume(g%numberOfNodes), g%wallDistance(g%numberOfNodes), &
         g%wallDistanceBoundaryNodeArray(g%numberOfNodes), g%seedNodeRegister(g%numberOfNodes))
    if (isFine) then
       allocate (g%coordinates(g%numberOfNodes, 3), g%coordinateMovement(g%numberOfNodes, 3))
       allocate (g%nodeVolumeP(g%numberOfNodes), g%nodeVolumeP2(g%numberOfNodes))
       do i = 1, g%numberOfNodes
          read (unit) idx, vol, volP, volP2, wd, wdn, x, xP, xP2
          g%nodeVolumeP(idx) = volP
          g%nodeVolumeP2(idx) = volP2
          if (timeScheme == 1) then
             g%coordinateMovement(idx, :) = x - xP
          else
             g%coordinateMovement(idx, :) = 1.5_c_double * x - 2.0_c_double * xP + 0.5_c_double * xP2
          end if
          g%nodeVolume(idx) = vol
          g%wallDistance(idx) = wd
          g%wallDistanceBoundaryNodeArray(idx) = wdn
          g%coordinates(idx, 1:3) = x
          g%seedNodeRegister(i) = i
       end do
       read (unit) ntrip
       if (ntrip /= 0) error stop "ERROR: wall trip lines are not on the GPU path"
       read (unit) ntripField
       read (unit) nComSides, nBComSides
       read (unit) nSend
       read (unit) nRecv
    else
       do i = 1, g%numberOfNodes
          read (unit) idx, vol, wd, wdn, seed
          g%nodeVolume(idx) = vol
          g%wallDistance(idx) = wd
          g%wallDistanceBoundaryNodeArray(idx) = wdn
          g%seedNodeRegister(i) = seed
       end do
       read (unit) nComSides, nBComSides
       read (unit) nSend
       read (unit) nRecv
       g%numberOfFineNodes = numberOfFineNodes
       allocate (g%prolongationArray(numberOfFineNodes), g%restrictionArray(numberOfFineNodes))
       g%prolongationArray = 0
       g%restrictionArray = 0.0_c_double
       read (unit) nIntFine
       read (unit) nIntCoarse
       do i = 1, nIntFine
          read (unit) f, c
          g%prolongationArray(f) = c
       end do
       read (unit) nCom
       if (nCom /= 0) error stop "ERROR: multi-domain prolongation lists are not on the GPU path"
       read (unit) nIntFine
       read (unit) nIntCoarse
       do i = 1, nIntFine
          read (unit) f, rw
          g%restrictionArray(f) = rw
       end do
       read (unit) nCom
       ! readComParallelizationData for one domain: total, then the
       ! domain's count (no register records when both are zero)
       read (unit) nCom
       read (unit) nSend
       read (unit) nCom
       read (unit) nRecv
       if (nSend /= 0 .or. nRecv /= 0) error stop "ERROR: multi-domain agglomeration lists are not on the GPU path"
    end if
  end subroutine readGridComputationData

  ! non-owning C view of one level (index_base 1: Fortran numbering)
  function level_view(g, isFine, moving) result(v)
    type(GridLevelData), intent(in), target :: g
    logical, intent(in) :: isFine
    logical, intent(in), optional :: moving  ! pass the moving-mesh arrays (flowType 2)
    type(flite_level_view) :: v
    v%index_base = 1
    v%n_nodes = g%numberOfNodes
    v%n_sides = g%numberOfSides
    v%n_bsides = g%numberOfBoundarySides
    v%n_bnodes = g%numberOfBoundaryNodes
    v%side_index = c_loc(g%sideIndexArray)
    v%side_weights = c_loc(g%sideWeightsArray)
    v%bside_index = c_loc(g%bsideIndexArray)
    v%bside_weights = c_loc(g%bsideWeightsArray)
    v%node_volume = c_loc(g%nodeVolume)
    v%wall_distance = c_loc(g%wallDistance)
    v%node_indicator_register = c_loc(g%nodeIndicatorRegister)
    v%node_indicator_array = c_loc(g%nodeIndicatorArray)
    if (isFine) then
       v%coordinates = c_loc(g%coordinates)
       v%n_fine = 0
       v%prolongation = c_null_ptr
       v%restriction = c_null_ptr
    else
       v%coordinates = c_null_ptr
       v%n_fine = g%numberOfFineNodes
       v%prolongation = c_loc(g%prolongationArray)
       v%restriction = c_loc(g%restrictionArray)
    end if
    if (present(moving)) then
       if (moving) then
          v%side_gcl = c_loc(g%GCLWeightsArray)
          v%bside_gcl = c_loc(g%bGCLWeightsArray)
          if (isFine) then
             v%node_volume_p = c_loc(g%nodeVolumeP)
             v%node_volume_p2 = c_loc(g%nodeVolumeP2)
             v%coordinate_movement = c_loc(g%coordinateMovement)
          end if
       end if
    end if
  end function level_view

  ! writeGridResults (DEF:15331-15345): N; rho, (u(:,j)/rho, j=2..nvar)
  subroutine writeGridResults(unit, u, n, nvar)
    integer, intent(in) :: unit, n, nvar
    real(c_double), intent(in) :: u(n, nvar)
    integer :: i, j
    write (unit) n
    write (unit) (u(i, 1), i=1, n), ((u(i, j)/u(i, 1), i=1, n), j=2, nvar)
  end subroutine writeGridResults

  ! start-up field of setUpInitialField (DEF:3744-3767, 3853-3858)
  subroutine readStartupField(unit, u, n, nvar)
    integer, intent(in) :: unit, n, nvar
    real(c_double), intent(out) :: u(n, nvar)
    integer :: i, j, number
    read (unit) number
    if (number /= n) error stop "ERROR: Startup file is not compatible"
    read (unit) (u(i, 1), i=1, n), ((u(i, j), i=1, n), j=2, nvar)
    do j = 2, nvar
       u(:, j) = u(:, 1)*u(:, j)
    end do
  end subroutine readStartupField

  ! makeplot (mgnsg3d_mg_par.F90:1599-1707): merge the per-domain result
  ! files <case>.res_<d> into <case>.unk through the preprocessor's plot
  ! register plotreg.reg, a record (N, numberOfDomains), then per domain its
  ! local node count and one (local, global) record per node.  The values
  ! are copied as the result files hold them.
  subroutine makePlot(caseName, nvar)
    character(len=*), intent(in) :: caseName
    integer, intent(in) :: nvar
    integer :: numberOfNodes, numberOfDomains, numberOfLocalNodes, numberOfToNodes, i, j, k
    integer, allocatable :: transArray(:, :)
    real(c_double), allocatable :: u(:, :), unk(:, :)
    character(len=16) :: dom
    open (15, file='plotreg.reg', form='unformatted', status='old')
    read (15) numberOfNodes, numberOfDomains
    allocate (u(numberOfNodes, nvar))
    u = 0.0_c_double
    do i = 1, numberOfDomains
       read (15) numberOfLocalNodes
       allocate (transArray(numberOfLocalNodes, 2))
       do j = 1, numberOfLocalNodes
          read (15) transArray(j, :)
       end do
       write (dom, '(I0)') i
       open (16, file=trim(caseName)//".res_"//trim(dom), form='unformatted', status='old')
       read (16) numberOfToNodes
       allocate (unk(numberOfToNodes, nvar))
       read (16) ((unk(j, k), j=1, numberOfToNodes), k=1, nvar)
       close (16)
       do j = 1, numberOfLocalNodes
          if (transArray(j, 1) < 1 .or. transArray(j, 1) > numberOfToNodes .or. &
              transArray(j, 2) < 1 .or. transArray(j, 2) > numberOfNodes) &
               error stop "ERROR: plotreg.reg does not match the result files"
          u(transArray(j, 2), :) = unk(transArray(j, 1), :)
       end do
       deallocate (unk, transArray)
    end do
    close (15)
    open (17, file=trim(caseName)//".unk", form='unformatted', status='replace')
    write (17) numberOfNodes
    write (17) ((u(i, j), i=1, numberOfNodes), j=1, nvar)
    close (17)
    deallocate (u)
  end subroutine makePlot

end module flite_gridio
