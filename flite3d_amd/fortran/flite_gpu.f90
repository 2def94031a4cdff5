! flite_gpu -- ISO_C_BINDING interface of the MI355X flow-solver hot path
! (include/flite3d_amd.h) for a Fortran host.
!
! The reference driver (SolverSource/mgnsg3d_mg_par.F90) calls the GridSolver
! module procedures of SolverSource/mgnsg3d_def_par.F90 on host arrays held in
! grp(:) and ivd.  With this module the driver keeps its control flow, its
! input files and its outputs, and calls the gpu_* procedures below in place
! of the GridSolver ones:
!
!   reference call (oneCycle, MG:939-1186)              replacement
!   doTimeIterations(grp(i),ivd,n,diss,turb,it,ts)  ->  gpu_doTimeIterations(i,n,diss,turb,it)
!   restrictGrid(grp(i),grp(i+1),ivd,i,turb)        ->  gpu_restrictGrid(i,turb)
!   prolongateGrid(grp(i+1),grp(i),ivd)             ->  gpu_prolongateGrid(i+1)
!   mapFineToCoarse(grp(i)%u,grp(i+1),grp(i+1)%u)   ->  gpu_mapFineToCoarse(i)
!   the whole of oneCycle(cycleNo,ts)               ->  gpu_oneCycle(cycleNo)
!   writeResiduals' sums + calculateLiftAndDrag     ->  gpu_monitor(res,maxres,maxind,forces)
!
! Errors: every entry point returns 0 or a negative FLITE_E_* code;
! flite_check turns a failure into the reference's behaviour on errors, a
! STOP with an "ERROR:" message.
module flite_gpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: FLITE_OK = 0, FLITE_E_ARG = -1, FLITE_E_HIP = -2, &
       FLITE_E_NAN = -3, FLITE_E_UNSUPPORTED = -4, FLITE_E_STATE = -5

  integer(c_int), parameter :: FLITE_FIELD_U = 0, FLITE_FIELD_UPREV = 1, FLITE_FIELD_RHS = 2, &
       FLITE_FIELD_DISSIPATION = 3, FLITE_FIELD_PRESSURE = 4, FLITE_FIELD_DT = 5, &
       FLITE_FIELD_DT_TURB = 6, FLITE_FIELD_LAMINAR_VISCOSITY = 7, FLITE_FIELD_VORTICITY = 8, &
       FLITE_FIELD_DIVERGENCE = 9, FLITE_FIELD_SOURCE = 10, FLITE_FIELD_TURB_DIFFUSION = 11, &
       FLITE_FIELD_GRADIENT = 12, FLITE_FIELD_WALL_STRESS = 13, FLITE_FIELD_UOLD = 14, &
       FLITE_FIELD_UOLD2 = 15

  ! flite_params: the InputVariablesData scalars the hot path uses, after
  ! readInputVariables' post-processing (DEF:386-415)
  type, bind(C) :: flite_params
     real(c_double) :: gamma
     real(c_double) :: mach_number
     real(c_double) :: reynolds_number
     real(c_double) :: prandtl_number
     real(c_double) :: turbulent_prandtl_number
     real(c_double) :: alpha, beta
     real(c_double) :: inflow_field(6, 2)      ! ivd%inflowField(6,2)
     real(c_double) :: inflow_temperature
     integer(c_int) :: walls_are_isentropic
     integer(c_int) :: turbulence_model
     integer(c_int) :: flux_type
     integer(c_int) :: hllc_flux_order
     integer(c_int) :: automatic_clipping
     integer(c_int) :: gradient_method
     real(c_double) :: hllc_clipping
     integer(c_int) :: boundary_term
     real(c_double) :: cfl_number
     real(c_double) :: turbulent_cfl_number
     real(c_double) :: rk_coefficients(5)
     integer(c_int) :: rk_stages
     real(c_double) :: max_increment_factor
     real(c_double) :: minimum_density
     real(c_double) :: max_turbulence_value
     real(c_double) :: turb_k
     real(c_double) :: second_order_dissipation_factor
     real(c_double) :: fourth_order_dissipation_factor
     integer(c_int) :: multigrid_scheme
     integer(c_int) :: number_of_relaxation_steps
     integer(c_int) :: number_of_turbulence_grids
     integer(c_int) :: number_of_turbulence_steps
     real(c_double) :: prolongation_relaxation
     real(c_double) :: turb_prolongation_relaxation
     real(c_double) :: multigrid_bc_relaxation
     real(c_double) :: reference_area
     real(c_double) :: groundplane
     real(c_double) :: datum(3)
     integer(c_int) :: coarse_fp32
     real(c_double) :: coarse_grid_dissipation_factor
     integer(c_int) :: dissipation_scheme
     integer(c_int) :: coarse_grid_dissipation_scheme
     integer(c_int) :: use_matrix_dissipation
     real(c_double) :: hartens_correction_factor
     integer(c_int) :: flow_type
     integer(c_int) :: time_scheme
     integer(c_int) :: add_time_step
     real(c_double) :: physical_time_step
  end type flite_params

  ! flite_level_view: non-owning pointers to one level's arrays as the
  ! .sol reader holds them (column-major, 1-based indices => index_base 1)
  type, bind(C) :: flite_level_view
     integer(c_int) :: index_base
     integer(c_int) :: n_nodes
     integer(c_int) :: n_sides
     integer(c_int) :: n_bsides
     integer(c_int) :: n_bnodes
     type(c_ptr) :: side_index               ! sideIndexArray(E,2)
     type(c_ptr) :: side_weights             ! sideWeightsArray(E,3)
     type(c_ptr) :: bside_index              ! brp%sideIndexArray(Eb,3)
     type(c_ptr) :: bside_weights            ! brp%sideWeightsArray(Eb,3)
     type(c_ptr) :: node_volume              ! nodeVolume(N)
     type(c_ptr) :: coordinates              ! coordinates(N,3), null on coarse levels
     type(c_ptr) :: wall_distance            ! wallDistance(N)
     type(c_ptr) :: node_indicator_register  ! nodeIndicatorRegister(-20:0)
     type(c_ptr) :: node_indicator_array     ! nodeIndicatorArray(Nb)
     integer(c_int) :: n_fine
     type(c_ptr) :: prolongation             ! pod%prolongationArray(n_fine)
     type(c_ptr) :: restriction              ! rod%restrictionArray(n_fine)
     ! moving meshes (flowType 2); null on a fixed mesh
     type(c_ptr) :: side_gcl = c_null_ptr            ! GCLWeightsArray(E)
     type(c_ptr) :: bside_gcl = c_null_ptr           ! brp%GCLWeightsArray(Eb,2)
     type(c_ptr) :: node_volume_p = c_null_ptr       ! nodeVolumeP(N), level 1
     type(c_ptr) :: node_volume_p2 = c_null_ptr      ! nodeVolumeP2(N), level 1
     type(c_ptr) :: coordinate_movement = c_null_ptr ! coordinateMovement(N,3), level 1
  end type flite_level_view

  interface
     integer(c_int) function flite_ctx_create(p, n_levels, n_batch, ctx) bind(C, name="flite_ctx_create")
       import :: c_int, c_ptr, flite_params
       type(flite_params), intent(in) :: p
       integer(c_int), value :: n_levels, n_batch
       type(c_ptr), intent(out) :: ctx
     end function flite_ctx_create

     integer(c_int) function flite_level_geometry(ctx, level, v) bind(C, name="flite_level_geometry")
       import :: c_int, c_ptr, flite_level_view
       type(c_ptr), value :: ctx
       integer(c_int), value :: level
       type(flite_level_view), intent(in) :: v
     end function flite_level_geometry

     integer(c_int) function flite_level_upload(ctx, level, v) bind(C, name="flite_level_upload")
       import :: c_int, c_ptr, flite_level_view
       type(c_ptr), value :: ctx
       integer(c_int), value :: level
       type(flite_level_view), intent(in) :: v
     end function flite_level_upload

     subroutine flite_ctx_destroy(ctx) bind(C, name="flite_ctx_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine flite_ctx_destroy

     type(c_ptr) function flite_last_error() bind(C, name="flite_last_error")
       import :: c_ptr
     end function flite_last_error

     integer(c_int) function flite_state_set(ctx, level, u, ld, ncomp, batch) bind(C, name="flite_state_set")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, ld, ncomp, batch
       real(c_double), intent(in) :: u(*)
     end function flite_state_set

     integer(c_int) function flite_state_get(ctx, level, u, ld, ncomp, batch) bind(C, name="flite_state_get")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, ld, ncomp, batch
       real(c_double), intent(out) :: u(*)
     end function flite_state_get

     integer(c_int) function flite_init_freestream(ctx) bind(C, name="flite_init_freestream")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function flite_init_freestream

     integer(c_int) function flite_init_from_state(ctx) bind(C, name="flite_init_from_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function flite_init_from_state

     integer(c_int) function flite_advance_physical_time(ctx) bind(C, name="flite_advance_physical_time")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function flite_advance_physical_time

     integer(c_int) function flite_do_time_iterations(ctx, level, relaxation_steps, calculate_dissipation, &
          calculate_turbulence, iteration_number) bind(C, name="flite_do_time_iterations")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, relaxation_steps, calculate_dissipation, calculate_turbulence, &
            iteration_number
     end function flite_do_time_iterations

     integer(c_int) function flite_restrict(ctx, fine_level, calculate_turbulence) bind(C, name="flite_restrict")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fine_level, calculate_turbulence
     end function flite_restrict

     integer(c_int) function flite_prolongate(ctx, coarse_level) bind(C, name="flite_prolongate")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: coarse_level
     end function flite_prolongate

     integer(c_int) function flite_map_u_fine_to_coarse(ctx, fine_level) bind(C, name="flite_map_u_fine_to_coarse")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fine_level
     end function flite_map_u_fine_to_coarse

     integer(c_int) function flite_cycle(ctx, cycle_number) bind(C, name="flite_cycle")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: cycle_number
     end function flite_cycle

     integer(c_int) function flite_monitor(ctx, res, maxres, maxidx, forces) bind(C, name="flite_monitor")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: res(7), maxres(7), forces(9)
       integer(c_int), intent(out) :: maxidx(7)
     end function flite_monitor

     integer(c_int) function flite_field_get(ctx, level, field, out, ld) bind(C, name="flite_field_get")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, field, ld
       real(c_double), intent(out) :: out(*)
     end function flite_field_get

     integer(c_int) function flite_field_set(ctx, level, field, in, ld) bind(C, name="flite_field_set")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, field, ld
       real(c_double), intent(in) :: in(*)
     end function flite_field_set

     integer(c_int) function flite_field_ncomp(field) bind(C, name="flite_field_ncomp")
       import :: c_int
       integer(c_int), value :: field
     end function flite_field_ncomp

     integer(c_int) function flite_apply(ctx, level, op) bind(C, name="flite_apply")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: level, op
     end function flite_apply

     integer(c_int) function flite_tangent_apply(ctx, du, ld_du, out, ld_out) bind(C, name="flite_tangent_apply")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: du(*)
       integer(c_int), value :: ld_du, ld_out
       real(c_double), intent(out) :: out(*)
     end function flite_tangent_apply

     integer(c_int) function flite_adjoint_apply(ctx, lam, ld_lam, out, ld_out) bind(C, name="flite_adjoint_apply")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: lam(*)
       integer(c_int), value :: ld_lam, ld_out
       real(c_double), intent(out) :: out(*)
     end function flite_adjoint_apply

     integer(c_int) function flite_drag_state_gradient(ctx, drag, dJdU, ld) bind(C, name="flite_drag_state_gradient")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: drag
       real(c_double), intent(out) :: dJdU(*)
       integer(c_int), value :: ld
     end function flite_drag_state_gradient

     integer(c_int) function flite_metric_adjoint(ctx, lam, ld_lam, gw, ld_w, gwb, ld_wb) &
          bind(C, name="flite_metric_adjoint")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: lam(*)
       integer(c_int), value :: ld_lam
       real(c_double), intent(out) :: gw(*)
       integer(c_int), value :: ld_w
       real(c_double), intent(out) :: gwb(*)
       integer(c_int), value :: ld_wb
     end function flite_metric_adjoint

     integer(c_int) function flite_drag_metric_gradient(ctx, drag, gwb, ld_wb) bind(C, name="flite_drag_metric_gradient")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: drag
       real(c_double), intent(out) :: gwb(*)
       integer(c_int), value :: ld_wb
     end function flite_drag_metric_gradient

     ! gwd may be c_null_ptr without Spalart-Allmaras
     integer(c_int) function flite_mesh_adjoint(ctx, lam, ld_lam, gw, ld_w, gwb, ld_wb, gvol, gxyz, ld_x, gwd) &
          bind(C, name="flite_mesh_adjoint")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: lam(*)
       integer(c_int), value :: ld_lam
       real(c_double), intent(out) :: gw(*)
       integer(c_int), value :: ld_w
       real(c_double), intent(out) :: gwb(*)
       integer(c_int), value :: ld_wb
       real(c_double), intent(out) :: gvol(*), gxyz(*)
       integer(c_int), value :: ld_x
       type(c_ptr), value :: gwd
     end function flite_mesh_adjoint

     integer(c_int) function flite_drag_mesh_gradient(ctx, drag, gw, ld_w, gwb, ld_wb, gvol) &
          bind(C, name="flite_drag_mesh_gradient")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: drag
       real(c_double), intent(out) :: gw(*)
       integer(c_int), value :: ld_w
       real(c_double), intent(out) :: gwb(*)
       integer(c_int), value :: ld_wb
       real(c_double), intent(out) :: gvol(*)
     end function flite_drag_mesh_gradient

     integer(c_int) function flite_set_coarse_precision(ctx, fp32) bind(C, name="flite_set_coarse_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fp32
     end function flite_set_coarse_precision

     integer(c_int) function flite_autotune(ctx, repeats, lanes_out) bind(C, name="flite_autotune")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: repeats
       integer(c_int), intent(out) :: lanes_out(*)
     end function flite_autotune

     integer(c_int) function flite_batch_select(ctx, batch) bind(C, name="flite_batch_select")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: batch
     end function flite_batch_select

     integer(c_int) function flite_sync(ctx) bind(C, name="flite_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function flite_sync
  end interface

  ! the context the gpu_* procedures act on (the reference keeps its grids in
  ! the module globals grp(:) and ivd in the same way, MG:27-28)
  type(c_ptr), save :: gpu_ctx = c_null_ptr

contains

  ! text of the last failure on this thread
  function flite_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = flite_last_error()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate (character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = s(i)
    end do
  end function flite_error_message

  subroutine flite_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= FLITE_OK) then
       write (*, '(A)') "ERROR: "//where//": "//flite_error_message()
       error stop "ERROR: GPU flow solver call failed"
    end if
  end subroutine flite_check

  integer(c_int) function c_flag(flag)
    logical, intent(in) :: flag
    c_flag = 0
    if (flag) c_flag = 1
  end function c_flag

  ! doTimeIterations (DEF:966-1000)
  subroutine gpu_doTimeIterations(gridNumber, relaxationSteps, calculateDissipation, &
       calculateTurbulence, iterationNumber)
    integer, intent(in) :: gridNumber, relaxationSteps, iterationNumber
    logical, intent(in) :: calculateDissipation, calculateTurbulence
    call flite_check(flite_do_time_iterations(gpu_ctx, gridNumber, relaxationSteps, &
         c_flag(calculateDissipation), c_flag(calculateTurbulence), iterationNumber), "doTimeIterations")
  end subroutine gpu_doTimeIterations

  ! restrictGrid (DEF:12141-12430), grid gridNumber -> gridNumber+1
  subroutine gpu_restrictGrid(gridNumber, calculateTurbulence)
    integer, intent(in) :: gridNumber
    logical, intent(in) :: calculateTurbulence
    call flite_check(flite_restrict(gpu_ctx, gridNumber, c_flag(calculateTurbulence)), "restrictGrid")
  end subroutine gpu_restrictGrid

  ! prolongateGrid (DEF:12432-12570), grid coarseGridNumber -> coarseGridNumber-1
  subroutine gpu_prolongateGrid(coarseGridNumber)
    integer, intent(in) :: coarseGridNumber
    call flite_check(flite_prolongate(gpu_ctx, coarseGridNumber), "prolongateGrid")
  end subroutine gpu_prolongateGrid

  ! mapFineToCoarse of u as oneCycle calls it (DEF:12572-12690)
  subroutine gpu_mapFineToCoarse(fineGridNumber)
    integer, intent(in) :: fineGridNumber
    call flite_check(flite_map_u_fine_to_coarse(gpu_ctx, fineGridNumber), "mapFineToCoarse")
  end subroutine gpu_mapFineToCoarse

  ! oneCycle (MG:907-1196)
  subroutine gpu_oneCycle(cycleNo)
    integer, intent(in) :: cycleNo
    call flite_check(flite_cycle(gpu_ctx, cycleNo), "oneCycle")
  end subroutine gpu_oneCycle

  ! writeResiduals' reductions (MG:1302-1445) and calculateLiftAndDrag
  ! (DEF:14710-14870): res(j) = sqrt(sum (u_j-uprev_j)^2), maxres/maxind,
  ! forces = lift, drag, latF, frictionDrag, latM, longM, ...
  subroutine gpu_monitor(res, maxres, maxind, forces)
    real(c_double), intent(out) :: res(7), maxres(7), forces(9)
    integer(c_int), intent(out) :: maxind(7)
    call flite_check(flite_monitor(gpu_ctx, res, maxres, maxind, forces), "writeResiduals")
  end subroutine gpu_monitor

end module flite_gpu
