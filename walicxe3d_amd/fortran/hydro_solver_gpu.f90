!===============================================================================
! Drop-in GPU replacement of Walicxe3D's module `hydro_solver`
! (source/hydro_solver.f90): same module name, same public subroutines and
! interfaces -
!   hydroSolver()          source/hydro_solver.f90:39-121
!   getTimestep(dt_glob)   source/hydro_solver.f90:129-211
!   doViscousStep()        source/hydro_solver.f90:217-250
! A Walicxe3D build links this file (and walicxe_hip.f90) instead of
! source/hydro_solver.f90 plus libwalicxe_hip.so; the rest of the code is
! unchanged.  The integrator then runs on the GPU:
!   - the GPU context is created on the first call from the `parameters`
!     module and the cooling table in `globals`;
!   - the block list is (re)sent whenever localBlocks changes (AMR);
!   - plain form (default; results match the CPU path bit for bit): each
!     call uploads U and the interior PRIM left by updatePrims/cooling, runs
!     getTimestep/Godunov/doViscousStep on the device and downloads U;
!   - fused forms (hydro_gpu_set_form, or WALICXE_GPU_FORM=fused|fused_fp64
!     in the environment): one call is one device step of the fused stage
!     kernels (getTimestep with its ramp and output clamp, both stages and
!     the viscosity), then U comes back.  The state stays on the device
!     between calls: it is sent again only after the block list changed,
!     after cooling (which the reference's main loop applies on the host
!     right after hydroSolver) or after hydro_gpu_state_changed;
! so updatePrims, cooling, admesh and output keep working on the host
! arrays as before.
! Supported: SOLVER_HLL1/HLL/HLLC/HLLE/HLLD (the fused forms: HLL/HLLC for
! hydro, HLLE/HLLD for MHD), COOL_NONE/COOL_TABLE, one MPI rank per GPU.
! Passive scalars: the plain FP64 form carries one or two for hydro (without
! wind sources) with the flow variables; otherwise a run whose passive
! scalars are not all zero stops with a message (zero passives stay zero
! under the reference's update, so they are left as they are).  userBoundary effects
! other than periodic faces (handled by the library) must be registered
! through hydro_gpu_set_winds (spherical sources) or hydro_gpu_set_plane_wind.
! With user_source_terms on, get_user_source_terms runs on the device only as
! a uniform acceleration registered through hydro_gpu_set_gravity (plain
! form); without one the run stops with a message.
!===============================================================================
module hydro_solver

  use iso_c_binding
  use walicxe_hip
  implicit none

  private
  public :: hydroSolver, getTimestep, doViscousStep, hydro_gpu_set_winds, hydro_gpu_release
  public :: hydro_gpu_set_plane_wind, hydro_gpu_set_gravity
  public :: hydro_gpu_set_form, hydro_gpu_state_changed

  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int), save :: gpu_form = -1      ! WX_FORM_*, fixed at the first call
  logical, save :: dev_current = .false.    ! the device holds the state of the host arrays
  integer(c_int), allocatable, target, save :: mesh_bids(:)
  integer(c_int), allocatable, target, save :: lbuf(:)
  real(c_double), allocatable, target, save :: ubuf(:,:,:,:,:)
  real(c_double), allocatable, target, save :: tab_T(:), tab_C(:)
  real(c_double), allocatable, target, save :: wind_rec(:)
  integer, save :: nwind = 0
  integer(c_int), save :: pw_plane = 0         ! plane wind (hydro_gpu_set_plane_wind)
  real(c_double), save :: pw_state(4) = 0.0d0  ! rho, vel, temp, mu
  integer, save :: nvg = 0                     ! variables on the device: neqhydro+neqmhd (+ passives)
  integer(c_int), save :: grav_on = 0          ! uniform acceleration (hydro_gpu_set_gravity)
  real(c_double), save :: grav_g(3) = 0.0d0    ! gx, gy, gz (code units)

contains

  !> Spherical wind sources re-imposed at every boundary call
  !! (source/winds.f90:123): 8 values per source, xc,yc,zc,radius,mdot,vinf,
  !! temp,mu in cgs.  Call before the first hydroSolver.
  subroutine hydro_gpu_set_winds(n, rec)
    integer, intent(in) :: n
    real(c_double), intent(in) :: rec(8, n)
    if (allocated(wind_rec)) deallocate(wind_rec)
    allocate(wind_rec(8*max(n,1)))
    if (n > 0) wind_rec(1:8*n) = reshape(rec, [8*n])
    nwind = n
  end subroutine hydro_gpu_set_winds

  !> The plane wind of a userBoundary that calls imposePlaneWind
  !! (source/winds.f90:250-340): plane (PLANE_LEFT=1 .. PLANE_TOP=6; 0
  !! removes it), rho, vel, temp, mu in cgs.  Plain FP64 form and hydro only
  !! (the library refuses it otherwise).  Takes effect at the next
  !! hydroSolver.
  subroutine hydro_gpu_set_plane_wind(plane, rho, vel, temp, mu)
    integer, intent(in) :: plane
    real(c_double), intent(in) :: rho, vel, temp, mu
    pw_plane = plane
    pw_state = [rho, vel, temp, mu]
    if (c_associated(ctx)) call check(wx_set_plane_wind(ctx, pw_plane, pw_state(1), pw_state(2), &
                                                        pw_state(3), pw_state(4)), "wx_set_plane_wind")
  end subroutine hydro_gpu_set_plane_wind

  !> The source term of a get_user_source_terms (source/sources.f90:124)
  !! that adds a uniform acceleration g (code units): rho*g to the momenta,
  !! rho*(v.g) to the energy.  Plain FP64 form (the library refuses it
  !! otherwise).  Takes effect at the next hydroSolver.
  subroutine hydro_gpu_set_gravity(g)
    real(c_double), intent(in) :: g(3)
    grav_on = 1
    grav_g = g
    if (c_associated(ctx)) call check(wx_set_uniform_gravity(ctx, grav_on, grav_g(1), grav_g(2), grav_g(3)), &
                                      "wx_set_uniform_gravity")
  end subroutine hydro_gpu_set_gravity

  subroutine hydro_gpu_release()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = wx_finalize(ctx)
    ctx = c_null_ptr
    if (allocated(mesh_bids)) deallocate(mesh_bids)
    dev_current = .false.
  end subroutine hydro_gpu_release

  !> Form of the device update (WX_FORM_PLAIN_FP64, WX_FORM_FUSED_FP32 or
  !! WX_FORM_FUSED_FP64); call before the first hydroSolver.
  subroutine hydro_gpu_set_form(form)
    integer, intent(in) :: form
    if (c_associated(ctx)) call check(-1_c_int, "hydro_gpu_set_form (after the first hydroSolver)")
    if (form /= WX_FORM_PLAIN_FP64 .and. form /= WX_FORM_FUSED_FP32 .and. form /= WX_FORM_FUSED_FP64) &
      call check(-1_c_int, "hydro_gpu_set_form (unknown form)")
    gpu_form = form
  end subroutine hydro_gpu_set_form

  !> The host arrays U were changed by other code than cooling or admesh:
  !! the fused forms send them to the device again at the next call.
  subroutine hydro_gpu_state_changed()
    dev_current = .false.
  end subroutine hydro_gpu_state_changed

  !> The form from WALICXE_GPU_FORM unless hydro_gpu_set_form chose one.
  subroutine choose_form()
    character(len=32) :: v
    integer :: l, st
    if (gpu_form >= 0) return
    gpu_form = WX_FORM_PLAIN_FP64
    call get_environment_variable("WALICXE_GPU_FORM", v, l, st)
    if (st /= 0) return
    select case (trim(v))
    case ("", "plain", "plain_fp64")
      gpu_form = WX_FORM_PLAIN_FP64
    case ("fused", "fused_fp32")
      gpu_form = WX_FORM_FUSED_FP32
    case ("fused_fp64")
      gpu_form = WX_FORM_FUSED_FP64
    case default
      call check(-1_c_int, "WALICXE_GPU_FORM (plain, fused or fused_fp64)")
    end select
  end subroutine choose_form

  !> Whether the device carries the passive scalars: the plain FP64 form,
  !! hydro, one or two passives, no wind sources (whose state leaves the
  !! passives unset).
  logical function carry_passives()
    use parameters
    carry_passives = npassive >= 1 .and. npassive <= 2 .and. .not. mhd .and. &
                     gpu_form == WX_FORM_PLAIN_FP64 .and. nwind == 0 .and. pw_plane == 0
  end function carry_passives

  !> Passive scalars the device does not carry: all-zero passives stay zero
  !! under the reference's update, so they are left untouched; any other
  !! value stops the run.
  subroutine check_passives()
    use parameters
    use globals
    use clean_quit, only : clean_abort
    integer :: nb
    if (npassive <= 0 .or. carry_passives()) return
    do nb = 1, nbMaxProc
      if (localBlocks(nb) /= -1) then
        if (any(U(nb, firstpas:neqtot, :, :, :) /= 0.0)) then
          write(logu,'(a,i0,a)') "passive scalars are not advected on the GPU (non-zero passive in block ", &
                                 localBlocks(nb), ")"
          call clean_abort (ERROR_GENERIC)
        end if
      end if
    end do
  end subroutine check_passives

  subroutine check(rc, what)
    use clean_quit, only : clean_abort
    use globals, only : logu
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc < 0) then
      write(logu,'(a,a,a,i0,a,a)') "GPU call ", what, " failed (", rc, "): ", trim(wx_error_text())
      call clean_abort (255)
    end if
  end subroutine check

  !> Context and mesh for the current block list.
  subroutine ensure_context()
    use parameters
    use globals
    type(wx_config), target :: cfg
    integer :: nb, nact, l
    integer(c_int), allocatable :: bids(:)
    real(c_double), allocatable :: dxyz(:)
    integer(c_int) :: rc

    if (.not. c_associated(ctx)) then
      nvg = neqhydro + neqmhd
      if (carry_passives()) nvg = neqtot
      cfg%nv = nvg
      cfg%mhd = merge(1, 0, mhd)
      cfg%solver = solver_type
      cfg%limiter = limiter_type
      cfg%eight_wave = merge(1, 0, eight_wave)
      cfg%cooling = merge(1, 0, cooling_type == COOL_TABLE)
      cfg%eos = eos_type
      cfg%bc = [bc_left, bc_right, bc_front, bc_back, bc_bottom, bc_top]
      cfg%gamma = gamma
      cfg%cv = cv
      cfg%cfl = CFL
      cfg%visc_eta = visc_eta
      cfg%cooling_limit = cooling_limit
      cfg%mu0 = mu0
      cfg%mui = mui
      cfg%ion_thres = ion_thres
      cfg%d_sc = d_sc
      cfg%e_sc = e_sc
      cfg%t_sc = t_sc
      cfg%p_sc = p_sc
      cfg%l_sc = l_sc
      cfg%v_sc = v_sc
      cfg%ntab = 0
      if (cooling_type == COOL_TABLE) then
        allocate(tab_T(nptsT), tab_C(nptsT))
        tab_T(:) = cooltable(1, 1:nptsT)
        tab_C(:) = cooltable(2, 1:nptsT)
        cfg%ntab = nptsT
        cfg%tab_logT = c_loc(tab_T)
        cfg%tab_loglam = c_loc(tab_C)
      else if (cooling_type /= COOL_NONE) then
        call check(-1_c_int, "init (only COOL_NONE and COOL_TABLE run on the GPU)")
      end if
      cfg%nwind = nwind
      if (nwind > 0) cfg%winds = c_loc(wind_rec)
      cfg%form = gpu_form
      if (gpu_form /= WX_FORM_PLAIN_FP64) then
        ! the reference's main loop cools on the host after hydroSolver
        ! (source/main.f90:91-100): the fused step leaves cooling to it
        cfg%cooling = 0
        cfg%ntab = 0
      end if
      call check(wx_init(cfg, ctx), "wx_init")
      if (pw_plane /= 0) call check(wx_set_plane_wind(ctx, pw_plane, pw_state(1), pw_state(2), pw_state(3), &
                                                      pw_state(4)), "wx_set_plane_wind")
      if (grav_on /= 0) call check(wx_set_uniform_gravity(ctx, grav_on, grav_g(1), grav_g(2), grav_g(3)), &
                                   "wx_set_uniform_gravity")
      allocate(ubuf(nbMaxProc, nvg, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(lbuf(nbMaxProc))
    end if

    nact = count(localBlocks(1:nbMaxProc) /= -1)
    allocate(bids(nact))
    nact = 0
    do nb = 1, nbMaxProc
      if (localBlocks(nb) /= -1) then
        nact = nact + 1
        bids(nact) = localBlocks(nb)
      end if
    end do
    if (allocated(mesh_bids)) then
      if (size(mesh_bids) == nact) then
        if (all(mesh_bids == bids)) return
      end if
      deallocate(mesh_bids)
    end if
    allocate(mesh_bids(nact))
    mesh_bids(:) = bids(:)
    dev_current = .false.
    allocate(dxyz(3*maxlev))
    do l = 1, maxlev
      dxyz(l) = dx(l)
      dxyz(maxlev + l) = dy(l)
      dxyz(2*maxlev + l) = dz(l)
    end do
    rc = wx_set_mesh(ctx, [int(nbrootx,c_int), int(nbrooty,c_int), int(nbrootz,c_int)], &
                     int(maxlev,c_int), dxyz, int(nact,c_int), mesh_bids)
    call check(rc, "wx_set_mesh")
  end subroutine ensure_context

  !> U and the interior PRIM of the host arrays to the device.
  subroutine to_device()
    use parameters
    use globals
    lbuf(:) = localBlocks(1:nbMaxProc)
    ubuf(:,:,:,:,:) = U(:, 1:nvg, :, :, :)
    call check(wx_upload_state(ctx, c_loc(ubuf), WX_LAYOUT_FORTRAN, int(nbMaxProc,c_int), c_loc(lbuf)), &
               "wx_upload_state")
    ubuf(:,:,:,:,:) = PRIM(:, 1:nvg, :, :, :)
    call check(wx_upload_prims(ctx, c_loc(ubuf), WX_LAYOUT_FORTRAN, int(nbMaxProc,c_int), c_loc(lbuf)), &
               "wx_upload_prims")
  end subroutine to_device

  !> U of the host arrays to the device (fused forms: PRIM is derived there).
  subroutine to_device_u()
    use parameters
    use globals
    lbuf(:) = localBlocks(1:nbMaxProc)
    ubuf(:,:,:,:,:) = U(:, 1:nvg, :, :, :)
    call check(wx_upload_state(ctx, c_loc(ubuf), WX_LAYOUT_FORTRAN, int(nbMaxProc,c_int), c_loc(lbuf)), &
               "wx_upload_state")
  end subroutine to_device_u

  !> Device U (interior) back into the host array.
  subroutine to_host()
    use parameters
    use globals
    ubuf(:,:,:,:,:) = U(:, 1:nvg, :, :, :)
    call check(wx_download_state(ctx, c_loc(ubuf), c_null_ptr, WX_LAYOUT_FORTRAN, int(nbMaxProc,c_int), &
               c_loc(lbuf)), "wx_download_state")
    U(:, 1:nvg, :, :, :) = ubuf(:,:,:,:,:)
  end subroutine to_host

  !> High-level wrapper for the numerical integrator (GPU).
  subroutine hydroSolver ()
    use parameters
    use globals
    use clean_quit, only : clean_abort
    real :: dt_new
    integer(c_int) :: order, itc, noc
    real(c_double) :: dtc, tc

    if (solver_type == SOLVER_LAX) then
      write(logu,'(a)') "SOLVER_LAX does not run on the GPU"
      call clean_abort (ERROR_GENERIC)
    end if
    if (user_source_terms .and. grav_on == 0) then
      ! get_user_source_terms (source/sources.f90:124) is problem code the
      ! device update does not call; a uniform acceleration is registered
      ! through hydro_gpu_set_gravity
      write(logu,'(a)') "user_source_terms: get_user_source_terms does not run on the GPU (see hydro_gpu_set_gravity)"
      call clean_abort (ERROR_GENERIC)
    end if
    call choose_form()
    call check_passives()
    call ensure_context()
    if (gpu_form /= WX_FORM_PLAIN_FP64) then
      ! one device step: dt (CFL minimum, start-up ramp, output clamp), both
      ! stages and the viscosity; time advances as doViscousStep does
      if (.not. dev_current .or. cooling_type /= COOL_NONE) call to_device_u()
      itc = it - 1
      dtc = dt
      tc = time
      noc = nextout
      call check(wx_run(ctx, 1_c_int, itc, dtc, tc, noc, real(dtout, c_double), real(t_sc, c_double)), "wx_run")
      dt = dtc
      time = tc
      if (noc /= nextout) dumpout = .true.
      call to_host()
      dev_current = .true.
      return
    end if
    call to_device()

    call getTimestep (dt_new)
    if ((dt.gt.0.0).and.(abs(dt-dt_new)/dt.gt.50)) then
      write(logu,'(a)') "Abnormal timestep change!"
      write(logu,'(a,es12.5)') "Previous timestep:", dt
      write(logu,'(a,es12.5)') "New timestep:     ", dt_new
    end if
    dt = dt_new

    order = 2
    if (solver_type == SOLVER_HLL1) order = 1
    call check(wx_godunov(ctx, order, dt), "wx_godunov")
    call doViscousStep()
    call to_host()

  end subroutine hydroSolver

  !> Global timestep: CFL minimum on the device, then the start-up ramp and
  !! the output clamp on the host as in the reference.
  subroutine getTimestep (dt_glob)
    use parameters
    use globals
    real, intent(out) :: dt_glob
    real(c_double) :: dt_raw
    real :: dt_loc

    call check(wx_get_timestep(ctx, dt_raw), "wx_get_timestep")
    dt_loc = dt_raw
    if (it <= 10) then
      dt_loc = (2.**(-real(11-it)))*dt_loc
    end if
#ifdef MPIP
    call MPI_ALLREDUCE(dt_loc, dt_glob, 1, mpi_real_kind, mpi_min, mpi_comm_world, ierr)
#else
    dt_glob = dt_loc
#endif
    if ((time + dt_glob) >= nextout*dtout/t_sc) then
      dt_glob = nextout*dtout/t_sc - time
      dumpout = .true.
    end if
  end subroutine getTimestep

  !> Advances the time and applies the artificial viscosity (device).
  subroutine doViscousStep ()
    use globals, only : time, dt
    time = time + dt
    call check(wx_viscous_step(ctx), "wx_viscous_step")
  end subroutine doViscousStep

end module hydro_solver
