!===============================================================================
! walicxe_hip - ISO_C_BINDING interface to libwalicxe_hip.so
! (include/walicxe_hip.h).  Plain C types only; every function returns the
! library status (0 ok, >0 floored cells, <0 error).
!===============================================================================
module walicxe_hip

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WX_OK = 0
  integer(c_int), parameter :: WX_EIO = -6
  integer(c_int), parameter :: WX_FORM_PLAIN_FP64 = 0
  integer(c_int), parameter :: WX_FORM_FUSED_FP32 = 1
  integer(c_int), parameter :: WX_FORM_FUSED_FP64 = 2
  integer(c_int), parameter :: WX_LAYOUT_BLOCKS = 0
  integer(c_int), parameter :: WX_LAYOUT_FORTRAN = 1
  integer(c_int), parameter :: WX_U = 0
  integer(c_int), parameter :: WX_UP = 1
  integer(c_int), parameter :: WX_CELLS_ALL = 0
  integer(c_int), parameter :: WX_CELLS_PHYS = 1
  integer(c_int), parameter :: WX_CELLS_GHOST = 2
  integer(c_int), parameter :: WX_FLAG_NONE = 0
  integer(c_int), parameter :: WX_FLAG_REFINE = 1
  integer(c_int), parameter :: WX_FLAG_COARSE = 2

  ! struct wx_config
  type, bind(C) :: wx_config
    integer(c_int) :: nv, mhd, solver, limiter, eight_wave, cooling, eos
    integer(c_int) :: bc(6)
    real(c_double) :: gamma, cv, cfl, visc_eta
    real(c_double) :: cooling_limit, mu0, mui, ion_thres
    real(c_double) :: d_sc, e_sc, t_sc, p_sc, l_sc, v_sc
    integer(c_int) :: ntab
    type(c_ptr) :: tab_logT = c_null_ptr
    type(c_ptr) :: tab_loglam = c_null_ptr
    integer(c_int) :: nwind
    type(c_ptr) :: winds = c_null_ptr
    integer(c_int) :: form
  end type wx_config

  interface

    integer(c_int) function wx_init(cfg, ctx) bind(C, name="wx_init")
      import :: c_int, c_ptr, wx_config
      type(wx_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function wx_finalize(ctx) bind(C, name="wx_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function wx_last_error() bind(C, name="wx_last_error")
      import :: c_ptr
    end function

    integer(c_int) function wx_set_mesh(ctx, nroot, maxlev, dxyz, nblk, bids) bind(C, name="wx_set_mesh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: nroot(3)
      integer(c_int), value :: maxlev, nblk
      real(c_double), intent(in) :: dxyz(*)
      integer(c_int), intent(in) :: bids(*)
    end function

    integer(c_int) function wx_remesh(ctx, nblk, bids) bind(C, name="wx_remesh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nblk
      integer(c_int), intent(in) :: bids(*)
    end function

    integer(c_int) function wx_mark_blocks(ctx, refine_thres, coarse_thres, flags, maxgrad) &
        bind(C, name="wx_mark_blocks")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, maxgrad
      real(c_double), value :: refine_thres, coarse_thres
      integer(c_int), intent(out) :: flags(*)
    end function

    integer(c_int) function wx_hilbert_keys(nroot, maxlev, nblk, bids, keys) bind(C, name="wx_hilbert_keys")
      import :: c_int
      integer(c_int), intent(in) :: nroot(3)
      integer(c_int), value :: maxlev, nblk
      integer(c_int), intent(in) :: bids(*)
      integer(c_int), intent(out) :: keys(*)
    end function

    integer(c_int) function wx_write_bin(ctx, path, npassive) bind(C, name="wx_write_bin")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: npassive
    end function

    integer(c_int) function wx_read_bin(path, nv, npassive, cap, nblk, bids, u) bind(C, name="wx_read_bin")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nv, npassive, cap
      integer(c_int), intent(out) :: nblk
      integer(c_int), intent(out) :: bids(*)
      real(c_double), intent(out) :: u(*)
    end function

    integer(c_int) function wx_write_state(path, time, t_sc, it, noutput, datadir) bind(C, name="wx_write_state")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*), datadir(*)
      real(c_double), value :: time, t_sc
      integer(c_int), value :: it, noutput
    end function

    integer(c_int) function wx_read_state(path, t_sc, time, it, nextout, datadir, len) bind(C, name="wx_read_state")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), value :: t_sc
      real(c_double), intent(out) :: time
      integer(c_int), intent(out) :: it, nextout
      character(kind=c_char), intent(out) :: datadir(*)
      integer(c_int), value :: len
    end function

    integer(c_int) function wx_write_vtk(ctx, datadir, noutput, npassive, units) bind(C, name="wx_write_vtk")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: datadir(*)
      integer(c_int), value :: noutput, npassive, units
    end function

    integer(c_int) function wx_record_ghosts(ctx) bind(C, name="wx_record_ghosts")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function wx_set_plane_wind(ctx, plane, rho, vel, temp, mu) bind(C, name="wx_set_plane_wind")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: plane
      real(c_double), value :: rho, vel, temp, mu
    end function

    integer(c_int) function wx_set_uniform_gravity(ctx, on, gx, gy, gz) bind(C, name="wx_set_uniform_gravity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      real(c_double), value :: gx, gy, gz
    end function

    integer(c_int) function wx_admesh_plan(nroot, maxlev, nblk, bids, flags, flags_out, cap, nblk_new, bids_new) &
        bind(C, name="wx_admesh_plan")
      import :: c_int, c_ptr
      integer(c_int), intent(in) :: nroot(3)
      integer(c_int), value :: maxlev, nblk, cap
      integer(c_int), intent(in) :: bids(*), flags(*)
      type(c_ptr), value :: flags_out
      integer(c_int), intent(out) :: nblk_new
      integer(c_int), intent(out) :: bids_new(*)
    end function

    integer(c_int) function wx_admesh(ctx, refine_thres, coarse_thres, cap, nblk_new, bids_new) &
        bind(C, name="wx_admesh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: refine_thres, coarse_thres
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: nblk_new
      integer(c_int), intent(out) :: bids_new(*)
    end function

    integer(c_int) function wx_upload_state(ctx, u, layout, nbmax, local_blocks) bind(C, name="wx_upload_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, local_blocks
      integer(c_int), value :: layout, nbmax
    end function

    integer(c_int) function wx_upload_prims(ctx, p, layout, nbmax, local_blocks) bind(C, name="wx_upload_prims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, p, local_blocks
      integer(c_int), value :: layout, nbmax
    end function

    integer(c_int) function wx_download_state(ctx, u, prim, layout, nbmax, local_blocks) &
        bind(C, name="wx_download_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, prim, local_blocks
      integer(c_int), value :: layout, nbmax
    end function

    integer(c_int) function wx_get_timestep(ctx, dt_raw) bind(C, name="wx_get_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt_raw
    end function

    integer(c_int) function wx_godunov(ctx, order, dt) bind(C, name="wx_godunov")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: order
      real(c_double), value :: dt
    end function

    integer(c_int) function wx_viscous_step(ctx) bind(C, name="wx_viscous_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function wx_update_prims(ctx) bind(C, name="wx_update_prims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function wx_cooling(ctx, dt) bind(C, name="wx_cooling")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    integer(c_int) function wx_step(ctx, it, dt, time, nextout, dtout, t_sc) bind(C, name="wx_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(inout) :: it, nextout
      real(c_double), intent(inout) :: dt, time
      real(c_double), value :: dtout, t_sc
    end function

    integer(c_int) function wx_run(ctx, nsteps, it, dt, time, nextout, dtout, t_sc) bind(C, name="wx_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      integer(c_int), intent(inout) :: it, nextout
      real(c_double), intent(inout) :: dt, time
      real(c_double), value :: dtout, t_sc
    end function

    integer(c_int) function wx_boundary(ctx, depth, which) bind(C, name="wx_boundary")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: depth, which
    end function

    integer(c_int) function wx_calc_prims(ctx, which, cells) bind(C, name="wx_calc_prims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, cells
    end function

  end interface

contains

  !> Text of the library's last error.
  function wx_error_text() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    msg = ''
    call c_f_pointer(wx_last_error(), s, [256])
    do n = 1, 256
      if (s(n) == c_null_char) exit
      msg(n:n) = s(n)
    end do
  end function wx_error_text

end module walicxe_hip
