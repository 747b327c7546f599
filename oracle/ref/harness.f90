!===============================================================================
! Test-oracle driver built against the unmodified Walicxe3D reference sources.
! It replaces only the reference's `program Walicxe3D` (source/main.f90:32-134)
! so that the state can be written out at fixed step counts.
!
! Control file `harness.nml` (namelist /ctl/):
!   mode   = 'run'  : initmain, basegrid, initflow, loadBalance, updatePrims
!                     (source/main.f90:53-68), then `nsteps` iterations of
!                     hydroSolver, updatePrims, cooling (source/main.f90:81-100);
!                     admesh/loadBalance/output are not called so the block
!                     topology stays the one built from ic.bin.  The output
!                     clamp of getTimestep advances `nextout` as main.f90 does
!                     after writing an output.
!   mode   = 'unit' : evaluates one reference routine on every record of
!                     unit_in.bin and writes unit_out.bin (op codes below).
!   amr    = .true. : after cooling, every step also runs the reference's
!                     mesh update as main.f90:103-106 does (admesh, then
!                     loadBalance), writing first the markByPhysical flag of
!                     every active block (flags_NNNNN.bin: nact, then bID and
!                     flag pairs) and afterwards the active bIDs in slot order
!                     (blocks_NNNNN.bin: nact, then bIDs), the order in which
!                     the state dumps of that step are written.
!   bindump = .true. : at every dumped step also the reference's own output
!                     routines writeBin and writeState (source/output.f90:
!                     156-268) with the step as output number
!                     (Blocks000.NNNN.bin, State.NNNN.dat)
!   vtkdump = .true. : at every dumped step also the reference's VTK output
!                     as writeOutput does it (updatePrims, write3DVTKBlocks,
!                     write3DVTKGrid: Blocks000.NNNN.vtk, Grid.NNNN.vtk,
!                     master.visit); updatePrims replaces the interior PRIM,
!                     so later steps are those of a run with VTK output
!   nsteps, ndump, dumps(:)  - steps after which U is written (0 = initial)
!   tstop  - if > 0, stop (and write the state) after the first step that
!            reaches time >= tstop
!
! Output of 'run': topo.bin, steps.txt, u_NNNNN.bin, prim_NNNNN.bin
!
! This file is test infrastructure; it is not part of the GPU product.
!===============================================================================
program wx_oracle_harness

  use parameters
  use globals
  use init
  use hydro_core
  use hydro_solver, only : hydroSolver
#ifdef WX_DROPIN
  use hydro_solver, only : hydro_gpu_set_plane_wind, hydro_gpu_set_gravity
  use userconds, only : have_pwind, pwind, have_grav, grav
#endif
  use coolingModule, only : cooling, find_coef
  use load_balance, only : loadBalance
  use amr, only : meshlevel, markByPhysical, admesh, getHilbertKey
  use output, only : writeBin, writeState, write3DVTKBlocks, write3DVTKGrid
  use HLL, only : primfhll
  use HLLC, only : primfhllc
#ifdef BFIELD
  use HLLE, only : primfhlle
  use HLLD, only : primfhlld
#endif
  implicit none

  character(len=8) :: mode = 'run'
  integer :: nsteps = 0, ndump = 0
  integer :: dumps(64) = -1
  integer :: op = 0
  real(8) :: tstop = -1.0
  logical :: amr = .false., bindump = .false., vtkdump = .false.
  namelist /ctl/ mode, nsteps, ndump, dumps, op, tstop, amr, bindump, vtkdump

  integer :: s

  open (unit=70, file='harness.nml', status='old', action='read')
  read (70, nml=ctl)
  close (70)

  if (trim(mode) == 'unit') then
    call unit_mode (op)
    stop
  end if

  call initmain ()
  call basegrid ()
  call initflow ()
  call loadBalance ()
  call updatePrims ()
  if (nextout == 0) nextout = 1
#ifdef WX_DROPIN
  ! the GPU drop-in does not call userBoundary or get_user_source_terms: a
  ! problem registers its plane wind and its uniform acceleration with it
  if (have_pwind) call hydro_gpu_set_plane_wind (pwind%plane, pwind%rho, pwind%vel, pwind%temp, pwind%mu)
  if (have_grav) call hydro_gpu_set_gravity (grav)
#endif

  ! multi-rank runs (timing): the master writes the step log only
  if (rank == master) then
    if (nProcs == 1) call write_topology ()
    open (unit=71, file='steps.txt', status='replace', action='write')
  end if
  if (nProcs == 1 .and. any(dumps(1:max(ndump,1)) == 0)) call write_state (0)
  if ((bindump .or. vtkdump) .and. any(dumps(1:max(ndump,1)) == 0)) call write_output_files (0)

  do s=1,nsteps
    it = it + 1
    call hydroSolver ()
    call updatePrims ()
    call cooling ()
    if (amr) then
      call write_flags (s)
      call admesh ()
      call loadBalance ()
      call write_blocks (s)
    end if
    if (rank == master) write (71,'(i8,3es26.17e3)') it, dt, time, nextout*dtout/t_sc
    if (dumpout) then
      nextout = nextout + 1
      dumpout = .false.
    end if
    if (nProcs == 1 .and. any(dumps(1:max(ndump,1)) == s)) call write_state (s)
    if ((bindump .or. vtkdump) .and. any(dumps(1:max(ndump,1)) == s)) call write_output_files (s)
    if (tstop > 0.0 .and. time >= tstop) then
      if (nProcs == 1 .and. .not. any(dumps(1:max(ndump,1)) == s)) call write_state (s)
      exit
    end if
  end do
  if (rank == master) close (71)
  call mpi_finalize (ierr)

contains

  !-----------------------------------------------------------------------------
  subroutine write_topology ()
    integer :: nb, lev, nact
    nact = 0
    do nb=1,nbMaxProc
      if (localBlocks(nb) /= -1) nact = nact + 1
    end do
    open (unit=72, file='topo.bin', access='stream', form='unformatted', status='replace')
    write (72) int(nact,4), int(maxlev,4), int(nbrootx,4), int(nbrooty,4), int(nbrootz,4)
    do nb=1,nbMaxProc
      if (localBlocks(nb) /= -1) then
        call meshlevel (localBlocks(nb), lev)
        write (72) int(localBlocks(nb),4), int(lev,4)
      end if
    end do
    write (72) dx(1:maxlev), dy(1:maxlev), dz(1:maxlev)
    write (72) gamma, cv, CFL, visc_eta, d_sc, e_sc, t_sc, p_sc, l_sc, v_sc
    close (72)
  end subroutine write_topology

  !-----------------------------------------------------------------------------
  ! the reference's checkpoint files for output number `step`
  subroutine write_output_files (step)
    integer, intent(in) :: step
    if (bindump) then
      call writeBin (step)
      if (rank == master) call writeState (step)
    end if
    if (vtkdump) then
      call updatePrims ()
      call write3DVTKBlocks (step)
      if (rank == master) call write3DVTKGrid (step)
    end if
  end subroutine write_output_files

  !-----------------------------------------------------------------------------
  ! markByPhysical (source/amr.f90:142-230) of every active block, before admesh
  subroutine write_flags (step)
    integer, intent(in) :: step
    character(len=32) :: fn
    integer :: nb, nact, flag
    nact = count(localBlocks(1:nbMaxProc) /= -1)
    write (fn,'(a,i5.5,a)') 'flags_', step, '.bin'
    open (unit=75, file=trim(fn), access='stream', form='unformatted', status='replace')
    write (75) int(nact,4)
    do nb=1,nbMaxProc
      if (localBlocks(nb) /= -1) then
        call markByPhysical (nb, localBlocks(nb), flag)
        write (75) int(localBlocks(nb),4), int(flag,4)
      end if
    end do
    close (75)
  end subroutine write_flags

  !-----------------------------------------------------------------------------
  ! Active bIDs in slot order (the order of write_state)
  subroutine write_blocks (step)
    integer, intent(in) :: step
    character(len=32) :: fn
    integer :: nb, nact
    nact = count(localBlocks(1:nbMaxProc) /= -1)
    write (fn,'(a,i5.5,a)') 'blocks_', step, '.bin'
    open (unit=76, file=trim(fn), access='stream', form='unformatted', status='replace')
    write (76) int(nact,4)
    do nb=1,nbMaxProc
      if (localBlocks(nb) /= -1) write (76) int(localBlocks(nb),4)
    end do
    close (76)
  end subroutine write_blocks

  !-----------------------------------------------------------------------------
  ! Interior U and PRIM of every active block, in slot order; per block the
  ! values run i fastest, then j, k, then variable.
  subroutine write_state (step)
    integer, intent(in) :: step
    character(len=32) :: fn
    integer :: nb, ieq
    write (fn,'(a,i5.5,a)') 'u_', step, '.bin'
    open (unit=73, file=trim(fn), access='stream', form='unformatted', status='replace')
    write (fn,'(a,i5.5,a)') 'prim_', step, '.bin'
    open (unit=74, file=trim(fn), access='stream', form='unformatted', status='replace')
    do nb=1,nbMaxProc
      if (localBlocks(nb) /= -1) then
        do ieq=1,neqtot
          write (73) U(nb,ieq,1:ncells_x,1:ncells_y,1:ncells_z)
          write (74) PRIM(nb,ieq,1:ncells_x,1:ncells_y,1:ncells_z)
        end do
      end if
    end do
    close (73)
    close (74)
  end subroutine write_state

  !-----------------------------------------------------------------------------
  ! op codes (records of NV-vectors unless stated):
  !  1 flow2prim   in u            out p, istat
  !  2 prim2flow   in p            out u
  !  3 prim2fluxes in p, dim       out f
  !  4 limiter     in pll,pl,pr,prr, lim   out pl,pr
  !  5 primfhll    in pL,pR        out f
  !  6 primfhllc   in pL,pR        out f
  !  7 primfhlle   in pL,pR        out f      (MHD build)
  !  8 primfhlld   in pL,pR        out f      (MHD build)
  !  9 sound/cfast in p            out cs, cfx, cfy, cfz, cfastX
  ! 10 find_coef   in logT         out coef   (needs a cooling table)
  ! 11 calcTemp    in p            out T
  ! 12 getHilbertKey in bID         out key (on the workload's mesh: needs ic.bin)
  subroutine unit_mode (opcode)
    integer, intent(in) :: opcode
    integer(4) :: n
    integer :: r, istat, lim, dimn, hk
    real(8) :: a(neqtot), b(neqtot), c(neqtot), d(neqtot), f(neqtot), x, y
    real :: cs, cfx, cfy, cfz, cfX1

    if (opcode == 10) then
      if (cooling_type == COOL_NONE) stop 6
      call loadcooldata_wrap ()
    end if
    if (opcode == 12) then
      call initmain ()
      call basegrid ()
    end if

    open (unit=75, file='unit_in.bin', access='stream', form='unformatted', status='old')
    open (unit=76, file='unit_out.bin', access='stream', form='unformatted', status='replace')
    read (75) n
    do r=1,n
      select case (opcode)
      case (1)
        read (75) a
        call flow2prim (a, b, istat)
        write (76) b, real(istat,8)
      case (2)
        read (75) a
        call prim2flow (a, b)
        write (76) b
      case (3)
        read (75) a, x
        dimn = nint(x)
        call prim2fluxes (a, dimn, b)
        write (76) b
      case (4)
        read (75) a, b, c, d, x
        lim = nint(x)
        call limiter (a, b, c, d, lim, neqtot)
        write (76) b, c
      case (5)
        read (75) a, b
        f = 0.0
        call primfhll (a, b, f)
        write (76) f
      case (6)
        read (75) a, b
        f = 0.0
        call primfhllc (a, b, f)
        write (76) f
#ifdef BFIELD
      case (7)
        read (75) a, b
        f = 0.0
        call primfhlle (a, b, f)
        write (76) f
      case (8)
        read (75) a, b
        f = 0.0
        call primfhlld (a, b, f)
        write (76) f
#endif
      case (9)
        read (75) a
        call sound (a, cs)
        cfx = 0.0; cfy = 0.0; cfz = 0.0; cfX1 = 0.0
#ifdef BFIELD
        call cfast (a(5), a(1), a(6), a(7), a(8), cfx, cfy, cfz)
        call cfastX (a, cfX1)
#endif
        write (76) real(cs,8), real(cfx,8), real(cfy,8), real(cfz,8), real(cfX1,8)
      case (10)
        read (75) x
        call find_coef (x, y)
        write (76) y
      case (11)
        read (75) a
        call calcTemp (a, x)
        write (76) x
      case (12)
        read (75) x
        call getHilbertKey (nint(x), hk)
        write (76) real(hk,8)
      case default
        stop 7
      end select
    end do
    close (75)
    close (76)
  end subroutine unit_mode

  subroutine loadcooldata_wrap ()
    use coolingModule, only : loadcooldata
    integer :: ierr2
    call mpi_init (ierr2)
    rank = 0
    call loadcooldata ()
  end subroutine loadcooldata_wrap

end program wx_oracle_harness
