!===============================================================================
! Prints the size of the bind(C) wx_config type and the byte offsets of a few
! of its fields, so that a test can compare them with the C struct.
!===============================================================================
program wx_layout_probe
  use iso_c_binding
  use walicxe_hip
  implicit none
  type(wx_config), target :: c
  integer(c_intptr_t) :: base
  base = transfer(c_loc(c), base)
  write(*,'(i0)') c_sizeof(c)
  write(*,'(i0)') transfer(c_loc(c%gamma), base) - base
  write(*,'(i0)') transfer(c_loc(c%ntab), base) - base
  write(*,'(i0)') transfer(c_loc(c%tab_logT), base) - base
  write(*,'(i0)') transfer(c_loc(c%winds), base) - base
  write(*,'(i0)') transfer(c_loc(c%form), base) - base
end program wx_layout_probe
