! Prints the size and selected field offsets of the interoperable types, for
! comparison with the C compiler's layout of tc_params, tc_stats and the
! equalize and stats records.
program tc_layout
  use iso_c_binding
  use tracer_hip
  implicit none
  type(tc_params), target :: p
  type(tc_stats), target :: s
  type(tc_equalize_info), target :: e
  type(tc_stats_rec), target :: r
  integer(c_intptr_t) :: b0, b1
  b0 = transfer(c_loc(p), b0)
  write(*, '(a,i0)') 'sizeof_params ', c_sizeof(p)
  b1 = transfer(c_loc(p%multi), b1);      write(*, '(a,i0)') 'off_multi ', b1 - b0
  b1 = transfer(c_loc(p%fixes), b1);      write(*, '(a,i0)') 'off_fixes ', b1 - b0
  b1 = transfer(c_loc(p%aspect), b1);     write(*, '(a,i0)') 'off_aspect ', b1 - b0
  b1 = transfer(c_loc(p%D1), b1);         write(*, '(a,i0)') 'off_D1 ', b1 - b0
  b1 = transfer(c_loc(p%Dint), b1);       write(*, '(a,i0)') 'off_Dint ', b1 - b0
  b1 = transfer(c_loc(p%Trrss_base), b1); write(*, '(a,i0)') 'off_Trrss_base ', b1 - b0
  b0 = transfer(c_loc(s), b0)
  write(*, '(a,i0)') 'sizeof_stats ', c_sizeof(s)
  b1 = transfer(c_loc(s%residual_last), b1); write(*, '(a,i0)') 'off_residual_last ', b1 - b0
  b1 = transfer(c_loc(s%mg_cycles), b1);     write(*, '(a,i0)') 'off_mg_cycles ', b1 - b0
  b1 = transfer(c_loc(s%q_step), b1);        write(*, '(a,i0)') 'off_q_step ', b1 - b0
  b0 = transfer(c_loc(e), b0)
  write(*, '(a,i0)') 'sizeof_equalize ', c_sizeof(e)
  b1 = transfer(c_loc(e%deficit), b1);       write(*, '(a,i0)') 'off_deficit ', b1 - b0
  b1 = transfer(c_loc(e%count), b1);         write(*, '(a,i0)') 'off_count ', b1 - b0
  b0 = transfer(c_loc(r), b0)
  write(*, '(a,i0)') 'sizeof_stats_row ', c_sizeof(r)
  b1 = transfer(c_loc(r%empty_cells), b1);   write(*, '(a,i0)') 'off_empty_cells ', b1 - b0
  b1 = transfer(c_loc(r%ent), b1);           write(*, '(a,i0)') 'off_ent ', b1 - b0
end program tc_layout
