! Prints the size of the Fortran tgx_soap_hypers and the byte offset of each
! member, for comparison with the C struct (tests/test_fortran.py).
program layout_check
  use iso_c_binding
  use tgx_mod
  implicit none
  type(tgx_soap_hypers), target :: h
  integer(c_intptr_t) :: b
  b = transfer(c_loc(h), b)
  write (*, '(A,I0)') 'size ', c_sizeof(h)
  write (*, '(A,I0)') 'n_species ', transfer(c_loc(h%n_species), b) - b
  write (*, '(A,I0)') 'l_max ', transfer(c_loc(h%l_max), b) - b
  write (*, '(A,I0)') 'alpha_max ', transfer(c_loc(h%alpha_max), b) - b
  write (*, '(A,I0)') 'rcut_hard ', transfer(c_loc(h%rcut_hard), b) - b
  write (*, '(A,I0)') 'rcut_soft ', transfer(c_loc(h%rcut_soft), b) - b
  write (*, '(A,I0)') 'central_weight ', transfer(c_loc(h%central_weight), b) - b
  write (*, '(A,I0)') 'nf ', transfer(c_loc(h%nf), b) - b
end program layout_check
