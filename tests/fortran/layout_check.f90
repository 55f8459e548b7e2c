! Prints the size of the Fortran mirrors of the C structs and the byte
! offset of each member, for comparison with tgx.h (tests/test_fortran.py).
program layout_check
  use iso_c_binding
  use tgx_mod
  implicit none
  type(tgx_soap_hypers), target :: h
  type(tgx_da_options), target :: o
  type(tgx_mc_stats), target :: st
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
  b = transfer(c_loc(o), b)
  write (*, '(A,I0)') 'da.size ', c_sizeof(o)
  write (*, '(A,I0)') 'da.n_train_trials ', transfer(c_loc(o%n_train_trials), b) - b
  write (*, '(A,I0)') 'da.train_every ', transfer(c_loc(o%train_every), b) - b
  write (*, '(A,I0)') 'da.epochs ', transfer(c_loc(o%epochs), b) - b
  write (*, '(A,I0)') 'da.batch ', transfer(c_loc(o%batch), b) - b
  write (*, '(A,I0)') 'da.lr ', transfer(c_loc(o%lr), b) - b
  write (*, '(A,I0)') 'da.max_rmse ', transfer(c_loc(o%max_rmse), b) - b
  b = transfer(c_loc(st), b)
  write (*, '(A,I0)') 'st.size ', c_sizeof(st)
  write (*, '(A,I0)') 'st.n_accepted ', transfer(c_loc(st%n_accepted), b) - b
  write (*, '(A,I0)') 'st.n_stage1_rejected ', transfer(c_loc(st%n_stage1_rejected), b) - b
  write (*, '(A,I0)') 'st.n_gap_evaluations ', transfer(c_loc(st%n_gap_evaluations), b) - b
  write (*, '(A,I0)') 'st.n_rebuilds ', transfer(c_loc(st%n_rebuilds), b) - b
  write (*, '(A,I0)') 'st.train_rmse ', transfer(c_loc(st%train_rmse), b) - b
  write (*, '(A,I0)') 'st.da_rmse ', transfer(c_loc(st%da_rmse), b) - b
  write (*, '(A,I0)') 'st.fallback ', transfer(c_loc(st%fallback), b) - b
end program layout_check
