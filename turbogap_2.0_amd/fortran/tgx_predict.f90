! predict driver on the tgx engine: energies, forces and virial of one
! structure (the role of "turbogap predict", turbogap_main.f90:483-1136).
! usage: tgx_predict model.gap structure.xyz [fp64|fp32|mixed]
! Output: total energy, virial, then "symbol x y z e_atom fx fy fz" per atom.
program tgx_predict
  use iso_c_binding
  use tgx_mod
  use tgx_io
  implicit none
  type(gap_model) :: m
  character(len=1024) :: fgap, fxyz, prec
  type(c_ptr) :: ctx
  integer :: n, i, k, mode
  real(c_double), allocatable :: pos(:), vel(:), e(:), f(:)
  integer(c_int), allocatable :: species(:)
  real(c_double) :: lattice(9), virial(9)

  call get_command_argument(1, fgap)
  call get_command_argument(2, fxyz)
  prec = 'fp64'
  if (command_argument_count() >= 3) call get_command_argument(3, prec)
  mode = TGX_FP64
  if (trim(prec) == 'fp32') mode = TGX_FP32
  if (trim(prec) == 'mixed') mode = TGX_MIXED

  call read_gap_file(trim(fgap), m)
  call read_xyz_frame(trim(fxyz), m, n, pos, species, lattice, vel)
  ctx = c_null_ptr
  if (tgx_create(0_c_int, int(mode, c_int), ctx) /= TGX_OK) then
     write (*, *) 'ERROR: no GPU context'
     stop 1
  end if
  call tgx_check(ctx, tgx_set_soap_model(ctx, m%h, int(m%dim, c_int), int(m%n_sparse, c_int), &
       m%Qs, m%alphas, m%delta, m%zeta), 'tgx_set_soap_model')
  call tgx_check(ctx, tgx_set_e0(ctx, int(m%h%n_species, c_int), m%e0), 'tgx_set_e0')
  do k = 1, m%n2b
     call tgx_check(ctx, tgx_add_2b_model(ctx, int(m%b2(k)%sp1, c_int), int(m%b2(k)%sp2, c_int), &
          m%b2(k)%rcut, m%b2(k)%buffer, m%b2(k)%delta, m%b2(k)%sigma, &
          int(m%b2(k)%n_sparse, c_int), m%b2(k)%Qs, m%b2(k)%alphas, m%b2(k)%cutoff), &
          'tgx_add_2b_model')
  end do
  allocate (e(n), f(3*n))
  call tgx_check(ctx, tgx_energy_forces(ctx, int(n, c_int), pos, species, lattice, 1_c_int, &
       e, f, virial), 'tgx_energy_forces')
  write (*, '(A,ES24.16)') 'energy ', sum(e)
  write (*, '(A,9ES24.16)') 'virial ', virial
  do i = 1, n
     write (*, '(A4,7ES24.16)') m%species(species(i)), pos(3*i - 2:3*i), e(i), f(3*i - 2:3*i)
  end do
  call tgx_check(ctx, tgx_destroy(ctx), 'tgx_destroy')
end program tgx_predict
