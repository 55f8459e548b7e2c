! predict driver on the tgx engine: energies, forces and virial of every
! structure of an extended-XYZ file (the role of "turbogap predict",
! turbogap_main.f90:483-1136, which loops over the frames of atoms_file).
! usage: tgx_predict model.gap structures.xyz [fp64|fp32|mixed] [out.xyz]
! Output: per frame, on stdout the total energy, the virial, then
! "symbol x y z e_atom fx fy fz" per atom; and every frame as an extended-XYZ
! frame (species, pos, forces, local_energy, Lattice, energy, virial,
! i_config) in out.xyz (default trajectory_out.xyz, the reference's name).
program tgx_predict
  use iso_c_binding
  use tgx_mod
  use tgx_io
  implicit none
  type(gap_model) :: m
  character(len=1024) :: fgap, fxyz, prec, fout
  type(c_ptr) :: ctx
  integer :: n, i, mode, uin, uout, frame
  logical :: ok
  real(c_double), allocatable :: pos(:), vel(:), e(:), f(:)
  integer(c_int), allocatable :: species(:)
  real(c_double) :: lattice(9), virial(9)

  call get_command_argument(1, fgap)
  call get_command_argument(2, fxyz)
  prec = 'fp64'
  if (command_argument_count() >= 3) call get_command_argument(3, prec)
  fout = 'trajectory_out.xyz'
  if (command_argument_count() >= 4) call get_command_argument(4, fout)
  mode = TGX_FP64
  if (trim(prec) == 'fp32') mode = TGX_FP32
  if (trim(prec) == 'mixed') mode = TGX_MIXED

  call read_gap_file(trim(fgap), m)
  ctx = c_null_ptr
  if (tgx_create(0_c_int, int(mode, c_int), ctx) /= TGX_OK) then
     write (*, *) 'ERROR: no GPU context'
     stop 1
  end if
  call load_model(ctx, m)
  open (newunit=uin, file=trim(fxyz), status='old', action='read')
  open (newunit=uout, file=trim(fout), status='replace', action='write')
  frame = 0
  do
     call read_xyz_next(uin, m, n, pos, species, lattice, vel, ok)
     if (.not. ok) exit
     frame = frame + 1
     if (allocated(e)) deallocate (e, f)
     allocate (e(n), f(3*n))
     call tgx_check(ctx, tgx_energy_forces(ctx, int(n, c_int), pos, species, lattice, 1_c_int, &
          e, f, virial), 'tgx_energy_forces')
     write (*, '(A,ES24.16)') 'energy ', sum(e)
     write (*, '(A,9ES24.16)') 'virial ', virial
     do i = 1, n
        write (*, '(A4,7ES24.16)') m%species(species(i)), pos(3*i - 2:3*i), e(i), f(3*i - 2:3*i)
     end do
     call write_extxyz(uout, m, n, species, pos, lattice, e, f, virial, frame)
  end do
  close (uin)
  close (uout)
  if (frame == 0) then
     write (*, *) 'ERROR: no frame in ', trim(fxyz)
     stop 1
  end if
  call tgx_check(ctx, tgx_destroy(ctx), 'tgx_destroy')
end program tgx_predict
