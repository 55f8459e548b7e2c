! NVE molecular dynamics driver on the tgx engine (the role of "turbogap md"
! with velocity Verlet, md/md_utils.f90:486-532, and the skin rebuild,
! md/md_interface.f90:1112-1151).  The whole run stays on the device.
! usage: tgx_md model.gap structure.xyz n_steps dt_fs skin_A [fp64|fp32|mixed]
! Velocities come from the xyz vx vy vz columns (zero if absent); masses are
! those of C/H/O/N by symbol.  Output: "step e_pot e_kin e_tot" lines.
program tgx_md
  use iso_c_binding
  use tgx_mod
  use tgx_io
  implicit none
  type(gap_model) :: m
  character(len=1024) :: fgap, fxyz, arg, prec
  type(c_ptr) :: ctx
  integer :: n, i, k, mode, n_steps, done, nreb
  real(c_double), allocatable :: pos(:), vel(:), f(:), mass(:), ep(:), ek(:)
  integer(c_int), allocatable :: species(:)
  real(c_double) :: lattice(9), dt, skin

  call get_command_argument(1, fgap)
  call get_command_argument(2, fxyz)
  call get_command_argument(3, arg)
  read (arg, *) n_steps
  call get_command_argument(4, arg)
  read (arg, *) dt
  call get_command_argument(5, arg)
  read (arg, *) skin
  prec = 'fp64'
  if (command_argument_count() >= 6) call get_command_argument(6, prec)
  mode = TGX_FP64
  if (trim(prec) == 'fp32') mode = TGX_FP32
  if (trim(prec) == 'mixed') mode = TGX_MIXED

  call read_gap_file(trim(fgap), m)
  call read_xyz_frame(trim(fxyz), m, n, pos, species, lattice, vel)
  allocate (mass(n), f(3*n), ep(n_steps + 1), ek(n_steps + 1))
  do i = 1, n
     select case (trim(m%species(species(i))))
     case ('H')
        mass(i) = 1.008d0
     case ('O')
        mass(i) = 15.999d0
     case ('N')
        mass(i) = 14.007d0
     case default
        mass(i) = 12.011d0
     end select
  end do
  ctx = c_null_ptr
  if (tgx_create(0_c_int, int(mode, c_int), ctx) /= TGX_OK) then
     write (*, *) 'ERROR: no GPU context'
     stop 1
  end if
  call load_model(ctx, m)
  call tgx_check(ctx, tgx_md_begin(ctx, int(n, c_int), species, lattice, mass, pos, vel, dt, skin, &
       int(n_steps + 1, c_int)), 'tgx_md_begin')
  call tgx_check(ctx, tgx_md_steps(ctx, int(n_steps, c_int)), 'tgx_md_steps')
  call tgx_check(ctx, tgx_md_read(ctx, pos, vel, f, ep, ek, done, nreb), 'tgx_md_read')
  call tgx_check(ctx, tgx_md_end(ctx), 'tgx_md_end')
  do k = 1, done + 1
     write (*, '(I8,3ES24.16)') k - 1, ep(k), ek(k), ep(k) + ek(k)
  end do
  write (*, '(A,I0)') 'rebuilds ', nreb
  call tgx_check(ctx, tgx_destroy(ctx), 'tgx_destroy')
end program tgx_md
