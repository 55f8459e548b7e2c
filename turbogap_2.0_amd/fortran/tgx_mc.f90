! Metropolis Monte Carlo driver on the tgx engine (the role of "turbogap mc"
! with single-atom moves, mc/mc_interface.f90:183-447 and mc/mc_utils.f90).
! usage: tgx_mc model.gap structure.xyz n_trials T_K move_max skin_A seed
!               [plain|da] [n_train] [fp64|fp32|mixed]
! move_max is the reference's mc_move_max (each component moves by up to
! move_max/2).  "da" trains the on-device surrogate on the first n_train
! trials (plain Metropolis) and screens the remaining ones with it (delayed
! acceptance).  Output: "accepted", "e_final", "rebuilds" and, for da,
! "stage1_rejected", "gap_evaluations", "train_rmse", "fallback" lines, then
! the final positions ("pos i x y z").
program tgx_mc
  use iso_c_binding
  use tgx_mod
  use tgx_io
  implicit none
  type(gap_model) :: m
  character(len=1024) :: fgap, fxyz, arg, prec, kind
  type(c_ptr) :: ctx
  integer :: n, i, k, mode, n_trials
  integer(c_int) :: nacc, nreb
  integer(c_int64_t) :: seed
  real(c_double), allocatable :: pos(:), vel(:), etr(:)
  integer(c_int), allocatable :: species(:)
  real(c_double) :: lattice(9), temp, move_max, skin
  type(tgx_da_options) :: opt
  type(tgx_mc_stats) :: st

  call get_command_argument(1, fgap)
  call get_command_argument(2, fxyz)
  call get_command_argument(3, arg)
  read (arg, *) n_trials
  call get_command_argument(4, arg)
  read (arg, *) temp
  call get_command_argument(5, arg)
  read (arg, *) move_max
  call get_command_argument(6, arg)
  read (arg, *) skin
  call get_command_argument(7, arg)
  read (arg, *) seed
  kind = 'plain'
  if (command_argument_count() >= 8) call get_command_argument(8, kind)
  if (command_argument_count() >= 9) then
     call get_command_argument(9, arg)
     read (arg, *) opt%n_train_trials
  end if
  prec = 'fp64'
  if (command_argument_count() >= 10) call get_command_argument(10, prec)
  mode = TGX_FP64
  if (trim(prec) == 'fp32') mode = TGX_FP32
  if (trim(prec) == 'mixed') mode = TGX_MIXED

  call read_gap_file(trim(fgap), m)
  call read_xyz_frame(trim(fxyz), m, n, pos, species, lattice, vel)
  allocate (etr(n_trials + 1))
  ctx = c_null_ptr
  if (tgx_create(0_c_int, int(mode, c_int), ctx) /= TGX_OK) then
     write (*, *) 'ERROR: no GPU context'
     stop 1
  end if
  call load_model(ctx, m)
  if (trim(kind) == 'da') then
     call tgx_check(ctx, tgx_surrogate_init(ctx, int(m%sb(1)%dim, c_int), seed), 'tgx_surrogate_init')
     opt%train_every = max(64, min(2048, opt%n_train_trials / 4))
     call tgx_check(ctx, tgx_mc_run_da(ctx, int(n, c_int), species, lattice, pos, &
          int(n_trials, c_int), temp, move_max, skin, seed, opt, etr, st), 'tgx_mc_run_da')
     nacc = int(st%n_accepted, c_int)
     nreb = int(st%n_rebuilds, c_int)
  else
     call tgx_check(ctx, tgx_mc_run(ctx, int(n, c_int), species, lattice, pos, &
          int(n_trials, c_int), temp, move_max, skin, seed, etr, nacc, nreb), 'tgx_mc_run')
  end if
  write (*, '(A,I0)') 'accepted ', nacc
  write (*, '(A,ES25.17)') 'e_final ', etr(n_trials + 1)
  write (*, '(A,I0)') 'rebuilds ', nreb
  if (trim(kind) == 'da') then
     write (*, '(A,I0)') 'stage1_rejected ', st%n_stage1_rejected
     write (*, '(A,I0)') 'gap_evaluations ', st%n_gap_evaluations
     write (*, '(A,ES25.17)') 'train_rmse ', st%train_rmse
     write (*, '(A,I0)') 'fallback ', st%fallback
  end if
  do i = 1, n
     write (*, '(A,I0,3ES25.17)') 'pos ', i, pos(3 * i - 2), pos(3 * i - 1), pos(3 * i)
  end do
  call tgx_check(ctx, tgx_destroy(ctx), 'tgx_destroy')
end program tgx_mc
