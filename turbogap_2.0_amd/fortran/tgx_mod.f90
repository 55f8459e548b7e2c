! iso_c_binding interface to the tgx C-ABI (include/tgx.h).
!
! This is the module TurboGAP's drivers use in place of their CPU hot path:
! build_neighbors_list (neighbors/neighbors_interface.f90:144), the SOAP/GAP
! term (gap/calculate_gap_soap.f90:171, gap/gap.f90:39), the 2-body term
! (gap/gap.f90:246) and the velocity-Verlet loop (md/md_utils.f90:486).  The
! argument conventions are those of the reference's own bind(C) layer
! (gpu/gpu_interface.f90:34-703): c_ptr contexts, value scalars, column-major
! arrays passed by address.  A non-zero return code is turned into the
! reference's behaviour (print the message, then stop) by tgx_check.
module tgx_mod
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: TGX_MAX_SPECIES = 8, TGX_MAX_SOAP_BLOCKS = 8
  integer(c_int), parameter, public :: TGX_OK = 0, TGX_ERR_ARG = 1, TGX_ERR_HIP = 2, &
       TGX_ERR_CAPACITY = 3, TGX_ERR_UNSUPPORTED = 4, TGX_ERR_STATE = 5
  integer(c_int), parameter, public :: TGX_FP64 = 0, TGX_FP32 = 1, TGX_MIXED = 2

  ! Mirror of tgx_soap_hypers (same field order and kinds as the C struct).
  type, bind(C), public :: tgx_soap_hypers
     integer(c_int32_t) :: n_species = 1
     integer(c_int32_t) :: l_max = 8
     integer(c_int32_t) :: radial_enhancement = 0
     integer(c_int32_t) :: central_species = 0
     integer(c_int32_t) :: basis = 0
     integer(c_int32_t) :: pad_ = 0
     integer(c_int32_t) :: alpha_max(TGX_MAX_SPECIES) = 8
     real(c_double) :: rcut_hard(TGX_MAX_SPECIES) = 5.0d0
     real(c_double) :: rcut_soft(TGX_MAX_SPECIES) = 4.5d0
     real(c_double) :: atom_sigma_r(TGX_MAX_SPECIES) = 0.5d0
     real(c_double) :: atom_sigma_t(TGX_MAX_SPECIES) = 0.5d0
     real(c_double) :: atom_sigma_r_scaling(TGX_MAX_SPECIES) = 0.0d0
     real(c_double) :: atom_sigma_t_scaling(TGX_MAX_SPECIES) = 0.0d0
     real(c_double) :: amplitude_scaling(TGX_MAX_SPECIES) = 1.0d0
     real(c_double) :: central_weight(TGX_MAX_SPECIES) = 1.0d0
     real(c_double) :: global_scaling(TGX_MAX_SPECIES) = 1.0d0
     real(c_double) :: nf(TGX_MAX_SPECIES) = 4.0d0
     ! compress_soap: number of entries of P and their arrays (1-based row,
     ! column, value; c_loc of arrays that live while the model is set)
     integer(c_int32_t) :: compress_P_nonzero = 0
     integer(c_int32_t) :: pad2_ = 0
     type(c_ptr) :: compress_P_i = c_null_ptr
     type(c_ptr) :: compress_P_j = c_null_ptr
     type(c_ptr) :: compress_P_el = c_null_ptr
  end type tgx_soap_hypers

  ! Delayed-acceptance options and Monte Carlo statistics (tgx_da_options,
  ! tgx_mc_stats in tgx.h).
  type, bind(C), public :: tgx_da_options
     integer(c_int32_t) :: n_train_trials = 20000
     integer(c_int32_t) :: train_every = 2048
     integer(c_int32_t) :: epochs = 3
     integer(c_int32_t) :: batch = 512
     real(c_double) :: lr = 2.0d-3
     real(c_double) :: max_rmse = 0.2d0
  end type tgx_da_options

  type, bind(C), public :: tgx_mc_stats
     integer(c_int64_t) :: n_accepted = 0
     integer(c_int64_t) :: n_stage1_rejected = 0
     integer(c_int64_t) :: n_gap_evaluations = 0
     integer(c_int64_t) :: n_rebuilds = 0
     real(c_double) :: train_rmse = 0
     real(c_double) :: da_rmse = 0
     integer(c_int32_t) :: fallback = 0
     integer(c_int32_t) :: pad_ = 0
  end type tgx_mc_stats

  public :: tgx_create, tgx_destroy, tgx_last_error, tgx_set_precision, tgx_set_autotune, &
       tgx_set_soap_model, tgx_add_soap_model, tgx_clear_soap_models, tgx_n_soap_models, &
       tgx_energy_forces_noalloc, &
       tgx_set_e0, tgx_add_2b_model, tgx_clear_2b_models, tgx_add_angle_3b, &
       tgx_add_core_pot, tgx_clear_extra_terms, tgx_neighbors_build, &
       tgx_soap_descriptors, tgx_gap_kernel_step, tgx_energy_forces, &
       tgx_energy_forces_batched, tgx_md_run, tgx_md_begin, tgx_md_steps, tgx_md_read, &
       tgx_md_end, tgx_sync, tgx_poly3_orthonormalization, tgx_stage_times, tgx_check, &
       tgx_error_string, tgx_kernel_times, tgx_set_timing, tgx_mc_run, tgx_set_list_buffer, tgx_mc_run_da, &
       tgx_surrogate_init, tgx_surrogate_set, tgx_surrogate_get, tgx_surrogate_predict, &
       tgx_surrogate_gradient, tgx_surrogate_fit, tgx_add_local_property, &
       tgx_clear_local_properties, tgx_local_property, tgx_set_vdw_ts, tgx_clear_vdw, &
       tgx_vdw_energy_forces

  interface
     integer(c_int) function tgx_create(device, precision, ctx) bind(C, name="tgx_create")
       import :: c_int, c_ptr
       integer(c_int), value :: device, precision
       type(c_ptr) :: ctx
     end function tgx_create

     integer(c_int) function tgx_destroy(ctx) bind(C, name="tgx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_destroy

     type(c_ptr) function tgx_last_error(ctx) bind(C, name="tgx_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function tgx_last_error

     integer(c_int) function tgx_set_precision(ctx, precision) bind(C, name="tgx_set_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precision
     end function tgx_set_precision

     ! cache_path: a C string (trim(path)//c_null_char) or c_null_ptr
     integer(c_int) function tgx_set_autotune(ctx, on, cache_path) bind(C, name="tgx_set_autotune")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
       type(c_ptr), value :: cache_path
     end function tgx_set_autotune

     integer(c_int) function tgx_set_soap_model(ctx, h, dim, n_sparse, Qs, alphas, delta, zeta) &
          bind(C, name="tgx_set_soap_model")
       import :: c_int, c_ptr, c_double, tgx_soap_hypers
       type(c_ptr), value :: ctx
       type(tgx_soap_hypers), intent(in) :: h
       integer(c_int), value :: dim, n_sparse
       real(c_double), intent(in) :: Qs(*), alphas(*)
       real(c_double), value :: delta, zeta
     end function tgx_set_soap_model

     integer(c_int) function tgx_add_soap_model(ctx, h, dim, n_sparse, Qs, alphas, delta, zeta) &
          bind(C, name="tgx_add_soap_model")
       import :: c_int, c_ptr, c_double, tgx_soap_hypers
       type(c_ptr), value :: ctx
       type(tgx_soap_hypers), intent(in) :: h
       integer(c_int), value :: dim, n_sparse
       real(c_double), intent(in) :: Qs(*), alphas(*)
       real(c_double), value :: delta, zeta
     end function tgx_add_soap_model

     integer(c_int) function tgx_clear_soap_models(ctx) bind(C, name="tgx_clear_soap_models")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_clear_soap_models

     integer(c_int) function tgx_n_soap_models(ctx) bind(C, name="tgx_n_soap_models")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_n_soap_models

     integer(c_int) function tgx_set_e0(ctx, n_species, e0) bind(C, name="tgx_set_e0")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_species
       real(c_double), intent(in) :: e0(*)
     end function tgx_set_e0

     integer(c_int) function tgx_add_2b_model(ctx, sp1, sp2, rcut, buffer, delta, sigma, &
          n_sparse, Qs, alphas, cutoff) bind(C, name="tgx_add_2b_model")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: sp1, sp2, n_sparse
       real(c_double), value :: rcut, buffer, delta, sigma
       real(c_double), intent(in) :: Qs(*), alphas(*), cutoff(*)
     end function tgx_add_2b_model

     integer(c_int) function tgx_clear_2b_models(ctx) bind(C, name="tgx_clear_2b_models")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_clear_2b_models

     integer(c_int) function tgx_add_angle_3b(ctx, sp_center, sp1, sp2, rcut, delta, sigma, &
          kernel, n_sparse, Qs, alphas, cutoff) bind(C, name="tgx_add_angle_3b")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: sp_center, sp1, sp2, kernel, n_sparse
       real(c_double), value :: rcut, delta
       real(c_double), intent(in) :: sigma(3), Qs(*), alphas(*), cutoff(*)
     end function tgx_add_angle_3b

     integer(c_int) function tgx_add_core_pot(ctx, sp1, sp2, n, x, V, yp1, ypn) &
          bind(C, name="tgx_add_core_pot")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: sp1, sp2, n
       real(c_double), intent(in) :: x(*), V(*)
       real(c_double), value :: yp1, ypn
     end function tgx_add_core_pot

     integer(c_int) function tgx_clear_extra_terms(ctx) bind(C, name="tgx_clear_extra_terms")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_clear_extra_terms

     ! local-property models (read_gap.f90:296-380, local_properties.f90:46-146)
     integer(c_int) function tgx_add_local_property(ctx, block, dim, n_sparse, Qs, alphas, &
          delta, zeta, V0, index_out) bind(C, name="tgx_add_local_property")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, dim, n_sparse
       real(c_double), intent(in) :: Qs(*), alphas(*)
       real(c_double), value :: delta, zeta, V0
       integer(c_int) :: index_out
     end function tgx_add_local_property

     integer(c_int) function tgx_clear_local_properties(ctx) &
          bind(C, name="tgx_clear_local_properties")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_clear_local_properties

     integer(c_int) function tgx_local_property(ctx, idx, n, pos, species, lattice, &
          rcut_list, V, cap, V_der, n_pairs) bind(C, name="tgx_local_property")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: idx, n
       real(c_double), intent(in) :: pos(*), lattice(*)
       integer(c_int), intent(in) :: species(*)
       real(c_double), value :: rcut_list
       real(c_double) :: V(*)
       integer(c_int64_t), value :: cap
       type(c_ptr), value :: V_der
       type(c_ptr), value :: n_pairs
     end function tgx_local_property

     ! Tkatchenko-Scheffler dispersion (vdw/vdw_interface.f90:40-385)
     integer(c_int) function tgx_set_vdw_ts(ctx, local_property, n_species, c6_ref, r0_ref, &
          alpha0_ref, rcut, buffer, rcut_inner, buffer_inner, sR, d) &
          bind(C, name="tgx_set_vdw_ts")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: local_property, n_species
       real(c_double), intent(in) :: c6_ref(*), r0_ref(*), alpha0_ref(*)
       real(c_double), value :: rcut, buffer, rcut_inner, buffer_inner, sR, d
     end function tgx_set_vdw_ts

     integer(c_int) function tgx_clear_vdw(ctx) bind(C, name="tgx_clear_vdw")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_clear_vdw

     integer(c_int) function tgx_vdw_energy_forces(ctx, n, pos, species, lattice, do_forces, &
          energies, forces, virial) bind(C, name="tgx_vdw_energy_forces")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, do_forces
       real(c_double), intent(in) :: pos(*), lattice(*)
       integer(c_int), intent(in) :: species(*)
       real(c_double) :: energies(*), forces(*), virial(*)
     end function tgx_vdw_energy_forces

     integer(c_int) function tgx_neighbors_build(ctx, n, pos, lattice, rcut, cap, n_neigh, &
          list, rjs, xyz, thetas, phis, n_pairs, indices) bind(C, name="tgx_neighbors_build")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: pos(*), lattice(*)
       real(c_double), value :: rcut
       integer(c_int64_t), value :: cap
       integer(c_int) :: n_neigh(*), list(*), indices(3)
       real(c_double) :: rjs(*), xyz(*), thetas(*), phis(*)
       integer(c_int64_t) :: n_pairs
     end function tgx_neighbors_build

     integer(c_int) function tgx_soap_descriptors(ctx, n, pos, species, lattice, soap, &
          n_sites, site_index) bind(C, name="tgx_soap_descriptors")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: pos(*), lattice(*)
       integer(c_int), intent(in) :: species(*)
       real(c_double) :: soap(*)
       integer(c_int) :: n_sites, site_index(*)
     end function tgx_soap_descriptors

     integer(c_int) function tgx_gap_kernel_step(ctx, n_sites, soap, energies, qss) &
          bind(C, name="tgx_gap_kernel_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_sites
       real(c_double), intent(in) :: soap(*)
       real(c_double) :: energies(*), qss(*)
     end function tgx_gap_kernel_step

     integer(c_int) function tgx_energy_forces(ctx, n, pos, species, lattice, do_forces, &
          e_atom, forces, virial) bind(C, name="tgx_energy_forces")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, do_forces
       real(c_double), intent(in) :: pos(*), lattice(*)
       integer(c_int), intent(in) :: species(*)
       real(c_double) :: e_atom(*), forces(*), virial(*)
     end function tgx_energy_forces

     integer(c_int) function tgx_energy_forces_noalloc(ctx, n, pos, species, lattice, do_forces, &
          e_atom, forces, virial) bind(C, name="tgx_energy_forces_noalloc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, do_forces
       real(c_double), intent(in) :: pos(*), lattice(*)
       integer(c_int), intent(in) :: species(*)
       real(c_double) :: e_atom(*), forces(*), virial(*)
     end function tgx_energy_forces_noalloc

     integer(c_int) function tgx_energy_forces_batched(ctx, n_struct, offsets, pos, species, &
          lattices, do_forces, e_atom, forces, virials) bind(C, name="tgx_energy_forces_batched")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_struct, do_forces
       integer(c_int), intent(in) :: offsets(*), species(*)
       real(c_double), intent(in) :: pos(*), lattices(*)
       real(c_double) :: e_atom(*), forces(*), virials(*)
     end function tgx_energy_forces_batched

     integer(c_int) function tgx_md_run(ctx, n, n_steps, dt, skin, species, lattice, mass, pos, &
          vel, forces, e_pot, e_kin, n_rebuilds) bind(C, name="tgx_md_run")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, n_steps
       real(c_double), value :: dt, skin
       integer(c_int), intent(in) :: species(*)
       real(c_double), intent(in) :: lattice(*), mass(*)
       real(c_double) :: pos(*), vel(*), forces(*), e_pot(*), e_kin(*)
       integer(c_int) :: n_rebuilds
     end function tgx_md_run

     integer(c_int) function tgx_md_begin(ctx, n, species, lattice, mass, pos, vel, dt, skin, &
          history) bind(C, name="tgx_md_begin")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, history
       integer(c_int), intent(in) :: species(*)
       real(c_double), intent(in) :: lattice(*), mass(*), pos(*), vel(*)
       real(c_double), value :: dt, skin
     end function tgx_md_begin

     integer(c_int) function tgx_md_steps(ctx, n_steps) bind(C, name="tgx_md_steps")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_steps
     end function tgx_md_steps

     integer(c_int) function tgx_md_read(ctx, pos, vel, forces, e_pot, e_kin, n_steps_done, &
          n_rebuilds) bind(C, name="tgx_md_read")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: pos(*), vel(*), forces(*), e_pot(*), e_kin(*)
       integer(c_int) :: n_steps_done, n_rebuilds
     end function tgx_md_read

     integer(c_int) function tgx_md_end(ctx) bind(C, name="tgx_md_end")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_md_end

     integer(c_int) function tgx_sync(ctx) bind(C, name="tgx_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tgx_sync

     integer(c_int) function tgx_poly3_orthonormalization(alpha_max, S, W) &
          bind(C, name="tgx_poly3_orthonormalization")
       import :: c_int, c_double
       integer(c_int), value :: alpha_max
       real(c_double) :: S(*), W(*)
     end function tgx_poly3_orthonormalization

     integer(c_int) function tgx_stage_times(ctx, ms7) bind(C, name="tgx_stage_times")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: ms7(7)
     end function tgx_stage_times

     integer(c_int) function tgx_set_list_buffer(ctx, buffer) bind(C, name="tgx_set_list_buffer")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: buffer
     end function tgx_set_list_buffer

     integer(c_int) function tgx_set_timing(ctx, on) bind(C, name="tgx_set_timing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function tgx_set_timing

     integer(c_int) function tgx_kernel_times(ctx, ms4) bind(C, name="tgx_kernel_times")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: ms4(4)
     end function tgx_kernel_times

     ! Single-atom-move Metropolis MC (calculate_mc_step "move",
     ! mc/mc_interface.f90:183-447); pos(3,n) updated in place.
     integer(c_int) function tgx_mc_run(ctx, n, species, lattice, pos, n_trials, temp, &
          move_max, skin, seed, e_trace, n_accepted, n_rebuilds) bind(C, name="tgx_mc_run")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, n_trials
       integer(c_int), intent(in) :: species(*)
       real(c_double), intent(in) :: lattice(9)
       real(c_double), intent(inout) :: pos(*)
       real(c_double), value :: temp, move_max, skin
       integer(c_int64_t), value :: seed
       real(c_double) :: e_trace(*)
       integer(c_int) :: n_accepted, n_rebuilds
     end function tgx_mc_run

     ! Delayed-acceptance Metropolis with the on-device surrogate (no
     ! reference counterpart); pos(3,n) updated in place.
     integer(c_int) function tgx_mc_run_da(ctx, n, species, lattice, pos, n_trials, temp, &
          move_max, skin, seed, opt, e_trace, stats) bind(C, name="tgx_mc_run_da")
       import :: c_int, c_ptr, c_double, c_int64_t, tgx_da_options, tgx_mc_stats
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, n_trials
       integer(c_int), intent(in) :: species(*)
       real(c_double), intent(in) :: lattice(9)
       real(c_double), intent(inout) :: pos(*)
       real(c_double), value :: temp, move_max, skin
       integer(c_int64_t), value :: seed
       type(tgx_da_options), intent(in) :: opt
       real(c_double) :: e_trace(*)
       type(tgx_mc_stats) :: stats
     end function tgx_mc_run_da

     ! The per-site surrogate network (no reference counterpart).
     integer(c_int) function tgx_surrogate_init(ctx, dim, seed) bind(C, name="tgx_surrogate_init")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: dim
       integer(c_int64_t), value :: seed
     end function tgx_surrogate_init

     integer(c_int) function tgx_surrogate_set(ctx, n_params, P, mu, sigma) &
          bind(C, name="tgx_surrogate_set")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_params
       real(c_double), intent(in) :: P(*)
       real(c_double), value :: mu, sigma
     end function tgx_surrogate_set

     integer(c_int) function tgx_surrogate_get(ctx, n_params, P, mu, sigma) &
          bind(C, name="tgx_surrogate_get")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_params
       real(c_double) :: P(*), mu, sigma
     end function tgx_surrogate_get

     integer(c_int) function tgx_surrogate_predict(ctx, n, q, e) &
          bind(C, name="tgx_surrogate_predict")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: q(*)
       real(c_double) :: e(*)
     end function tgx_surrogate_predict

     integer(c_int) function tgx_surrogate_gradient(ctx, n, q, y, grad, loss) &
          bind(C, name="tgx_surrogate_gradient")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: q(*), y(*)
       real(c_double) :: grad(*), loss
     end function tgx_surrogate_gradient

     integer(c_int) function tgx_surrogate_fit(ctx, n, q, y, epochs, batch, lr, seed, rmse) &
          bind(C, name="tgx_surrogate_fit")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, epochs, batch
       real(c_double), intent(in) :: q(*), y(*)
       real(c_double), value :: lr
       integer(c_int64_t), value :: seed
       real(c_double) :: rmse
     end function tgx_surrogate_fit

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Message of the last failed call on ctx as a Fortran string.
  function tgx_error_string(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: k, m
    p = tgx_last_error(ctx)
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    m = int(c_strlen(p))
    call c_f_pointer(p, chars, [m])
    allocate (character(len=m) :: msg)
    do k = 1, m
       msg(k:k) = chars(k)
    end do
  end function tgx_error_string

  ! Reference error behaviour (utils/error.f90:35-44): print, then stop.
  subroutine tgx_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= TGX_OK) then
       write (*, '(A,A,A,I0,A,A)') 'ERROR in ', where, ' (code ', rc, '): ', &
            tgx_error_string(ctx)
       stop 1
    end if
  end subroutine tgx_check

end module tgx_mod
