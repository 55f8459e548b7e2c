! Input and output for the Fortran drivers: TurboGAP's .gap model format
! for soap_turbo, distance_2b, angle_3b and core_pot blocks
! (gap/read_gap.f90:45-771, "keyword = value(s)" lines between gap_beg and
! gap_end; any number of soap_turbo blocks, as n_soap_turbo there), the
! sparse-set files (read_alphas_and_descriptors, read_gap.f90:773-875),
! extended-XYZ frames in (xyz/read_xyz.f90: Lattice="..." and
! species/pos[/vel] columns; one or several frames per file) and out
! (xyz/write_xyz.f90:177-442: species, pos, forces, local_energy columns,
! Lattice, energy and virial on the comment line).
module tgx_io
  use iso_c_binding
  use tgx_mod
  implicit none
  private

  type, public :: two_body_block
     integer :: sp1 = 1, sp2 = 1, n_sparse = 0
     real(c_double) :: rcut = 0, buffer = 0.5d0, delta = 1, sigma = 1
     real(c_double), allocatable :: Qs(:), alphas(:), cutoff(:)
  end type two_body_block

  ! one soap_turbo block: hyper-parameters and sparse set; with
  ! compress_soap the P transformation read from file_compress_soap
  type, public :: soap_block
     type(tgx_soap_hypers) :: h
     integer :: dim = 0, n_sparse = 0
     real(c_double) :: delta = 1, zeta = 2
     real(c_double), allocatable :: Qs(:), alphas(:)
     logical :: compress = .false.
     integer :: dim_compressed = 0
     integer(c_int32_t), allocatable :: cP_i(:), cP_j(:)
     real(c_double), allocatable :: cP_el(:)
  end type soap_block

  type, public :: angle_3b_block
     integer :: sp0 = 1, sp1 = 1, sp2 = 1, n_sparse = 0, kernel = 0
     real(c_double) :: rcut = 0, delta = 1, sigma(3) = 1
     real(c_double), allocatable :: Qs(:), alphas(:), cutoff(:)
  end type angle_3b_block

  type, public :: core_pot_block
     integer :: sp1 = 1, sp2 = 1, n = 0
     real(c_double) :: yp1 = 0, ypn = 0
     real(c_double), allocatable :: x(:), V(:)
  end type core_pot_block

  type, public :: gap_model
     integer :: n_species = 0
     character(len=8) :: species(TGX_MAX_SPECIES) = ''
     integer :: n_soap = 0
     type(soap_block) :: sb(TGX_MAX_SOAP_BLOCKS)
     real(c_double) :: e0(TGX_MAX_SPECIES) = 0
     integer :: n2b = 0
     type(two_body_block) :: b2(8)
     integer :: n3b = 0
     type(angle_3b_block) :: b3(8)
     integer :: ncore = 0
     type(core_pot_block) :: core(8)
  end type gap_model

  ! core_pot tables are cut at core_pot_cutoff and tapered by a cosine over
  ! the last core_pot_buffer (read_gap.f90:663-682); the defaults are the
  ! reference's (types.f90:55-56), input-file keywords there.
  real(c_double), public :: core_pot_cutoff = 1.0d10, core_pot_buffer = 0.0d0

  public :: read_gap_file, read_xyz_frame, read_xyz_next, species_index, load_model, &
       write_extxyz

contains

  integer function count_lines(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios
    count_lines = 0
    open (newunit=u, file=fname, status='old', action='read')
    do
       read (u, *, iostat=ios)
       if (ios /= 0) exit
       count_lines = count_lines + 1
    end do
    close (u)
  end function count_lines

  ! alphas file: one alpha per line; desc file: dim*n_sparse values, one per
  ! line, column-major Qs(dim, n_sparse); dim = lines(desc)/n_sparse.
  subroutine read_soap_sparse(fa, fd, m)
    character(len=*), intent(in) :: fa, fd
    type(soap_block), intent(inout) :: m
    integer :: u, k
    m%n_sparse = count_lines(fa)
    m%dim = count_lines(fd)/m%n_sparse
    allocate (m%alphas(m%n_sparse), m%Qs(m%dim*m%n_sparse))
    open (newunit=u, file=fa, status='old', action='read')
    do k = 1, m%n_sparse
       read (u, *) m%alphas(k)
    end do
    close (u)
    open (newunit=u, file=fd, status='old', action='read')
    do k = 1, m%dim*m%n_sparse
       read (u, *) m%Qs(k)
    end do
    close (u)
  end subroutine read_soap_sparse

  ! distance_2b: alphas file "alpha cutoff" per line, desc file one distance
  ! per line (read_gap.f90:829-848).
  subroutine read_2b_sparse(fa, fd, b)
    character(len=*), intent(in) :: fa, fd
    type(two_body_block), intent(inout) :: b
    integer :: u, k
    b%n_sparse = count_lines(fa)
    allocate (b%alphas(b%n_sparse), b%cutoff(b%n_sparse), b%Qs(b%n_sparse))
    open (newunit=u, file=fa, status='old', action='read')
    do k = 1, b%n_sparse
       read (u, *) b%alphas(k), b%cutoff(k)
    end do
    close (u)
    open (newunit=u, file=fd, status='old', action='read')
    do k = 1, b%n_sparse
       read (u, *) b%Qs(k)
    end do
    close (u)
  end subroutine read_2b_sparse

  ! angle_3b: alphas file "alpha cutoff" per line, desc file three values
  ! per sparse point, one per line; Qs kept (n_sparse, 3) column-major
  ! (read_gap.f90:850-872).
  subroutine read_3b_sparse(fa, fd, b)
    character(len=*), intent(in) :: fa, fd
    type(angle_3b_block), intent(inout) :: b
    integer :: u, k, j
    b%n_sparse = count_lines(fa)
    allocate (b%alphas(b%n_sparse), b%cutoff(b%n_sparse), b%Qs(3*b%n_sparse))
    open (newunit=u, file=fa, status='old', action='read')
    do k = 1, b%n_sparse
       read (u, *) b%alphas(k), b%cutoff(k)
    end do
    close (u)
    open (newunit=u, file=fd, status='old', action='read')
    do k = 1, b%n_sparse
       do j = 1, 3
          read (u, *) b%Qs((j - 1)*b%n_sparse + k)
       end do
    end do
    close (u)
  end subroutine read_3b_sparse

  ! core_pot file: "n yp1 ypn", then n lines "x V"; points beyond
  ! core_pot_cutoff dropped, the last core_pot_buffer tapered.
  subroutine read_core_pot(fname, b)
    character(len=*), intent(in) :: fname
    type(core_pot_block), intent(inout) :: b
    integer :: u, k, n, kept
    real(c_double) :: xk, vk
    real(c_double), allocatable :: x(:), V(:)
    open (newunit=u, file=fname, status='old', action='read')
    read (u, *) n, b%yp1, b%ypn
    allocate (x(n), V(n))
    kept = 0
    do k = 1, n
       read (u, *) xk, vk
       if (xk <= core_pot_cutoff) then
          kept = kept + 1
          x(kept) = xk
          if (xk <= core_pot_cutoff - core_pot_buffer) then
             V(kept) = vk
          else
             V(kept) = vk*0.5d0*(cos(acos(-1.d0)/core_pot_buffer*(xk - core_pot_cutoff + &
                  core_pot_buffer)) + 1.d0)
          end if
       end if
    end do
    close (u)
    b%n = kept
    allocate (b%x(kept), b%V(kept))
    b%x = x(1:kept)
    b%V = V(1:kept)
  end subroutine read_core_pot

  integer function species_index(m, sym)
    type(gap_model), intent(in) :: m
    character(len=*), intent(in) :: sym
    integer :: s
    species_index = 0
    do s = 1, m%n_species
       if (trim(m%species(s)) == trim(sym)) species_index = s
    end do
  end function species_index

  ! file_compress_soap (read_gap.f90:489-527): the first line holds n_max of
  ! every species, l_max and the compressed dimension; then either a line
  ! "P_transformation", lines "i j value" and a line "end_transformation",
  ! or (the older format) one kept component index per line, weight 1.
  subroutine read_compress_file(fname, ns, sb)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: ns
    type(soap_block), intent(inout) :: sb
    integer :: u, k, n, ios, idum(TGX_MAX_SPECIES + 1)
    character(len=1024) :: line
    open (newunit=u, file=fname, status='old', action='read')
    read (u, *) idum(1:ns + 1), sb%dim_compressed
    read (u, '(A)') line
    if (trim(adjustl(line)) == 'P_transformation') then
       n = 0
       do
          read (u, '(A)', iostat=ios) line
          if (ios /= 0) then
             write (*, *) 'ERROR: no end_transformation in ', trim(fname)
             stop 1
          end if
          if (trim(adjustl(line)) == 'end_transformation') exit
          n = n + 1
       end do
       rewind (u)
       read (u, *)
       read (u, *)
       allocate (sb%cP_i(n), sb%cP_j(n), sb%cP_el(n))
       do k = 1, n
          read (u, *) sb%cP_i(k), sb%cP_j(k), sb%cP_el(k)
       end do
    else
       n = sb%dim_compressed
       allocate (sb%cP_i(n), sb%cP_j(n), sb%cP_el(n))
       read (line, *) sb%cP_j(1)
       do k = 2, n
          read (u, *) sb%cP_j(k)
       end do
       do k = 1, n
          sb%cP_i(k) = k
          sb%cP_el(k) = 1.0d0
       end do
    end if
    close (u)
  end subroutine read_compress_file

  subroutine read_gap_file(fname, m)
    character(len=*), intent(in) :: fname
    type(gap_model), intent(inout) :: m
    character(len=1024) :: line, key, rest, fa, fd, fcomp
    character(len=64) :: kind, cmode, basis
    character(len=8) :: s1, s2, names(TGX_MAX_SPECIES)
    integer :: u, ios, ns, eq, k, ib
    character(len=8) :: s0
    character(len=16) :: ktype
    logical :: in_soap, in_2b, in_3b, in_core
    real(c_double) :: buffer(TGX_MAX_SPECIES)
    type(two_body_block) :: b
    type(angle_3b_block) :: b3
    type(core_pot_block) :: cp
    in_soap = .false.
    in_2b = .false.
    in_3b = .false.
    in_core = .false.
    ib = 0
    open (newunit=u, file=fname, status='old', action='read')
    do
       read (u, '(A)', iostat=ios) line
       if (ios /= 0) exit
       line = adjustl(line)
       if (line(1:7) == 'gap_beg') then
          read (line(8:), *) kind
          in_soap = trim(kind) == 'soap_turbo'
          in_2b = trim(kind) == 'distance_2b'
          in_3b = trim(kind) == 'angle_3b'
          in_core = trim(kind) == 'core_pot'
          if (in_3b) then
             b3 = angle_3b_block()
             ktype = 'exp'
             fa = ''
             fd = ''
          end if
          if (in_core) then
             cp = core_pot_block()
             fa = ''
          end if
          if (in_soap) then
             if (m%n_soap == TGX_MAX_SOAP_BLOCKS) then
                write (*, *) 'ERROR: more than ', TGX_MAX_SOAP_BLOCKS, ' soap_turbo blocks'
                stop 1
             end if
             m%n_soap = m%n_soap + 1
             ib = m%n_soap
             buffer = 0.5d0
             names = ''
             fa = ''
             fd = ''
             fcomp = 'none'
             cmode = 'none'
          end if
          if (in_2b) then
             b = two_body_block()
             fa = ''
             fd = ''
          end if
          cycle
       end if
       if (line(1:7) == 'gap_end') then
          if (in_soap) then
             associate (sb => m%sb(ib))
               do k = 1, sb%h%n_species
                  if (buffer(k) == 0.d0) then
                     sb%h%rcut_soft(k) = sb%h%rcut_hard(k)
                  else
                     sb%h%rcut_soft(k) = sb%h%rcut_hard(k) - buffer(k)
                  end if
               end do
               call read_soap_sparse(trim(fa), trim(fd), sb)
               if (sb%compress) then
                  ! the recipes of compress_mode live in soap_turbo
                  ! (get_compress_indices), which is not part of this engine
                  if (trim(fcomp) /= 'none') then
                     call read_compress_file(trim(fcomp), sb%h%n_species, sb)
                  else if (trim(cmode) /= 'none') then
                     write (*, *) 'ERROR: compress_mode = ', trim(cmode), &
                          ' is not implemented; give file_compress_soap'
                     stop 1
                  else
                     write (*, *) 'ERROR: compress_soap needs file_compress_soap'
                     stop 1
                  end if
                  if (sb%dim_compressed /= sb%dim) then
                     write (*, *) 'ERROR: compressed dimension ', sb%dim_compressed, &
                          ' but the sparse set has ', sb%dim
                     stop 1
                  end if
               end if
               ! every block on the model's species list (the first block's)
               if (ib == 1) then
                  m%n_species = sb%h%n_species
                  m%species = names
               else if (sb%h%n_species /= m%n_species .or. &
                        any(names(1:m%n_species) /= m%species(1:m%n_species))) then
                  write (*, *) 'ERROR: soap_turbo block ', ib, ' has another species list'
                  stop 1
               end if
             end associate
          else if (in_2b) then
             call read_2b_sparse(trim(fa), trim(fd), b)
             b%sp1 = species_index(m, s1)
             b%sp2 = species_index(m, s2)
             m%n2b = m%n2b + 1
             m%b2(m%n2b) = b
          else if (in_3b) then
             call read_3b_sparse(trim(fa), trim(fd), b3)
             b3%sp0 = species_index(m, s0)
             b3%sp1 = species_index(m, s1)
             b3%sp2 = species_index(m, s2)
             if (trim(ktype) == 'exp') then
                b3%kernel = 0
             else if (trim(ktype) == 'pol') then
                b3%kernel = 1
             else
                write (*, *) 'ERROR: angle_3b kernel_type must be exp or pol'
                stop 1
             end if
             m%n3b = m%n3b + 1
             m%b3(m%n3b) = b3
          else if (in_core) then
             call read_core_pot(trim(fa), cp)
             cp%sp1 = species_index(m, s1)
             cp%sp2 = species_index(m, s2)
             m%ncore = m%ncore + 1
             m%core(m%ncore) = cp
          end if
          in_soap = .false.
          in_2b = .false.
          in_3b = .false.
          in_core = .false.
          cycle
       end if
       eq = index(line, '=')
       if (eq == 0) cycle
       key = adjustl(line(1:eq - 1))
       rest = line(eq + 1:)
       if (in_soap) then
          associate (h => m%sb(ib)%h)
            ns = h%n_species
            select case (trim(key))
            case ('n_species')
               read (rest, *) h%n_species
            case ('species')
               read (rest, *) names(1:ns)
            case ('rcut')
               read (rest, *) h%rcut_hard(1:ns)
            case ('buffer')
               read (rest, *) buffer(1:ns)
            case ('atom_sigma_r')
               read (rest, *) h%atom_sigma_r(1:ns)
            case ('atom_sigma_t')
               read (rest, *) h%atom_sigma_t(1:ns)
            case ('atom_sigma_r_scaling')
               read (rest, *) h%atom_sigma_r_scaling(1:ns)
            case ('atom_sigma_t_scaling')
               read (rest, *) h%atom_sigma_t_scaling(1:ns)
            case ('amplitude_scaling')
               read (rest, *) h%amplitude_scaling(1:ns)
            case ('central_weight')
               read (rest, *) h%central_weight(1:ns)
            case ('global_scaling')
               read (rest, *) h%global_scaling(1:ns)
            case ('nf')
               read (rest, *) h%nf(1:ns)
            case ('n_max')
               read (rest, *) h%alpha_max(1:ns)
            case ('l_max')
               read (rest, *) h%l_max
            case ('radial_enhancement')
               read (rest, *) h%radial_enhancement
            case ('central_species')
               read (rest, *) h%central_species
            case ('zeta')
               read (rest, *) m%sb(ib)%zeta
            case ('delta')
               read (rest, *) m%sb(ib)%delta
            case ('alphas_sparse')
               read (rest, *) fa
            case ('desc_sparse')
               read (rest, *) fd
            case ('compress_soap')
               read (rest, *) m%sb(ib)%compress
            case ('file_compress_soap')
               read (rest, *) fcomp
            case ('compress_mode')
               read (rest, *) cmode
            case ('basis')
               read (rest, *) basis
               if (trim(basis) == 'poly3gauss') then
                  h%basis = 1
               else
                  if (trim(basis) /= 'poly3') write (*, *) &
                       'WARNING: basis ', trim(basis), ' not understood; using poly3'
                  h%basis = 0
               end if
            end select
          end associate
       else if (in_3b) then
          select case (trim(key))
          case ('species1', 'Z1', 'z1')
             read (rest, *) s1
          case ('species2', 'Z2', 'z2')
             read (rest, *) s2
          case ('species_center', 'Z_center', 'z_center')
             read (rest, *) s0
          case ('rcut')
             read (rest, *) b3%rcut
          case ('delta')
             read (rest, *) b3%delta
          case ('sigma')
             read (rest, *) b3%sigma
          case ('kernel_type')
             read (rest, *) ktype
          case ('alphas_sparse')
             read (rest, *) fa
          case ('desc_sparse')
             read (rest, *) fd
          end select
       else if (in_core) then
          select case (trim(key))
          case ('species1', 'Z1', 'z1')
             read (rest, *) s1
          case ('species2', 'Z2', 'z2')
             read (rest, *) s2
          case ('core_pot_file')
             read (rest, *) fa
          end select
       else if (in_2b) then
          select case (trim(key))
          case ('species1', 'Z1', 'z1')
             read (rest, *) s1
          case ('species2', 'Z2', 'z2')
             read (rest, *) s2
          case ('rcut')
             read (rest, *) b%rcut
          case ('delta')
             read (rest, *) b%delta
          case ('sigma')
             read (rest, *) b%sigma
          case ('alphas_sparse')
             read (rest, *) fa
          case ('desc_sparse')
             read (rest, *) fd
          end select
       end if
    end do
    close (u)
  end subroutine read_gap_file

  ! Hands the model to a context: the soap_turbo blocks in file order, e0,
  ! the distance_2b, angle_3b and core_pot blocks.
  subroutine load_model(ctx, m)
    type(c_ptr), intent(in) :: ctx
    type(gap_model), intent(in), target :: m
    type(tgx_soap_hypers) :: h
    integer :: k
    do k = 1, m%n_soap
       associate (sb => m%sb(k))
         h = sb%h
         if (sb%compress) then
            ! the library copies P when the block is set
            h%compress_P_nonzero = int(size(sb%cP_i), c_int32_t)
            h%compress_P_i = c_loc(m%sb(k)%cP_i)
            h%compress_P_j = c_loc(m%sb(k)%cP_j)
            h%compress_P_el = c_loc(m%sb(k)%cP_el)
         end if
         if (k == 1) then
            call tgx_check(ctx, tgx_set_soap_model(ctx, h, int(sb%dim, c_int), &
                 int(sb%n_sparse, c_int), sb%Qs, sb%alphas, sb%delta, sb%zeta), &
                 'tgx_set_soap_model')
         else
            call tgx_check(ctx, tgx_add_soap_model(ctx, h, int(sb%dim, c_int), &
                 int(sb%n_sparse, c_int), sb%Qs, sb%alphas, sb%delta, sb%zeta), &
                 'tgx_add_soap_model')
         end if
       end associate
    end do
    call tgx_check(ctx, tgx_set_e0(ctx, int(m%n_species, c_int), m%e0), 'tgx_set_e0')
    do k = 1, m%n2b
       call tgx_check(ctx, tgx_add_2b_model(ctx, int(m%b2(k)%sp1, c_int), &
            int(m%b2(k)%sp2, c_int), m%b2(k)%rcut, m%b2(k)%buffer, m%b2(k)%delta, &
            m%b2(k)%sigma, int(m%b2(k)%n_sparse, c_int), m%b2(k)%Qs, m%b2(k)%alphas, &
            m%b2(k)%cutoff), 'tgx_add_2b_model')
    end do
    call tgx_check(ctx, tgx_clear_extra_terms(ctx), 'tgx_clear_extra_terms')
    do k = 1, m%n3b
       associate (b => m%b3(k))
         call tgx_check(ctx, tgx_add_angle_3b(ctx, int(b%sp0, c_int), int(b%sp1, c_int), &
              int(b%sp2, c_int), b%rcut, b%delta, b%sigma, int(b%kernel, c_int), &
              int(b%n_sparse, c_int), b%Qs, b%alphas, b%cutoff), 'tgx_add_angle_3b')
       end associate
    end do
    do k = 1, m%ncore
       associate (b => m%core(k))
         call tgx_check(ctx, tgx_add_core_pot(ctx, int(b%sp1, c_int), int(b%sp2, c_int), &
              int(b%n, c_int), b%x, b%V, b%yp1, b%ypn), 'tgx_add_core_pot')
       end associate
    end do
  end subroutine load_model

  ! The next extended-XYZ frame on an open unit: the atom count, then a
  ! comment line with Lattice="ax ay az bx by bz cx cy cz", then lines
  ! "symbol x y z [vx vy vz]".  ok = .false. at the end of the file.
  subroutine read_xyz_next(u, m, n, pos, species, lattice, vel, ok)
    integer, intent(in) :: u
    type(gap_model), intent(in) :: m
    integer, intent(out) :: n
    real(c_double), allocatable, intent(inout) :: pos(:), vel(:)
    integer(c_int), allocatable, intent(inout) :: species(:)
    real(c_double), intent(out) :: lattice(9)
    logical, intent(out) :: ok
    character(len=4096) :: line
    character(len=8) :: sym
    integer :: i, p, q, ios
    real(c_double) :: v(6)
    ok = .false.
    n = 0
    do
       read (u, '(A)', iostat=ios) line
       if (ios /= 0) return
       if (len_trim(line) > 0) exit
    end do
    read (line, *) n
    read (u, '(A)') line
    p = index(line, 'Lattice="')
    if (p == 0) p = index(line, 'lattice="')
    q = index(line(p + 9:), '"')
    read (line(p + 9:p + 7 + q), *) lattice
    if (allocated(pos)) deallocate (pos)
    if (allocated(vel)) deallocate (vel)
    if (allocated(species)) deallocate (species)
    allocate (pos(3*n), vel(3*n), species(n))
    vel = 0
    do i = 1, n
       read (u, '(A)') line
       read (line, *, iostat=ios) sym, v(1:6)
       if (ios /= 0) then
          read (line, *) sym, v(1:3)
          v(4:6) = 0
       end if
       pos(3*i - 2:3*i) = v(1:3)
       vel(3*i - 2:3*i) = v(4:6)
       species(i) = species_index(m, sym)
       if (species(i) == 0) then
          write (*, *) 'ERROR: species ', trim(sym), ' not in the model'
          stop 1
       end if
    end do
    ok = .true.
  end subroutine read_xyz_next

  ! The first frame of a file.
  subroutine read_xyz_frame(fname, m, n, pos, species, lattice, vel)
    character(len=*), intent(in) :: fname
    type(gap_model), intent(in) :: m
    integer, intent(out) :: n
    real(c_double), allocatable, intent(out) :: pos(:), vel(:)
    integer(c_int), allocatable, intent(out) :: species(:)
    real(c_double), intent(out) :: lattice(9)
    integer :: u
    logical :: ok
    open (newunit=u, file=fname, status='old', action='read')
    call read_xyz_next(u, m, n, pos, species, lattice, vel, ok)
    close (u)
    if (.not. ok) then
       write (*, *) 'ERROR: no frame in ', trim(fname)
       stop 1
    end if
  end subroutine read_xyz_frame

  ! One extended-XYZ frame of results, in the layout of the reference's
  ! write_extxyz (xyz/write_xyz.f90:177-442) with the array properties
  ! species, pos, forces and local_energy and the scalar properties
  ! Lattice, energy, virial and i_config (the frame number in predict mode).
  subroutine write_extxyz(u, m, n, species, pos, lattice, e_atom, forces, virial, i_config)
    integer, intent(in) :: u, n, i_config
    type(gap_model), intent(in) :: m
    integer(c_int), intent(in) :: species(n)
    real(c_double), intent(in) :: pos(3*n), lattice(9), e_atom(n), forces(3*n), virial(9)
    character(len=16) :: cell(9), tmp
    integer :: i
    write (u, '(I8)') n
    write (u, '(1X,A)', advance='no') 'Properties=species:S:1:pos:R:3:forces:R:3:local_energy:R:1'
    do i = 1, 9
       write (cell(i), '(F16.10)') lattice(i)
    end do
    write (u, '(1X,11A)', advance='no') 'Lattice="', adjustl(cell(1)), cell(2:9), '"'
    write (tmp, '(F16.6)') sum(e_atom)
    write (u, '(1X,2A)', advance='no') 'energy=', trim(adjustl(tmp))
    do i = 1, 9
       write (cell(i), '(F16.8)') virial(i)
    end do
    write (u, '(1X,11A)', advance='no') 'virial="', adjustl(cell(1)), cell(2:9), '"'
    write (tmp, '(I10)') i_config
    write (u, '(1X,2A)') 'i_config=', trim(adjustl(tmp))
    do i = 1, n
       write (u, '(1X,A8,1X,F16.8,1X,F16.8,1X,F16.8,1X,F16.8,1X,F16.8,1X,F16.8,1X,F16.8)') &
            m%species(species(i)), pos(3*i - 2:3*i), forces(3*i - 2:3*i), e_atom(i)
    end do
    flush (u)
  end subroutine write_extxyz

end module tgx_io
