! Input readers for the Fortran drivers: the subset of TurboGAP's .gap model
! format used by soap_turbo and distance_2b blocks (gap/read_gap.f90:45-771,
! "keyword = value(s)" lines between gap_beg and gap_end), the sparse-set
! files (read_alphas_and_descriptors, read_gap.f90:773-875) and extended-XYZ
! frames (xyz/read_xyz.f90: Lattice="..." and species/pos[/vel] columns).
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

  type, public :: gap_model
     type(tgx_soap_hypers) :: h
     character(len=8) :: species(TGX_MAX_SPECIES) = ''
     integer :: dim = 0, n_sparse = 0
     real(c_double) :: delta = 1, zeta = 2
     real(c_double), allocatable :: Qs(:), alphas(:)
     real(c_double) :: e0(TGX_MAX_SPECIES) = 0
     integer :: n2b = 0
     type(two_body_block) :: b2(8)
  end type gap_model

  public :: read_gap_file, read_xyz_frame, species_index

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
    type(gap_model), intent(inout) :: m
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

  integer function species_index(m, sym)
    type(gap_model), intent(in) :: m
    character(len=*), intent(in) :: sym
    integer :: s
    species_index = 0
    do s = 1, m%h%n_species
       if (trim(m%species(s)) == trim(sym)) species_index = s
    end do
  end function species_index

  subroutine read_gap_file(fname, m)
    character(len=*), intent(in) :: fname
    type(gap_model), intent(inout) :: m
    character(len=1024) :: line, key, rest, fa, fd
    character(len=64) :: kind
    character(len=8) :: s1, s2
    integer :: u, ios, ns, eq, k
    logical :: in_soap, in_2b
    real(c_double) :: buffer(TGX_MAX_SPECIES)
    type(two_body_block) :: b
    in_soap = .false.
    in_2b = .false.
    buffer = 0.5d0
    open (newunit=u, file=fname, status='old', action='read')
    do
       read (u, '(A)', iostat=ios) line
       if (ios /= 0) exit
       line = adjustl(line)
       if (line(1:7) == 'gap_beg') then
          read (line(8:), *) kind
          in_soap = trim(kind) == 'soap_turbo'
          in_2b = trim(kind) == 'distance_2b'
          if (in_soap) then
             fa = ''
             fd = ''
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
             do k = 1, m%h%n_species
                if (buffer(k) == 0.d0) then
                   m%h%rcut_soft(k) = m%h%rcut_hard(k)
                else
                   m%h%rcut_soft(k) = m%h%rcut_hard(k) - buffer(k)
                end if
             end do
             call read_soap_sparse(trim(fa), trim(fd), m)
          else if (in_2b) then
             call read_2b_sparse(trim(fa), trim(fd), b)
             b%sp1 = species_index(m, s1)
             b%sp2 = species_index(m, s2)
             m%n2b = m%n2b + 1
             m%b2(m%n2b) = b
          end if
          in_soap = .false.
          in_2b = .false.
          cycle
       end if
       eq = index(line, '=')
       if (eq == 0) cycle
       key = adjustl(line(1:eq - 1))
       rest = line(eq + 1:)
       ns = m%h%n_species
       if (in_soap) then
          select case (trim(key))
          case ('n_species')
             read (rest, *) m%h%n_species
          case ('species')
             read (rest, *) m%species(1:ns)
          case ('rcut')
             read (rest, *) m%h%rcut_hard(1:ns)
          case ('buffer')
             read (rest, *) buffer(1:ns)
          case ('atom_sigma_r')
             read (rest, *) m%h%atom_sigma_r(1:ns)
          case ('atom_sigma_t')
             read (rest, *) m%h%atom_sigma_t(1:ns)
          case ('atom_sigma_r_scaling')
             read (rest, *) m%h%atom_sigma_r_scaling(1:ns)
          case ('atom_sigma_t_scaling')
             read (rest, *) m%h%atom_sigma_t_scaling(1:ns)
          case ('amplitude_scaling')
             read (rest, *) m%h%amplitude_scaling(1:ns)
          case ('central_weight')
             read (rest, *) m%h%central_weight(1:ns)
          case ('global_scaling')
             read (rest, *) m%h%global_scaling(1:ns)
          case ('nf')
             read (rest, *) m%h%nf(1:ns)
          case ('n_max')
             read (rest, *) m%h%alpha_max(1:ns)
          case ('l_max')
             read (rest, *) m%h%l_max
          case ('radial_enhancement')
             read (rest, *) m%h%radial_enhancement
          case ('central_species')
             read (rest, *) m%h%central_species
          case ('zeta')
             read (rest, *) m%zeta
          case ('delta')
             read (rest, *) m%delta
          case ('alphas_sparse')
             read (rest, *) fa
          case ('desc_sparse')
             read (rest, *) fd
          end select
       else if (in_2b) then
          select case (trim(key))
          case ('species1')
             read (rest, *) s1
          case ('species2')
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

  ! One extended-XYZ frame: Lattice="ax ay az bx by bz cx cy cz", then lines
  ! "symbol x y z [vx vy vz]".
  subroutine read_xyz_frame(fname, m, n, pos, species, lattice, vel)
    character(len=*), intent(in) :: fname
    type(gap_model), intent(in) :: m
    integer, intent(out) :: n
    real(c_double), allocatable, intent(out) :: pos(:), vel(:)
    integer(c_int), allocatable, intent(out) :: species(:)
    real(c_double), intent(out) :: lattice(9)
    character(len=4096) :: line
    character(len=8) :: sym
    integer :: u, i, p, q, ios
    real(c_double) :: v(6)
    open (newunit=u, file=fname, status='old', action='read')
    read (u, *) n
    read (u, '(A)') line
    p = index(line, 'Lattice="')
    if (p == 0) p = index(line, 'lattice="')
    q = index(line(p + 9:), '"')
    read (line(p + 9:p + 7 + q), *) lattice
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
    close (u)
  end subroutine read_xyz_frame

end module tgx_io
