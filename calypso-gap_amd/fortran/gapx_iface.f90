! ISO_C_BINDING interfaces to the gapx C ABI (include/gapx.h) and the state
! the Fortran facade keeps between calls: the model header read from
! ./gap_parameters and the device-resident model built from it.
module gapx_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GAPX_OK = 0
  integer(c_int), parameter :: GAPX_GRAD = 1
  integer(c_int), parameter :: GAPX_MIXED = 2
  integer(c_int), parameter :: GAPX_LAYOUT_F = 4
  integer(c_int), parameter :: GAPX_REF_NEIGHBOR_CAP = 8
  integer(c_int), parameter :: GAPX_BF16 = 16

  interface
    integer(c_int) function gapx_model_create(nspecies, z, w, nsf, sf_type, sf_alpha, sf_cut, sf_zeta, &
                                              des_len, nsparse, theta, mm, mm_layout, ld_mm, coeff, out) &
                                              bind(C, name="gapx_model_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nspecies, nsf, des_len, nsparse, mm_layout, ld_mm
      integer(c_int), intent(in) :: z(*), sf_type(*)
      real(c_double), intent(in) :: w(*), sf_alpha(*), sf_cut(*), sf_zeta(*), theta(*), mm(*), coeff(*)
      type(c_ptr), intent(out) :: out
    end function

    subroutine gapx_model_free(m) bind(C, name="gapx_model_free")
      import :: c_ptr
      type(c_ptr), value :: m
    end subroutine

    integer(c_int) function gapx_calc(m, na, species, lat, pos, rcut, flags, ene, force, stress6, variance) &
                                      bind(C, name="gapx_calc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: na, flags
      integer(c_int), intent(in) :: species(*)
      real(c_double), intent(in) :: lat(*), pos(*)
      real(c_double), value :: rcut
      real(c_double), intent(out) :: ene, force(*), stress6(*), variance
    end function

    integer(c_int) function gapx_calc_batch(m, nstruct, natoms, species, lat, pos, rcut, flags, ene, force, &
                                            stress6) bind(C, name="gapx_calc_batch")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nstruct, flags
      integer(c_int), intent(in) :: natoms(*), species(*)
      real(c_double), intent(in) :: lat(*), pos(*)
      real(c_double), value :: rcut
      real(c_double), intent(out) :: ene(*), force(*), stress6(*)
    end function

    integer(c_int) function gapx_descriptors(m, na, species, lat, pos, rcut, flags, x) &
                                             bind(C, name="gapx_descriptors")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: na, flags
      integer(c_int), intent(in) :: species(*)
      real(c_double), intent(in) :: lat(*), pos(*)
      real(c_double), value :: rcut
      real(c_double), intent(out) :: x(*)
    end function

    integer(c_int) function gapx_min_distance(na, lat, pos, rcut, flags, dmin) bind(C, name="gapx_min_distance")
      import :: c_int, c_double
      integer(c_int), value :: na, flags
      real(c_double), intent(in) :: lat(*), pos(*)
      real(c_double), value :: rcut
      real(c_double), intent(out) :: dmin
    end function

    integer(c_int) function gapx_car2acsf(m, na, max_neighbor, nf, pos, neighbor, neighbor_count, xx, dxdy, &
                                          strs, lgrad) bind(C, name="gapx_car2acsf")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: na, max_neighbor, nf, lgrad
      real(c_double), intent(in) :: pos(*), neighbor(*)
      integer(c_int), intent(in) :: neighbor_count(*)
      real(c_double), intent(out) :: xx(*), dxdy(*), strs(*)
    end function

    integer(c_int) function gapx_file_stamp(path, fsize, mtime) bind(C, name="gapx_file_stamp")
      import :: c_int, c_char, c_long_long
      character(kind=c_char), intent(in) :: path(*)
      integer(c_long_long), intent(out) :: fsize, mtime
    end function

    type(c_ptr) function gapx_last_error() bind(C, name="gapx_last_error")
      import :: c_ptr
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

  ! ---- state kept between calls ---------------------------------------------
  character(len=*), parameter :: model_file = "gap_parameters"
  type(c_ptr), save :: model = c_null_ptr
  integer(c_long_long), save :: stamp_size = -1, stamp_mtime = -1
  integer(c_int64_t), save :: model_hash = 0
  integer, save :: model_m = -1, model_d = -1
  integer, save :: hdr_nspecies = 0, hdr_nsf = 0
  ! bumped on every successful header read; each device model records the
  ! generation it was built from
  integer, save :: hdr_gen = 0, model_gen = -1
  ! header-only model for CAR2ACSF (species table and SF lines, one zero
  ! sparse point), rebuilt when the header changes
  type(c_ptr), save :: hmodel = c_null_ptr
  integer, save :: hmodel_gen = -1
  integer(c_int), allocatable, save :: hdr_z(:), hdr_type(:)
  real(c_double), allocatable, save :: hdr_w(:), hdr_alpha(:), hdr_cut(:), hdr_zeta(:)

contains

  ! Message of the calling thread's last gapx failure.
  function last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n, k
    p = gapx_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [max(n, 1)])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = cs(k)
    end do
  end function

  ! Precision from the environment: GAPX_PRECISION = fp64 (default) | mixed | mixed-bf16.
  integer(c_int) function precision_flags()
    character(len=32) :: v
    integer :: l, st
    precision_flags = 0
    call get_environment_variable("GAPX_PRECISION", v, l, st)
    if (st /= 0) return
    if (v(1:l) == "mixed") precision_flags = GAPX_MIXED
    if (v(1:l) == "mixed-bf16") precision_flags = ior(GAPX_MIXED, GAPX_BF16)
  end function

  ! 64-bit FNV-1a over the bit patterns of an array (model cache key).
  subroutine hash_doubles(h, a, n)
    integer(c_int64_t), intent(inout) :: h
    integer, intent(in) :: n
    real(c_double), intent(in) :: a(n)
    integer(c_int64_t), parameter :: prime = 1099511628211_c_int64_t
    integer(c_int64_t) :: bits
    integer :: k
    do k = 1, n
      bits = transfer(a(k), bits)
      h = ieor(h, bits)
      h = h * prime
    end do
  end subroutine

  ! Re-read the header of ./gap_parameters when its size or mtime changed
  ! (gap_calc.f90:75-83 species weights, wacsf.f90:40-56 SF table), with
  ! list-directed reads; an optional 4th column on an SF line is zeta.
  ! status: 0 ok, 1 file missing, 2 unreadable.
  subroutine load_header(status, changed)
    integer, intent(out) :: status
    logical, intent(out) :: changed
    integer(c_long_long) :: fsize, mtime
    integer :: u, ios, k, t
    character(len=512) :: line
    real(c_double) :: a, c, z
    changed = .false.
    status = 0
    if (gapx_file_stamp(model_file // c_null_char, fsize, mtime) /= GAPX_OK) then
      status = 1
      return
    end if
    if (fsize == stamp_size .and. mtime == stamp_mtime .and. allocated(hdr_z)) return
    open(newunit=u, file=model_file, status="old", action="read", iostat=ios)
    if (ios /= 0) then
      status = 2
      return
    end if
    call free_header()
    read(u, *, iostat=ios) hdr_nspecies
    if (ios /= 0 .or. hdr_nspecies <= 0) goto 90
    allocate(hdr_z(hdr_nspecies), hdr_w(hdr_nspecies))
    do k = 1, hdr_nspecies
      read(u, *, iostat=ios) hdr_z(k), hdr_w(k)
      if (ios /= 0) goto 90
    end do
    read(u, *, iostat=ios) hdr_nsf
    if (ios /= 0 .or. hdr_nsf <= 0) goto 90
    allocate(hdr_type(hdr_nsf), hdr_alpha(hdr_nsf), hdr_cut(hdr_nsf), hdr_zeta(hdr_nsf))
    do k = 1, hdr_nsf
      read(u, "(a)", iostat=ios) line
      if (ios /= 0) goto 90
      read(line, *, iostat=ios) t, a, c, z
      if (ios /= 0) then
        read(line, *, iostat=ios) t, a, c
        if (ios /= 0) goto 90
        z = 1.0_c_double
      end if
      hdr_type(k) = t
      hdr_alpha(k) = a
      hdr_cut(k) = c
      hdr_zeta(k) = z
    end do
    close(u)
    stamp_size = fsize
    stamp_mtime = mtime
    changed = .true.
    hdr_gen = hdr_gen + 1
    return
90  close(u)
    status = 2
    ! a header read that failed part-way leaves no header and no stamps, so
    ! the next call reads the file again from the start
    call free_header()
    stamp_size = -1
    stamp_mtime = -1
  end subroutine

  ! Every header array that is allocated, each under its own check (a failed
  ! read can leave some allocated and some not).
  subroutine free_header()
    if (allocated(hdr_z)) deallocate(hdr_z)
    if (allocated(hdr_w)) deallocate(hdr_w)
    if (allocated(hdr_type)) deallocate(hdr_type)
    if (allocated(hdr_alpha)) deallocate(hdr_alpha)
    if (allocated(hdr_cut)) deallocate(hdr_cut)
    if (allocated(hdr_zeta)) deallocate(hdr_zeta)
  end subroutine

  ! Device model for (header, theta, MM, coeff); rebuilt when the header file
  ! or any of the arrays changed.  status = gapx status code.
  subroutine ensure_model(nsparse, des_len, theta, mm, coeff, status)
    integer, intent(in) :: nsparse, des_len
    real(c_double), intent(in) :: theta(des_len), mm(nsparse, des_len), coeff(nsparse)
    integer, intent(out) :: status
    integer :: hstat
    logical :: changed
    integer(c_int64_t) :: h
    call load_header(hstat, changed)
    if (hstat == 1) then
      status = 2
      return
    else if (hstat /= 0) then
      status = 3
      return
    end if
    h = -3750763034362895579_c_int64_t
    call hash_doubles(h, theta, des_len)
    call hash_doubles(h, mm, nsparse * des_len)
    call hash_doubles(h, coeff, nsparse)
    if (model_gen == hdr_gen .and. c_associated(model) .and. h == model_hash .and. nsparse == model_m &
        .and. des_len == model_d) then
      status = GAPX_OK
      return
    end if
    if (c_associated(model)) call gapx_model_free(model)
    model = c_null_ptr
    status = gapx_model_create(int(hdr_nspecies, c_int), hdr_z, hdr_w, int(hdr_nsf, c_int), hdr_type, hdr_alpha, &
                               hdr_cut, hdr_zeta, int(des_len, c_int), int(nsparse, c_int), theta, mm, 1_c_int, &
                               int(nsparse, c_int), coeff, model)
    if (status == GAPX_OK) then
      model_hash = h
      model_m = nsparse
      model_d = des_len
      model_gen = hdr_gen
    end if
  end subroutine

  ! Header-only device model of ./gap_parameters for CAR2ACSF, which reads
  ! only the species table and the SF lines (wacsf.f90:40-56).  status: gapx
  ! status code (2: file missing, 3: unreadable header).
  subroutine ensure_header_model(status)
    integer, intent(out) :: status
    integer :: hstat
    logical :: changed
    real(c_double), allocatable :: theta(:), mm(:), coeff(:)
    call load_header(hstat, changed)
    if (hstat == 1) then
      status = 2
      return
    else if (hstat /= 0) then
      status = 3
      return
    end if
    if (hmodel_gen == hdr_gen .and. c_associated(hmodel)) then
      status = GAPX_OK
      return
    end if
    if (c_associated(hmodel)) call gapx_model_free(hmodel)
    hmodel = c_null_ptr
    allocate(theta(2 * hdr_nsf), mm(2 * hdr_nsf), coeff(1))
    theta = 1.0_c_double
    mm = 0.0_c_double
    coeff = 0.0_c_double
    status = gapx_model_create(int(hdr_nspecies, c_int), hdr_z, hdr_w, int(hdr_nsf, c_int), hdr_type, hdr_alpha, &
                               hdr_cut, hdr_zeta, int(2 * hdr_nsf, c_int), 1_c_int, theta, mm, 0_c_int, 0_c_int, &
                               coeff, hmodel)
    if (status == GAPX_OK) hmodel_gen = hdr_gen
  end subroutine

end module gapx_iface
