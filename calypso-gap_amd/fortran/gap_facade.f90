! Fortran host facade of the gapx evaluator: the external subroutines of
! libgap with the reference's names, argument lists and units, backed by the
! HIP kernels through the C ABI (module gapx_iface).
!
!   FGAP_CALC   gappy/libgap/gap_calc.f90:1-46    energy, forces, stress
!   FGAP_READ   gappy/libgap/gap_calc.f90:303-314 model reader (fixed caps)
!   FGET_BOND   gappy/libgap/get_bond.f90:4-12    minimum interatomic distance
!   CAR2ACSF    gappy/libgap/wacsf.f90:1-12       descriptors and their derivatives
!   WRITE_ARRAY_2DIM  gappy/libgap/wacsf.f90:798-810  a matrix as text
! and, with no reference counterpart:
!   FGAP_CALC_STATUS   FGAP_CALC returning a status instead of stopping
!   FGAP_CALC_BATCH    many structures in one call
!   FGAP_READ_DIMS / FGAP_READ2   the model reader without the fixed caps
!   FGAP_DESCRIPTORS   the wACSF matrix (CAR2ACSF's xx) of a structure
!
! Like the reference, FGAP_CALC takes theta/MM/coeff as arguments and reads
! the species weights and symmetry-function table from ./gap_parameters; the
! header is re-read only when the file's size or mtime changes, and the
! device model is rebuilt only when the header or the arrays change.

subroutine fgap_calc_status(na, species, lat, pos, ene, force, stress, variance, nsparsex, des_len, theta, mm, &
                            coeff, rcut, lgrad, mode, status)
  use gapx_iface
  implicit none
  integer, intent(in) :: na, nsparsex, des_len, mode
  integer, intent(in) :: species(na)
  double precision, intent(in) :: lat(3, 3), pos(na, 3)
  double precision, intent(in) :: theta(des_len), mm(nsparsex, des_len), coeff(nsparsex), rcut
  logical, intent(in) :: lgrad
  double precision, intent(out) :: ene, force(na, 3), stress(6), variance
  integer, intent(out) :: status
  integer(c_int) :: flags
  ene = 0.d0
  force = 0.d0
  stress = 0.d0
  variance = 0.d0
  call ensure_model(nsparsex, des_len, theta, mm, coeff, status)
  if (status /= GAPX_OK) return
  flags = ior(GAPX_LAYOUT_F, GAPX_REF_NEIGHBOR_CAP)
  if (lgrad) flags = ior(flags, GAPX_GRAD)
  if (mode == 1) flags = ior(flags, GAPX_MIXED)
  if (mode == 2) flags = ior(flags, ior(GAPX_MIXED, GAPX_BF16))
  if (mode < 0) flags = ior(flags, precision_flags())
  status = gapx_calc(model, int(na, c_int), species, lat, pos, rcut, flags, ene, force, stress, variance)
end subroutine fgap_calc_status

subroutine fgap_calc(na, species, lat, pos, ene, force, stress, variance, nsparsex, des_len, theta, mm, qmm, coeff, &
                     rcut, lgrad)
  use gapx_iface
  implicit none
  integer, intent(in) :: na
  integer, intent(in), dimension(na) :: species
  double precision, intent(in), dimension(3, 3) :: lat
  double precision, intent(in), dimension(na, 3) :: pos
  integer, intent(in) :: des_len
  integer, intent(in) :: nsparsex
  double precision, intent(in), dimension(des_len) :: theta
  double precision, intent(in), dimension(nsparsex, des_len) :: mm
  double precision, intent(in), dimension(nsparsex, nsparsex) :: qmm
  double precision, intent(in), dimension(nsparsex) :: coeff
  double precision, intent(in) :: rcut
  logical, intent(in) :: lgrad
  double precision, intent(out) :: ene
  double precision, intent(out) :: variance
  double precision, intent(out), dimension(na, 3) :: force
  double precision, intent(out), dimension(6) :: stress
  integer :: status
  ! qmm is accepted and ignored, as in the reference (gap_calc.f90:205-210)
  if (.false.) ene = qmm(1, 1)
  call fgap_calc_status(na, species, lat, pos, ene, force, stress, variance, nsparsex, des_len, theta, mm, coeff, &
                        rcut, lgrad, -1, status)
  if (status == GAPX_OK) return
  if (status == 2) then
    print *, "gap_parameters file does not exist!"
  else
    print *, "gapx: ", last_error()
  end if
  stop
end subroutine fgap_calc

subroutine fgap_calc_batch(nstruct, natoms, ntot, species, lat, pos, ene, force, stress, nsparsex, des_len, theta, &
                           mm, coeff, rcut, lgrad, mode, status)
  use gapx_iface
  implicit none
  integer, intent(in) :: nstruct, ntot, nsparsex, des_len, mode
  integer, intent(in) :: natoms(nstruct), species(ntot)
  double precision, intent(in) :: lat(3, 3, nstruct), pos(3, ntot)
  double precision, intent(in) :: theta(des_len), mm(nsparsex, des_len), coeff(nsparsex), rcut
  logical, intent(in) :: lgrad
  double precision, intent(out) :: ene(nstruct), force(3, ntot), stress(6, nstruct)
  integer, intent(out) :: status
  integer(c_int) :: flags
  double precision :: latc(3, 3, nstruct)
  integer :: s
  ene = 0.d0
  force = 0.d0
  stress = 0.d0
  if (sum(natoms) /= ntot) then
    status = 1
    return
  end if
  call ensure_model(nsparsex, des_len, theta, mm, coeff, status)
  if (status /= GAPX_OK) return
  ! lat(:,:,s) is Fortran LAT(3,3) (row k = lattice vector k); the batch entry
  ! takes lattice vectors as consecutive triples, i.e. the transpose in memory
  do s = 1, nstruct
    latc(:, :, s) = transpose(lat(:, :, s))
  end do
  flags = GAPX_REF_NEIGHBOR_CAP
  if (lgrad) flags = ior(flags, GAPX_GRAD)
  if (mode == 1) flags = ior(flags, GAPX_MIXED)
  if (mode == 2) flags = ior(flags, ior(GAPX_MIXED, GAPX_BF16))
  if (mode < 0) flags = ior(flags, precision_flags())
  status = gapx_calc_batch(model, int(nstruct, c_int), natoms, species, latc, pos, rcut, flags, ene, force, stress)
end subroutine fgap_calc_batch

subroutine fgap_descriptors(na, species, lat, pos, nsparsex, des_len, theta, mm, coeff, rcut, xx, status)
  use gapx_iface
  implicit none
  integer, intent(in) :: na, nsparsex, des_len
  integer, intent(in) :: species(na)
  double precision, intent(in) :: lat(3, 3), pos(na, 3)
  double precision, intent(in) :: theta(des_len), mm(nsparsex, des_len), coeff(nsparsex), rcut
  double precision, intent(out) :: xx(des_len, na)
  integer, intent(out) :: status
  xx = 0.d0
  call ensure_model(nsparsex, des_len, theta, mm, coeff, status)
  if (status /= GAPX_OK) return
  ! x is row = atom, column = feature in C order, which is xx(des_len, na) here
  status = gapx_descriptors(model, int(na, c_int), species, lat, pos, rcut, ior(GAPX_LAYOUT_F, 0_c_int), xx)
end subroutine fgap_descriptors

subroutine fgap_read(nsparsex, des_len, theta, mm, invcmm, coeff)
  implicit none
  integer, parameter :: nsf_max = 100
  integer, parameter :: nsparsex_max = 4000
  integer, intent(out) :: nsparsex
  integer, intent(out) :: des_len
  double precision, intent(out) :: theta(nsf_max)
  double precision, intent(out) :: mm(nsparsex_max, nsf_max)
  double precision, intent(out) :: invcmm(nsparsex_max, nsparsex_max)
  double precision, intent(out) :: coeff(nsparsex_max)
  integer :: status
  theta = 0.d0
  mm = 0.d0
  coeff = 0.d0
  invcmm = 0.d0
  call fgap_read_dims(nsparsex, des_len, status)
  if (status == 2) then
    print *, "gap_parameters file does not exist!"
    stop
  end if
  if (status /= 0) then
    print *, "gap_parameters: cannot read the model"
    stop
  end if
  ! the reference's fixed caps and messages (gap_calc.f90:341-350)
  if (nsparsex > nsparsex_max) then
    print *, 'The siez of sparse set large than nsparseX_max=', nsparsex_max, &
             "and you should modify it in gap_init.f90"
    stop
  end if
  if (des_len > nsf_max) then
    print *, 'The length of descriptors large than nsf_max=', nsf_max, &
             "and you should modify it in gap_init.f90"
    stop
  end if
  call read_arrays(nsparsex, des_len, theta, mm, nsparsex_max, coeff, status)
  if (status /= 0) then
    print *, "gap_parameters: cannot read the model"
    stop
  end if
end subroutine fgap_read

! Sizes of the model in ./gap_parameters (no caps).  status: 0, 2 missing, 3 unreadable.
subroutine fgap_read_dims(nsparsex, des_len, status)
  implicit none
  integer, intent(out) :: nsparsex, des_len, status
  integer :: u, ios, k, nspecies, nsf
  logical :: alive
  nsparsex = 0
  des_len = 0
  inquire(file="gap_parameters", exist=alive)
  if (.not. alive) then
    status = 2
    return
  end if
  status = 3
  open(newunit=u, file="gap_parameters", status="old", action="read", iostat=ios)
  if (ios /= 0) return
  read(u, *, iostat=ios) nspecies
  if (ios /= 0) goto 90
  do k = 1, nspecies
    read(u, *, iostat=ios)
    if (ios /= 0) goto 90
  end do
  read(u, *, iostat=ios) nsf
  if (ios /= 0) goto 90
  do k = 1, nsf
    read(u, *, iostat=ios)
    if (ios /= 0) goto 90
  end do
  read(u, *, iostat=ios) nsparsex, des_len
  if (ios /= 0) goto 90
  status = 0
90 close(u)
end subroutine fgap_read_dims

! Model arrays of ./gap_parameters in caller-sized arrays (no caps).
subroutine fgap_read2(nsparsex, des_len, theta, mm, coeff, status)
  implicit none
  integer, intent(in) :: nsparsex, des_len
  double precision, intent(out) :: theta(des_len), mm(nsparsex, des_len), coeff(nsparsex)
  integer, intent(out) :: status
  call read_arrays(nsparsex, des_len, theta, mm, nsparsex, coeff, status)
end subroutine fgap_read2

! Layout of gap_calc.f90:328-362: species table, SF table, sizes, three
! skipped records, theta, one record per sparse point, coefficients.
subroutine read_arrays(nsparsex, des_len, theta, mm, ldm, coeff, status)
  implicit none
  integer, intent(in) :: nsparsex, des_len, ldm
  double precision, intent(out) :: theta(*), mm(ldm, *), coeff(*)
  integer, intent(out) :: status
  integer :: u, ios, k, nspecies, nsf, m2, d2
  status = 3
  open(newunit=u, file="gap_parameters", status="old", action="read", iostat=ios)
  if (ios /= 0) then
    status = 2
    return
  end if
  read(u, *, iostat=ios) nspecies
  if (ios /= 0) goto 90
  do k = 1, nspecies
    read(u, *, iostat=ios)
    if (ios /= 0) goto 90
  end do
  read(u, *, iostat=ios) nsf
  if (ios /= 0) goto 90
  do k = 1, nsf
    read(u, *, iostat=ios)
    if (ios /= 0) goto 90
  end do
  read(u, *, iostat=ios) m2, d2
  if (ios /= 0 .or. m2 /= nsparsex .or. d2 /= des_len) goto 90
  read(u, *, iostat=ios)
  read(u, *, iostat=ios)
  read(u, *, iostat=ios)
  read(u, *, iostat=ios) theta(1:des_len)
  if (ios /= 0) goto 90
  do k = 1, nsparsex
    read(u, *, iostat=ios) mm(k, 1:des_len)
    if (ios /= 0) goto 90
  end do
  read(u, *, iostat=ios) coeff(1:nsparsex)
  if (ios /= 0) goto 90
  status = 0
90 close(u)
end subroutine read_arrays

subroutine fget_bond(na, lat, elements, pos, rcut, min_bond)
  use gapx_iface
  implicit none
  integer, intent(in) :: na
  double precision, intent(in), dimension(3, 3) :: lat
  integer, intent(in), dimension(na) :: elements
  double precision, intent(in), dimension(na, 3) :: pos
  double precision, intent(in) :: rcut
  double precision, intent(out) :: min_bond
  integer(c_int) :: st
  ! species do not enter the distance (get_bond.f90:47-78)
  if (.false.) min_bond = dble(elements(1))
  st = gapx_min_distance(int(na, c_int), lat, pos, rcut, GAPX_LAYOUT_F, min_bond)
  if (st /= GAPX_OK) then
    print *, "gapx: ", last_error()
    stop
  end if
end subroutine fget_bond

! CAR2ACSF (wacsf.f90:1-12): descriptors xx(nf, na) and, with lgrad, their
! dense derivatives dxdy(nf, na, na, 3) and strs(3, 3, nf, na) from the
! caller's neighbour table, with the SF table of ./gap_parameters; stops like
! the reference when the file is missing (wacsf.f90:41-44) and on any other
! error with the message of the failed call.
subroutine car2acsf(na, max_neighbor, nf, pos, neighbor, neighbor_count, xx, dxdy, strs, lgrad)
  use gapx_iface
  implicit none
  integer, intent(in) :: na, max_neighbor, nf
  double precision, intent(in), dimension(na, 3) :: pos
  double precision, intent(in), dimension(na, max_neighbor, 6) :: neighbor
  integer, intent(in), dimension(na) :: neighbor_count
  double precision, intent(out), dimension(nf, na) :: xx
  double precision, intent(out), dimension(nf, na, na, 3) :: dxdy
  double precision, intent(out), dimension(3, 3, nf, na) :: strs
  logical, intent(in) :: lgrad
  integer :: status
  integer(c_int) :: lg
  call ensure_header_model(status)
  if (status == 2) then
    print *, "gap_parameters file does not exist!"
    stop
  end if
  lg = 0
  if (lgrad) lg = 1
  if (status == GAPX_OK) status = gapx_car2acsf(hmodel, int(na, c_int), int(max_neighbor, c_int), int(nf, c_int), &
                                                pos, neighbor, neighbor_count, xx, dxdy, strs, lg)
  if (status /= GAPX_OK) then
    print *, "gapx: ", last_error()
    stop
  end if
end subroutine car2acsf

! write_array_2dim (wacsf.f90:798-810): a(n, m) as text, one row per line,
! each value in F20.10, to the file named by name (leading and trailing
! blanks removed).
subroutine write_array_2dim(n, m, a, name)
  implicit none
  integer, intent(in) :: n, m
  double precision, intent(in), dimension(n, m) :: a
  character(len=*), intent(in) :: name
  integer :: u, r, c
  open(newunit=u, file=trim(adjustl(name)), status="replace", action="write")
  do r = 1, n
    do c = 1, m
      write(u, "(F20.10)", advance="no") a(r, c)
    end do
    write(u, *)
  end do
  close(u)
end subroutine write_array_2dim

! Message of the last failed gapx call (for callers of the *_status entries).
subroutine fgap_last_error(msg)
  use gapx_iface
  implicit none
  character(len=512), intent(out) :: msg
  msg = last_error()
end subroutine fgap_last_error
