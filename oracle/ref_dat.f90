! Driver of the reference's text writers (modGlobalMetod.f90:186-306), linked
! against the flang build of the reference modules (oracle/build_ref.sh).
! TEST INFRASTRUCTURE ONLY.
!
! Usage: ref_dat in.bin prefix
!   in.bin: int32 n; real64 v(n); int32 iv(n)
! Writes prefix_vc.dat   guardarVectorC of cmplx(v(2i-1), v(2i)), i = 1..n/2
!        prefix_mr.dat   guardarMatrizR of reshape(v, [n/3, 3]) (n/3 rows)
!        prefix_vr.dat   guardarVectorR of v
!        prefix_vi.dat   guardarVectorI of iv
program ref_dat
  use modGlobalParam
  use modGlobalMetod
  implicit none
  character(len=512) :: fin, pre
  integer :: u, n, i
  real (kind = 8), allocatable :: v(:), m(:,:)
  integer (kind = 4), allocatable :: iv(:)
  complex (kind = 8), allocatable :: c(:)
  call get_command_argument(1, fin)
  call get_command_argument(2, pre)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n
  allocate(v(n), iv(n))
  read(u) v, iv
  close(u)
  allocate(c(n / 2), m(n / 3, 3))
  do i = 1, n / 2
    c(i) = cmplx(v(2 * i - 1), v(2 * i), kind = 8)
  end do
  m = reshape(v(1:3 * (n / 3)), [n / 3, 3])
  call guardarVectorC(c, trim(pre) // '_vc.dat')
  call guardarMatrizR(m, trim(pre) // '_mr.dat')
  call guardarVectorR(v, trim(pre) // '_vr.dat')
  call guardarVectorI(iv, trim(pre) // '_vi.dat')
end program
