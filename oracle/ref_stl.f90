! Reads an ASCII STL with the reference's reader (obtener_p_t,
! modReadStl.f90:1378-1533) and writes what it returns as a stream file:
!   int32 np, nt; real64 p(3,np); int32 t(3,nt); real64 normal(3,nt)
! TEST INFRASTRUCTURE ONLY (linked by oracle/build_ref.sh).
program ref_stl
  use modGlobalParam
  use modReadStl
  implicit none
  character(len=512) :: fin, fout
  real(kind=8), allocatable :: p(:,:), n(:,:)
  integer(kind=4), allocatable :: t(:,:)
  integer(kind=4) :: np, nt
  integer :: u
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call obtener_p_t(trim(fin), p, t, np, nt, n)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) np, nt
  write(u) p(:, 1:np), t(:, 1:nt), n(:, 1:nt)
  close(u)
end program
