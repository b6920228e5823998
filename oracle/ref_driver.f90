! Driver linked against the flang build of the reference modules
! (oracle/build_ref.sh).  TEST INFRASTRUCTURE ONLY.
!
! Reads a triangle mesh and run parameters from a stream file, runs the
! reference's own routines and writes their results as a stream file:
!   parametros_rwg / hallar_geometria_rwg (modRWG.f90), inicializar_MLFMA and
!   MatxVec_MLFMA (modMLFMA.f90), ilu0C / lusol0C / CGSold (modIterativo.f90),
!   inicializar_RCS / sigmaRCS (modRCS.f90).
! The plane wave is generar_onda's formula (main.f90:1488-1525) restated here,
! because that routine is internal to the reference's main program.
!
! Input stream (native endianness):
!   int32 np, nt; real64 p(3,np); int32 t(3,nt); real64 normal(3,nt)
!   real64 lambda, ladomax, errmax; int32 ntheta, use_prec (1: ILU0 in CGSold)
!   real64 dir(3); complex128 pol(3), ctte; real64 theta(ntheta), phi
! Output stream:
!   int32 ne, nnz
!   int32 e_t(2,ne), e_po(2,ne); real64 e_long(ne)
!   complex128 y1(ne)        MatxVec_MLFMA(v), v(i) = i/1.3823 + j 57.382/i
!   complex128 y2(ne)        MatxVec_MLFMA(b)
!   int32 ia(ne+1), ja(nnz); complex64 z(nnz)
!   int32 ierr; complex64 alu(nnz+1); int32 jlu(nnz+1), ju(ne)
!   complex128 lus(ne)       lusol0C applied to b
!   complex128 b(ne)
!   int32 nmatvec            matvec calls made by CGSold
!   complex128 x(ne)         CGSold solution from guess 0 (ILU0 if use_prec)
!   real64 sigma(ntheta)     sigmaRCS(theta(i), phi) of that solution
module drv_counter
  use modMLFMA
  implicit none
  integer :: nmv = 0
contains
  function counted_matvec(vecX) result(res)
    complex (kind = 8), intent(in), dimension(:) :: vecX
    complex (kind = 8), dimension(size(vecX,1)) :: res
    nmv = nmv + 1
    res = MatxVec_MLFMA(vecX)
  end function
end module

program ref_driver
  use modGlobalParam
  use modGlobalMetod
  use modRWG
  use modMLFMA
  use modIterativo
  use modRCS
  use drv_counter
  implicit none
  character(len=512) :: fin, fout
  integer :: u, i, nth, ierr, use_prec
  integer (kind = 4), target :: num_p, num_t, num_e
  real (kind = 8), allocatable, target :: p_coord(:,:), t_normal(:,:), e_long(:), e_centro(:,:)
  real (kind = 8), allocatable, target :: t_area(:), t_baric(:,:), t_baric_sub(:,:,:)
  integer (kind = 4), allocatable, target :: t_p(:,:), e_p(:,:), e_t(:,:), e_po(:,:)
  real (kind = 8) :: lam, lmax, errmax, dir(3), phi, ud(3)
  real (kind = 8), allocatable :: theta(:), sig(:)
  complex (kind = 8), target :: pol(3), ctte
  complex (kind = 8), allocatable, target :: v(:), y1(:), y2(:), b(:), x(:), guess(:), lus(:)
  complex (kind = 4), allocatable :: alu(:)
  integer (kind = 4), allocatable :: jlu(:), ju(:), iw(:)
  complex (kind = 8) :: e1(3)
  real (kind = 8) :: bar(3)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) num_p, num_t
  allocate(p_coord(3,num_p), t_p(3,num_t), t_normal(3,num_t))
  read(u) p_coord, t_p, t_normal
  read(u) lam, lmax, errmax, nth, use_prec
  allocate(theta(nth), sig(nth))
  read(u) dir, pol, ctte, theta, phi
  close(u)

  call setlambda(lam)
  conf%criterio_ladomax = lmax
  call parametros_rwg(p_coord, t_p, num_t, e_p, e_long, e_t, e_po, num_e, e_centro)
  call hallar_geometria_rwg(p_coord, t_p, num_p, num_t, t_area, t_baric, t_baric_sub)
  call inicializar_MLFMA(e_centro, num_e, e_t, e_po, e_long, t_area, t_baric, t_baric_sub, &
                         p_coord, t_p, e_p, num_p, num_t, t_normal)

  allocate(v(num_e), y1(num_e), y2(num_e), b(num_e), x(num_e), guess(num_e), lus(num_e))
  do i = 1, num_e
    v(i) = cmplx(dble(i)/1.3823d0, 57.382d0/dble(i), kind=8)
  end do
  y1 = MatxVec_MLFMA(v)

  ! plane wave, EFIE term of generar_onda (alphaCFIE = 1)
  ud = dir/sqrt(dot_product(dir, dir))
  do i = 1, num_e
    bar = t_baric(:, e_t(1,i))
    e1 = pol*ctte*exp((-jj)*(w*sqrt(mu*epsi))*dot_product(ud, bar))
    b(i) = sum(e1*(bar - p_coord(:, e_po(1,i))))
    bar = t_baric(:, e_t(2,i))
    e1 = pol*ctte*exp((-jj)*(w*sqrt(mu*epsi))*dot_product(ud, bar))
    b(i) = b(i) + sum(e1*(p_coord(:, e_po(2,i)) - bar))
    b(i) = b(i)*0.5*e_long(i)
  end do
  y2 = MatxVec_MLFMA(b)

  allocate(alu(n_esparcida_MLFMA + 1), jlu(n_esparcida_MLFMA + 1), ju(num_e), iw(num_e))
  alu = 0.
  jlu = 0
  ju = 0
  iw = 0
  call ilu0C(num_e, Zesparcida_MLFMA, ja_MLFMA, ia_MLFMA, alu, jlu, ju, iw, ierr)
  call lusol0C(num_e, b, lus, alu, jlu, ju)

  guess = 0.
  nmv = 0
  if (use_prec == 1) then
    call CGSold(counted_matvec, b, guess, errmax, num_e, x, prec_alu=alu, prec_jlu=jlu, prec_ju=ju)
  else
    call CGSold(counted_matvec, b, guess, errmax, num_e, x)
  end if
  call inicializar_RCS(e_centro, x, e_long, t_baric, e_t, num_e, num_t, pol, ctte)
  do i = 1, nth
    sig(i) = sigmaRCS(theta(i), phi)
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) num_e, n_esparcida_MLFMA
  write(u) e_t(:, 1:num_e), e_po(:, 1:num_e), e_long
  write(u) y1, y2
  write(u) ia_MLFMA, ja_MLFMA, Zesparcida_MLFMA
  write(u) ierr, alu, jlu, ju
  write(u) lus, b
  write(u) nmv
  write(u) x, sig
  close(u)
end program
