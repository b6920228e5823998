! Fortran host program over modMLFMA / modIterativo / modRCS: the flow of the
! reference's testMLFMA (main.f90:1092-1154) on a synthetic surface.
!   mesh -> RWG -> inicializar_MLFMA -> plane wave -> ilu0C -> CGSold ->
!   inicializar_RCS / sigmaRCS
! Usage: lw_solve subdiv ax ay az lambda ladomax errsolver precond ntheta out
!   precond 1 = ILU0, 0 = none.  Writes a stream file:
!   int32 ne, iterations; real64 relres; complex128 y(ne) = MatxVec(v),
!   v(i) = i/1.3823 + j 57.382/i; complex128 I(ne); real64 sigma(ntheta) in
!   the phi = 0 cut at theta = (i-1) pi/(ntheta-1).
program lw_solve
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  use modMLFMA
  use modIterativo
  use modRCS
  implicit none
  character(len=256) :: arg, fout
  integer :: subdiv, prec, nth, i, u, ierr
  real(kind=dp) :: ax, ay, az, lam, lmax, errsolver
  integer(kind=il), target :: num_p, num_t, num_e
  real(kind=dp), allocatable, target :: p_coord(:,:), t_normal(:,:), e_long(:), e_centro(:,:)
  real(kind=dp), allocatable, target :: t_area(:), t_baric(:,:), t_baric_sub(:,:,:)
  integer(kind=il), allocatable, target :: t_p(:,:), e_p(:,:), e_t(:,:), e_po(:,:)
  integer(c_int), allocatable :: ep_c(:), et_c(:), epo_c(:)
  real(c_double), allocatable :: el_c(:), ec_c(:)
  complex(kind=dp), allocatable :: b(:), guess(:), Icur(:), v(:), y(:)
  complex(kind=dp_prec), allocatable, target :: alu(:)
  integer(kind=il), allocatable :: jlu(:), ju(:), iw(:)
  real(kind=dp), allocatable :: sig(:)
  complex(kind=dp) :: pol(3), ctte
  real(kind=dp) :: dir(3), th
  integer(c_int) :: rc, np_c, nt_c, ne_c

  call get_command_argument(1, arg); read(arg, *) subdiv
  call get_command_argument(2, arg); read(arg, *) ax
  call get_command_argument(3, arg); read(arg, *) ay
  call get_command_argument(4, arg); read(arg, *) az
  call get_command_argument(5, arg); read(arg, *) lam
  call get_command_argument(6, arg); read(arg, *) lmax
  call get_command_argument(7, arg); read(arg, *) errsolver
  call get_command_argument(8, arg); read(arg, *) prec
  call get_command_argument(9, arg); read(arg, *) nth
  call get_command_argument(10, fout)

  ! geometry (cap_geom, main.f90:1430-1468, on a generated surface)
  rc = lw_mesh_icosphere_size(subdiv, np_c, nt_c)
  if (rc /= LW_OK) call lw_fail('mesh size', rc)
  num_p = np_c
  num_t = nt_c
  allocate(p_coord(3, num_p), t_p(3, num_t), t_normal(3, num_t))
  rc = lw_mesh_icosphere(subdiv, ax, ay, az, p_coord, t_p, t_normal)
  if (rc /= LW_OK) call lw_fail('mesh', rc)
  allocate(ep_c(6*num_t), et_c(6*num_t), epo_c(6*num_t), el_c(3*num_t), ec_c(9*num_t))
  ne_c = lw_rwg_build(num_p, p_coord, num_t, t_p, ep_c, et_c, epo_c, el_c, ec_c)
  if (ne_c <= 0) call lw_fail('rwg', ne_c)
  num_e = ne_c
  allocate(e_p(2, num_e), e_t(2, num_e), e_po(2, num_e), e_long(num_e), e_centro(3, num_e))
  e_p = reshape(ep_c(1:2*num_e), [2, num_e])
  e_t = reshape(et_c(1:2*num_e), [2, num_e])
  e_po = reshape(epo_c(1:2*num_e), [2, num_e])
  e_long = el_c(1:num_e)
  e_centro = reshape(ec_c(1:3*num_e), [3, num_e])
  allocate(t_area(num_t), t_baric(3, num_t), t_baric_sub(3, 9, num_t))
  rc = lw_tri_geometry(num_p, p_coord, num_t, t_p, t_area, t_baric, t_baric_sub)
  if (rc /= LW_OK) call lw_fail('geometry', rc)

  call setlambda(lam)
  conf%criterio_ladomax = lmax
  call inicializar_MLFMA(e_centro, num_e, e_t, e_po, e_long, t_area, t_baric, t_baric_sub, p_coord, t_p, e_p, &
                         num_p, num_t, t_normal)

  allocate(v(num_e), y(num_e), b(num_e), guess(num_e), Icur(num_e))
  do i = 1, num_e
    v(i) = cmplx(dble(i)/1.3823d0, 57.382d0/dble(i), kind=dp)
  end do
  y = MatxVec_MLFMA(v)

  ! x-polarised wave travelling along -z (generar_onda, main.f90:1488-1525)
  dir = [0._dp, 0._dp, -1._dp]
  pol = [(1._dp, 0._dp), (0._dp, 0._dp), (0._dp, 0._dp)]
  ctte = (1._dp, 0._dp)
  rc = lw_plane_wave(lw_mlfma_handle, dir, ctte, pol, b)
  if (rc /= LW_OK) call lw_fail('plane wave', rc)
  guess = (0._dp, 0._dp)

  ierr = -1
  if (prec == 1) then
    ! crear_precond, ILU0 branch (main.f90:884-903)
    allocate(alu(n_esparcida_MLFMA + 1), jlu(n_esparcida_MLFMA + 1), ju(num_e), iw(num_e))
    call ilu0C(num_e, Zesparcida_MLFMA, ja_MLFMA, ia_MLFMA, alu, jlu, ju, iw, ierr)
  end if
  if (ierr == 0) then
    call CGSold(MatxVec_MLFMA, b, guess, errsolver, num_e, Icur, 1, alu, jlu, ju)
  else
    call CGSold(MatxVec_MLFMA, b, guess, errsolver, num_e, Icur, 1)
  end if

  allocate(sig(nth))
  call inicializar_RCS(e_centro, Icur, e_long, t_baric, e_t, num_e, num_t, pol, ctte)
  do i = 1, nth
    th = pi*dble(i - 1)/dble(max(nth - 1, 1))
    sig(i) = sigmaRCS(th, 0._dp)
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(num_e, 4), int(lw_cgs_last_iterations, 4)
  write(u) lw_cgs_last_relres
  write(u) y, Icur, sig
  close(u)
  call destruir_MLFMA()
end program lw_solve
