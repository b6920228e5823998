! Fortran host program over modMLFMA / modMonostatica: the reference's
! monostatic RCS run (gen_rcs_monostatica, main.f90:967-1089) on a synthetic
! surface, with the angles solved through the batched solver.
!   mesh -> RWG -> inicializar_MLFMA -> gen_rcs_monostatica_lote per
!   polarisation -> guardarMatrizR tables
! Usage: lw_rcsmono subdiv ax ay az lambda ladomax errsolver n_pasos polariz
!                   testname network out
!   polariz VV, HH or VH (VV then HH, as the reference's loop); network: the
!   trained network file, or - for the reference's warm start.
! Writes testname_RCSMONO<pol>.dat per polarisation (the reference's file
! name and f16.8 table, main.f90:1064) and a stream file: int32 ne, npol,
! n_pasos; then per polarisation real64 mat_rcs_mono(n_pasos, 2), int32
! iterations(n_pasos), int32 stats(5).
program lw_rcsmono
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  use modMLFMA
  use modMonostatica
  implicit none
  character(len=256) :: arg, fout, tname, red
  character(len=2) :: pols(2)
  integer :: subdiv, npol, ip, u
  integer(kind=il) :: n_pasos
  real(kind=dp) :: ax, ay, az, lam, lmax, errsolver
  integer(kind=il), target :: num_p, num_t, num_e
  real(kind=dp), allocatable, target :: p_coord(:,:), t_normal(:,:), e_long(:), e_centro(:,:)
  real(kind=dp), allocatable, target :: t_area(:), t_baric(:,:), t_baric_sub(:,:,:)
  integer(kind=il), allocatable, target :: t_p(:,:), e_p(:,:), e_t(:,:), e_po(:,:)
  integer(c_int), allocatable :: ep_c(:), et_c(:), epo_c(:)
  real(c_double), allocatable :: el_c(:), ec_c(:)
  real(kind=dp), allocatable :: mat(:,:)
  integer(kind=il), allocatable :: its(:)
  complex(kind=dp) :: ctte
  integer(c_int) :: rc, np_c, nt_c, ne_c

  call get_command_argument(1, arg); read(arg, *) subdiv
  call get_command_argument(2, arg); read(arg, *) ax
  call get_command_argument(3, arg); read(arg, *) ay
  call get_command_argument(4, arg); read(arg, *) az
  call get_command_argument(5, arg); read(arg, *) lam
  call get_command_argument(6, arg); read(arg, *) lmax
  call get_command_argument(7, arg); read(arg, *) errsolver
  call get_command_argument(8, arg); read(arg, *) n_pasos
  call get_command_argument(9, arg)
  call get_command_argument(10, tname)
  call get_command_argument(11, red)
  call get_command_argument(12, fout)
  select case (trim(arg))
  case ('VV', 'HH')
    npol = 1
    pols(1) = trim(arg)
  case ('VH')
    npol = 2
    pols = ['VV', 'HH']
  case default
    print *, 'polariz: VV, HH or VH'
    error stop 1
  end select

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

  ctte = (1._dp, 0._dp)
  allocate(mat(n_pasos, 2), its(n_pasos))
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(num_e, 4), int(npol, 4), int(n_pasos, 4)
  do ip = 1, npol
    if (trim(red) == '-') then
      call gen_rcs_monostatica_lote(n_pasos, pols(ip), ctte, errsolver, mat, its)
    else
      call gen_rcs_monostatica_lote(n_pasos, pols(ip), ctte, errsolver, mat, its, red=trim(red))
    end if
    print *, 'RCS monostatica ', pols(ip), ': ', n_pasos, ' angulos, iteraciones ', sum(its), &
             ', reinicios/robos ', lw_sweep_stats(1:2)
    ! guardarMatrizR(mat_rcs_mono, trim(conf%test_name) // '_RCSMONO' // pol // '.dat')
    rc = lw_dat_write_matrix_r(trim(tname) // '_RCSMONO' // pols(ip) // '.dat' // c_null_char, &
                               int(n_pasos, c_int), 2_c_int, mat)
    if (rc /= LW_OK) call lw_fail('guardarMatrizR', rc)
    write(u) mat, int(its, 4), int(lw_sweep_stats, 4)
  end do
  close(u)
  call destruir_MLFMA()
end program lw_rcsmono
