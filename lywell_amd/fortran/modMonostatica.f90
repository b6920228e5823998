! The monostatic RCS sweep of the reference's main program
! (gen_rcs_monostatica, main.f90:967-1089) as a module routine over the
! batched solver.  The reference solves the n_pasos incidence angles
! angphi = (ang-1) pi/n_pasos one after another with CGSold, each from the
! previous angle's current times the phase correction of main.f90:1017-1022,
! and fills mat_rcs_mono(ang,:) = (angphi in degrees, 10 log10 sigmaRCS(pi/2,
! angphi)).  gen_rcs_monostatica_lote fills the same table for one
! polarisation with the angles run as chains through the multi-RHS solver
! (lw_sweep_monostatic): every column restarts as soon as it has converged,
! from the reference's warm start or, given a trained network, from the
! learned guess of two solved neighbours.  A caller with tipo_polarizac = 'VH'
! calls it for 'VV' and then 'HH', as the reference's loop does.
!
! The operator is the one inicializar_MLFMA set up (modMLFMA); the far field
! runs in complex(8), the solves stop at the same relative residual errmax as
! CGSold would, without a preconditioner.
module modMonostatica
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  use modMLFMA, only: lw_mlfma_handle, lw_mlfma_ne
  implicit none
  private
  public :: gen_rcs_monostatica_lote, lw_sweep_nrhs, lw_sweep_maxit, lw_sweep_stats

  ! columns of the batch (64 or 128), iteration cap per angle, and the counts
  ! of the last sweep (restarts, steals, angles started from zero / one / two
  ! solved neighbours)
  integer :: lw_sweep_nrhs = 64
  integer :: lw_sweep_maxit = 100000
  integer(c_int) :: lw_sweep_stats(5) = 0

contains

  subroutine gen_rcs_monostatica_lote(n_pasos, tipo_polarizac, ctte_onda, errmax, mat_rcs_mono, iteraciones, red)
    integer(kind=il), intent(in) :: n_pasos
    character(len=*), intent(in) :: tipo_polarizac           ! 'VV' or 'HH'
    complex(kind=dp), intent(in) :: ctte_onda
    real(kind=dp), intent(in) :: errmax
    real(kind=dp), intent(out) :: mat_rcs_mono(n_pasos, 2)
    integer(kind=il), intent(out), optional :: iteraciones(n_pasos)
    character(len=*), intent(in), optional :: red             ! trained network file (.npz)
    type(c_ptr) :: bt, net
    integer(c_int) :: rc, pol
    integer(c_int), allocatable :: its(:), conv(:)
    real(c_double), allocatable :: sig(:)
    real(kind=dp) :: dephi, angphi
    integer :: ang

    if (.not. c_associated(lw_mlfma_handle)) call lw_fail('gen_rcs_monostatica_lote: no operator', LW_ERR_STATE)
    select case (trim(tipo_polarizac))
    case ('VV')
      pol = LW_POL_VV
    case ('HH')
      pol = LW_POL_HH
    case default
      call lw_fail('gen_rcs_monostatica_lote: polarisation ' // trim(tipo_polarizac) // ' (VV or HH)', LW_ERR_ARG)
    end select
    net = c_null_ptr
    if (present(red)) then
      rc = lw_guess_load(trim(red) // c_null_char, net)
      if (rc /= LW_OK) call lw_fail('gen_rcs_monostatica_lote: network ' // trim(red), rc)
    end if
    rc = lw_batch_create(lw_mlfma_handle, int(lw_sweep_nrhs, c_int), LW_BATCH_FAR_FP64, bt)
    if (rc /= LW_OK) call lw_fail('gen_rcs_monostatica_lote: batch', rc)
    allocate(its(n_pasos), conv(n_pasos), sig(n_pasos))
    rc = lw_sweep_monostatic(lw_mlfma_handle, bt, net, int(n_pasos, c_int), pi, pol, ctte_onda, errmax, &
                             int(lw_sweep_maxit, c_int), sig, its, conv, c_null_ptr, lw_sweep_stats)
    if (rc /= LW_OK) call lw_fail('gen_rcs_monostatica_lote: sweep', rc)
    rc = lw_batch_destroy(bt)
    if (c_associated(net)) rc = lw_mlp_destroy(net)

    ! the table of main.f90:1040-1050
    dephi = 1.*pi/n_pasos
    do ang = 1, n_pasos
      angphi = (ang - 1)*dephi
      mat_rcs_mono(ang, 1) = angphi*180./pi
      mat_rcs_mono(ang, 2) = 10*log10(sig(ang))
    end do
    if (present(iteraciones)) iteraciones = its
    if (any(conv == 0)) print *, 'gen_rcs_monostatica_lote: ', count(conv == 0), ' angles stopped at the cap'
  end subroutine
end module modMonostatica
