! modRCS on the GPU: inicializar_RCS / sigmaRCS (modRCS.f90:23-105) over the
! lywell C ABI.  The geometry is the one already uploaded by
! inicializar_MLFMA; the current vector and wave parameters are copied here.
module modRCS
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  use modMLFMA, only: lw_mlfma_handle, lw_mlfma_ne
  implicit none
  private
  public :: inicializar_RCS, sigmaRCS

contains

  subroutine inicializar_RCS(arg_centros_comun, arg_Vi, arg_long_l, arg_baric_t, arg_t_comun, arg_num_lcom, &
                             arg_num_t, arg_polariz, arg_ctteOnda)
    real(kind=dp), dimension(:,:), target :: arg_centros_comun
    complex(kind=dp), dimension(:), intent(in), target :: arg_Vi
    real(kind=dp), dimension(:), intent(in), target :: arg_long_l
    real(kind=dp), dimension(:,:), intent(in), target :: arg_baric_t
    integer(kind=il), dimension(:,:), intent(in), target :: arg_t_comun
    integer(kind=il), intent(in), target :: arg_num_lcom
    integer(kind=il), intent(in), target :: arg_num_t
    complex(kind=dp), dimension(3) :: arg_polariz
    complex(kind=dp) :: arg_ctteOnda
    integer(c_int) :: rc
    if (.not. c_associated(lw_mlfma_handle) .or. arg_num_lcom /= lw_mlfma_ne) &
      call lw_fail('inicializar_RCS: geometry not on the device (call inicializar_MLFMA)', LW_ERR_STATE)
    rc = lw_rcs_init(lw_mlfma_handle, arg_Vi, arg_polariz, arg_ctteOnda)
    if (rc /= LW_OK) call lw_fail('inicializar_RCS', rc)
  end subroutine

  function sigmaRCS(theta, phi) result(sigma)
    real(kind=dp), intent(in) :: theta, phi
    real(kind=dp) :: sigma
    real(c_double) :: th(1), ph(1), s(1)
    integer(c_int) :: rc
    th(1) = theta
    ph(1) = phi
    rc = lw_rcs_sigma(lw_mlfma_handle, 1, th, ph, s)
    if (rc /= LW_OK) call lw_fail('sigmaRCS', rc)
    sigma = s(1)
  end function
end module modRCS
