! modMLFMA on the GPU: the public interface of the reference's module
! (modMLFMA.f90:12) over the lywell C ABI.
!   inicializar_MLFMA  modMLFMA.f90:113   builds and uploads the operator
!   MatxVec_MLFMA      modMLFMA.f90:1901  y = Z x on the GPU
!   Zesparcida_MLFMA, ia_MLFMA, ja_MLFMA, n_esparcida_MLFMA   (:30-32)
!                      host copies of the near-field CSR, as the reference
!                      exposes them to crear_precond
!   destruir_MLFMA     modMLFMA.f90:85
! lambda and conf%criterio_ladomax are read from modGlobalParam as in the
! reference.  There is one operator per program, as there.
module modMLFMA
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  implicit none
  private
  public :: inicializar_MLFMA, MatxVec_MLFMA, Zesparcida_MLFMA, ia_MLFMA, ja_MLFMA, n_esparcida_MLFMA, &
            destruir_MLFMA, lw_mlfma_handle, lw_mlfma_ne

  integer(kind=il) :: n_esparcida_MLFMA = 0
  integer(kind=il), allocatable, dimension(:) :: ia_MLFMA, ja_MLFMA
  complex(kind=dp_near), allocatable, target, dimension(:) :: Zesparcida_MLFMA
  type(c_ptr) :: lw_mlfma_handle = c_null_ptr
  integer :: lw_mlfma_ne = 0

contains

  subroutine inicializar_MLFMA(arg_e_centro, arg_num_e, arg_e_t, arg_e_po, arg_e_long, arg_t_area, arg_t_baric, &
                               arg_t_baric_sub, arg_p_coord, arg_t_p, arg_e_p, arg_num_p, arg_num_t, arg_t_normal)
    real(kind=dp), dimension(:,:), target :: arg_e_centro
    integer(kind=il), target :: arg_num_e
    real(kind=dp), dimension(:,:), intent(in), target :: arg_t_baric
    real(kind=dp), dimension(:,:,:), intent(in), target :: arg_t_baric_sub
    real(kind=dp), dimension(:), intent(in), target :: arg_t_area, arg_e_long
    integer(kind=il), dimension(:,:), intent(in), target :: arg_e_t, arg_e_po
    real(kind=dp), target, dimension(:,:) :: arg_p_coord
    integer(kind=il), target, dimension(:,:) :: arg_t_p, arg_e_p
    integer(kind=il), target :: arg_num_p, arg_num_t
    real(kind=dp), target, dimension(:,:) :: arg_t_normal
    integer(c_int) :: rc, flags
    integer(c_long_long) :: nnz

    call destruir_MLFMA()
    flags = 0
    if (dp_near == 8) flags = LW_NEAR_FP64
    rc = lw_mlfma_init(lambda, conf%criterio_ladomax, flags, arg_num_e, arg_e_centro, arg_e_t, arg_e_po, &
                       arg_e_long, arg_num_t, arg_t_area, arg_t_baric, arg_t_baric_sub, arg_num_p, arg_p_coord, &
                       arg_t_p, arg_e_p, arg_t_normal, lw_mlfma_handle)
    if (rc == LW_ERR_OCTREE) then
      ! the reference's behaviour (modMLFMA.f90:195-198)
      print *, 'El dispersor es demasiado pequeno electricamente para aplicar MLFMA'
      stop "El programa ha finalizado"
    end if
    if (rc /= LW_OK) call lw_fail('inicializar_MLFMA', rc)
    lw_mlfma_ne = arg_num_e
    nnz = lw_near_nnz(lw_mlfma_handle)
    n_esparcida_MLFMA = int(nnz, il)
    allocate(ia_MLFMA(arg_num_e + 1), ja_MLFMA(n_esparcida_MLFMA), Zesparcida_MLFMA(n_esparcida_MLFMA))
    if (dp_near == 8) then
      rc = lw_near_csr64(lw_mlfma_handle, ia_MLFMA, ja_MLFMA, c_loc(Zesparcida_MLFMA(1)))
    else
      rc = lw_near_csr(lw_mlfma_handle, ia_MLFMA, ja_MLFMA, c_loc(Zesparcida_MLFMA(1)))
    end if
    if (rc /= LW_OK) call lw_fail('inicializar_MLFMA', rc)
  end subroutine

  function MatxVec_MLFMA(vecX) result(res)
    complex(kind=dp), intent(in), dimension(:) :: vecX
    complex(kind=dp), dimension(size(vecX,1)) :: res
    integer(c_int) :: rc
    if (.not. c_associated(lw_mlfma_handle)) call lw_fail('MatxVec_MLFMA: not initialised', LW_ERR_STATE)
    if (size(vecX) /= lw_mlfma_ne) call lw_fail('MatxVec_MLFMA: vector length', LW_ERR_ARG)
    rc = lw_matvec(lw_mlfma_handle, vecX, res)
    if (rc /= LW_OK) call lw_fail('MatxVec_MLFMA', rc)
  end function

  subroutine destruir_MLFMA()
    integer(c_int) :: rc
    if (c_associated(lw_mlfma_handle)) then
      rc = lw_mlfma_destroy(lw_mlfma_handle)
      lw_mlfma_handle = c_null_ptr
    end if
    if (allocated(ia_MLFMA)) deallocate(ia_MLFMA)
    if (allocated(ja_MLFMA)) deallocate(ja_MLFMA)
    if (allocated(Zesparcida_MLFMA)) deallocate(Zesparcida_MLFMA)
    n_esparcida_MLFMA = 0
    lw_mlfma_ne = 0
  end subroutine
end module modMLFMA
