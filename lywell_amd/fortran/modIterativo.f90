! modIterativo on the GPU: the hot-path routines of the reference's module
! with their F77-style dummy arguments, over the lywell C ABI.
!   CGSold   modIterativo.f90:331-493   right-preconditioned CGS; with
!            MatxVec_MLFMA as the operator the whole loop runs on the GPU
!   ilu0C    modIterativo.f90:1805-1976 ILU(0) in MSR form, of the MLFMA near
!            field or of any CSR matrix with ascending columns
!   lusol0C  modIterativo.f90:504-560   forward/backward solve with the factor
! Only the MLFMA operator of modMLFMA is accepted: there is no CPU path, and
! CGSold stops with a message for any other operator.
module modIterativo
  use iso_c_binding
  use modGlobalParam
  use lw_capi
  use modMLFMA, only: MatxVec_MLFMA, Zesparcida_MLFMA, lw_mlfma_handle, lw_mlfma_ne
  implicit none
  private
  public :: CGSold, ilu0C, lusol0C, lw_cgs_maxit, lw_cgs_last_iterations, lw_cgs_last_relres

  ! iteration cap (the reference's loop has none) and the last run's outcome
  integer :: lw_cgs_maxit = 100000
  integer :: lw_cgs_last_iterations = 0
  real(kind=dp) :: lw_cgs_last_relres = 0._dp
  ! the factor last produced by ilu0C, so CGSold can tell it is already on
  ! the device
  type(c_ptr) :: factor_on_device = c_null_ptr

  abstract interface
    function matvec_iface(vecX) result(res)
      complex(kind=8), intent(in), dimension(:) :: vecX
      complex(kind=8), dimension(size(vecX,1)) :: res
    end function
  end interface

contains

  subroutine CGSold(func_MatxVec, b, guess, errmax, n, x, mensaje, prec_alu, prec_jlu, prec_ju)
    complex(kind=dp), intent(in), dimension(n) :: guess, b
    complex(kind=dp), intent(out), dimension(n) :: x
    real(kind=dp), intent(in) :: errmax
    integer(kind=il), intent(in) :: n
    integer(kind=il), optional, intent(in) :: mensaje
    complex(kind=dp_prec), optional, intent(in), target, dimension(:) :: prec_alu
    integer(kind=il), optional, intent(in), dimension(:) :: prec_jlu
    integer(kind=il), optional, intent(in), dimension(:) :: prec_ju
    procedure(matvec_iface) :: func_MatxVec
    procedure(matvec_iface), pointer :: op
    integer(c_int) :: rc, use_prec, iters, conv
    real(c_double) :: relres

    op => func_MatxVec
    if (.not. associated(op, MatxVec_MLFMA)) then
      print *, 'CGSold: only MatxVec_MLFMA runs on the GPU path'
      error stop 'lywell: unsupported operator'
    end if
    if (.not. c_associated(lw_mlfma_handle) .or. n /= lw_mlfma_ne) &
      call lw_fail('CGSold: operator not initialised for this size', LW_ERR_STATE)
    use_prec = 0
    if (present(prec_alu) .and. present(prec_jlu) .and. present(prec_ju)) then
      use_prec = 1
      if (.not. c_associated(factor_on_device, c_loc(prec_alu(1)))) then
        if (dp_prec == 8) then
          rc = lw_ilu0_set(lw_mlfma_handle, c_null_ptr, c_loc(prec_alu(1)), prec_jlu, prec_ju)
        else
          rc = lw_ilu0_set(lw_mlfma_handle, c_loc(prec_alu(1)), c_null_ptr, prec_jlu, prec_ju)
        end if
        if (rc /= LW_OK) call lw_fail('CGSold: preconditioner upload', rc)
        factor_on_device = c_loc(prec_alu(1))
      end if
    end if
    rc = lw_cgs(lw_mlfma_handle, b, guess, errmax, use_prec, lw_cgs_maxit, x, iters, relres, conv)
    if (rc /= LW_OK) call lw_fail('CGSold', rc)
    lw_cgs_last_iterations = iters
    lw_cgs_last_relres = relres
    if (present(mensaje)) then
      print '(a,i0,a,es12.5)', ' CGS iteraciones: ', iters, ' residuo relativo ', relres
      if (conv == 0) print *, 'CGSold: no convergence within ', lw_cgs_maxit, ' iterations'
    end if
  end subroutine

  subroutine ilu0C(n, a, ja, ia, alu, jlu, ju, iw, ierr)
    integer(kind=il) :: n
    complex(kind=dp_near), target :: a(*)
    complex(kind=dp_prec), target :: alu(*)
    integer(kind=il) :: ia(n+1), ja(*), jlu(*), ju(*), iw(n)
    integer(kind=il) :: ierr
    integer(c_int) :: rc, piv, flags

    iw(1:n) = 0
    if (.not. (c_associated(lw_mlfma_handle) .and. n == lw_mlfma_ne .and. allocated(Zesparcida_MLFMA))) then
      call ilu0_any()
      return
    end if
    if (.not. c_associated(c_loc(a(1)), c_loc(Zesparcida_MLFMA(1)))) then
      call ilu0_any()
      return
    end if
    rc = lw_ilu0(lw_mlfma_handle, piv)
    if (rc == LW_ERR_PIVOT) then
      ierr = piv
      return
    end if
    if (rc /= LW_OK) call lw_fail('ilu0C', rc)
    if (dp_prec == 8) then
      rc = lw_ilu0_get(lw_mlfma_handle, c_null_ptr, c_loc(alu(1)), jlu, ju)
    else
      rc = lw_ilu0_get(lw_mlfma_handle, c_loc(alu(1)), c_null_ptr, jlu, ju)
    end if
    if (rc /= LW_OK) call lw_fail('ilu0C', rc)
    factor_on_device = c_loc(alu(1))
    ierr = 0

  contains

    ! any other CSR matrix: factored on the device without an operator; the
    ! factor stays with the caller (CGSold and lusol0C upload it)
    subroutine ilu0_any()
      flags = 0
      if (dp_prec == 8) flags = LW_NEAR_FP64
      if (dp_near /= dp_prec) call lw_fail('ilu0C: the matrix and the factor must share a precision', LW_ERR_ARG)
      rc = lw_ilu0_csr(n, ia, ja, c_loc(a(1)), flags, c_loc(alu(1)), jlu, ju, piv)
      if (rc == LW_ERR_PIVOT) then
        ierr = piv
        return
      end if
      if (rc /= LW_OK) call lw_fail('ilu0C', rc)
      factor_on_device = c_null_ptr
      ierr = 0
    end subroutine
  end subroutine

  subroutine lusol0C(n, y, x, alu, jlu, ju)
    integer(kind=il) :: n
    complex(kind=dp_prec), target :: alu(*)
    integer(kind=il) :: jlu(*), ju(*)
    complex(kind=dp) :: x(n), y(n)
    integer(c_int) :: rc
    if (.not. c_associated(lw_mlfma_handle) .or. n /= lw_mlfma_ne) &
      call lw_fail('lusol0C: operator not initialised for this size', LW_ERR_STATE)
    if (.not. c_associated(factor_on_device, c_loc(alu(1)))) then
      if (dp_prec == 8) then
        rc = lw_ilu0_set(lw_mlfma_handle, c_null_ptr, c_loc(alu(1)), jlu, ju)
      else
        rc = lw_ilu0_set(lw_mlfma_handle, c_loc(alu(1)), c_null_ptr, jlu, ju)
      end if
      if (rc /= LW_OK) call lw_fail('lusol0C: factor upload', rc)
      factor_on_device = c_loc(alu(1))
    end if
    rc = lw_lusol0(lw_mlfma_handle, y, x)
    if (rc /= LW_OK) call lw_fail('lusol0C', rc)
  end subroutine
end module modIterativo
