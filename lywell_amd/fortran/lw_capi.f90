! ISO_C_BINDING interfaces of the lywell C ABI (include/lywell.h).
! Complex(8) buffers are passed as complex(c_double_complex) arrays, which
! have the interleaved (re, im) layout the C side expects.
module lw_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LW_OK = 0, LW_ERR_ARG = -1, LW_ERR_HIP = -2, LW_ERR_OCTREE = -3, &
                               LW_ERR_NOMEM = -4, LW_ERR_STATE = -5, LW_ERR_PIVOT = -6, LW_ERR_NODEV = -7
  integer(c_int), parameter :: LW_NEAR_FP64 = 1
  integer(c_int), parameter :: LW_BATCH_FAR_FP64 = 1, LW_BATCH_FP64_RESIDUAL = 2
  integer(c_int), parameter :: LW_POL_VV = 0, LW_POL_HH = 1

  interface
    function lw_last_error() result(p) bind(C, name="lw_last_error")
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function lw_mesh_icosphere_size(subdiv, np, nt) result(rc) bind(C, name="lw_mesh_icosphere_size")
      import :: c_int
      integer(c_int), value :: subdiv
      integer(c_int), intent(out) :: np, nt
      integer(c_int) :: rc
    end function

    function lw_mesh_icosphere(subdiv, ax, ay, az, p_coord, t_p, t_normal) result(rc) &
        bind(C, name="lw_mesh_icosphere")
      import :: c_int, c_double
      integer(c_int), value :: subdiv
      real(c_double), value :: ax, ay, az
      real(c_double), intent(out) :: p_coord(*), t_normal(*)
      integer(c_int), intent(out) :: t_p(*)
      integer(c_int) :: rc
    end function

    function lw_rwg_build(np, p_coord, nt, t_p, e_p, e_t, e_po, e_long, e_centro) result(ne) &
        bind(C, name="lw_rwg_build")
      import :: c_int, c_double
      integer(c_int), value :: np, nt
      real(c_double), intent(in) :: p_coord(*)
      integer(c_int), intent(in) :: t_p(*)
      integer(c_int), intent(out) :: e_p(*), e_t(*), e_po(*)
      real(c_double), intent(out) :: e_long(*), e_centro(*)
      integer(c_int) :: ne
    end function

    function lw_tri_geometry(np, p_coord, nt, t_p, t_area, t_baric, t_baric_sub) result(rc) &
        bind(C, name="lw_tri_geometry")
      import :: c_int, c_double
      integer(c_int), value :: np, nt
      real(c_double), intent(in) :: p_coord(*)
      integer(c_int), intent(in) :: t_p(*)
      real(c_double), intent(out) :: t_area(*), t_baric(*), t_baric_sub(*)
      integer(c_int) :: rc
    end function

    function lw_mlfma_init(lambda, ladomax, flags, num_e, e_centro, e_t, e_po, e_long, num_t, t_area, &
                           t_baric, t_baric_sub, num_p, p_coord, t_p, e_p, t_normal, h) result(rc) &
        bind(C, name="lw_mlfma_init")
      import :: c_int, c_double, c_ptr
      real(c_double), value :: lambda, ladomax
      integer(c_int), value :: flags, num_e, num_t, num_p
      real(c_double), intent(in) :: e_centro(*), e_long(*), t_area(*), t_baric(*), t_baric_sub(*), &
                                    p_coord(*), t_normal(*)
      integer(c_int), intent(in) :: e_t(*), e_po(*), t_p(*), e_p(*)
      type(c_ptr), intent(out) :: h
      integer(c_int) :: rc
    end function

    function lw_mlfma_destroy(h) result(rc) bind(C, name="lw_mlfma_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function lw_matvec(h, x, y) result(rc) bind(C, name="lw_matvec")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: x(*)
      complex(c_double_complex), intent(out) :: y(*)
      integer(c_int) :: rc
    end function

    function lw_near_nnz(h) result(n) bind(C, name="lw_near_nnz")
      import :: c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_long_long) :: n
    end function

    ! z: address of a complex(4) (lw_near_csr) or complex(8) (lw_near_csr64) array
    function lw_near_csr(h, ia, ja, z) result(rc) bind(C, name="lw_near_csr")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, z
      integer(c_int), intent(out) :: ia(*), ja(*)
      integer(c_int) :: rc
    end function

    function lw_near_csr64(h, ia, ja, z) result(rc) bind(C, name="lw_near_csr64")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, z
      integer(c_int), intent(out) :: ia(*), ja(*)
      integer(c_int) :: rc
    end function

    function lw_ilu0(h, ierr) result(rc) bind(C, name="lw_ilu0")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: ierr
      integer(c_int) :: rc
    end function

    function lw_ilu0_csr(n, ia, ja, a, flags, alu, jlu, ju, ierr) result(rc) bind(C, name="lw_ilu0_csr")
      import :: c_int, c_ptr
      integer(c_int), value :: n, flags
      integer(c_int), intent(in) :: ia(*), ja(*)
      type(c_ptr), value :: a, alu
      integer(c_int), intent(out) :: jlu(*), ju(*), ierr
      integer(c_int) :: rc
    end function

    function lw_ilu0_get(h, alu, alu64, jlu, ju) result(rc) bind(C, name="lw_ilu0_get")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, alu, alu64
      integer(c_int), intent(out) :: jlu(*), ju(*)
      integer(c_int) :: rc
    end function

    function lw_ilu0_set(h, alu, alu64, jlu, ju) result(rc) bind(C, name="lw_ilu0_set")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, alu, alu64
      integer(c_int), intent(in) :: jlu(*), ju(*)
      integer(c_int) :: rc
    end function

    function lw_lusol0(h, y, x) result(rc) bind(C, name="lw_lusol0")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: y(*)
      complex(c_double_complex), intent(out) :: x(*)
      integer(c_int) :: rc
    end function

    function lw_cgs(h, b, guess, errmax, use_prec, maxit, x, iters, relres, converged) result(rc) &
        bind(C, name="lw_cgs")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: b(*), guess(*)
      real(c_double), value :: errmax
      integer(c_int), value :: use_prec, maxit
      complex(c_double_complex), intent(out) :: x(*)
      integer(c_int), intent(out) :: iters, converged
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function lw_plane_wave(h, dir, ctte, pol, b) result(rc) bind(C, name="lw_plane_wave")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      real(c_double), intent(in) :: dir(3)
      complex(c_double_complex), intent(in) :: ctte, pol(3)
      complex(c_double_complex), intent(out) :: b(*)
      integer(c_int) :: rc
    end function

    function lw_rcs_init(h, current, pol, ctte) result(rc) bind(C, name="lw_rcs_init")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: current(*), pol(3), ctte
      integer(c_int) :: rc
    end function

    function lw_rcs_sigma(h, n, theta, phi, sigma) result(rc) bind(C, name="lw_rcs_sigma")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: theta(*), phi(*)
      real(c_double), intent(out) :: sigma(*)
      integer(c_int) :: rc
    end function
    ! ---- result files (guardarMatrizR / guardarVectorC, modGlobalMetod.f90:192-306)
    function lw_dat_write_matrix_r(path, nrows, ncols, m) result(rc) bind(C, name="lw_dat_write_matrix_r")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nrows, ncols
      real(c_double), intent(in) :: m(*)
      integer(c_int) :: rc
    end function

    function lw_dat_write_vector_c(path, n, v) result(rc) bind(C, name="lw_dat_write_vector_c")
      import :: c_int, c_char, c_double_complex
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: n
      complex(c_double_complex), intent(in) :: v(*)
      integer(c_int) :: rc
    end function

    ! ---- multi-RHS form: blocks are complex(8) (nrhs, num_e) arrays; a
    ! ---- c_null_ptr guess is a zero guess
    function lw_batch_create(h, nrhs, flags, b) result(rc) bind(C, name="lw_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nrhs, flags
      type(c_ptr), intent(out) :: b
      integer(c_int) :: rc
    end function

    function lw_batch_destroy(b) result(rc) bind(C, name="lw_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function

    function lw_batch_nrhs(b) result(n) bind(C, name="lw_batch_nrhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: n
    end function

    function lw_batch_set_replacement(b, period) result(rc) bind(C, name="lw_batch_set_replacement")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: period
      integer(c_int) :: rc
    end function

    function lw_batch_matvec(b, x, y) result(rc) bind(C, name="lw_batch_matvec")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: b
      complex(c_double_complex), intent(in) :: x(*)
      complex(c_double_complex), intent(out) :: y(*)
      integer(c_int) :: rc
    end function

    function lw_batch_cgs(b, rhs, guess, errmax, use_prec, maxit, x, iters, relres, converged) result(rc) &
        bind(C, name="lw_batch_cgs")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: b, guess
      complex(c_double_complex), intent(in) :: rhs(*)
      real(c_double), value :: errmax
      integer(c_int), value :: use_prec, maxit
      complex(c_double_complex), intent(out) :: x(*)
      integer(c_int), intent(out) :: iters(*), converged(*)
      real(c_double), intent(out) :: relres(*)
      integer(c_int) :: rc
    end function

    function lw_batch_cgs_start(b, rhs, guess, errmax, use_prec, maxit) result(rc) &
        bind(C, name="lw_batch_cgs_start")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: b, guess
      complex(c_double_complex), intent(in) :: rhs(*)
      real(c_double), value :: errmax
      integer(c_int), value :: use_prec, maxit
      integer(c_int) :: rc
    end function

    function lw_batch_cgs_advance(b, iterations, state, relres) result(rc) bind(C, name="lw_batch_cgs_advance")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: iterations
      integer(c_int), intent(out) :: state(4, *)
      real(c_double), intent(out) :: relres(*)
      integer(c_int) :: rc
    end function

    function lw_batch_column(b, col, x) result(rc) bind(C, name="lw_batch_column")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: b
      integer(c_int), value :: col
      complex(c_double_complex), intent(out) :: x(*)
      integer(c_int) :: rc
    end function

    function lw_batch_restart_column(b, col, rhs, guess) result(rc) bind(C, name="lw_batch_restart_column")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: b, guess
      integer(c_int), value :: col
      complex(c_double_complex), intent(in) :: rhs(*)
      integer(c_int) :: rc
    end function

    function lw_batch_plane_wave(b, dirs, pols, ctte, rhs) result(rc) bind(C, name="lw_batch_plane_wave")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: b
      real(c_double), intent(in) :: dirs(3, *)
      complex(c_double_complex), intent(in) :: pols(3, *), ctte
      complex(c_double_complex), intent(out) :: rhs(*)
      integer(c_int) :: rc
    end function

    function lw_batch_rcs(b, cur, pols, ctte, theta, phi, sigma) result(rc) bind(C, name="lw_batch_rcs")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: b
      complex(c_double_complex), intent(in) :: cur(*), pols(3, *), ctte
      real(c_double), intent(in) :: theta(*), phi(*)
      real(c_double), intent(out) :: sigma(*)
      integer(c_int) :: rc
    end function

    ! ---- learned initial guess
    function lw_guess_load(path, m) result(rc) bind(C, name="lw_guess_load")
      import :: c_int, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      type(c_ptr), intent(out) :: m
      integer(c_int) :: rc
    end function

    function lw_mlp_destroy(m) result(rc) bind(C, name="lw_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int) :: rc
    end function

    function lw_guess_predict(m, h, ncol, dir, pol, dir1, x1, dir2, x2, guess) result(rc) &
        bind(C, name="lw_guess_predict")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: m, h
      integer(c_int), value :: ncol
      real(c_double), intent(in) :: dir(3, *), pol(3, *), dir1(3, *), dir2(3, *)
      complex(c_double_complex), intent(in) :: x1(*), x2(*)
      complex(c_double_complex), intent(out) :: guess(*)
      integer(c_int) :: rc
    end function

    ! ---- monostatic sweep (gen_rcs_monostatica, main.f90:967-1089)
    function lw_sweep_monostatic(h, b, net, n, span, pol, ctte, errmax, maxit, sigma, iterations, converged, &
                                 currents, stats) result(rc) bind(C, name="lw_sweep_monostatic")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h, b, net, currents
      integer(c_int), value :: n, pol, maxit
      real(c_double), value :: span, errmax
      complex(c_double_complex), intent(in) :: ctte
      real(c_double), intent(out) :: sigma(*)
      integer(c_int), intent(out) :: iterations(*), converged(*), stats(5)
      integer(c_int) :: rc
    end function
  end interface

contains

  ! Message of the last failed call, as a Fortran string.
  function lw_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = lw_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  subroutine lw_fail(where, rc)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    print *, trim(where) // ': lywell error ', rc, ' ', lw_error_text()
    error stop 'lywell: GPU path failed (no CPU fallback)'
  end subroutine
end module lw_capi
