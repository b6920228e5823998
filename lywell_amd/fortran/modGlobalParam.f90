! Kinds, physical constants and run configuration read by the modMLFMA,
! modIterativo and modRCS modules.  A host program that already has the
! reference's modGlobalParam (modGlobalParam.f90:1-123) links against that
! one instead: the names below are the ones those modules read from it and
! carry the same meaning, so either module satisfies them.
module modGlobalParam
  implicit none
  integer, parameter :: dp = kind(0.d0)
  integer, parameter :: dp_near = 4   ! near-field storage (complex(4) as shipped)
  integer, parameter :: dp_prec = 4   ! ILU0 factor storage
  integer, parameter :: il = 4

  real(kind=dp), parameter :: mu = 1.2566370614E-6_dp
  real(kind=dp), parameter :: epsi = 8.8541878176E-12_dp
  real(kind=dp), parameter :: pi = 3.141592653589793_dp
  complex(kind=dp), parameter :: jj = (0._dp, 1._dp)

  real(kind=dp) :: lambda = 1._dp, w = 0._dp, kappa = 0._dp

  ! the part of the reference's run configuration the hot path reads
  type :: config
    real(kind=dp) :: criterio_ladomax = 0.5_dp
    real(kind=dp) :: errorSolverCGS = 1.e-3_dp
  end type
  type(config) :: conf

contains

  ! wavelength setter: angular frequency and wavenumber follow from lambda
  subroutine setlambda(v)
    real(kind=dp), intent(in) :: v
    lambda = v
    w = 2._dp*pi*(1._dp/sqrt(mu*epsi))/lambda
    kappa = 2._dp*pi/lambda
  end subroutine
end module modGlobalParam
