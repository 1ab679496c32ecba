! reference_types_mirror.f90 -- the parts of MCTomo's own modules that the
! m_likelihood drop-in (m_likelihood.f90) reads, under their module and type
! names, so that m_likelihood.f90 compiles unchanged against either these or
! the sampler's own modules:
!   m_utils        ii10                         (src/utils.f90)
!   m_exception    exception_raiseError         (src/m_exception.f90:223-228)
!   m_settings     T_GRID, T_MOD                (src/settings.f90:20-33)
!   like_settings  T_DATA, T_LIKE_BASE, T_LIKE_SET (src/likelihood_settings.f90:19-79)
!   run_info       T_RUN_INFO (model and noise part) (src/run_time_info.f90:27-57)
!   cgal_delaunay  d3                           (src/cgal_delaunay_wrapper.f90:54-57)
! Only the components the likelihood uses are declared; a sampler build links
! its own modules instead of this file.
module m_utils
    use iso_c_binding
    implicit none
    integer, parameter :: ii10 = c_double
end module m_utils

module m_exception
    implicit none
contains
    subroutine exception_raiseError(msg)
        character(len=*), intent(in) :: msg
        write(*, '(a,a)') 'Error: ', trim(msg)
        stop 1
    end subroutine exception_raiseError
end module m_exception

module m_settings
    use iso_c_binding
    use m_utils, only : ii10
    implicit none
    type T_GRID
        integer(c_int)    :: nx, ny, nz
        real( kind=ii10 ) :: xmin, xmax
        real( kind=ii10 ) :: ymin, ymax
        real( kind=ii10 ) :: zmin, zmax
        real( kind=ii10 ) :: dx, dy, dz
        real( kind=ii10 ) :: waterDepth
        real( kind=ii10 ) :: scaling
    end type T_GRID
    type T_MOD
        real( kind=ii10 ), dimension(:,:,:), allocatable :: vp, vs, rho
    end type T_MOD
end module m_settings

module like_settings
    use m_utils, only : ii10
    use m_settings, only : T_GRID
    implicit none
    type T_DATA
        integer :: datatype
        integer :: nrays, np, nmodes
        integer :: nsrc, nrev
        real( kind=ii10 ), dimension(:,:), allocatable   :: src
        real( kind=ii10 ), dimension(:,:), allocatable   :: rev
        integer, dimension(:,:,:), allocatable           :: raystat
        real( kind=ii10 ), dimension(:,:,:), allocatable :: ttime
        real( kind=ii10 ), dimension(:), allocatable     :: freqs
    end type T_DATA
    type T_LIKE_BASE
        real( kind=ii10 ) :: like, misfit, unweighted_misfit
        real( kind=ii10 ), dimension(:,:,:), allocatable :: vel, gvel
        real( kind=ii10 ), dimension(:,:,:), allocatable :: phaseTime
        real( kind=ii10 ), dimension(:), allocatable :: grads
        real( kind=ii10 ), dimension(:,:), allocatable :: srdist, sigma
    end type T_LIKE_BASE
    type T_LIKE_SET
        integer :: datatype
        integer :: sigdep
        integer :: nthreads
        type(T_GRID) :: grid
        integer :: isStraight
        real( kind=ii10 ) :: dPhaseVel
        integer :: raylov, phaseGroup
        real( kind=ii10 ) :: tol
        integer :: gridx, gridy
        integer :: sgref, sgdic, sgext
        integer :: order
        real( kind=ii10 ) :: band
        integer :: uar
        integer :: dynamic
    end type T_LIKE_SET
end module like_settings

module run_info
    use iso_c_binding
    implicit none
    type T_RUN_INFO
        integer(c_int) :: num_bad_model, num_bad_ray
        integer(c_size_t) :: ncells
        real(c_double), dimension(:,:), allocatable :: points
        real(c_double), dimension(:,:), allocatable :: parameters
        integer, dimension(:,:,:), allocatable :: sites_id
        real(c_double), dimension(:), allocatable :: bnoise0, bnoise1
        real(c_double), dimension(:), allocatable :: snoise0, snoise1
        real(c_double), dimension(:,:), allocatable :: locations
    end type T_RUN_INFO
end module run_info

module cgal_delaunay
    use iso_c_binding
    implicit none
    type, bind(C) :: d3
        real(c_double) :: x, y, z
    end type d3
end module cgal_delaunay
