! m_likelihood.f90 -- the sampler's likelihood module over libmctomo_hip.so.
!
! The same module name, public names and argument lists as MCTomo's
! m_likelihood (src/likelihood.f90:3-191): like_setup, likelihood,
! noise_likelihood and location_likelihood, with T_LIKE as there plus one
! component, the device contexts.  Replacing src/likelihood.f90 with this file
! (and linking include/mctomo_hip.f90 and the library) moves the forward model
! to the GPU under the unchanged sampler.
!
! The contexts hold device memory and the immutable data only: like_copy =
! like (src/mcmc_loc2.f90:201) copies the handles, and every call here passes
! the chain's state in (the Voronoi model, noise, locations, and for
! noise_likelihood the traveltimes of like%likelihoods(i)), so the restore on
! reject (:554-567) stays correct.  Results are written into the same
! T_LIKE_BASE arrays the reference fills (phaseTime, srdist, sigma).  Soft
! rejects set like = huge and count RTI%num_bad_model / num_bad_ray as
! surf_likelihood does (src/likelihood_surf.F90:161-164,212-224,341-345);
! errors stop through exception_raiseError.  The perturbed box is not needed:
! every call evaluates the whole grid, which gives the box call's result.
! update_lgP_grads (HMC gradients) has no device form here.
module m_likelihood

    use iso_c_binding
    use m_exception,   only : exception_raiseError
    use m_utils,       only : ii10
    use m_settings,    only : T_MOD
    use run_info,      only : T_RUN_INFO
    use like_settings, only : T_DATA, T_LIKE_BASE, T_LIKE_SET
    use cgal_delaunay, only : d3
    use mctomo_hip

    implicit none

    private

    public :: T_LIKE, like_setup
    public :: likelihood, noise_likelihood, location_likelihood

    type T_LIKE
        real(kind=ii10) :: like, misfit, unweighted_misfit
        real(kind=ii10), dimension(:), allocatable :: grads
        type(T_LIKE_BASE), dimension(:), allocatable :: likelihoods
        ! device contexts: (1) body waves, (2) surface waves
        type(c_ptr), dimension(2) :: gpu = [c_null_ptr, c_null_ptr]
    end type

contains

    subroutine like_setup(like, dat, like_set, ncell_max)
        implicit none
        type(T_LIKE), intent(out) :: like
        type(T_DATA), dimension(:), intent(in) :: dat
        type(T_LIKE_SET), intent(in) :: like_set
        integer, intent(in) :: ncell_max

        type(mct_grid) :: g
        integer :: i, n

        like%like = huge(like%like)
        like%misfit = huge(like%misfit)
        like%unweighted_misfit = huge(like%unweighted_misfit)
        allocate(like%grads(2*ncell_max + 4*dat(1)%nsrc))
        n = 1
        if (like_set%datatype == 3) n = 2
        allocate(like%likelihoods(n))
        do i = 1, n
            call base_setup(like%likelihoods(i), dat(i), like_set)
        enddo

        g%nx = like_set%grid%nx
        g%ny = like_set%grid%ny
        g%nz = like_set%grid%nz
        g%xmin = like_set%grid%xmin
        g%ymin = like_set%grid%ymin
        g%zmin = like_set%grid%zmin
        g%dx = like_set%grid%dx
        g%dy = like_set%grid%dy
        g%dz = like_set%grid%dz
        g%scaling = like_set%grid%scaling

        select case (like_set%datatype)
        case (0, 1)
            like%gpu(1) = body_context(g, dat(1), like_set, like_set%datatype, ncell_max)
        case (2)
            like%gpu(2) = surf_context(g, dat(1), like_set, 2, ncell_max)
        case (3)
            like%gpu(1) = body_context(g, dat(1), like_set, 3, ncell_max)
            like%gpu(2) = surf_context(g, dat(2), like_set, 3, ncell_max)
        case default
            call exception_raiseError('mctomo_hip: unknown datatype')
        end select

    end subroutine like_setup

    ! the arrays likeBaseSetup allocates (src/likelihood_settings.f90:82-164)
    subroutine base_setup(lb, dat, like_set)
        type(T_LIKE_BASE), intent(out) :: lb
        type(T_DATA), intent(in) :: dat
        type(T_LIKE_SET), intent(in) :: like_set
        lb%like = huge(lb%like)
        lb%misfit = huge(lb%misfit)
        lb%unweighted_misfit = huge(lb%unweighted_misfit)
        allocate(lb%phaseTime(dat%nrev, dat%nsrc, dat%np))
        allocate(lb%srdist(dat%nsrc*dat%nrev, dat%np))
        allocate(lb%sigma(dat%nsrc*dat%nrev, dat%np))
        lb%phaseTime = 0
        lb%srdist = 0
        lb%sigma = 0
        if (dat%datatype == 2) then
            allocate(lb%vel(dat%np, like_set%grid%ny + 2, like_set%grid%nx + 2))
            lb%vel = 0
        endif
    end subroutine base_setup

    function body_context(g, dat, like_set, datatype, ncell_max) result(ctx)
        type(mct_grid), intent(in) :: g
        type(T_DATA), intent(in) :: dat
        type(T_LIKE_SET), intent(in) :: like_set
        integer, intent(in) :: datatype, ncell_max
        type(c_ptr) :: ctx
        type(mct_body_setup) :: s
        real(c_double), allocatable, target :: rev(:,:), ttime(:,:,:)
        integer(c_int), allocatable, target :: raystat(:,:,:)
        integer(c_int) :: rc
        allocate(rev(3, dat%nrev), ttime(dat%nsrc*dat%nrev, 3, dat%np), raystat(dat%nsrc*dat%nrev, 2, dat%np))
        rev = dat%rev(1:3, :)
        ttime = dat%ttime
        raystat = dat%raystat
        s%datatype = datatype
        s%order = like_set%order
        s%sigdep = like_set%sigdep
        s%nsrc = dat%nsrc
        s%nrev = dat%nrev
        s%rev = c_loc(rev)
        s%ttime = c_loc(ttime)
        s%raystat = c_loc(raystat)
        s%nrays = dat%nrays
        s%max_cells = ncell_max
        s%max_chains = 1
        s%dynamic = like_set%dynamic
        ctx = c_null_ptr
        rc = mct_body_create(g, s, ctx)
        if (rc /= MCT_OK) call exception_raiseError('mctomo_hip: mct_body_create failed')
    end function body_context

    function surf_context(g, dat, like_set, datatype, ncell_max) result(ctx)
        type(mct_grid), intent(in) :: g
        type(T_DATA), intent(in) :: dat
        type(T_LIKE_SET), intent(in) :: like_set
        integer, intent(in) :: datatype, ncell_max
        type(c_ptr) :: ctx
        type(mct_surf_setup) :: s
        real(c_double), allocatable, target :: src(:,:), rev(:,:), ttime(:,:,:), freqs(:)
        integer(c_int), allocatable, target :: raystat(:,:,:)
        integer(c_int) :: rc
        allocate(src(2, dat%nsrc), rev(2, dat%nrev), freqs(dat%np))
        allocate(ttime(dat%nsrc*dat%nrev, 3, dat%np), raystat(dat%nsrc*dat%nrev, 2, dat%np))
        src = dat%src(1:2, :)
        rev = dat%rev(1:2, :)
        freqs = dat%freqs
        ttime = dat%ttime
        raystat = dat%raystat
        s%raylov = like_set%raylov
        s%phase_group = like_set%phaseGroup
        s%dphase = like_set%dPhaseVel
        s%np = dat%np
        s%freqs = c_loc(freqs)
        s%nsrc = dat%nsrc
        s%nrev = dat%nrev
        s%src = c_loc(src)
        s%rev = c_loc(rev)
        s%ttime = c_loc(ttime)
        s%raystat = c_loc(raystat)
        s%sigdep = like_set%sigdep
        s%nrays = dat%nrays
        s%gridx = like_set%gridx
        s%gridy = like_set%gridy
        s%sgref = like_set%sgref
        s%sgdic = like_set%sgdic
        s%sgext = like_set%sgext
        s%order = like_set%order
        s%band = like_set%band
        s%max_cells = ncell_max
        s%max_chains = 1
        s%straight = like_set%isStraight
        s%datatype = datatype
        s%water_depth = like_set%grid%waterDepth
        ctx = c_null_ptr
        rc = mct_surf_create(g, s, ctx)
        if (rc /= MCT_OK) call exception_raiseError('mctomo_hip: mct_surf_create failed')
    end function surf_context

    ! a status from the library: > 0 the model is rejected (like = huge and the
    ! sampler's counters), < 0 an error
    logical function rejected(rc, lb, RTI)
        integer(c_int), intent(in) :: rc
        type(T_LIKE_BASE), intent(inout) :: lb
        type(T_RUN_INFO), intent(inout) :: RTI
        character(len=16) :: code
        rejected = .false.
        if (rc > 0) then
            lb%like = huge(lb%like)
            if (rc == MCT_REJECT_RAY) then
                RTI%num_bad_ray = RTI%num_bad_ray + 1
            else
                RTI%num_bad_model = RTI%num_bad_model + 1
            endif
            rejected = .true.
        elseif (rc < 0) then
            write(code, '(i0)') rc
            call exception_raiseError('mctomo_hip: likelihood failed with status ' // trim(code))
        endif
    end function rejected

    subroutine body_part(ctx, RTI, lb)
        type(c_ptr), intent(in) :: ctx
        type(T_RUN_INFO), intent(inout), target :: RTI
        type(T_LIKE_BASE), intent(inout), target :: lb
        real(c_double), target :: out(3)
        integer(c_int) :: rc
        rc = mct_body_likelihood(ctx, c_loc(RTI%points), c_loc(RTI%parameters), int(RTI%ncells, c_int), &
                                 c_loc(RTI%locations), c_loc(RTI%bnoise0), c_loc(RTI%bnoise1), c_loc(out), &
                                 c_loc(lb%phaseTime), c_loc(lb%srdist), c_loc(lb%sigma))
        if (rejected(rc, lb, RTI)) return
        lb%like = out(1)
        lb%misfit = out(2)
        lb%unweighted_misfit = out(3)
    end subroutine body_part

    subroutine surf_part(ctx, RTI, lb)
        type(c_ptr), intent(in) :: ctx
        type(T_RUN_INFO), intent(inout), target :: RTI
        type(T_LIKE_BASE), intent(inout), target :: lb
        real(c_double), target :: out(3)
        integer(c_int) :: rc
        rc = mct_surf_likelihood(ctx, c_loc(RTI%points), c_loc(RTI%parameters), int(RTI%ncells, c_int), &
                                 c_loc(RTI%snoise0), c_loc(RTI%snoise1), c_loc(out), c_loc(lb%phaseTime), &
                                 c_loc(lb%srdist), c_loc(lb%sigma), c_null_ptr)
        if (rejected(rc, lb, RTI)) return
        lb%like = out(1)
        lb%misfit = out(2)
        lb%unweighted_misfit = out(3)
    end subroutine surf_part

    subroutine sum_parts(like, n)
        type(T_LIKE), intent(inout) :: like
        integer, intent(in) :: n
        if (n == 1) then
            like%like = like%likelihoods(1)%like
            like%misfit = like%likelihoods(1)%misfit
            like%unweighted_misfit = like%likelihoods(1)%unweighted_misfit
        else
            like%like = like%likelihoods(1)%like + like%likelihoods(2)%like
            like%misfit = like%likelihoods(1)%misfit + like%likelihoods(2)%misfit
            like%unweighted_misfit = like%likelihoods(1)%unweighted_misfit + like%likelihoods(2)%unweighted_misfit
        endif
    end subroutine sum_parts

    subroutine likelihood(dat, model, RTI, perturbed_box, like_set, like)
        implicit none
        type(T_DATA), dimension(:), intent(in)      :: dat
        type(T_MOD), intent(inout)                  :: model
        type(T_RUN_INFO), intent(inout), target     :: RTI
        type(d3), dimension(2), intent(in)          :: perturbed_box
        type(T_LIKE_SET), intent(in)                :: like_set
        type(T_LIKE), intent(inout), target         :: like

        select case (like_set%datatype)
        case (0, 1)
            call body_part(like%gpu(1), RTI, like%likelihoods(1))
            call sum_parts(like, 1)
        case (2)
            call surf_part(like%gpu(2), RTI, like%likelihoods(1))
            call sum_parts(like, 1)
        case (3)
            call body_part(like%gpu(1), RTI, like%likelihoods(1))
            call surf_part(like%gpu(2), RTI, like%likelihoods(2))
            call sum_parts(like, 2)
        end select

    end subroutine likelihood

    subroutine noise_likelihood(dat, RTI, like_set, like)
        implicit none
        type(T_DATA), dimension(:), intent(in)      :: dat
        type(T_RUN_INFO), intent(inout), target     :: RTI
        type(T_LIKE_SET), intent(in)                :: like_set
        type(T_LIKE), intent(inout), target         :: like
        real(c_double), target :: out(3)
        integer(c_int) :: rc
        integer :: i, ib, is

        ib = 0
        is = 0
        select case (like_set%datatype)
        case (0, 1)
            ib = 1
        case (2)
            is = 1
        case (3)
            ib = 1
            is = 2
        end select
        if (ib > 0) then
            rc = mct_body_noise_likelihood_from(like%gpu(1), c_loc(like%likelihoods(ib)%phaseTime), &
                                                c_loc(RTI%locations), c_loc(RTI%bnoise0), c_loc(RTI%bnoise1), &
                                                c_loc(out), c_loc(like%likelihoods(ib)%sigma))
            if (.not. rejected(rc, like%likelihoods(ib), RTI)) then
                like%likelihoods(ib)%like = out(1)
                like%likelihoods(ib)%misfit = out(2)
                like%likelihoods(ib)%unweighted_misfit = out(3)
            endif
        endif
        if (is > 0) then
            rc = mct_surf_noise_likelihood_from(like%gpu(2), c_loc(like%likelihoods(is)%phaseTime), &
                                                c_loc(like%likelihoods(is)%srdist), c_loc(RTI%snoise0), &
                                                c_loc(RTI%snoise1), c_loc(out), c_loc(like%likelihoods(is)%sigma))
            if (.not. rejected(rc, like%likelihoods(is), RTI)) then
                like%likelihoods(is)%like = out(1)
                like%likelihoods(is)%misfit = out(2)
                like%likelihoods(is)%unweighted_misfit = out(3)
            endif
        endif
        i = 1
        if (like_set%datatype == 3) i = 2
        call sum_parts(like, i)

    end subroutine noise_likelihood

    subroutine location_likelihood(dat, model, RTI, perturbed_box, like_set, like)
        implicit none
        type(T_DATA), dimension(:), intent(in)      :: dat
        type(T_MOD), intent(inout)                  :: model
        type(T_RUN_INFO), intent(inout), target     :: RTI
        type(d3), dimension(2), intent(in)          :: perturbed_box
        type(T_LIKE_SET), intent(in)                :: like_set
        type(T_LIKE), intent(inout), target         :: like
        real(c_double), target :: out(3)
        integer(c_int) :: rc

        if (like_set%datatype == 2) &
            call exception_raiseError('Error. Surface wave tomography cannot change source locations!')
        if (like_set%dynamic == 0) then
            ! body_location_likelihood traces again (src/likelihood_body.F90:80-99)
            call body_part(like%gpu(1), RTI, like%likelihoods(1))
        else
            ! the receivers' fields the context kept from its last likelihood call
            ! (src/likelihood_body.F90:100-149), interpolated at the new locations
            rc = mct_body_location_likelihood(like%gpu(1), c_loc(RTI%locations), c_loc(RTI%bnoise0), &
                                              c_loc(RTI%bnoise1), c_loc(out))
            if (.not. rejected(rc, like%likelihoods(1), RTI)) then
                like%likelihoods(1)%like = out(1)
                like%likelihoods(1)%misfit = out(2)
                like%likelihoods(1)%unweighted_misfit = out(3)
            endif
        endif
        if (like_set%datatype == 3) then
            call sum_parts(like, 2)
        else
            call sum_parts(like, 1)
        endif

    end subroutine location_likelihood

end module m_likelihood
