! A sampler-side caller of the m_likelihood drop-in: reads one chain's data,
! model and noise from a list-directed file, calls like_setup, likelihood,
! noise_likelihood (with the noise changed) and, for body data,
! location_likelihood (with the locations moved), as src/mcmc_loc2.f90 does,
! and prints the results.  Built and run by tests/test_gpu_fortran_likelihood.py.
program mlike_driver
    use iso_c_binding
    use m_utils, only : ii10
    use m_settings, only : T_MOD
    use like_settings, only : T_DATA, T_LIKE_SET
    use run_info, only : T_RUN_INFO
    use cgal_delaunay, only : d3
    use m_likelihood
    implicit none
    type(T_DATA), allocatable :: dat(:)
    type(T_LIKE_SET) :: ls
    type(T_RUN_INFO) :: RTI
    type(T_MOD) :: model
    type(T_LIKE) :: like, like_copy
    type(d3) :: box(2)
    integer :: u, ndat, i, n, nd, ncells, nloc, nb, ns
    character(len=512) :: path

    call get_command_argument(1, path)
    open(newunit=u, file=trim(path), status='old')
    read(u, *) ls%datatype, ndat
    read(u, *) ls%grid%nx, ls%grid%ny, ls%grid%nz, ls%grid%xmin, ls%grid%xmax, ls%grid%ymin, ls%grid%ymax, &
               ls%grid%zmin, ls%grid%zmax, ls%grid%dx, ls%grid%dy, ls%grid%dz, ls%grid%waterDepth, ls%grid%scaling
    read(u, *) ls%sigdep, ls%isStraight, ls%raylov, ls%phaseGroup, ls%gridx, ls%gridy, ls%sgref, ls%sgdic, &
               ls%sgext, ls%order, ls%dynamic
    read(u, *) ls%dPhaseVel, ls%band
    read(u, *) ncells
    RTI%ncells = ncells
    allocate(RTI%points(3, ncells), RTI%parameters(3, ncells))
    read(u, *) RTI%points
    read(u, *) RTI%parameters
    allocate(dat(ndat))
    do i = 1, ndat
        read(u, *) dat(i)%datatype, dat(i)%np, dat(i)%nsrc, dat(i)%nrev, dat(i)%nrays, nd
        n = dat(i)%nsrc * dat(i)%nrev
        allocate(dat(i)%src(nd, dat(i)%nsrc), dat(i)%rev(3, dat(i)%nrev), dat(i)%freqs(dat(i)%np))
        allocate(dat(i)%ttime(n, 3, dat(i)%np), dat(i)%raystat(n, 2, dat(i)%np))
        read(u, *) dat(i)%src
        read(u, *) dat(i)%rev
        read(u, *) dat(i)%freqs
        read(u, *) dat(i)%ttime
        read(u, *) dat(i)%raystat
    enddo
    read(u, *) nloc
    allocate(RTI%locations(4, max(nloc, 1)))
    RTI%locations = 0
    if (nloc > 0) read(u, *) RTI%locations
    read(u, *) nb
    allocate(RTI%bnoise0(max(nb, 1)), RTI%bnoise1(max(nb, 1)))
    RTI%bnoise0 = 0
    RTI%bnoise1 = 0
    if (nb > 0) read(u, *) RTI%bnoise0, RTI%bnoise1
    read(u, *) ns
    allocate(RTI%snoise0(max(ns, 1)), RTI%snoise1(max(ns, 1)))
    RTI%snoise0 = 0
    RTI%snoise1 = 0
    if (ns > 0) read(u, *) RTI%snoise0, RTI%snoise1
    close(u)
    RTI%num_bad_model = 0
    RTI%num_bad_ray = 0

    call like_setup(like, dat, ls, ncells)
    call likelihood(dat, model, RTI, box, ls, like)
    write(*, '(a,3es25.16)') 'LIKE ', like%like, like%misfit, like%unweighted_misfit
    do i = 1, size(like%likelihoods)
        write(*, '(a,i0,3es25.16)') 'PART ', i, like%likelihoods(i)%like, like%likelihoods(i)%misfit, &
            like%likelihoods(i)%unweighted_misfit
    enddo
    ! the sampler keeps like_copy; a noise proposal re-evaluates the misfit
    like_copy = like
    RTI%bnoise0 = RTI%bnoise0 * 1.5d0 + 0.01d0
    RTI%bnoise1 = RTI%bnoise1 * 1.5d0 + 0.01d0
    RTI%snoise0 = RTI%snoise0 * 1.5d0 + 0.01d0
    RTI%snoise1 = RTI%snoise1 * 1.5d0 + 0.01d0
    call noise_likelihood(dat, RTI, ls, like_copy)
    write(*, '(a,3es25.16)') 'NOISE ', like_copy%like, like_copy%misfit, like_copy%unweighted_misfit
    if (ls%datatype /= 2) then
        RTI%locations(1:3, :) = RTI%locations(1:3, :) + 0.05d0
        RTI%locations(4, :) = RTI%locations(4, :) + 0.02d0
        call location_likelihood(dat, model, RTI, box, ls, like_copy)
        write(*, '(a,3es25.16)') 'LOCATION ', like_copy%like, like_copy%misfit, like_copy%unweighted_misfit
    endif
    write(*, '(a,2i8)') 'BAD ', RTI%num_bad_model, RTI%num_bad_ray
end program mlike_driver
