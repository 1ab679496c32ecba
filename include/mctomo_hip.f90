! mctomo_hip.f90 -- Fortran iso_c_binding interfaces to libmctomo_hip.so.
!
! The declarations mirror include/mctomo_hip.h one to one.  Arrays are passed
! as type(c_ptr) from c_loc of the Fortran-owned arrays, as
! src/fastMarching_wrapper.f90:51-74 does for fastMarching3D; scalars by value.
module mctomo_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MCT_OK = 0, MCT_REJECT_MODEL = 1, MCT_REJECT_RAY = 2
    integer(c_int), parameter :: MCT_ERR_ARG = -1, MCT_ERR_HIP = -2, MCT_ERR_NOMEM = -3
    integer(c_int), parameter :: MCT_ERR_SIGMA = -4, MCT_ERR_NOCONV = -5, MCT_ERR_OUTSIDE = -6
    integer(c_int), parameter :: MCT_ERR_IO = -7

    ! T_GRID (src/settings.f90:20-28), zmin and dz scaled
    type, bind(C) :: mct_grid
        integer(c_int) :: nx, ny, nz
        real(c_double) :: xmin, ymin, zmin
        real(c_double) :: dx, dy, dz
        real(c_double) :: scaling
    end type

    type, bind(C) :: mct_body_setup
        integer(c_int) :: datatype, order, sigdep, nsrc, nrev
        type(c_ptr) :: rev, ttime, raystat
        integer(c_int) :: nrays, max_cells, max_chains
        integer(c_int) :: dynamic
    end type

    type, bind(C) :: mct_surf_setup
        integer(c_int) :: raylov, phase_group
        real(c_double) :: dphase
        integer(c_int) :: np
        type(c_ptr) :: freqs
        integer(c_int) :: nsrc, nrev
        type(c_ptr) :: src, rev, ttime, raystat
        integer(c_int) :: sigdep, nrays, gridx, gridy, sgref, sgdic, sgext, order
        real(c_double) :: band
        integer(c_int) :: max_cells, max_chains
        integer(c_int) :: straight
        integer(c_int) :: datatype
        real(c_double) :: water_depth
    end type

    interface

        function mct_version() bind(C, name="mct_version")
            import
            type(c_ptr) :: mct_version
        end function

        ! kdtree_to_grid (src/mcmc_loc2.f90:2002-2082)
        function mct_voronoi_to_grid(grid, points, params, ncells, box, pm, sites_id, vp, vs, rho) &
                bind(C, name="mct_voronoi_to_grid")
            import
            type(mct_grid), intent(in) :: grid
            type(c_ptr), value :: points, params
            integer(c_int), value :: ncells
            type(c_ptr), value :: box, pm, sites_id, vp, vs, rho
            integer(c_int) :: mct_voronoi_to_grid
        end function

        ! surfdisp96 as surfmodes calls it (surfmodes/surfmodes.f90:79-95)
        function mct_surfdisp96_batch(nmod, nl_max, nlayer, layers, np, periods, iwave, igr, dphase, &
                phase, group, ierr) bind(C, name="mct_surfdisp96_batch")
            import
            integer(c_int), value :: nmod, nl_max
            type(c_ptr), value :: nlayer, layers
            integer(c_int), value :: np
            type(c_ptr), value :: periods
            integer(c_int), value :: iwave, igr
            real(c_double), value :: dphase
            type(c_ptr), value :: phase, group, ierr
            integer(c_int) :: mct_surfdisp96_batch
        end function

        ! modrays' phase-time path (fm2d/fm2dray_cartesian.f90:67-480)
        function mct_modrays(nvx, nvz, goxd, gozd, dvxd, dvzd, np, velv, gdx, gdz, asgr, sgdl, sgs, &
                fom, nsrc, scx, scz, nrc, rcx, rcz, srs, ttime, method) bind(C, name="mct_modrays")
            import
            integer(c_int), value :: nvx, nvz
            real(c_double), value :: goxd, gozd, dvxd, dvzd
            integer(c_int), value :: np
            type(c_ptr), value :: velv
            integer(c_int), value :: gdx, gdz, asgr, sgdl, sgs, fom, nsrc
            type(c_ptr), value :: scx, scz
            integer(c_int), value :: nrc
            type(c_ptr), value :: rcx, rcz, srs, ttime
            integer(c_int), value :: method
            integer(c_int) :: mct_modrays
        end function

        ! fastMarching3D keeps the reference's own binding
        ! (src/fastMarching_wrapper.f90:18-29); it is exported under that name.

        function mct_body_create(grid, setup, ctx) bind(C, name="mct_body_create")
            import
            type(mct_grid), intent(in) :: grid
            type(mct_body_setup), intent(in) :: setup
            type(c_ptr) :: ctx
            integer(c_int) :: mct_body_create
        end function

        function mct_destroy(ctx) bind(C, name="mct_destroy")
            import
            type(c_ptr), value :: ctx
            integer(c_int) :: mct_destroy
        end function

        function mct_body_likelihood(ctx, points, params, ncells, locations, bnoise0, bnoise1, out, &
                phase_time, srdist, sigma) bind(C, name="mct_body_likelihood")
            import
            type(c_ptr), value :: ctx, points, params
            integer(c_int), value :: ncells
            type(c_ptr), value :: locations, bnoise0, bnoise1, out, phase_time, srdist, sigma
            integer(c_int) :: mct_body_likelihood
        end function

        function mct_surf_create(grid, setup, ctx) bind(C, name="mct_surf_create")
            import
            type(mct_grid), intent(in) :: grid
            type(mct_surf_setup), intent(in) :: setup
            type(c_ptr) :: ctx
            integer(c_int) :: mct_surf_create
        end function

        function mct_surf_destroy(ctx) bind(C, name="mct_surf_destroy")
            import
            type(c_ptr), value :: ctx
            integer(c_int) :: mct_surf_destroy
        end function

        function mct_surf_likelihood(ctx, points, params, ncells, snoise0, snoise1, out, phase_time, &
                srdist, sigma, phase_vel) bind(C, name="mct_surf_likelihood")
            import
            type(c_ptr), value :: ctx, points, params
            integer(c_int), value :: ncells
            type(c_ptr), value :: snoise0, snoise1, out, phase_time, srdist, sigma, phase_vel
            integer(c_int) :: mct_surf_likelihood
        end function

        ! likelihood with a perturbed box (src/likelihood_surf.F90:155-270)
        function mct_surf_likelihood_box(ctx, points, params, ncells, box, snoise0, snoise1, out, &
                phase_time, srdist, sigma, phase_vel) bind(C, name="mct_surf_likelihood_box")
            import
            type(c_ptr), value :: ctx, points, params
            integer(c_int), value :: ncells
            type(c_ptr), value :: box, snoise0, snoise1, out, phase_time, srdist, sigma, phase_vel
            integer(c_int) :: mct_surf_likelihood_box
        end function

        ! the sampler kept the last proposal (src/mcmc_loc2.f90:549-567)
        function mct_surf_accept(ctx, chain) bind(C, name="mct_surf_accept")
            import
            type(c_ptr), value :: ctx
            integer(c_int), value :: chain
            integer(c_int) :: mct_surf_accept
        end function

        ! noise_likelihood (src/likelihood.f90:131-159)
        function mct_body_noise_likelihood(ctx, locations, bnoise0, bnoise1, out) &
                bind(C, name="mct_body_noise_likelihood")
            import
            type(c_ptr), value :: ctx, locations, bnoise0, bnoise1, out
            integer(c_int) :: mct_body_noise_likelihood
        end function

        ! location_likelihood with dynamic /= 0 (src/likelihood.f90:162-191)
        function mct_body_location_likelihood(ctx, locations, bnoise0, bnoise1, out) &
                bind(C, name="mct_body_location_likelihood")
            import
            type(c_ptr), value :: ctx, locations, bnoise0, bnoise1, out
            integer(c_int) :: mct_body_location_likelihood
        end function

        function mct_surf_noise_likelihood(ctx, snoise0, snoise1, out) &
                bind(C, name="mct_surf_noise_likelihood")
            import
            type(c_ptr), value :: ctx, snoise0, snoise1, out
            integer(c_int) :: mct_surf_noise_likelihood
        end function

        ! noise_likelihood over the caller's like%phaseTime (and like%srdist)
        function mct_body_noise_likelihood_from(ctx, phase_time, locations, bnoise0, bnoise1, out, sigma) &
                bind(C, name="mct_body_noise_likelihood_from")
            import
            type(c_ptr), value :: ctx, phase_time, locations, bnoise0, bnoise1, out, sigma
            integer(c_int) :: mct_body_noise_likelihood_from
        end function

        function mct_surf_noise_likelihood_from(ctx, phase_time, srdist, snoise0, snoise1, out, sigma) &
                bind(C, name="mct_surf_noise_likelihood_from")
            import
            type(c_ptr), value :: ctx, phase_time, srdist, snoise0, snoise1, out, sigma
            integer(c_int) :: mct_surf_noise_likelihood_from
        end function

        ! datatype 3: the joint body + surface context (src/likelihood.f90:83-88)
        function mct_joint_create(grid, body, surf, ctx) bind(C, name="mct_joint_create")
            import
            type(mct_grid), intent(in) :: grid
            type(mct_body_setup), intent(in) :: body
            type(mct_surf_setup), intent(in) :: surf
            type(c_ptr) :: ctx
            integer(c_int) :: mct_joint_create
        end function

        function mct_joint_destroy(ctx) bind(C, name="mct_joint_destroy")
            import
            type(c_ptr), value :: ctx
            integer(c_int) :: mct_joint_destroy
        end function

        function mct_joint_parts(ctx, body, surf) bind(C, name="mct_joint_parts")
            import
            type(c_ptr), value :: ctx
            type(c_ptr) :: body, surf
            integer(c_int) :: mct_joint_parts
        end function

        function mct_joint_likelihood(ctx, points, params, ncells, locations, bnoise0, bnoise1, snoise0, &
                snoise1, out, out_parts) bind(C, name="mct_joint_likelihood")
            import
            type(c_ptr), value :: ctx, points, params
            integer(c_int), value :: ncells
            type(c_ptr), value :: locations, bnoise0, bnoise1, snoise0, snoise1, out, out_parts
            integer(c_int) :: mct_joint_likelihood
        end function

        ! data files: read_sources / read_times_3 (src/likelihood_settings.f90:198-420),
        ! write_sources, write_atimes (src/read_write.f90:484-530); paths are
        ! NUL-terminated character(kind=c_char) strings
        function mct_read_sources(path, bnd, nbnd, cap, n, ndim, src) bind(C, name="mct_read_sources")
            import
            character(kind=c_char), dimension(*) :: path
            type(c_ptr), value :: bnd, src
            integer(c_int), value :: nbnd
            integer(c_long), value :: cap
            integer(c_int) :: n, ndim
            integer(c_int) :: mct_read_sources
        end function

        function mct_read_times(path, nsrc, nrev, cap_np, np, freqs, ttime, raystat, nrays) &
                bind(C, name="mct_read_times")
            import
            character(kind=c_char), dimension(*) :: path
            integer(c_int), value :: nsrc, nrev, cap_np
            integer(c_int) :: np, nrays
            type(c_ptr), value :: freqs, ttime, raystat
            integer(c_int) :: mct_read_times
        end function

        function mct_write_sources(path, n, ndim, src) bind(C, name="mct_write_sources")
            import
            character(kind=c_char), dimension(*) :: path
            integer(c_int), value :: n, ndim
            type(c_ptr), value :: src
            integer(c_int) :: mct_write_sources
        end function

        function mct_write_times(path, np, freqs, nsrc, nrev, raystat, phase_time, sigma, locations) &
                bind(C, name="mct_write_times")
            import
            character(kind=c_char), dimension(*) :: path
            integer(c_int), value :: np, nsrc, nrev
            type(c_ptr), value :: freqs, raystat, phase_time, sigma, locations
            integer(c_int) :: mct_write_times
        end function

        ! like%gvel after a call with phase_group = 1 (src/likelihood_surf.F90:227-229)
        function mct_surf_group_velocities(ctx, chain, group_vel) &
                bind(C, name="mct_surf_group_velocities")
            import
            type(c_ptr), value :: ctx, group_vel
            integer(c_int), value :: chain
            integer(c_int) :: mct_surf_group_velocities
        end function

    end interface

end module mctomo_hip
