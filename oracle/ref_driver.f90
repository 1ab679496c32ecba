! ref_driver.f90 -- C-callable harness around the reference's own Fortran
! routines, compiled in place from /root/reference into oracle/_ref/.
! TEST INFRASTRUCTURE ONLY: it produces golden vectors that pin the C oracle.
! It calls the reference code exactly as the reference's callers do; it
! re-implements nothing.
module mct_ref_driver
    use iso_c_binding
    implicit none
contains

    ! kdtree_to_grid's use of kdtree2 (src/mcmc_loc2.f90:2030-2078) for a list
    ! of query points: q(3,nq) -> idx(nq), 1-based site numbers.
    subroutine ref_kdtree_nearest(ncells, points, nq, q, idx) bind(C, name="ref_kdtree_nearest")
        use kdtree2_precision_module, only : kdkind
        use kdtree2_module, only : kdtree2, kdtree2_result, kdtree2_create, &
            kdtree2_destroy, kdtree2_n_nearest
        integer(c_int), value :: ncells, nq
        real(c_double), intent(in), target :: points(3, ncells)
        real(c_double), intent(in) :: q(3, nq)
        integer(c_int), intent(out) :: idx(nq)
        type(kdtree2), pointer :: tree
        type(kdtree2_result), dimension(1) :: results
        real(kdkind), dimension(3) :: qv
        integer :: i
        tree => kdtree2_create(points(:, 1:ncells), sort=.false., rearrange=.true.)
        do i = 1, nq
            qv = q(:, i)
            call kdtree2_n_nearest(tp=tree, qv=qv, nn=1, results=results)
            idx(i) = results(1)%idx
        enddo
        call kdtree2_destroy(tree)
    end subroutine

    ! surfdisp96 exactly as surfmodes calls it (surfmodes/surfmodes.f90:79-95):
    ! REAL*4 layer model, flat earth, fundamental mode.
    subroutine ref_surfdisp96(nlayer, thkm, vpm, vsm, rhom, iwave, igr, kmax, t, dphase, &
                              cp, cg, ierr) bind(C, name="ref_surfdisp96")
        integer(c_int), value :: nlayer, iwave, igr, kmax
        real(c_float), intent(in) :: thkm(nlayer), vpm(nlayer), vsm(nlayer), rhom(nlayer)
        real(c_double), intent(in) :: t(kmax)
        real(c_double), value :: dphase
        real(c_double), intent(out) :: cp(kmax), cg(kmax)
        integer(c_int), intent(out) :: ierr
        real(4) :: d4(200), a4(200), b4(200), r4(200)
        real(8) :: tt(60)
        integer :: ie
        d4 = 0; a4 = 0; b4 = 0; r4 = 0; tt = 0
        d4(1:nlayer) = thkm; a4(1:nlayer) = vpm; b4(1:nlayer) = vsm; r4(1:nlayer) = rhom
        tt(1:kmax) = t
        call surfdisp96(d4, a4, b4, r4, nlayer, 0, iwave, 1, igr, kmax, tt, dphase, cp, cg, ie)
        ierr = ie
    end subroutine

    ! modrays (fm2d/fm2dray_cartesian.f90:67-480) for one period, phase times
    ! only (uar = 1, no ray tracing): velv(nvz+2,nvx+2) is like%vel(i,:,:).
    subroutine ref_modrays(nsrc, scx, scz, nrc, rcx, rcz, srs, nvx, nvz, goxd, gozd, dvxd, dvzd, &
                           velv, gdx, gdz, asgr, sgdl, sgs, fom, snb, ttime, crazy) &
                           bind(C, name="ref_modrays")
        use m_fm2d, only : modrays, T_RAY
        integer(c_int), value :: nsrc, nrc, nvx, nvz, gdx, gdz, asgr, sgdl, sgs, fom
        real(c_double), intent(in) :: scx(nsrc), scz(nsrc), rcx(nrc), rcz(nrc)
        integer(c_int), intent(in) :: srs(nrc*nsrc, 2)
        real(c_double), value :: goxd, gozd, dvxd, dvzd, snb
        real(c_double), intent(in) :: velv(nvz+2, nvx+2)
        real(c_double), intent(inout) :: ttime(nrc, nsrc)
        integer(c_int), intent(out) :: crazy
        type(T_RAY), allocatable :: rays(:)
        integer :: cr, srsl(nrc*nsrc, 2)
        real(8) :: vv(nvz+2, nvx+2)
        allocate(rays(nrc*nsrc))
        rays%srcid = 0
        rays%revid = 0
        rays%npoints = 0
        cr = 0
        srsl = srs
        vv = velv
        call modrays(nsrc, scx, scz, nrc, rcx, rcz, srsl, 0, nvx, nvz, goxd, gozd, dvxd, dvzd, vv, &
                     gdx, gdz, asgr, sgdl, sgs, 6371.0d0, fom, snb, ttime, rays, cr, 1)
        crazy = cr
    end subroutine

    ! modrays with ray tracing (uar = 0, as surf_likelihood calls it for group
    ! velocities, src/likelihood_surf.F90:245,311-320): phase times plus the
    ! rays rpaths stores at slot srs(:,2) (the pair index, as read_times_3
    ! sets dat%raystat(:,2,:)).  npts(slot) and pts(2, maxpts, slot) return the
    ! ray points (the first maxpts of a longer ray).
    subroutine ref_modrays_rays(nsrc, scx, scz, nrc, rcx, rcz, srs, nvx, nvz, goxd, gozd, dvxd, dvzd, &
                                velv, gdx, gdz, asgr, sgdl, sgs, fom, snb, ttime, maxpts, npts, pts, crazy) &
                                bind(C, name="ref_modrays_rays")
        use m_fm2d, only : modrays, T_RAY
        integer(c_int), value :: nsrc, nrc, nvx, nvz, gdx, gdz, asgr, sgdl, sgs, fom, maxpts
        real(c_double), intent(in) :: scx(nsrc), scz(nsrc), rcx(nrc), rcz(nrc)
        integer(c_int), intent(in) :: srs(nrc*nsrc, 2)
        real(c_double), value :: goxd, gozd, dvxd, dvzd, snb
        real(c_double), intent(in) :: velv(nvz+2, nvx+2)
        real(c_double), intent(inout) :: ttime(nrc, nsrc)
        integer(c_int), intent(out) :: npts(nrc*nsrc)
        real(c_double), intent(out) :: pts(2, maxpts, nrc*nsrc)
        integer(c_int), intent(out) :: crazy
        type(T_RAY), allocatable :: rays(:)
        integer :: cr, q, n, srsl(nrc*nsrc, 2)
        real(8) :: vv(nvz+2, nvx+2)
        allocate(rays(nrc*nsrc))
        rays%srcid = 0
        rays%revid = 0
        rays%npoints = 0
        cr = 0
        srsl = srs
        vv = velv
        call modrays(nsrc, scx, scz, nrc, rcx, rcz, srsl, 0, nvx, nvz, goxd, gozd, dvxd, dvzd, vv, &
                     gdx, gdz, asgr, sgdl, sgs, 6371.0d0, fom, snb, ttime, rays, cr, 0)
        crazy = cr
        pts = 0
        do q = 1, nrc*nsrc
            npts(q) = rays(q)%npoints
            if (rays(q)%npoints > 0 .and. allocated(rays(q)%points)) then
                n = min(rays(q)%npoints, maxpts)
                pts(:, 1:n, q) = rays(q)%points(:, 1:n)
            endif
        enddo
    end subroutine

    ! read_data (src/likelihood_settings.f90:166-196) for one data set:
    ! datatype 2 reads the surface-wave files, 0 the body-wave files.  Paths
    ! are NUL-terminated.  Sizes come back first (cap_* = 0), then the arrays:
    ! src (ndims, cap_src), rev (ndimr, cap_rev), freqs (cap_np),
    ! ttime (cap_pair, 3, cap_np), raystat (cap_pair, 2, cap_np).
    subroutine ref_read_data(datatype, fsrc, frev, fdat, cap_src, cap_rev, cap_np, cap_pair, &
                             sizes, src, rev, freqs, ttime, raystat) bind(C, name="ref_read_data")
        use like_settings, only : T_DATA, T_LIKE_SET, read_data
        integer(c_int), value :: datatype, cap_src, cap_rev, cap_np, cap_pair
        character(kind=c_char), intent(in) :: fsrc(*), frev(*), fdat(*)
        integer(c_int), intent(out) :: sizes(7)
        real(c_double), intent(out) :: src(*), rev(*), freqs(*), ttime(*)
        integer(c_int), intent(out) :: raystat(*)
        type(T_DATA) :: dat(1)
        type(T_LIKE_SET) :: ls
        integer :: ns, nr, np, npair, ds, dr, i, j, k, o
        ls%datatype = datatype
        if (datatype == 2) then
            ls%ssources_file = cstr(fsrc); ls%sreceivers_file = cstr(frev); ls%sdata_file = cstr(fdat)
        else
            ls%bsources_file = cstr(fsrc); ls%breceivers_file = cstr(frev); ls%bdata_file = cstr(fdat)
        endif
        call read_data(dat, ls)
        ds = size(dat(1)%src, 1); ns = size(dat(1)%src, 2)
        dr = size(dat(1)%rev, 1); nr = size(dat(1)%rev, 2)
        np = dat(1)%np; npair = ns * nr
        sizes = [ns, ds, nr, dr, np, dat(1)%nrays, dat(1)%nmodes]
        if (cap_src < ns .or. cap_rev < nr .or. cap_np < np .or. cap_pair < npair) return
        o = 0
        do i = 1, ns
            do j = 1, ds
                o = o + 1; src(o) = dat(1)%src(j, i)
            enddo
        enddo
        o = 0
        do i = 1, nr
            do j = 1, dr
                o = o + 1; rev(o) = dat(1)%rev(j, i)
            enddo
        enddo
        freqs(1:np) = dat(1)%freqs(1:np)
        do k = 1, np
            do j = 1, 3
                do i = 1, npair
                    ttime(i + (j - 1) * npair + (k - 1) * 3 * npair) = dat(1)%ttime(i, j, k)
                enddo
            enddo
            do j = 1, 2
                do i = 1, npair
                    raystat(i + (j - 1) * npair + (k - 1) * 2 * npair) = dat(1)%raystat(i, j, k)
                enddo
            enddo
        enddo
    end subroutine

    ! The reference's surface-wave forward stages composed as surf_likelihood
    ! composes them (src/likelihood_surf.F90:185-330), OpenMP over columns and
    ! over periods as there: surfmodes' surfdisp96 call per column on the layers
    ! convert_to_layer produced (lay(4,maxl,ncol) = thick, vp, vs, rho; nl(ncol)),
    ! the like%vel padding, modrays per period (uar = 1).  pvel(np,ncol) and
    ! ptime(nrev,nsrc,np); nbad counts columns with ierr /= 0 (then no modrays).
    subroutine ref_surf_forward(nthreads, ncol, maxl, nl, lay, np, freqs, iwave, dphase, nx, ny, &
                                xmin, ymin, dx, dy, nsrc, scx, scz, nrev, rcx, rcz, srs, gdx, gdz, asgr, &
                                sgdl, sgs, fom, snb, pvel, ptime, nbad) bind(C, name="ref_surf_forward")
        use m_fm2d, only : modrays, T_RAY
        use omp_lib
        integer(c_int), value :: nthreads, ncol, maxl, np, iwave, nx, ny, nsrc, nrev
        integer(c_int), value :: gdx, gdz, asgr, sgdl, sgs, fom
        integer(c_int), intent(in) :: nl(ncol), srs(nsrc*nrev, 2, np)
        real(c_double), intent(in) :: lay(4, maxl, ncol), freqs(np)
        real(c_double), value :: dphase, xmin, ymin, dx, dy, snb
        real(c_double), intent(in) :: scx(nsrc), scz(nsrc), rcx(nrev), rcz(nrev)
        real(c_double), intent(out) :: pvel(np, ncol), ptime(nrev, nsrc, np)
        integer(c_int), intent(out) :: nbad
        real(4) :: d4(200), a4(200), b4(200), r4(200)
        real(8) :: tt(60), cg(60)
        real(8), allocatable :: velv(:,:,:), vv(:,:)
        type(T_RAY), allocatable :: rays(:)
        integer :: c, n, ie, i, ix, iy, cr, srsl(nsrc*nrev, 2)
        if (nthreads > 0) call omp_set_num_threads(nthreads)
        tt = 0
        tt(1:np) = dble(1/freqs)
        nbad = 0
        !$omp parallel do private(c, n, d4, a4, b4, r4, cg, ie) reduction(+:nbad) schedule(dynamic, 16)
        do c = 1, ncol
            n = nl(c)
            d4 = 0; a4 = 0; b4 = 0; r4 = 0
            d4(1:n) = real(lay(1, 1:n, c), 4)
            a4(1:n) = real(lay(2, 1:n, c), 4)
            b4(1:n) = real(lay(3, 1:n, c), 4)
            r4(1:n) = real(lay(4, 1:n, c), 4)
            call surfdisp96(d4, a4, b4, r4, n, 0, iwave, 1, 0, np, tt, dphase, pvel(:, c), cg, ie)
            if (ie /= 0) nbad = nbad + 1
        enddo
        !$omp end parallel do
        if (nbad > 0) return
        allocate(velv(np, ny + 2, nx + 2))
        do ix = 1, nx + 2
            do iy = 1, ny + 2
                velv(:, iy, ix) = pvel(:, min(max(iy - 1, 1), ny) + ny * (min(max(ix - 1, 1), nx) - 1))
            enddo
        enddo
        !$omp parallel do private(i, vv, rays, cr, srsl) schedule(dynamic, 1)
        do i = 1, np
            allocate(vv(ny + 2, nx + 2), rays(nsrc * nrev))
            vv = velv(i, :, :)
            rays%srcid = 0
            rays%revid = 0
            rays%npoints = 0
            cr = 0
            srsl = srs(:, :, i)
            call modrays(nsrc, scx, scz, nrev, rcx, rcz, srsl, 0, nx, ny, xmin, ymin, dx, dy, vv, &
                         gdx, gdz, asgr, sgdl, sgs, 6371.0d0, fom, snb, ptime(:, :, i), rays, cr, 1)
            deallocate(vv, rays)
        enddo
        !$omp end parallel do
        deallocate(velv)
    end subroutine

    function cstr(c) result(f)
        character(kind=c_char), intent(in) :: c(*)
        character(len=100) :: f
        integer :: i
        f = ' '
        do i = 1, 100
            if (c(i) == c_null_char) exit
            f(i:i) = c(i)
        enddo
    end function

end module mct_ref_driver
