"""ctypes bindings to oracle/liboracle.so (C restatement of the reference) and,
when built, oracle/_ref/libmctref.so (the reference's own Fortran, compiled in
place).  Test infrastructure only."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_D = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load liboracle.so (build it with `make -C oracle`)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run make -C oracle")
        _lib = ctypes.CDLL(path)
    return _lib


def ref():
    """Load oracle/_ref/libmctref.so, or return None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libmctref.so")
        if not os.path.exists(path):
            return None
        _ref = ctypes.CDLL(path)
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- nearest site
def kdtree_to_grid(grid, points, params, box=None, pm=None, vp=None, vs=None, rho=None):
    """grid: dict nx,ny,nz,xmin,ymin,zmin,dx,dy,dz (z already scaled).
    points, params: (ncells,3) arrays (Fortran (3,ncells) storage).
    Returns sites_id, vp, vs, rho as (nx,ny,nz) C arrays == Fortran (nz,ny,nx)."""
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    points = c64(points)
    params = c64(params)
    sid = np.zeros((nx, ny, nz), np.int32)
    vp = np.ones((nx, ny, nz)) if vp is None else c64(vp).copy()
    vs = np.ones((nx, ny, nz)) if vs is None else c64(vs).copy()
    rho = np.ones((nx, ny, nz)) if rho is None else c64(rho).copy()
    L = lib()
    L.orc_kdtree_to_grid.restype = ctypes.c_int
    rc = L.orc_kdtree_to_grid(
        ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
        ctypes.c_double(grid["xmin"]), ctypes.c_double(grid["ymin"]), ctypes.c_double(grid["zmin"]),
        ctypes.c_double(grid["dx"]), ctypes.c_double(grid["dy"]), ctypes.c_double(grid["dz"]),
        _p(points), _p(params), ctypes.c_int(points.shape[0]),
        _p(None if box is None else c64(box)), _p(None if pm is None else c64(pm)),
        _p(sid), _p(vp), _p(vs), _p(rho))
    assert rc == 0
    return sid, vp, vs, rho


def kdtree_nearest(points, q):
    points = c64(points)
    q = c64(q)
    return lib().orc_kdtree_nearest(_p(points), ctypes.c_int(points.shape[0]), _p(q))


def brute_nearest(points, q):
    points = c64(points)
    q = c64(q)
    sd = ctypes.c_double()
    n = ctypes.c_int()
    i = lib().orc_brute_nearest(_p(points), ctypes.c_int(points.shape[0]), _p(q),
                                ctypes.byref(sd), ctypes.byref(n))
    return i, sd.value, n.value


def ref_kdtree_nearest(points, queries):
    """Reference kdtree2 (compiled Fortran) 1-NN for each query row."""
    R = ref()
    if R is None:
        return None
    points = c64(points)
    queries = c64(queries)
    out = np.zeros(queries.shape[0], np.int32)
    R.ref_kdtree_nearest(ctypes.c_int(points.shape[0]), _p(points),
                         ctypes.c_int(queries.shape[0]), _p(queries), _p(out))
    return out


# ---------------------------------------------------------------- 3-D eikonal
def fast_marching_3d(vel, dx, dy, dz, xmin, ymin, zmin, src, rev, order, field=False):
    """vel: (nx,ny,nz) C array (== Fortran (nz,ny,nx)). rev: (nrev,3).
    Returns ttime (nrev,) and, if field, the (nx,ny,nz) traveltime field."""
    vel = c64(vel)
    nx, ny, nz = vel.shape
    rev = c64(rev).reshape(-1, 3)
    tt = np.zeros(rev.shape[0])
    tf = np.zeros_like(vel) if field else None
    L = lib()
    L.orc_fastMarching3D(ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
                         ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(dz),
                         ctypes.c_double(xmin), ctypes.c_double(ymin), ctypes.c_double(zmin),
                         _p(vel), _p(c64(src)), _p(rev), ctypes.c_int(rev.shape[0]),
                         ctypes.c_int(order), _p(tt), _p(tf))
    return (tt, tf) if field else tt


def body_traveltimes(vel, grid, locations, rev, order, nthreads=0):
    """All sources (locations (nsrc,4)) -> phase_time (nsrc,nrev) [Fortran (nrev,nsrc)]."""
    vel = c64(vel)
    nx, ny, nz = vel.shape
    loc = c64(locations)
    rev = c64(rev)
    out = np.zeros((loc.shape[0], rev.shape[0]))
    lib().orc_body_traveltimes(
        ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
        ctypes.c_double(grid["dx"]), ctypes.c_double(grid["dy"]), ctypes.c_double(grid["dz"]),
        ctypes.c_double(grid["xmin"]), ctypes.c_double(grid["ymin"]), ctypes.c_double(grid["zmin"]),
        ctypes.c_double(grid.get("scaling", 1.0)), _p(vel), ctypes.c_int(loc.shape[0]), _p(loc),
        ctypes.c_int(rev.shape[0]), _p(rev), ctypes.c_int(order), _p(out), ctypes.c_int(nthreads))
    return out


def body_misfit(phase_time, ttime, raystat, locations, sigdep=0, bnoise0=None, bnoise1=None,
                nrays=None):
    """phase_time (np,nsrc,nrev); ttime (np,3,npair); raystat (np,2,npair) int32;
    locations (nsrc,4).  Returns (like, misfit, unweighted), srdist, sigma, rc."""
    pt = c64(phase_time)
    npp, nsrc, nrev = pt.shape
    tt = c64(ttime)
    rs = np.ascontiguousarray(raystat, dtype=np.int32)
    loc = c64(locations)
    b0 = c64(np.zeros(npp) if bnoise0 is None else bnoise0)
    b1 = c64(np.zeros(npp) if bnoise1 is None else bnoise1)
    if nrays is None:
        nrays = int((rs[:, 0, :] == 1).sum())
    srd = np.zeros((npp, nsrc * nrev))
    sig = np.zeros((npp, nsrc * nrev))
    out = np.zeros(3)
    rc = lib().orc_body_misfit(ctypes.c_int(nsrc), ctypes.c_int(nrev), ctypes.c_int(npp), _p(pt),
                               _p(tt), _p(rs), _p(loc), ctypes.c_int(sigdep), _p(b0), _p(b1),
                               ctypes.c_int(nrays), _p(srd), _p(sig), _p(out))
    return out, srd, sig, rc


def interpolate3d(grid, field, point):
    """Interpolate3d (src/likelihood_body.F90:976-1007).  grid: dict with zmin/dz
    scaled and scaling; field (nx,ny,nz) C array; point (3,) unscaled."""
    f = c64(field)
    nx, ny, nz = f.shape
    L = lib()
    L.orc_interpolate3d.restype = ctypes.c_double
    return L.orc_interpolate3d(ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
                               ctypes.c_double(grid["xmin"]), ctypes.c_double(grid["ymin"]),
                               ctypes.c_double(grid["zmin"]), ctypes.c_double(grid["dx"]),
                               ctypes.c_double(grid["dy"]), ctypes.c_double(grid["dz"]),
                               ctypes.c_double(grid.get("scaling", 1.0)), _p(f), _p(c64(point)))


# ---------------------------------------------------------------- dispersion
_F = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def surfdisp96(thk, vp, vs, rho, periods, iwave, igr, dphase=1e-3, ref=False):
    """surfdisp96 on a REAL*4 layer model.  Returns cp, cg, ierr, n_dltar
    (n_dltar is None for the reference build)."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    thk, vp, vs, rho = f32(thk), f32(vp), f32(vs), f32(rho)
    t = c64(periods)
    n, k = thk.shape[0], t.shape[0]
    cp = np.zeros(k)
    cg = np.zeros(k)
    if ref:
        R = globals()["ref"]()
        if R is None:
            return None
        ierr = ctypes.c_int()
        R.ref_surfdisp96(ctypes.c_int(n), _p(thk), _p(vp), _p(vs), _p(rho), ctypes.c_int(iwave),
                         ctypes.c_int(igr), ctypes.c_int(k), _p(t), ctypes.c_double(dphase),
                         _p(cp), _p(cg), ctypes.byref(ierr))
        return cp, cg, ierr.value, None
    nd = ctypes.c_long()
    ierr = lib().orc_surfdisp96(_p(thk), _p(vp), _p(vs), _p(rho), ctypes.c_int(n), ctypes.c_int(iwave),
                                ctypes.c_int(igr), ctypes.c_int(k), _p(t), ctypes.c_double(dphase),
                                _p(cp), _p(cg), ctypes.byref(nd))
    return cp, cg, ierr, nd.value


# ---------------------------------------------------------------- 2-D eikonal
def fm2d(velv, goxd, gozd, dvxd, dvzd, scx, scz, rcx, rcz, srs=None, gdx=1, gdz=1, asgr=1,
         sgdl=4, sgs=8, fom=1, snb=0.5, ttime=None, field=False, ref=False):
    """modrays phase times for one period.  velv: (nvx+2, nvy+2) C array
    (= Fortran velv(nvz+2,nvx+2)).  Returns ttime (nsrc, nrc) [, field]."""
    velv = c64(velv)
    nvx, nvz = velv.shape[0] - 2, velv.shape[1] - 2
    scx, scz, rcx, rcz = c64(scx), c64(scz), c64(rcx), c64(rcz)
    nsrc, nrc = scx.shape[0], rcx.shape[0]
    if srs is None:
        srs = np.ones(nsrc * nrc, np.int32)
    srs = np.ascontiguousarray(srs, dtype=np.int32).reshape(-1)
    tt = np.zeros((nsrc, nrc)) if ttime is None else c64(ttime).copy()
    if ref:
        R = globals()["ref"]()
        if R is None:
            return None
        srs2 = np.ascontiguousarray(np.concatenate([srs, srs]), dtype=np.int32)
        cr = ctypes.c_int()
        R.ref_modrays(nsrc, _p(scx), _p(scz), nrc, _p(rcx), _p(rcz), _p(srs2), nvx, nvz,
                      ctypes.c_double(goxd), ctypes.c_double(gozd), ctypes.c_double(dvxd),
                      ctypes.c_double(dvzd), _p(velv), gdx, gdz, asgr, sgdl, sgs, fom,
                      ctypes.c_double(snb), _p(tt), ctypes.byref(cr))
        return tt
    nnx, nnz = (nvx - 1) * gdx + 1, (nvz - 1) * gdz + 1
    tf = np.zeros((nsrc, nnx, nnz)) if field else None
    rc = lib().orc_fm2d(nvx, nvz, ctypes.c_double(goxd), ctypes.c_double(gozd),
                        ctypes.c_double(dvxd), ctypes.c_double(dvzd), _p(velv), gdx, gdz, asgr,
                        sgdl, sgs, fom, ctypes.c_double(snb), nsrc, _p(scx), _p(scz), nrc,
                        _p(rcx), _p(rcz), _p(srs), _p(tt), _p(tf))
    if rc:
        raise RuntimeError("orc_fm2d status %d" % rc)
    return (tt, tf) if field else tt


def fm2d_rays(velv, goxd, gozd, dvxd, dvzd, scx, scz, rcx, rcz, srs=None, gvel=None, gdx=1, gdz=1,
              asgr=1, sgdl=4, sgs=8, fom=1, snb=0.5, maxpts=4096, ref=False):
    """modrays with ray tracing (uar = 0) for one period.  gvel: (nvx, nvy) C
    array (= Fortran gvel(nvz,nvx)) or None.  Returns a dict: tt (nsrc, nrc)
    phase times, npts (nsrc*nrc), pts (nsrc*nrc, maxpts, 2), crazy, and from
    the oracle also rlen / gtime (nsrc, nrc).  ref=True runs the reference's own
    modrays/rpaths (oracle/_ref) instead (no rlen / gtime)."""
    velv = c64(velv)
    nvx, nvz = velv.shape[0] - 2, velv.shape[1] - 2
    scx, scz, rcx, rcz = c64(scx), c64(scz), c64(rcx), c64(rcz)
    nsrc, nrc = scx.shape[0], rcx.shape[0]
    if srs is None:
        srs = np.ones(nsrc * nrc, np.int32)
    srs = np.ascontiguousarray(srs, dtype=np.int32).reshape(-1)
    tt = np.zeros((nsrc, nrc))
    npts = np.zeros(nsrc * nrc, np.int32)
    pts = np.zeros((nsrc * nrc, maxpts, 2))
    cr = ctypes.c_int()
    if ref:
        R = globals()["ref"]()
        if R is None:
            return None
        slot = np.where(srs != 0, np.arange(1, nsrc * nrc + 1), 0).astype(np.int32)
        srs2 = np.ascontiguousarray(np.concatenate([srs, slot]), dtype=np.int32)
        R.ref_modrays_rays(nsrc, _p(scx), _p(scz), nrc, _p(rcx), _p(rcz), _p(srs2), nvx, nvz,
                           ctypes.c_double(goxd), ctypes.c_double(gozd), ctypes.c_double(dvxd),
                           ctypes.c_double(dvzd), _p(velv), gdx, gdz, asgr, sgdl, sgs, fom,
                           ctypes.c_double(snb), _p(tt), maxpts, _p(npts), _p(pts), ctypes.byref(cr))
        return dict(tt=tt, npts=npts, pts=pts, crazy=cr.value)
    gv = None if gvel is None else c64(gvel)
    rlen = np.zeros((nsrc, nrc))
    gt = np.zeros((nsrc, nrc))
    rc = lib().orc_fm2d_rays(nvx, nvz, ctypes.c_double(goxd), ctypes.c_double(gozd),
                             ctypes.c_double(dvxd), ctypes.c_double(dvzd), _p(velv), gdx, gdz, asgr,
                             sgdl, sgs, fom, ctypes.c_double(snb), nsrc, _p(scx), _p(scz), nrc,
                             _p(rcx), _p(rcz), _p(srs), _p(tt), None, 1, _p(gv), _p(gt), _p(rlen),
                             ctypes.byref(cr), maxpts, _p(npts), _p(pts))
    if rc:
        raise RuntimeError("orc_fm2d_rays status %d" % rc)
    return dict(tt=tt, npts=npts, pts=pts, crazy=cr.value, rlen=rlen, gtime=gt)


# ---------------------------------------------------------------- data files
def ref_read_data(datatype, sources_file, receivers_file, data_file):
    """read_data (src/likelihood_settings.f90:166-196) from the reference's own
    compiled Fortran: dict src (n,ndim), rev, freqs, ttime (np,3,npair), raystat
    (np,2,npair), nrays, nmodes; None where oracle/_ref is not built."""
    R = globals()["ref"]()
    if R is None:
        return None
    f = [os.fsencode(os.fspath(x)) + b"\0" for x in (sources_file, receivers_file, data_file)]
    sz = np.zeros(7, np.int32)
    z = np.zeros(1)
    zi = np.zeros(1, np.int32)
    R.ref_read_data(int(datatype), f[0], f[1], f[2], 0, 0, 0, 0, _p(sz), _p(z), _p(z), _p(z), _p(z), _p(zi))
    ns, ds, nr, dr, npp, nrays, nmodes = (int(v) for v in sz)
    npair = ns * nr
    src, rev, fr = np.zeros((ns, ds)), np.zeros((nr, dr)), np.zeros(npp)
    tt, rs = np.zeros((npp, 3, npair)), np.zeros((npp, 2, npair), np.int32)
    R.ref_read_data(int(datatype), f[0], f[1], f[2], ns, nr, npp, npair, _p(sz), _p(src), _p(rev), _p(fr),
                    _p(tt), _p(rs))
    return dict(src=src, rev=rev, freqs=fr, ttime=tt, raystat=rs, nrays=nrays, nmodes=nmodes)


def ref_surf_forward(grid, points, params, freqs, src, rev, raystat, raylov=1, dphase=1e-3, gridx=1,
                     gridy=1, sgref=1, sgdic=4, sgext=8, order=1, band=0.5, nthreads=0):
    """The surface-wave forward model with the reference's own compiled stages
    doing the work: nearest site and convert_to_layer from the oracle, then
    surfdisp96 per column and modrays per period from oracle/_ref (OpenMP over
    columns and over periods, as surf_likelihood).  Returns (nbad, pvel
    (nx,ny,np), phase_time (np,nsrc,nrev)) or None where oracle/_ref is not built."""
    R = globals()["ref"]()
    if R is None:
        return None
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    _, _, vs, _ = kdtree_to_grid(grid, points, params)
    ncol = nx * ny
    maxl = nz
    lay = np.zeros((ncol, maxl, 4))
    nl = np.zeros(ncol, np.int32)
    if lib().orc_column_layers_all(ctypes.c_long(ncol), nz, _p(c64(vs)), ctypes.c_double(grid["dz"]),
                                   ctypes.c_double(grid.get("scaling", 1.0)), ctypes.c_double(0.0), maxl,
                                   _p(lay), _p(nl)):
        raise RuntimeError("layer count above nz")
    fr = c64(freqs).reshape(-1)
    npp = fr.shape[0]
    s, r = c64(src).reshape(-1, 2), c64(rev).reshape(-1, 2)
    nsrc, nrev = s.shape[0], r.shape[0]
    rs = np.ascontiguousarray(raystat, dtype=np.int32).reshape(npp, 2, nsrc * nrev)
    pv = np.zeros((ncol, npp))
    pt = np.zeros((npp, nsrc, nrev))
    nbad = ctypes.c_int()
    R.ref_surf_forward(int(nthreads), ncol, maxl, _p(nl), _p(lay), npp, _p(fr), 2 if raylov == 1 else 1,
                       ctypes.c_double(dphase), nx, ny, ctypes.c_double(grid["xmin"]), ctypes.c_double(grid["ymin"]),
                       ctypes.c_double(grid["dx"]), ctypes.c_double(grid["dy"]), nsrc, _p(c64(s[:, 0])),
                       _p(c64(s[:, 1])), nrev, _p(c64(r[:, 0])), _p(c64(r[:, 1])), _p(rs), gridx, gridy, sgref,
                       sgdic, sgext, order, ctypes.c_double(band), _p(pv), _p(pt), ctypes.byref(nbad))
    return nbad.value, pv.reshape(nx, ny, npp), pt


# ---------------------------------------------------------------- surface likelihood
def surf_likelihood(grid, points, params, freqs, src, rev, ttime, raystat, raylov=1, dphase=1e-3,
                    sigdep=0, snoise0=None, snoise1=None, nrays=None, gridx=1, gridy=1, sgref=1,
                    sgdic=4, sgext=8, order=1, band=0.5, nthreads=0, phase_group=0, group_vel=None,
                    straight=0, datatype=2, water=0.0):
    """orc_surf_likelihood.  grid: dict nx..dz (+ scaling); water: T_GRID%waterDepth
    (a water layer on every column above it, 0 for none).  Returns (rc, out(3),
    phase_time (np,nsrc,nrev), srdist, sigma (np,npair), phase_vel (nx,ny,np),
    sites_id (nx,ny,nz)).  phase_group=1: group times along rays; group_vel, an
    (nx,ny,np) array, receives the group velocities.  datatype 3: the layers take
    the sites' own vp and rho (the surface half of the joint likelihood)."""
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    pts, prm = c64(points), c64(params)
    fr = c64(freqs).reshape(-1)
    npp = fr.shape[0]
    s, r = c64(src).reshape(-1, 2), c64(rev).reshape(-1, 2)
    nsrc, nrev = s.shape[0], r.shape[0]
    npair = nsrc * nrev
    tt = c64(ttime).reshape(npp, 3, npair)
    rs = np.ascontiguousarray(raystat, dtype=np.int32).reshape(npp, 2, npair)
    if nrays is None:
        nrays = int((rs[:, 0, :] == 1).sum())
    s0 = c64(np.zeros(npp) if snoise0 is None else np.broadcast_to(snoise0, (npp,)))
    s1 = c64(np.zeros(npp) if snoise1 is None else np.broadcast_to(snoise1, (npp,)))
    out = np.zeros(3)
    pt = np.zeros((npp, nsrc, nrev))
    sd = np.zeros((npp, npair))
    sg = np.zeros((npp, npair))
    pv = np.zeros((nx, ny, npp))
    sid = np.zeros((nx, ny, nz), np.int32)
    L = lib()
    rc = L.orc_surf_likelihood(
        ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
        ctypes.c_double(grid["xmin"]), ctypes.c_double(grid["ymin"]), ctypes.c_double(grid["zmin"]),
        ctypes.c_double(grid["dx"]), ctypes.c_double(grid["dy"]), ctypes.c_double(grid["dz"]),
        ctypes.c_double(grid.get("scaling", 1.0)), _p(pts), _p(prm), ctypes.c_int(pts.shape[0]),
        ctypes.c_int(npp), _p(fr), ctypes.c_int(raylov), ctypes.c_double(dphase),
        ctypes.c_int(nsrc), _p(s), ctypes.c_int(nrev), _p(r), _p(tt), _p(rs), ctypes.c_int(sigdep),
        _p(s0), _p(s1), ctypes.c_int(nrays), ctypes.c_int(gridx), ctypes.c_int(gridy),
        ctypes.c_int(sgref), ctypes.c_int(sgdic), ctypes.c_int(sgext), ctypes.c_int(order),
        ctypes.c_double(band), ctypes.c_int(nthreads), _p(out), _p(pt), _p(sd), _p(sg), _p(pv), _p(sid),
        ctypes.c_int(phase_group), _p(group_vel), ctypes.c_int(straight), ctypes.c_int(datatype),
        ctypes.c_double(water))
    return rc, out, pt, sd, sg, pv, sid
