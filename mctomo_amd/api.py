"""ctypes bindings of libmctomo_hip.so and the Python mirror of the reference's
likelihood interface for the hot path (see the package docstring)."""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK = 0
REJECT_MODEL = 1
REJECT_RAY = 2
_ERRORS = {
    -1: "invalid argument",
    -2: "HIP runtime error",
    -3: "device allocation failed",
    -4: "The noise level is 0!",
    -5: "eikonal solver did not converge",
    -6: "source or receiver outside the grid",
    -7: "data file could not be opened, read or written",
}


class MCTError(RuntimeError):
    """Fatal status from the library (the reference raises exception_raiseError)."""

    def __init__(self, code, where=""):
        self.code = code
        super().__init__("%s: %s (status %d)" % (where, _ERRORS.get(code, "error"), code))


def lib_path():
    return os.path.join(HERE, "libmctomo_hip.so")


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("scaling", ctypes.c_double)]


class _BodySetup(ctypes.Structure):
    _fields_ = [("datatype", ctypes.c_int), ("order", ctypes.c_int), ("sigdep", ctypes.c_int),
                ("nsrc", ctypes.c_int), ("nrev", ctypes.c_int),
                ("rev", ctypes.c_void_p), ("ttime", ctypes.c_void_p), ("raystat", ctypes.c_void_p),
                ("nrays", ctypes.c_int), ("max_cells", ctypes.c_int), ("max_chains", ctypes.c_int),
                ("dynamic", ctypes.c_int)]


def lib():
    """Load libmctomo_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError("libmctomo_hip.so is not built: run `python -m mctomo_amd._build`")
        L = ctypes.CDLL(p)
        L.mct_version.restype = ctypes.c_char_p
        for name in ("fastMarching3D", "mct_voronoi_to_grid", "mct_body_create", "mct_destroy",
                     "mct_body_likelihood", "mct_body_likelihood_batch_dev", "mct_sync",
                     "mct_last_timings", "mct_set_kernel_timing", "mct_set_tuning", "mct_dev_alloc", "mct_dev_free", "mct_copy_h2d",
                     "mct_copy_d2h", "mct_copy_d2d", "mct_surfdisp96_batch", "mct_modrays", "mct_surf_create",
                     "mct_surf_destroy", "mct_surf_likelihood", "mct_surf_likelihood_batch_dev",
                     "mct_surf_sync", "mct_surf_last_timings", "mct_surf_dev_arrays",
                     "mct_surf_set_method", "mct_surf_dispersion_work", "mct_surf_dispersion_solved", "mct_surf_fm2d_stats",
                     "mct_surf_set_dispersion_mode", "mct_surf_split_parts", "mct_surf_last_timings2d", "mct_surf_stream", "mct_serve_create",
                     "mct_serve_submit", "mct_serve_wait", "mct_serve_stats", "mct_serve_destroy",
                     "mct_body_noise_likelihood", "mct_body_noise_likelihood_batch_dev",
                     "mct_surf_noise_likelihood", "mct_surf_noise_likelihood_batch_dev",
                     "mct_surf_likelihood_box", "mct_surf_box_columns", "mct_surf_likelihood_box_batch_dev",
                     "mct_surf_accept", "mct_surf_accept_batch", "mct_surf_group_velocities",
                     "mct_read_sources", "mct_read_times", "mct_write_sources", "mct_write_times",
                     "mct_body_location_likelihood", "mct_body_location_likelihood_batch_dev",
                     "mct_serve_create_mix", "mct_serve_submit_to", "mct_serve_memory", "mct_surf_device_bytes",
                     "mct_surf_set_tuning", "mct_surf_get_tuning", "mct_surf_likelihood_batch_graph",
                     "mct_fm2d_last_handover", "mct_surf_fm2d_handover", "mct_dispersion_columns",
                     "mct_body_stream", "mct_joint_create", "mct_joint_destroy", "mct_joint_parts",
                     "mct_joint_likelihood", "mct_joint_likelihood_batch_dev", "mct_joint_sync",
                     "mct_body_noise_likelihood_from", "mct_surf_noise_likelihood_from",
                     "mct_modrays_dev_noalloc", "mct_noalloc_scratch_bytes", "mct_serve_set_fill",
                     "mct_body_set_order1"):
            getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


def version():
    return lib().mct_version().decode()


# ---------------------------------------------------------------- data files
def _path(p):
    return os.fsencode(os.fspath(p))


def read_sources(path, bnd=None):
    """read_sources (src/likelihood_settings.f90:198-228): (n, ndim) array.
    bnd: (min, max) per coordinate, or None."""
    n, nd = ctypes.c_int(), ctypes.c_int()
    _check(lib().mct_read_sources(_path(path), None, 0, ctypes.c_long(0), ctypes.byref(n), ctypes.byref(nd),
                                  None), "mct_read_sources")
    out = np.zeros((n.value, nd.value))
    b = None if bnd is None else _f64(bnd).reshape(-1)
    _check(lib().mct_read_sources(_path(path), _p(b), 0 if b is None else b.size, ctypes.c_long(out.size),
                                  ctypes.byref(n), ctypes.byref(nd), _p(out)), "mct_read_sources")
    return out


def read_times(path, nsrc, nrev):
    """read_times_3 (src/likelihood_settings.f90:332-420): returns (freqs (np),
    ttime (np,3,nsrc*nrev), raystat (np,2,nsrc*nrev), nrays)."""
    npp = ctypes.c_int()
    _check(lib().mct_read_times(_path(path), int(nsrc), int(nrev), 0, ctypes.byref(npp), None, None, None, None),
           "mct_read_times")
    k = npp.value
    fr = np.zeros(k)
    tt = np.zeros((k, 3, nsrc * nrev))
    rs = np.zeros((k, 2, nsrc * nrev), np.int32)
    nr = ctypes.c_int()
    _check(lib().mct_read_times(_path(path), int(nsrc), int(nrev), k, ctypes.byref(npp), _p(fr), _p(tt), _p(rs),
                                ctypes.byref(nr)), "mct_read_times")
    return fr, tt, rs, nr.value


def read_data(sources_file, receivers_file, data_file):
    """read_data for one data set (src/likelihood_settings.f90:166-196):
    (src, rev, freqs, ttime, raystat, nrays)."""
    src = read_sources(sources_file)
    rev = read_sources(receivers_file)
    return (src, rev) + read_times(data_file, src.shape[0], rev.shape[0])


def write_sources(path, src):
    a = _f64(src)
    _check(lib().mct_write_sources(_path(path), a.shape[0], a.shape[1], _p(a)), "mct_write_sources")


def write_times(path, freqs, raystat, phase_time, sigma, locations=None):
    """write_atimes (src/read_write.f90:484-530).  phase_time (np,nsrc,nrev),
    sigma (np,nsrc*nrev), raystat (np,2,nsrc*nrev), locations (nsrc,4) or None."""
    fr = _f64(freqs).reshape(-1)
    pt = _f64(phase_time)
    sg = _f64(sigma)
    rs = np.ascontiguousarray(raystat, dtype=np.int32)
    loc = None if locations is None else _f64(locations)
    _check(lib().mct_write_times(_path(path), fr.shape[0], _p(fr), pt.shape[1], pt.shape[2], _p(rs), _p(pt),
                                 _p(sg), _p(loc)), "mct_write_times")


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc, where):
    if rc < 0:
        raise MCTError(rc, where)
    return rc


class DeviceArray:
    """A device allocation made through the library (hipMalloc), with the
    numpy shape/dtype it mirrors.  Used to hand device pointers to the _dev
    entry points without any GPU framework in the process."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().mct_dev_alloc(ctypes.byref(p), ctypes.c_int64(self.nbytes)), "mct_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().mct_copy_h2d(ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data),
                                  ctypes.c_int64(self.nbytes)), "mct_copy_h2d")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().mct_copy_d2h(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr),
                                  ctypes.c_int64(self.nbytes)), "mct_copy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().mct_dev_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Grid:
    """T_GRID (src/settings.f90:20-28); zmin and dz are the scaled values."""
    nx: int
    ny: int
    nz: int
    xmin: float
    ymin: float
    zmin: float
    dx: float
    dy: float
    dz: float
    scaling: float = 1.0

    def c(self):
        return _Grid(self.nx, self.ny, self.nz, self.xmin, self.ymin, self.zmin,
                     self.dx, self.dy, self.dz, self.scaling)

    @classmethod
    def from_bounds(cls, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, scaling=1.0):
        """grid_setup (src/settings.f90:100-123): scale z, then the spacings."""
        zmin = zmin * scaling
        zmax = zmax * scaling
        return cls(nx, ny, nz, xmin, ymin, zmin, (xmax - xmin) / (nx - 1),
                   (ymax - ymin) / (ny - 1), (zmax - zmin) / (nz - 1), scaling)


def fast_marching_3d(vel, dx, dy, dz, xmin, ymin, zmin, src, rev, order, field=False):
    """fastMarching3D: vel (nx,ny,nz) C array (= Fortran (nz,ny,nx)), src (3,),
    rev (nrev,3).  Returns traveltimes (nrev,) [and the (nx,ny,nz) field]."""
    vel = _f64(vel)
    nx, ny, nz = vel.shape
    rev = _f64(rev).reshape(-1, 3)
    tt = np.zeros(rev.shape[0])
    tf = np.zeros_like(vel) if field else None
    rc = lib().fastMarching3D(nx, ny, nz, ctypes.c_double(dx), ctypes.c_double(dy),
                              ctypes.c_double(dz), ctypes.c_double(xmin), ctypes.c_double(ymin),
                              ctypes.c_double(zmin), _p(vel), _p(_f64(src)), _p(rev),
                              rev.shape[0], int(order), _p(tt), _p(tf))
    _check(rc, "fastMarching3D")
    return (tt, tf) if field else tt


def surfdisp96_batch(models, periods, iwave, igr, dphase=1e-3):
    """surfdisp96 as surfmodes calls it, for many layer models at once.
    models: list of (thick, vp, vs, rho) arrays (cast to REAL*4).
    iwave 1 Love, 2 Rayleigh; igr 0 phase, 1 group.
    Returns phase (nmod,np), group (nmod,np), ierr (nmod,)."""
    nmod = len(models)
    nl = np.array([len(m[0]) for m in models], np.int32)
    nlm = int(nl.max())
    lay = np.zeros((nmod, 4, nlm), np.float32)
    for i, m in enumerate(models):
        for j in range(4):
            lay[i, j, :nl[i]] = np.asarray(m[j], dtype=np.float32)
    per = _f64(periods)
    ph = np.zeros((nmod, per.shape[0]))
    gr = np.zeros((nmod, per.shape[0]))
    ie = np.zeros(nmod, np.int32)
    rc = lib().mct_surfdisp96_batch(nmod, nlm, _p(nl), _p(lay), per.shape[0], _p(per), int(iwave),
                                    int(igr), ctypes.c_double(dphase), _p(ph), _p(gr), _p(ie))
    _check(rc, "mct_surfdisp96_batch")
    return ph, gr, ie


def dispersion_columns(vs, dz, periods, iwave, igr, mode=3, dphase=1e-3, scaling=1.0, water=0.0):
    """The surface likelihood's dispersion stage on explicit Vs columns
    (check_model, convert_to_layer, surfdisp96).  vs: (ncol, nz).  mode 1 lane
    per column, 2 wave per column, 3 column queue; water: T_GRID%waterDepth
    (0: no water layer).  Returns phase (ncol, np), group (ncol, np), ierr
    (ncol,), bad_model."""
    vs = _f64(vs)
    ncol, nz = vs.shape
    per = _f64(periods)
    ph = np.zeros((ncol, per.shape[0]))
    gr = np.zeros((ncol, per.shape[0]))
    ie = np.zeros(ncol, np.int32)
    bad = np.zeros(1, np.int32)
    rc = lib().mct_dispersion_columns(ncol, nz, ctypes.c_double(dz), ctypes.c_double(scaling),
                                      ctypes.c_double(water), _p(vs),
                                      per.shape[0], _p(per), int(iwave), int(igr), ctypes.c_double(dphase),
                                      int(mode), _p(ph), _p(gr), _p(ie), _p(bad))
    _check(rc, "mct_dispersion_columns")
    return ph, gr, ie, int(bad[0])


def modrays(velv, goxd, gozd, dvxd, dvzd, scx, scz, rcx, rcz, srs=None, gdx=1, gdz=1, asgr=1,
            sgdl=4, sgs=8, fom=1, ttime=None, method=0):
    """modrays phase times (fm2d/fm2dray_cartesian.f90:67-480) on the GPU.
    velv: (np, nvx+2, nvz+2) or (nvx+2, nvz+2) C array, i.e. Fortran
    velv(0:nvz+1, 0:nvx+1) per period; srs (np, nsrc, nrc) or (nsrc, nrc)
    marks the pairs to compute.  Returns ttime (np, nsrc, nrc) (or (nsrc, nrc)
    for a single grid); unmarked entries keep the values of ``ttime``.
    method 0: parallel sweeps with the heap loop near the source; 1: the
    sequential heap loop per (period, source)."""
    velv = _f64(velv)
    single = velv.ndim == 2
    if single:
        velv = velv[None]
    npd, nvx, nvz = velv.shape[0], velv.shape[1] - 2, velv.shape[2] - 2
    scx, scz, rcx, rcz = _f64(scx), _f64(scz), _f64(rcx), _f64(rcz)
    nsrc, nrc = scx.shape[0], rcx.shape[0]
    if srs is None:
        srs = np.ones((npd, nsrc, nrc), np.int32)
    srs = np.ascontiguousarray(np.broadcast_to(np.asarray(srs, np.int32).reshape(-1, nsrc, nrc),
                                               (npd, nsrc, nrc)))
    tt = np.zeros((npd, nsrc, nrc)) if ttime is None else \
        np.ascontiguousarray(np.broadcast_to(np.asarray(ttime, np.float64).reshape(-1, nsrc, nrc),
                                             (npd, nsrc, nrc)))
    rc = lib().mct_modrays(nvx, nvz, ctypes.c_double(goxd), ctypes.c_double(gozd),
                           ctypes.c_double(dvxd), ctypes.c_double(dvzd), npd, _p(velv), gdx, gdz,
                           asgr, sgdl, sgs, fom, nsrc, _p(scx), _p(scz), nrc, _p(rcx), _p(rcz),
                           _p(srs), _p(tt), int(method))
    _check(rc, "mct_modrays")
    return tt[0] if single else tt


def modrays_last_handover():
    """Problems the calling thread's last modrays() call handed from the
    parallel path to the sequential solver."""
    n = ctypes.c_longlong()
    _check(lib().mct_fm2d_last_handover(ctypes.byref(n)), "mct_fm2d_last_handover")
    return n.value


def kdtree_to_grid(grid, points, params, box=None, pm=None, vp=None, vs=None, rho=None):
    """kdtree_to_grid: nearest Voronoi site of each node (1-based ids) and the
    site's (vp, vs, rho).  Arrays are (nx,ny,nz) C (= Fortran (nz,ny,nx))."""
    g = grid.c()
    pts = _f64(points)
    prm = _f64(params)
    shape = (grid.nx, grid.ny, grid.nz)
    sid = np.zeros(shape, np.int32)
    vp = np.ones(shape) if vp is None else _f64(vp).copy()
    vs = np.ones(shape) if vs is None else _f64(vs).copy()
    rho = np.ones(shape) if rho is None else _f64(rho).copy()
    rc = lib().mct_voronoi_to_grid(ctypes.byref(g), _p(pts), _p(prm), pts.shape[0],
                                   _p(None if box is None else _f64(box)),
                                   _p(None if pm is None else _f64(pm)),
                                   _p(sid), _p(vp), _p(vs), _p(rho))
    _check(rc, "mct_voronoi_to_grid")
    return sid, vp, vs, rho


class BodyLikelihood:
    """Body-wave likelihood context: like_setup + likelihood for datatype 0/1.

    rev (nrev,3); ttime (np,3,nsrc*nrev) = (time, sigma, -) per pair with the
    receiver index fastest (nrr = k + (j-1)*nrev); raystat (np,2,nsrc*nrev).
    """

    def __init__(self, grid, rev, ttime, raystat, nsrc, datatype=0, order=0, sigdep=0,
                 nrays=None, max_cells=1024, max_chains=1, dynamic=0):
        self.grid = grid
        self.np = 1 if datatype == 0 else 2
        self.nsrc = int(nsrc)
        self.rev = _f64(rev).reshape(-1, 3)
        self.nrev = self.rev.shape[0]
        self.ttime = _f64(ttime).reshape(self.np, 3, self.nsrc * self.nrev)
        self.raystat = np.ascontiguousarray(raystat, dtype=np.int32).reshape(self.np, 2, -1)
        if nrays is None:
            nrays = int((self.raystat[:, 0, :] == 1).sum())
        self.nrays = nrays
        self.sigdep = sigdep
        self.max_cells = max_cells
        self.max_chains = max_chains
        st = _BodySetup(datatype, order, sigdep, self.nsrc, self.nrev,
                        self.rev.ctypes.data, self.ttime.ctypes.data, self.raystat.ctypes.data,
                        nrays, max_cells, max_chains, int(dynamic))
        self._g = grid.c()
        h = ctypes.c_void_p()
        _check(lib().mct_body_create(ctypes.byref(self._g), ctypes.byref(st), ctypes.byref(h)),
               "mct_body_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mct_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def likelihood(self, points, params, locations, bnoise0=None, bnoise1=None, arrays=False):
        """One chain.  Returns (status, (like, misfit, unweighted)[, phase_time
        (np,nsrc,nrev), srdist (np,npair), sigma (np,npair)])."""
        pts = _f64(points)
        prm = _f64(params)
        loc = _f64(locations)
        b0 = None if bnoise0 is None else _f64(np.atleast_1d(bnoise0))
        b1 = None if bnoise1 is None else _f64(np.atleast_1d(bnoise1))
        out = np.zeros(3)
        pt = sd = sg = None
        if arrays:
            npair = self.nsrc * self.nrev
            pt = np.zeros((self.np, self.nsrc, self.nrev))
            sd = np.zeros((self.np, npair))
            sg = np.zeros((self.np, npair))
        rc = lib().mct_body_likelihood(self._h, _p(pts), _p(prm), pts.shape[0], _p(loc),
                                       _p(b0), _p(b1), _p(out), _p(pt), _p(sd), _p(sg))
        _check(rc, "mct_body_likelihood")
        if arrays:
            return rc, out, pt, sd, sg
        return rc, out

    def likelihood_batch_dev(self, nchain, points_ptr, params_ptr, ncells_ptr, locations_ptr,
                             noise_ptr, out_ptr, status_ptr):
        """Batched call on device pointers (e.g. torch tensors' data_ptr())."""
        rc = lib().mct_body_likelihood_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(points_ptr), ctypes.c_void_p(params_ptr),
            ctypes.c_void_p(ncells_ptr), ctypes.c_void_p(locations_ptr),
            ctypes.c_void_p(noise_ptr) if noise_ptr else None, ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr) if status_ptr else None)
        return _check(rc, "mct_body_likelihood_batch_dev")

    def noise_likelihood(self, bnoise0, bnoise1, locations=None):
        """noise_likelihood: the last call's traveltimes under new noise
        parameters (and origin times).  Returns (status, out[3])."""
        b0 = _f64(np.atleast_1d(bnoise0))
        b1 = _f64(np.atleast_1d(bnoise1))
        loc = None if locations is None else _f64(locations)
        out = np.zeros(3)
        rc = lib().mct_body_noise_likelihood(self._h, _p(loc), _p(b0), _p(b1), _p(out))
        _check(rc, "mct_body_noise_likelihood")
        return rc, out

    def location_likelihood(self, locations, bnoise0, bnoise1):
        """location_likelihood with dynamic != 0: the last call's receiver
        fields interpolated at new source locations (nsrc, 4).  Returns
        (status, out[3])."""
        b0 = _f64(np.atleast_1d(bnoise0))
        b1 = _f64(np.atleast_1d(bnoise1))
        loc = _f64(locations)
        out = np.zeros(3)
        rc = lib().mct_body_location_likelihood(self._h, _p(loc), _p(b0), _p(b1), _p(out))
        _check(rc, "mct_body_location_likelihood")
        return rc, out

    def location_likelihood_batch_dev(self, nchain, locations_ptr, noise_ptr, out_ptr, status_ptr):
        rc = lib().mct_body_location_likelihood_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(locations_ptr), ctypes.c_void_p(noise_ptr),
            ctypes.c_void_p(out_ptr), ctypes.c_void_p(status_ptr) if status_ptr else None)
        _check(rc, "mct_body_location_likelihood_batch_dev")
        return rc

    def noise_likelihood_batch_dev(self, nchain, locations_ptr, noise_ptr, out_ptr, status_ptr):
        rc = lib().mct_body_noise_likelihood_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(locations_ptr) if locations_ptr else None,
            ctypes.c_void_p(noise_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr) if status_ptr else None)
        return _check(rc, "mct_body_noise_likelihood_batch_dev")

    def sync(self):
        _check(lib().mct_sync(self._h), "mct_sync")

    def set_tuning(self, band_tiles=0.0, t64=-1):
        """Eikonal band advance per round (tiles) and arrival-time precision."""
        _check(lib().mct_set_tuning(self._h, ctypes.c_float(band_tiles), int(t64)), "mct_set_tuning")

    def set_order1(self, method):
        """How order 1 is solved: "heap" replays the reference's SFMM heap loop
        exactly (the default), "parallel" relaxes the same update to its fixed
        point with the tiled solver (faster, about 1e-2 relative)."""
        m = {"heap": 0, "parallel": 1}[method]
        _check(lib().mct_body_set_order1(self._h, m), "mct_body_set_order1")

    def set_kernel_timing(self, on):
        _check(lib().mct_set_kernel_timing(self._h, int(bool(on))), "mct_set_kernel_timing")

    def timings(self):
        t = np.zeros(6)
        lib().mct_last_timings(self._h, _p(t))
        return {"nearest_ms": t[0], "eikonal_ms": t[1], "misfit_ms": t[2], "rounds": int(t[3]),
                "round_kernels_ms": t[4], "round_launches": int(t[5])}


class _SurfSetup(ctypes.Structure):
    _fields_ = [("raylov", ctypes.c_int), ("phase_group", ctypes.c_int), ("dphase", ctypes.c_double),
                ("np", ctypes.c_int), ("freqs", ctypes.c_void_p), ("nsrc", ctypes.c_int),
                ("nrev", ctypes.c_int), ("src", ctypes.c_void_p), ("rev", ctypes.c_void_p),
                ("ttime", ctypes.c_void_p), ("raystat", ctypes.c_void_p), ("sigdep", ctypes.c_int),
                ("nrays", ctypes.c_int), ("gridx", ctypes.c_int), ("gridy", ctypes.c_int),
                ("sgref", ctypes.c_int), ("sgdic", ctypes.c_int), ("sgext", ctypes.c_int),
                ("order", ctypes.c_int), ("band", ctypes.c_double), ("max_cells", ctypes.c_int),
                ("max_chains", ctypes.c_int), ("straight", ctypes.c_int), ("datatype", ctypes.c_int),
                ("water_depth", ctypes.c_double)]


def _surf_setup(grid, freqs, src, rev, ttime, raystat, raylov=1, dphase=1e-3, sigdep=0, nrays=None,
                gridx=1, gridy=1, sgref=1, sgdic=4, sgext=8, order=1, band=0.5, max_cells=1024,
                max_chains=1, phase_group=0, straight=0, datatype=2, water_depth=0.0):
    """A filled mct_surf_setup and the arrays it points to (keep them alive)."""
    fr = _f64(freqs).reshape(-1)
    npd = fr.shape[0]
    sr, rv = _f64(src).reshape(-1, 2), _f64(rev).reshape(-1, 2)
    npair = sr.shape[0] * rv.shape[0]
    tt = _f64(ttime).reshape(npd, 3, npair)
    rs = np.ascontiguousarray(raystat, dtype=np.int32).reshape(npd, 2, npair)
    if nrays is None:
        nrays = int((rs[:, 0, :] == 1).sum())
    st = _SurfSetup(int(raylov), int(phase_group), float(dphase), npd, fr.ctypes.data, sr.shape[0], rv.shape[0],
                    sr.ctypes.data, rv.ctypes.data, tt.ctypes.data, rs.ctypes.data, int(sigdep), int(nrays),
                    int(gridx), int(gridy), int(sgref), int(sgdic), int(sgext), int(order), float(band),
                    int(max_cells), int(max_chains), int(straight), int(datatype), float(water_depth))
    return st, (fr, sr, rv, tt, rs)


class SurfLikelihood:
    """Surface-wave likelihood context: like_setup + likelihood for datatype 2
    (src/likelihood_surf.F90:110-405).

    freqs (np) Hz; src (nsrc,2) and rev (nrev,2) as (x, y); ttime
    (np,3,nsrc*nrev) = (time, sigma, -) with the receiver index fastest;
    raystat (np,2,nsrc*nrev).  Settings as LIKE_SET: raylov (1 Rayleigh,
    0 Love), phase_group (0 phase, 1 group velocities: group delays along
    rays), dphase (dPhaseVel), gridx/gridy, sgref/sgdic/sgext, order, band.
    datatype 2: vp and rho follow from Vs; 3: the surface half of the joint
    likelihood, vp and rho are the sites' own (params columns 0 and 2).
    """

    def __init__(self, grid, freqs, src, rev, ttime, raystat, raylov=1, dphase=1e-3, sigdep=0,
                 nrays=None, gridx=1, gridy=1, sgref=1, sgdic=4, sgext=8, order=1, band=0.5,
                 max_cells=1024, max_chains=1, phase_group=0, straight=0, datatype=2, water_depth=0.0):
        self.grid = grid
        self.freqs = _f64(freqs).reshape(-1)
        self.np = self.freqs.shape[0]
        self.src = _f64(src).reshape(-1, 2)
        self.rev = _f64(rev).reshape(-1, 2)
        self.nsrc, self.nrev = self.src.shape[0], self.rev.shape[0]
        npair = self.nsrc * self.nrev
        self.ttime = _f64(ttime).reshape(self.np, 3, npair)
        self.raystat = np.ascontiguousarray(raystat, dtype=np.int32).reshape(self.np, 2, npair)
        if nrays is None:
            nrays = int((self.raystat[:, 0, :] == 1).sum())
        self.nrays, self.sigdep = nrays, sigdep
        self.max_cells, self.max_chains = max_cells, max_chains
        self.phase_group = int(phase_group)
        st = _SurfSetup(int(raylov), self.phase_group, float(dphase), self.np, self.freqs.ctypes.data, self.nsrc,
                        self.nrev, self.src.ctypes.data, self.rev.ctypes.data, self.ttime.ctypes.data,
                        self.raystat.ctypes.data, int(sigdep), int(nrays), int(gridx), int(gridy),
                        int(sgref), int(sgdic), int(sgext), int(order), float(band), int(max_cells),
                        int(max_chains), int(straight), int(datatype), float(water_depth))
        self._g = grid.c()
        h = ctypes.c_void_p()
        _check(lib().mct_surf_create(ctypes.byref(self._g), ctypes.byref(st), ctypes.byref(h)),
               "mct_surf_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mct_surf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def likelihood(self, points, params, snoise0=None, snoise1=None, arrays=False):
        """One chain.  Returns (status, (like, misfit, unweighted)[, phase_time
        (np,nsrc,nrev), srdist (np,npair), sigma (np,npair), phase_vel (nx,ny,np)])."""
        pts = _f64(points)
        prm = _f64(params)
        s0 = None if snoise0 is None else _f64(np.broadcast_to(np.atleast_1d(snoise0), (self.np,)))
        s1 = None if snoise1 is None else _f64(np.broadcast_to(np.atleast_1d(snoise1), (self.np,)))
        out = np.zeros(3)
        pt = sd = sg = pv = None
        if arrays:
            npair = self.nsrc * self.nrev
            pt = np.zeros((self.np, self.nsrc, self.nrev))
            sd = np.zeros((self.np, npair))
            sg = np.zeros((self.np, npair))
            pv = np.zeros((self.grid.nx, self.grid.ny, self.np))
        rc = lib().mct_surf_likelihood(self._h, _p(pts), _p(prm), pts.shape[0], _p(s0), _p(s1),
                                       _p(out), _p(pt), _p(sd), _p(sg), _p(pv))
        _check(rc, "mct_surf_likelihood")
        if arrays:
            return rc, out, pt, sd, sg, pv
        return rc, out

    def likelihood_batch_dev(self, nchain, points_ptr, params_ptr, ncells_ptr, noise_ptr, out_ptr,
                             status_ptr):
        """Batched call on device pointers (see DeviceArray)."""
        rc = lib().mct_surf_likelihood_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(points_ptr), ctypes.c_void_p(params_ptr),
            ctypes.c_void_p(ncells_ptr), ctypes.c_void_p(noise_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr))
        return _check(rc, "mct_surf_likelihood_batch_dev")

    def likelihood_batch_graph(self, nchain, points_ptr, params_ptr, ncells_ptr, noise_ptr, out_ptr,
                               status_ptr):
        """likelihood_batch_dev replayed from a HIP graph captured on the first call
        with these arguments."""
        rc = lib().mct_surf_likelihood_batch_graph(
            self._h, int(nchain), ctypes.c_void_p(points_ptr), ctypes.c_void_p(params_ptr),
            ctypes.c_void_p(ncells_ptr), ctypes.c_void_p(noise_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr))
        return _check(rc, "mct_surf_likelihood_batch_graph")

    def likelihood_box(self, points, params, box, snoise0=None, snoise1=None, arrays=False):
        """likelihood() of a proposal that changed the model only inside box =
        (x0, y0, z0, x1, y1, z1): dispersion only for the box's columns, the
        others from the accepted state (see accept).  Same returns as likelihood."""
        pts, prm = _f64(points), _f64(params)
        bx = _f64(np.asarray(box, dtype=np.float64).reshape(6))
        s0 = None if snoise0 is None else _f64(np.broadcast_to(np.atleast_1d(snoise0), (self.np,)))
        s1 = None if snoise1 is None else _f64(np.broadcast_to(np.atleast_1d(snoise1), (self.np,)))
        out = np.zeros(3)
        pt = sd = sg = pv = None
        if arrays:
            npair = self.nsrc * self.nrev
            pt = np.zeros((self.np, self.nsrc, self.nrev))
            sd = np.zeros((self.np, npair))
            sg = np.zeros((self.np, npair))
            pv = np.zeros((self.grid.nx, self.grid.ny, self.np))
        rc = lib().mct_surf_likelihood_box(self._h, _p(pts), _p(prm), pts.shape[0], _p(bx), _p(s0), _p(s1),
                                           _p(out), _p(pt), _p(sd), _p(sg), _p(pv))
        _check(rc, "mct_surf_likelihood_box")
        if arrays:
            return rc, out, pt, sd, sg, pv
        return rc, out

    def box_columns(self, box):
        """(ix0, ix1, iy0, iy1), 1-based inclusive: the columns a box call recomputes."""
        bx = _f64(np.asarray(box, dtype=np.float64).reshape(6))
        cols = np.zeros(4, np.int32)
        _check(lib().mct_surf_box_columns(self._h, _p(bx), _p(cols)), "mct_surf_box_columns")
        return tuple(int(v) for v in cols)

    def likelihood_box_batch_dev(self, nchain, points_ptr, params_ptr, ncells_ptr, cols, noise_ptr, out_ptr,
                                 status_ptr):
        """Batched box call; cols (nchain, 4) int32 host column ranges."""
        c = np.ascontiguousarray(cols, dtype=np.int32).reshape(-1, 4)
        rc = lib().mct_surf_likelihood_box_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(points_ptr), ctypes.c_void_p(params_ptr),
            ctypes.c_void_p(ncells_ptr), _p(c), ctypes.c_void_p(noise_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr))
        return _check(rc, "mct_surf_likelihood_box_batch_dev")

    def accept(self, chain=0):
        """The last call's proposal for chain was accepted: its phase velocities
        become the state box calls start from."""
        _check(lib().mct_surf_accept(self._h, int(chain)), "mct_surf_accept")

    def accept_batch(self, accepted):
        a = np.ascontiguousarray(accepted, dtype=np.int32)
        _check(lib().mct_surf_accept_batch(self._h, a.shape[0], _p(a)), "mct_surf_accept_batch")

    def group_velocities(self, chain=0):
        """Group velocities of chain's last call, (nx, ny, np)."""
        out = np.zeros((self.grid.nx, self.grid.ny, self.np))
        _check(lib().mct_surf_group_velocities(self._h, int(chain), _p(out)), "mct_surf_group_velocities")
        return out

    def noise_likelihood(self, snoise0, snoise1):
        """noise_likelihood: the last call's phase times under new noise
        parameters.  Returns (status, out[3])."""
        s0 = _f64(np.broadcast_to(np.atleast_1d(snoise0), (self.np,)))
        s1 = _f64(np.broadcast_to(np.atleast_1d(snoise1), (self.np,)))
        out = np.zeros(3)
        rc = lib().mct_surf_noise_likelihood(self._h, _p(s0), _p(s1), _p(out))
        _check(rc, "mct_surf_noise_likelihood")
        return rc, out

    def noise_likelihood_batch_dev(self, nchain, noise_ptr, out_ptr, status_ptr):
        rc = lib().mct_surf_noise_likelihood_batch_dev(
            self._h, int(nchain), ctypes.c_void_p(noise_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(status_ptr))
        return _check(rc, "mct_surf_noise_likelihood_batch_dev")

    def sync(self):
        _check(lib().mct_surf_sync(self._h), "mct_surf_sync")

    def timings(self):
        """Stage times (ms) of the last synchronised call: nearest, dispersion, 2-D, misfit."""
        t = np.zeros(4)
        lib().mct_surf_last_timings(self._h, _p(t))
        return t

    def set_method(self, fm2d_method):
        """2-D traveltimes: 0 parallel sweeps (default), 1 sequential heap loop."""
        _check(lib().mct_surf_set_method(self._h, int(fm2d_method)), "mct_surf_set_method")

    def dispersion_work(self, nchain=1):
        """(dltar evaluations, evaluations x (layers-1)) of the last call's first nchain chains."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().mct_surf_dispersion_work(self._h, int(nchain), ctypes.byref(a), ctypes.byref(b)),
               "mct_surf_dispersion_work")
        return a.value, b.value

    def dispersion_solved(self, nchain=1):
        """(columns solved, their evaluations x (layers-1)) of the last call's first
        nchain chains; columns identical to an earlier one of the chain are copied."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().mct_surf_dispersion_solved(self._h, int(nchain), ctypes.byref(a), ctypes.byref(b)),
               "mct_surf_dispersion_solved")
        return a.value, b.value

    def timings2d(self):
        """(near-source ms, sweep ms) of the last synchronised call's 2-D stage."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().mct_surf_last_timings2d(self._h, ctypes.byref(a), ctypes.byref(b)), "mct_surf_last_timings2d")
        return a.value, b.value

    def split_parts(self, nchain):
        """(parts, chains of the first part) a batch of nchain chains runs in."""
        n = ctypes.c_int()
        p = lib().mct_surf_split_parts(self._h, int(nchain), ctypes.byref(n))
        _check(min(p, 0), "mct_surf_split_parts")
        return p, n.value

    def set_dispersion_mode(self, mode):
        """Dispersion launch shape: 0 by size, 1 lane per column, 2 wave per column,
        3 lane per column with columns taken from a queue."""
        _check(lib().mct_surf_set_dispersion_mode(self._h, int(mode)), "mct_surf_set_dispersion_mode")

    TUNE_DSP_WAVES, TUNE_SWEEP_KIND, TUNE_SWEEP_THREADS, TUNE_FM2D_CHUNK, TUNE_DSP_DEDUP = 1, 2, 3, 4, 5

    def set_tuning(self, key, value):
        """Launch shape (results do not depend on it): TUNE_DSP_WAVES, TUNE_SWEEP_KIND
        (0 default, 1 strip, 2 anti-diagonal), TUNE_SWEEP_THREADS, TUNE_FM2D_CHUNK,
        TUNE_DSP_DEDUP (1: identical column profiles solved once, 0: every column)."""
        _check(lib().mct_surf_set_tuning(self._h, int(key), ctypes.c_int64(int(value))), "mct_surf_set_tuning")

    def get_tuning(self, key):
        v = ctypes.c_int64()
        _check(lib().mct_surf_get_tuning(self._h, int(key), ctypes.byref(v)), "mct_surf_get_tuning")
        return v.value

    def fm2d_stats(self):
        """(slots, mean rounds, mean diagonal steps, mean node updates) of the last 2-D chunk."""
        t = np.zeros(4)
        _check(lib().mct_surf_fm2d_stats(self._h, _p(t)), "mct_surf_fm2d_stats")
        return tuple(t)

    HANDOVER_REASONS = ("no time", "tie", "early", "rising key", "final key", "band heap", "no convergence")

    def fm2d_handover(self, reasons=False):
        """(problems that failed the heap-order check, all problems handed to
        the sequential solver) in the last 2-D chunk; with reasons, also a dict
        of problems per reason of the check."""
        t = np.zeros(9, np.int64)
        _check(lib().mct_surf_fm2d_handover(self._h, _p(t)), "mct_surf_fm2d_handover")
        if reasons:
            return int(t[0]), int(t[1]), dict(zip(self.HANDOVER_REASONS, (int(x) for x in t[2:])))
        return int(t[0]), int(t[1])

    def dev_arrays(self):
        """Device addresses of the phase_time and phase-velocity buffers."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        lib().mct_surf_dev_arrays(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


class JointLikelihood:
    """Joint body + surface-wave likelihood (datatype 3, src/likelihood.f90:83-88):
    body data (P and S) and surface data on one grid, the surface layers taking
    the sites' own vp and rho.  body: dict(rev, ttime, raystat, nsrc[, order,
    sigdep, nrays]); surf: dict(freqs, src, rev, ttime, raystat[, raylov,
    order, ...]) as SurfLikelihood takes them."""

    def __init__(self, grid, body, surf, max_cells=1024, max_chains=1):
        self.grid = grid
        b = dict(body)
        self.nsrc = int(b["nsrc"])
        self._brev = _f64(b["rev"]).reshape(-1, 3)
        nrev = self._brev.shape[0]
        self._btt = _f64(b["ttime"]).reshape(2, 3, self.nsrc * nrev)
        self._brs = np.ascontiguousarray(b["raystat"], dtype=np.int32).reshape(2, 2, -1)
        nrays = b.get("nrays")
        if nrays is None:
            nrays = int((self._brs[:, 0, :] == 1).sum())
        bst = _BodySetup(3, int(b.get("order", 0)), int(b.get("sigdep", 0)), self.nsrc, nrev,
                         self._brev.ctypes.data, self._btt.ctypes.data, self._brs.ctypes.data,
                         int(nrays), int(max_cells), int(max_chains), 0)
        sk = dict(surf)
        args = [sk.pop(k) for k in ("freqs", "src", "rev", "ttime", "raystat")]
        sst, self._keep = _surf_setup(grid, *args, max_cells=max_cells, max_chains=max_chains, datatype=3, **sk)
        self.np = sst.np
        self.max_cells, self.max_chains = max_cells, max_chains
        self._g = grid.c()
        h = ctypes.c_void_p()
        _check(lib().mct_joint_create(ctypes.byref(self._g), ctypes.byref(bst), ctypes.byref(sst),
                                      ctypes.byref(h)), "mct_joint_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mct_joint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def likelihood(self, points, params, locations, bnoise0=None, bnoise1=None, snoise0=None, snoise1=None):
        """One chain.  Returns (status, (like, misfit, unweighted), body triple,
        surface triple)."""
        pts, prm, loc = _f64(points), _f64(params), _f64(locations)
        out = np.zeros(3)
        parts = np.zeros(6)
        z2, zp = np.zeros(2), np.zeros(self.np)
        rc = lib().mct_joint_likelihood(self._h, _p(pts), _p(prm), pts.shape[0], _p(loc),
                                        _p(z2 if bnoise0 is None else _f64(bnoise0)),
                                        _p(z2 if bnoise1 is None else _f64(bnoise1)),
                                        _p(zp if snoise0 is None else _f64(snoise0)),
                                        _p(zp if snoise1 is None else _f64(snoise1)), _p(out), _p(parts))
        if rc < 0:
            _check(rc, "mct_joint_likelihood")
        return rc, out, parts[:3], parts[3:]

    def likelihood_batch_dev(self, nchain, points_ptr, params_ptr, ncells_ptr, locations_ptr, bnoise_ptr,
                             snoise_ptr, out_ptr, status_ptr):
        _check(lib().mct_joint_likelihood_batch_dev(self._h, int(nchain), ctypes.c_void_p(points_ptr),
                                                    ctypes.c_void_p(params_ptr), ctypes.c_void_p(ncells_ptr),
                                                    ctypes.c_void_p(locations_ptr), ctypes.c_void_p(bnoise_ptr),
                                                    ctypes.c_void_p(snoise_ptr), ctypes.c_void_p(out_ptr),
                                                    ctypes.c_void_p(status_ptr)), "mct_joint_likelihood_batch_dev")

    def sync(self):
        _check(lib().mct_joint_sync(self._h), "mct_joint_sync")


class SurfServer:
    """Continuous-batching server for surface-wave likelihoods (one grid and
    data set): nlanes contexts of max_batch chains each, fed by a dispatcher
    thread in the library.  submit() returns a ticket at once; wait() blocks
    until that evaluation is done and returns (status, out[3], latency_ms)."""

    def __init__(self, grid, freqs, src, rev, ttime, raystat, max_batch=16, nlanes=2, **kw):
        self.nps = [len(freqs)]
        st, keep = _surf_setup(grid, freqs, src, rev, ttime, raystat, **kw)
        self._keep = keep
        self._g = grid.c()
        h = ctypes.c_void_p()
        _check(lib().mct_serve_create(ctypes.byref(self._g), ctypes.byref(st), int(max_batch), int(nlanes),
                                      ctypes.byref(h)), "mct_serve_create")
        self._h = h

    @classmethod
    def mix(cls, models, max_lanes=2, hbm_budget=0):
        """A server over several models: models is a list of dicts with keys
        grid, freqs, src, rev, ttime, raystat, max_batch and any SurfLikelihood
        setting.  Lanes are admitted within hbm_budget device bytes (0: no
        limit).  submit(..., model=m) names the model."""
        self = cls.__new__(cls)
        setups, keeps, grids = [], [], []
        for m in models:
            m = dict(m)
            g = m.pop("grid")
            mb = m.pop("max_batch", 16)
            st, keep = _surf_setup(g, m.pop("freqs"), m.pop("src"), m.pop("rev"), m.pop("ttime"),
                                   m.pop("raystat"), **m)
            setups.append((st, mb))
            keeps.append(keep)
            grids.append(g.c())
        n = len(models)
        ST = (_SurfSetup * n)(*[s for s, _ in setups])
        GR = (type(grids[0]) * n)(*grids)
        MB = (ctypes.c_int * n)(*[int(b) for _, b in setups])
        self.nps = [int(s.np) for s, _ in setups]
        self._keep = (keeps, ST, GR, MB)
        h = ctypes.c_void_p()
        _check(lib().mct_serve_create_mix(n, GR, ST, MB, int(max_lanes), ctypes.c_int64(int(hbm_budget)),
                                          ctypes.byref(h)), "mct_serve_create_mix")
        self._h = h
        return self

    def submit(self, points, params, snoise0=None, snoise1=None, model=0):
        pts, prm = _f64(points), _f64(params)
        known = 0 <= int(model) < len(self.nps)      # the library rejects other models
        npm = self.nps[model] if known else 1
        s0 = None if snoise0 is None or not known else _f64(np.broadcast_to(np.atleast_1d(snoise0), (npm,)))
        s1 = None if snoise1 is None or not known else _f64(np.broadcast_to(np.atleast_1d(snoise1), (npm,)))
        t = ctypes.c_int64()
        _check(lib().mct_serve_submit_to(self._h, int(model), _p(pts), _p(prm), pts.shape[0], _p(s0), _p(s1),
                                         ctypes.byref(t)), "mct_serve_submit_to")
        return t.value

    def set_fill(self, fill_ms):
        """Longest wait (ms) of a partial batch while another lane runs; < 0:
        launch whatever is pending as soon as a lane is free."""
        _check(lib().mct_serve_set_fill(self._h, ctypes.c_double(float(fill_ms))), "mct_serve_set_fill")

    def memory(self):
        """(device bytes in use, peak bytes, lanes, lane releases)."""
        u, pk, r = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        ln = ctypes.c_int32()
        _check(lib().mct_serve_memory(self._h, ctypes.byref(u), ctypes.byref(pk), ctypes.byref(ln), ctypes.byref(r)),
               "mct_serve_memory")
        return u.value, pk.value, ln.value, r.value

    def wait(self, ticket):
        out = np.zeros(3)
        st, lat = ctypes.c_int32(), ctypes.c_double()
        _check(lib().mct_serve_wait(self._h, ctypes.c_int64(ticket), _p(out), ctypes.byref(st), ctypes.byref(lat)),
               "mct_serve_wait")
        return st.value, out, lat.value

    def stats(self):
        b, e = ctypes.c_int64(), ctypes.c_int64()
        rc = lib().mct_serve_stats(self._h, ctypes.byref(b), ctypes.byref(e))
        return b.value, e.value, rc

    def close(self):
        if getattr(self, "_h", None):
            lib().mct_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
