"""The allocation-free 2-D entry (mct_modrays_dev_noalloc) against the
allocating one (mct_modrays) through the C ABI: the same phase times bit for
bit on two calls in a row, the second smaller than the first (the scratch
keeps the first call's contents), for the parallel and sequential methods; a
request too large for the scratch is MCT_ERR_SCRATCH with nothing launched."""
import ctypes

import numpy as np
import pytest

import mctomo_amd as mct

pytestmark = pytest.mark.gpu


def case(seed, nv, npd, ns, nr, fom):
    rng = np.random.default_rng(seed)
    h = 0.5
    velv = rng.uniform(2.6, 3.6, (npd, nv + 2, nv + 2))
    L = (nv - 1) * h
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, L - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, L, nr)
    srs = (rng.random((npd, ns, nr)) < 0.8).astype(np.int32)
    return dict(velv=velv, h=h, sx=sx, sz=sz, rx=rx, rz=rz, srs=srs, fom=fom)


def noalloc(c, method):
    L = mct.lib()
    npd, nv2 = c["velv"].shape[0], c["velv"].shape[1]
    ns, nr = c["sx"].shape[0], c["rx"].shape[0]
    dv = [mct.DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in
          (c["velv"], c["sx"], c["sz"], c["rx"], c["rz"], c["srs"])]
    tt0 = np.full((npd, ns, nr), -1.0)
    dt = mct.DeviceArray.from_numpy(tt0)
    st = mct.DeviceArray.from_numpy(np.array([7], np.int32))
    vp = ctypes.c_void_p
    rc = L.mct_modrays_dev_noalloc(nv2 - 2, nv2 - 2, ctypes.c_double(0.0), ctypes.c_double(0.0),
                                   ctypes.c_double(c["h"]), ctypes.c_double(c["h"]), npd, vp(dv[0].ptr), 1, 1, 1, 4,
                                   8, c["fom"], ns, vp(dv[1].ptr), vp(dv[2].ptr), nr, vp(dv[3].ptr),
                                   vp(dv[4].ptr), vp(dv[5].ptr), vp(dt.ptr), vp(st.ptr), method, None)
    assert rc == 0, rc
    assert L.mct_copy_d2h(vp(tt0.ctypes.data), vp(dt.ptr), ctypes.c_int64(tt0.nbytes)) == 0   # waits
    return tt0, int(st.numpy()[0])


def allocating(c, method):
    return mct.modrays(c["velv"], 0.0, 0.0, c["h"], c["h"], c["sx"], c["sz"], c["rx"], c["rz"], srs=c["srs"],
                       fom=c["fom"], ttime=np.full(c["srs"].shape, -1.0), method=method)


@pytest.mark.parametrize("method", [0, 1])
def test_noalloc_equals_allocating_entry_on_two_calls(method):
    big = case(1, 61, 4, 6, 9, 1)
    small = case(2, 33, 2, 3, 5, 0)
    for c in (big, small):
        got, st = noalloc(c, method)
        want = allocating(c, method)
        assert st == 0
        np.testing.assert_array_equal(got, want)


def test_noalloc_rejects_requests_beyond_its_scratch():
    L = mct.lib()
    b = ctypes.c_int64()
    assert L.mct_noalloc_scratch_bytes(ctypes.byref(b)) == 0 and b.value > 0
    # velocities of 20 periods on a 2001 x 2001 coarse grid alone exceed the scratch
    n = 2001
    assert 8 * n * n * 20 > b.value
    d = mct.DeviceArray((8,), np.float64)
    i = mct.DeviceArray((8,), np.int32)
    vp = ctypes.c_void_p
    rc = L.mct_modrays_dev_noalloc(n, n, ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.1),
                                   ctypes.c_double(0.1), 20, vp(d.ptr), 1, 1, 1, 4, 8, 0, 1, vp(d.ptr), vp(d.ptr),
                                   1, vp(d.ptr), vp(d.ptr), vp(i.ptr), vp(d.ptr), vp(i.ptr), 0, None)
    assert rc == -8
