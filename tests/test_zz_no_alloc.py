"""hh_validate_and_replace_noalloc (the entry point that allocates nothing)
against hh_validate_and_replace on the same grids: flags and the replaced u,
v bit-identical, over two calls in a row with the second on a smaller grid
(the shared scratch keeps the first call's contents), with and without
alternatives and a grid mask; grids beyond the scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def scene(rng, ny, nx, nalt=2, masked=False):
    u = (2.0 + 0.1 * rng.standard_normal((ny, nx))).astype(np.float32)
    v = (-1.0 + 0.1 * rng.standard_normal((ny, nx))).astype(np.float32)
    bad = rng.random((ny, nx)) < 0.08
    u[bad] += rng.uniform(-5, 5, bad.sum()).astype(np.float32)
    pr = rng.uniform(1.0, 3.0, (ny, nx)).astype(np.float32)
    cm = rng.uniform(0.1, 1.0, (ny, nx)).astype(np.float32)
    au = np.full((ny, nx, nalt), np.nan, np.float32)
    av = np.full((ny, nx, nalt), np.nan, np.float32)
    au[..., 0] = u + rng.uniform(-3, 3, (ny, nx)).astype(np.float32)
    av[..., 0] = v + rng.uniform(-0.2, 0.2, (ny, nx)).astype(np.float32)
    if nalt > 1:
        au[..., 1] = (2.0 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
        av[..., 1] = (-1.0 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
    m = None
    if masked:
        m = np.zeros((ny, nx), bool)
        m[ny // 3:ny // 2, nx // 4:nx // 2] = True
        for a in (u, v, pr, cm):
            a[m] = np.nan
    return u, v, pr, cm, au, av, m


class HipBuffers:
    """Device buffers of the test from the HIP runtime the library is bound
    to: the runtime's functions are looked up through the library's own
    handle (dlsym searches its dependencies), so a second HIP runtime in the
    process (torch's) is never the one used."""

    def __init__(self):
        import __graft_entry__ as ge

        self.rt = ge.load_package()._lib.load()
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        self.live.append(p)
        return p.value

    def get(self, ptr, shape, dtype):
        out = np.empty(int(np.prod(shape)), dtype=dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2) == 0
        return out.reshape(shape, order="F")

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def run_noalloc(hh, ctx, p, u, v, pr, cm, au, av, m, force):
    ny, nx = u.shape
    hb = HipBuffers()
    try:
        du, dv, dpr, dcm = hb.put(u), hb.put(v), hb.put(pr), hb.put(cm)
        dau = hb.put(au) if au is not None else None
        dav = hb.put(av) if av is not None else None
        dm = hb.put(m.astype(np.uint8)) if m is not None else None
        dout = hb.put(np.full((ny, nx), 7, np.uint8))
        ctx.validate_and_replace_noalloc(p, ny, nx, du, dv, dpr, dcm, dau, dav, dm, dout, force_replace=force)
        ctx.synchronize()
        return (hb.get(du, (ny, nx), np.float32), hb.get(dv, (ny, nx), np.float32),
                hb.get(dout, (ny, nx), np.uint8).astype(bool))
    finally:
        hb.free()


@pytest.mark.parametrize("case", ["alternatives", "mask", "plain", "no_replace"])
def test_noalloc_matches_validate_and_replace(hh, ctx, case):
    rng = np.random.default_rng(21)
    p = hh.PIVParameters(uod_neighborhood=2, replace_outliers=case != "no_replace")
    for call, (ny, nx) in enumerate(((61, 53), (23, 19))):  # second call smaller
        u, v, pr, cm, au, av, m = scene(rng, ny, nx, masked=case == "mask")
        if case in ("plain", "no_replace"):
            au = av = None
        force = case != "no_replace"
        ru, rv, rf = ctx.validate_and_replace(p, u, v, pr, cm, au, av, force_replace=force, mask=m)
        gu, gv, gf = run_noalloc(hh, ctx, p, u, v, pr, cm, au, av, m, force)
        assert np.array_equal(gf, rf), (case, call)
        assert np.array_equal(gu, ru, equal_nan=True) and np.array_equal(gv, rv, equal_nan=True), (case, call)
        # the scene exercises the chain: flags remain, or vectors were
        # substituted or replaced
        assert rf.any() or not np.array_equal(ru, u, equal_nan=True)


def test_noalloc_rejects_grids_beyond_the_scratch(hh, ctx):
    ny, nx = 1025, 1024  # one row more than the 1 << 20 cells of scratch
    hb = HipBuffers()
    try:
        bufs = [hb.put(np.zeros((ny, nx), np.float32)) for _ in range(4)]
        out = hb.put(np.zeros((ny, nx), np.uint8))
        with pytest.raises(hh.ArgumentError, match="scratch"):
            ctx.validate_and_replace_noalloc(hh.PIVParameters(), ny, nx, *bufs, d_outliers=out)
    finally:
        hb.free()
