"""ZQ3D_SolvePoissonRedBlack_noalloc (device pointers, static scratch, no allocation) against
ZQ3D_SolvePoissonRedBlack_ex (host pointers, its own workspace): the same kernels, so the same bits.
Two calls in a row per mode, the second on a smaller grid, so the second call runs on scratch that
still holds the first call's contents."""
import ctypes as C

import numpy as np
import pytest

from util import faces, occupancy

torch = pytest.importorskip("torch")


def _noalloc(L, u, v, w, occ, opts, stream=None):
    D, H, W = occ.shape
    dev = [torch.from_numpy(a.copy()).cuda() for a in (u, v, w)]
    docc = torch.from_numpy(occ.copy()).cuda()
    torch.cuda.synchronize()
    st = L.ProjStats()
    rc = L.lib().ZQ3D_SolvePoissonRedBlack_noalloc(*(C.c_void_p(t.data_ptr()) for t in dev),
                                                   C.c_void_p(docc.data_ptr()), W, H, D, C.byref(opts),
                                                   C.byref(st), stream)
    L.check(rc, "ZQ3D_SolvePoissonRedBlack_noalloc")
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in dev], st.as_dict()


def _ex(L, u, v, w, occ, opts):
    D, H, W = occ.shape
    out = [a.copy() for a in (u, v, w)]
    st = L.ProjStats()
    ms = L.lib().ZQ3D_SolvePoissonRedBlack_ex(*(a.ctypes.data_as(C.c_void_p) for a in out),
                                              occ.ctypes.data_as(C.c_void_p), W, H, D, C.byref(opts), C.byref(st))
    L.check(ms, "ZQ3D_SolvePoissonRedBlack_ex")
    return out, st.as_dict()


CASES = [
    # (shape, occupancy fraction, walls, options)
    ((48, 40, 44), 0.1, False, dict(open_mask=63, closed_mean=0, omega=1.8, iters=60, mixed_precision=0,
                                    max_outer=0, tol=0.0)),
    ((21, 17, 13), 0.1, False, dict(open_mask=8, closed_mean=0, omega=1.7, iters=35, mixed_precision=0,
                                    max_outer=0, tol=0.0)),
    ((64, 64, 64), 0.05, True, dict(open_mask=0, closed_mean=1, omega=1.8, iters=40, mixed_precision=1,
                                    max_outer=300, tol=1e-6)),
    ((30, 26, 22), 0.1, True, dict(open_mask=0, closed_mean=1, omega=1.7, iters=25, mixed_precision=1,
                                   max_outer=300, tol=1e-7)),
]


@pytest.mark.gpu
def test_noalloc_matches_ex_on_consecutive_calls(rng):
    from zq_smokesimulation_amd import _lib as L
    for (W, H, D), frac, walls, o in CASES:
        print("noalloc vs ex", (W, H, D), "mixed" if o["mixed_precision"] else "fp32", flush=True)
        u, v, w = faces(rng, W, H, D)
        occ = occupancy(rng, W, H, D, frac, shell=walls)
        opts = L.SorOpts(**o)
        want, st_want = _ex(L, u, v, w, occ, opts)
        got, st_got = _noalloc(L, u, v, w, occ, opts)
        for a, b, n in zip(got, want, "uvw"):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "%s differs at %s" % (n, (W, H, D))
        assert st_got == st_want
        if o["mixed_precision"]:
            assert st_got["max_residual"] <= o["tol"]


@pytest.mark.gpu
def test_noalloc_on_a_torch_stream(rng):
    from zq_smokesimulation_amd import _lib as L
    W, H, D = 33, 29, 31
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    opts = L.SorOpts(open_mask=63, closed_mean=0, omega=1.8, iters=50, mixed_precision=0, max_outer=0, tol=0.0)
    want, _ = _ex(L, u, v, w, occ, opts)
    s = torch.cuda.Stream()
    dev = [torch.from_numpy(a.copy()).cuda() for a in (u, v, w)]
    docc = torch.from_numpy(occ.copy()).cuda()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        rc = L.lib().ZQ3D_SolvePoissonRedBlack_noalloc(*(C.c_void_p(t.data_ptr()) for t in dev),
                                                       C.c_void_p(docc.data_ptr()), W, H, D, C.byref(opts), None,
                                                       C.c_void_p(s.cuda_stream))
    L.check(rc, "ZQ3D_SolvePoissonRedBlack_noalloc")
    s.synchronize()
    for t, b, n in zip(dev, want, "uvw"):
        assert np.array_equal(t.cpu().numpy().view(np.uint32), b.view(np.uint32)), n
