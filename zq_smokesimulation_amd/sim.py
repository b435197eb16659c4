"""Device-resident smoke simulation (Part 3 of the C ABI) with a numpy front end."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

# hash noise shared by the product and its tests: lowbias32 finalizer on (index, seed)
_M = 0xFFFFFFFF


def _hash_u32(x):
    x = np.asarray(x, np.uint64) & _M
    x ^= x >> 16
    x = (x * 0x7FEB352D) & _M
    x ^= x >> 15
    x = (x * 0x846CA68B) & _M
    x ^= x >> 16
    return x


def noise_field(W, H, D, seed, amp):
    """Uniform [-amp, amp] per cell from a counter-based hash of (cell index, seed)."""
    idx = np.arange(W * H * D, dtype=np.uint64)
    s = int(_hash_u32(np.uint64((seed + 0x632BE5AB) & _M)))
    x = _hash_u32(((idx * 0x9E3779B9) & _M) ^ np.uint64(s))
    u01 = (x >> np.uint64(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    out = np.float32(amp) * (np.float32(2.0) * u01 - np.float32(1.0))
    return out.astype(np.float32).reshape(D, H, W)


def default_config(W, H, D):
    cfg = _lib.Config()
    check(_lib.lib().zq3d_config_default(C.byref(cfg), W, H, D), "zq3d_config_default")
    return cfg


class SmokeSim3D:
    """One simulation with every field resident in HBM.

    config: a _lib.Config, or keyword overrides applied to default_config(W, H, D).
    Keys of the nested projection options are prefixed "sor_" (sor_omega, sor_iters, ...).
    """

    def __init__(self, W, H, D, **overrides):
        cfg = default_config(W, H, D)
        for k, val in overrides.items():
            if k.startswith("sor_"):
                setattr(cfg.sor, k[4:], val)
            else:
                setattr(cfg, k, val)
        self.cfg = cfg
        self.W, self.H, self.D = W, H, D
        h = C.c_void_p()
        check(_lib.lib().zq3d_create(C.byref(cfg), C.byref(h)), "zq3d_create")
        self._h = h
        self.last_stats = None

    def close(self):
        if self._h:
            _lib.lib().zq3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- field I/O --------------------------------------------------------------
    def _shape(self, field):
        W, H, D = self.W, self.H, self.D
        return {
            _lib.FIELD_U: ((D, H, W + 1), np.float32), _lib.FIELD_V: ((D, H + 1, W), np.float32),
            _lib.FIELD_W: ((D + 1, H, W), np.float32), _lib.FIELD_T: ((D, H, W), np.float32),
            _lib.FIELD_RHO: ((D, H, W), np.float32), _lib.FIELD_OCC: ((D, H, W), np.uint8),
            _lib.FIELD_P64: ((D + 2, H + 2, W + 2), np.float64), _lib.FIELD_VEL4: ((D, H, W, 4), np.float32),
        }[field]

    def upload(self, field, arr):
        shape, dt = self._shape(field)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape != shape:
            raise ValueError("field %d expects shape %s, got %s" % (field, shape, a.shape))
        check(_lib.lib().zq3d_upload(self._h, field, a.ctypes.data_as(C.c_void_p), a.nbytes), "zq3d_upload")

    def download(self, field):
        shape, dt = self._shape(field)
        a = np.empty(shape, dt)
        check(_lib.lib().zq3d_download(self._h, field, a.ctypes.data_as(C.c_void_p), a.nbytes), "zq3d_download")
        return a

    def device_ptr(self, field):
        p = C.c_void_p()
        n = C.c_size_t()
        check(_lib.lib().zq3d_device_ptr(self._h, field, C.byref(p), C.byref(n)), "zq3d_device_ptr")
        return p.value, n.value

    def set_state(self, u=None, v=None, w=None, T=None, rho=None, occ=None):
        for f, a in ((_lib.FIELD_U, u), (_lib.FIELD_V, v), (_lib.FIELD_W, w), (_lib.FIELD_T, T),
                     (_lib.FIELD_RHO, rho), (_lib.FIELD_OCC, occ)):
            if a is not None:
                self.upload(f, a)

    def get_state(self):
        return {k: self.download(f) for k, f in (("u", _lib.FIELD_U), ("v", _lib.FIELD_V), ("w", _lib.FIELD_W),
                                                 ("T", _lib.FIELD_T), ("rho", _lib.FIELD_RHO),
                                                 ("occ", _lib.FIELD_OCC))}

    # ---- sources and obstacles ----------------------------------------------------
    def set_box_sources(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, np.float32).reshape(-1, 8))
        check(_lib.lib().zq3d_set_box_sources(self._h, len(b), b.ctypes.data_as(C.c_void_p)), "set_box_sources")
        self._boxes = b

    def set_sphere_sources(self, spheres):
        s = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 6))
        check(_lib.lib().zq3d_set_sphere_sources(self._h, len(s), s.ctypes.data_as(C.c_void_p)),
              "set_sphere_sources")
        self._spheres = s

    def stamp_sphere_obstacle(self, cx, cy, cz, r, vel=(0.0, 0.0, 0.0), shell_mask=_lib.OPEN_ALL):
        check(_lib.lib().zq3d_stamp_sphere_obstacle(self._h, cx, cy, cz, r, vel[0], vel[1], vel[2], shell_mask),
              "stamp_sphere_obstacle")

    # ---- stages -------------------------------------------------------------------
    def add_force_attenuate(self, extra_force_dev=None):
        check(_lib.lib().zq3d_add_force_attenuate(self._h, extra_force_dev), "add_force_attenuate")

    def advect_velocity(self):
        check(_lib.lib().zq3d_advect_velocity(self._h), "advect_velocity")

    def project(self):
        st = _lib.ProjStats()
        check(_lib.lib().zq3d_project(self._h, C.byref(st)), "project")
        self.last_stats = st.as_dict()
        return self.last_stats

    def advect_scalar(self):
        check(_lib.lib().zq3d_advect_scalar(self._h), "advect_scalar")

    def reinject(self):
        check(_lib.lib().zq3d_reinject(self._h), "reinject")

    def step(self, extra_force_dev=None, stats=True):
        st = _lib.ProjStats()
        check(_lib.lib().zq3d_step(self._h, extra_force_dev, C.byref(st) if stats else None), "step")
        self.last_stats = st.as_dict() if stats else None
        return self.last_stats

    def mac_to_regular(self):
        check(_lib.lib().zq3d_mac_to_regular(self._h), "mac_to_regular")

    def synchronize(self):
        check(_lib.lib().zq3d_synchronize(self._h), "synchronize")

    def set_profiling(self, on=True):
        check(_lib.lib().zq3d_set_profiling(self._h, int(on)), "set_profiling")

    def timing(self):
        t = _lib.Timing()
        check(_lib.lib().zq3d_get_timing(self._h, C.byref(t)), "get_timing")
        return t.as_dict()

    def autotune_sor(self, iters=20):
        """Time every fused-sweep tile shape on this grid and keep the fastest; returns the report."""
        rep = _lib.TuneReport()
        check(_lib.lib().zq3d_autotune_sor(self._h, int(iters), C.byref(rep)), "autotune_sor")
        return rep.as_dict()

    def sor_kernel(self):
        return _lib.lib().zq3d_sor_kernel(self._h).decode()

    def set_sor_shape(self, ty, zc):
        check(_lib.lib().zq3d_set_sor_shape(self._h, int(ty), int(zc)), "set_sor_shape")

    def set_guide(self, coeff, gu=None, gv=None, gw=None):
        """Guiding control of the coarse projection: coarse-grid guide faces (None: off).  On an octree
        simulation any gu other than None turns on the special-region projections, which read no guide
        values (ZQ_SmokeOctreeGrid3D::ApplyGuidingVelocity's CUDA branch)."""
        if gu is None:
            check(_lib.lib().zq3d_set_guide(self._h, 0.0, None, None, None), "set_guide")
            self._guide = None
            return
        if self.cfg.octree:
            flag = np.zeros(1, np.float32)
            check(_lib.lib().zq3d_set_guide(self._h, float(coeff), flag.ctypes.data_as(C.c_void_p), None, None),
                  "set_guide")
            self._guide = None
            return
        arrs = [np.ascontiguousarray(a, np.float32) for a in (gu, gv, gw)]
        s = self.cfg.sub_size
        Wc, Hc, Dc = self.W // s, self.H // s, self.D // s
        if [a.shape for a in arrs] != [(Dc, Hc, Wc + 1), (Dc, Hc + 1, Wc), (Dc + 1, Hc, Wc)]:
            raise ValueError("guide faces must have the coarse grid's face shapes")
        check(_lib.lib().zq3d_set_guide(self._h, float(coeff), *(a.ctypes.data_as(C.c_void_p) for a in arrs)),
              "set_guide")
        self._guide = arrs

    def set_coarse_sweep(self, k, zc):
        """Coarse SOR schedule of the coarse + sub-grid projection (k == 0: one launch per colour)."""
        check(_lib.lib().zq3d_set_coarse_sweep(self._h, int(k), int(zc)), "set_coarse_sweep")

    def coarse_sweep(self):
        k, zc = C.c_int(), C.c_int()
        check(_lib.lib().zq3d_get_coarse_sweep(self._h, C.byref(k), C.byref(zc)), "coarse_sweep")
        return k.value, zc.value

    def set_stream(self, stream_handle):
        check(_lib.lib().zq3d_set_stream(self._h, stream_handle), "set_stream")
