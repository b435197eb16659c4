"""Ensembles of independent smoke grids of uneven sizes (Part 6 of the C ABI).

Every member has its own grid and physics (a zq3d_config) and steps bit for bit as a SmokeSim3D of
the same configuration with the plain FP32 projection; all members advance together, one launch
per stage over all of them, with the step captured once in a hipGraph and replayed.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .sim import default_config

_SHAPES = {_lib.FIELD_U: lambda W, H, D: ((D, H, W + 1), np.float32),
           _lib.FIELD_V: lambda W, H, D: ((D, H + 1, W), np.float32),
           _lib.FIELD_W: lambda W, H, D: ((D + 1, H, W), np.float32),
           _lib.FIELD_T: lambda W, H, D: ((D, H, W), np.float32),
           _lib.FIELD_RHO: lambda W, H, D: ((D, H, W), np.float32),
           _lib.FIELD_OCC: lambda W, H, D: ((D, H, W), np.uint8)}


def member_config(W, H, D, **overrides):
    """A zq3d_config for one member: default_config(W, H, D) with keyword overrides ("sor_" prefix
    for the projection options)."""
    cfg = default_config(W, H, D)
    for k, val in overrides.items():
        if k.startswith("sor_"):
            setattr(cfg.sor, k[4:], val)
        else:
            setattr(cfg, k, val)
    return cfg


class SmokeEnsemble:
    def __init__(self, configs):
        self.configs = list(configs)
        arr = (_lib.Config * len(self.configs))(*self.configs)
        h = C.c_void_p()
        check(_lib.lib().zq3d_ens_create(len(self.configs), arr, C.byref(h)), "zq3d_ens_create")
        self._h = h
        self.dims = [(c.width, c.height, c.depth) for c in self.configs]

    def close(self):
        if self._h:
            _lib.lib().zq3d_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.configs)

    def cells(self):
        return int(_lib.lib().zq3d_ens_cells(self._h))

    def set_sources(self, member, boxes=(), spheres=()):
        b = np.ascontiguousarray(np.asarray(boxes, np.float32).reshape(-1))
        s = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1))
        check(_lib.lib().zq3d_ens_set_sources(self._h, member, len(b) // 8, b.ctypes.data_as(C.c_void_p) if len(b) else None,
                                              len(s) // 6, s.ctypes.data_as(C.c_void_p) if len(s) else None),
              "zq3d_ens_set_sources")

    def upload(self, member, field, arr):
        shape, dt = _SHAPES[field](*self.dims[member])
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape != shape:
            raise ValueError("member %d field %d expects %s, got %s" % (member, field, shape, a.shape))
        check(_lib.lib().zq3d_ens_upload(self._h, member, field, a.ctypes.data_as(C.c_void_p), a.nbytes),
              "zq3d_ens_upload")

    def download(self, member, field):
        shape, dt = _SHAPES[field](*self.dims[member])
        a = np.empty(shape, dt)
        check(_lib.lib().zq3d_ens_download(self._h, member, field, a.ctypes.data_as(C.c_void_p), a.nbytes),
              "zq3d_ens_download")
        return a

    def set_state(self, member, u=None, v=None, w=None, T=None, rho=None, occ=None):
        for f, a in ((_lib.FIELD_U, u), (_lib.FIELD_V, v), (_lib.FIELD_W, w), (_lib.FIELD_T, T),
                     (_lib.FIELD_RHO, rho), (_lib.FIELD_OCC, occ)):
            if a is not None:
                self.upload(member, f, a)

    def get_state(self, member):
        return {k: self.download(member, f) for k, f in (("u", _lib.FIELD_U), ("v", _lib.FIELD_V),
                                                         ("w", _lib.FIELD_W), ("T", _lib.FIELD_T),
                                                         ("rho", _lib.FIELD_RHO), ("occ", _lib.FIELD_OCC))}

    def step(self, steps=1, graph=True):
        check(_lib.lib().zq3d_ens_step(self._h, int(steps), 1 if graph else 0), "zq3d_ens_step")

    def synchronize(self):
        check(_lib.lib().zq3d_ens_synchronize(self._h), "zq3d_ens_synchronize")

    def autotune(self, iters=10):
        rep = _lib.EnsTune()
        check(_lib.lib().zq3d_ens_autotune(self._h, int(iters), C.byref(rep)), "zq3d_ens_autotune")
        return rep.as_dict()

    def set_sweep(self, kind, tbx, tby, blocks):
        """Selects an autotuner candidate: kind "half_sweeps" (0) or "split_one_launch" (1)."""
        k = {"half_sweeps": 0, "split_one_launch": 1}.get(kind, kind)
        check(_lib.lib().zq3d_ens_set_sweep(self._h, int(k), int(tbx), int(tby), int(blocks)), "zq3d_ens_set_sweep")
