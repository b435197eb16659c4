"""Reference config files (ZQ_SmokeSimulationPara3D::LoadFromFile) from Python, host only."""
import ctypes as C

import numpy as np

from . import _lib

SOLVERS = ["CLOSED_POISSON", "CLOSED_FLUX", "OPEN_POISSON", "OPEN_FLUX", "COARSE_OPEN_POISSON_SOR",
           "COARSE_CLOSED_POISSON_SOR", "COARSE_OPEN_FLUX_SOR", "COARSE_CLOSED_FLUX_SOR", "TOTAL_OPEN_POISSON_SOR",
           "TOTAL_CLOSED_POISSON_SOR", "TOTAL_OPEN_FLUX_SOR", "TOTAL_CLOSED_FLUX_SOR", "CLOSED_OCTREE_POISSON",
           "CLOSED_OCTREE_FLUX", "OPEN_OCTREE_POISSON", "OPEN_OCTREE_FLUX", "OPEN_OCTREE_POISSON_SOR",
           "CLOSED_OCTREE_POISSON_SOR"]


def load_scene(path):
    """Parse a config; returns the zq3d_scene structure, or raises ValueError when rejected."""
    sc = _lib.Scene()
    rc = _lib.lib().zq3d_scene_load(str(path).encode(), C.byref(sc))
    if rc != 0:
        raise ValueError("config %s rejected" % path)
    return sc


def initial_occupancy(path):
    sc = load_scene(path)
    W, H, D = sc.config.width, sc.config.height, sc.config.depth
    occ = np.zeros((D, H, W), np.uint8)
    rc = _lib.lib().zq3d_scene_init_occupancy(str(path).encode(), occ.ctypes.data_as(C.c_void_p), occ.nbytes)
    if rc != 0:
        raise ValueError("ZQ_InitFineGrid3D failed for %s" % path)
    return occ


def apply_moving_objects(path, frames, u, v, w):
    """ZQ_ApplyMovingObject3D after `frames` object updates; returns occupancy, updates u, v, w."""
    sc = load_scene(path)
    W, H, D = sc.config.width, sc.config.height, sc.config.depth
    occ = np.zeros((D, H, W), np.uint8)
    for a in (u, v, w):
        if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
            raise TypeError("faces must be C-contiguous float32")
    rc = _lib.lib().zq3d_scene_apply_moving_objects(str(path).encode(), int(frames), occ.ctypes.data_as(C.c_void_p),
                                                    u.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                                    w.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError("moving objects of %s could not be applied" % path)
    return occ


def read_di3(path, header_only=False):
    """A .di3 density file (ZQ_SmokeGrid3D::ExportDensity's layout): returns (width, height, depth,
    channels) and, unless header_only, the samples as a float32 array (depth, height, width[, channels])."""
    dims = (C.c_int * 4)()
    rc = _lib.lib().zq3d_di3_read(str(path).encode(), C.byref(dims), None, 0)
    if rc != 0:
        raise ValueError("%s is not a whole .di3 file" % path)
    W, H, D, nch = (int(x) for x in dims)
    if header_only:
        return (W, H, D, nch), None
    data = np.empty(W * H * D * nch, np.float32)
    rc = _lib.lib().zq3d_di3_read(str(path).encode(), C.byref(dims), data.ctypes.data_as(C.c_void_p), data.size)
    if rc != 0:
        raise ValueError("%s could not be read" % path)
    shape = (D, H, W) if nch == 1 else (D, H, W, nch)
    return (W, H, D, nch), data.reshape(shape)
