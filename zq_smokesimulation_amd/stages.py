"""numpy front end of the reference-compatible stage entry points (Part 1/2 of the C ABI).

Each function takes numpy arrays in the reference layout (x fastest: a cell
field is (D, H, W), u faces (D, H, W+1), v (D, H+1, W), w (D+1, H, W), the
occupancy uint8/bool (D, H, W)), updates them in place like the reference's
host wrappers, and returns the elapsed milliseconds reported by the library.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _p(a, dtype):
    if a.dtype != dtype or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous %s array, got %s" % (np.dtype(dtype), a.dtype))
    return a.ctypes.data_as(C.c_void_p)


def _occ(occ):
    if occ.dtype == np.bool_:
        return _p(occ, np.bool_)
    return _p(occ, np.uint8)


def _dims_from_occ(occ):
    D, H, W = occ.shape
    return W, H, D


def _check_faces(u, v, w, W, H, D):
    if u.shape != (D, H, W + 1) or v.shape != (D, H + 1, W) or w.shape != (D + 1, H, W):
        raise ValueError("face arrays do not match the grid %dx%dx%d" % (W, H, D))


def prepare_advection(W, H, D, voxel_size, steps, dtt):
    """ZQ_Cuda_Prepare_Advection (ZQ_CUDA_Advection3D.h:23)."""
    _lib.lib().ZQ_Cuda_Prepare_Advection(W, H, D, voxel_size, steps, dtt)


def advect_velocity(u, v, w, occ, adv_type=_lib.ADV_VEL_INREG_OUTREG, is_open=True):
    """Advect_Velocity (ZQ_CUDA_Advection3D.h:31); call prepare_advection first."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    return check(_lib.lib().Advect_Velocity(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ),
                                            bool(is_open), int(adv_type)), "Advect_Velocity")


def advect_scalar(u, v, w, occ, T_in, rho_in, T_out, rho_out, adv_type=_lib.ADV_SCA_REG):
    """Advect_Scalar (ZQ_CUDA_Advection3D.h:43-44)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    for a in (T_in, rho_in, T_out, rho_out):
        if a.shape != (D, H, W):
            raise ValueError("scalar field shape mismatch")
    return check(_lib.lib().Advect_Scalar(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ),
                                          _p(T_in, np.float32), _p(rho_in, np.float32), _p(T_out, np.float32),
                                          _p(rho_out, np.float32), int(adv_type)), "Advect_Scalar")


def add_force(u, v, w, occ, T, h, dt, buoy, confine, Tamb, force_type=_lib.ADD_FORCE_ENTIRE):
    """AddForce3D (ZQ_CUDA_AddForce3D.h:12-14)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    return check(_lib.lib().AddForce3D(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ),
                                       _p(T, np.float32), h, dt, buoy, confine, Tamb, W, H, D, int(force_type)),
                 "AddForce3D")


def attenuation(u, v, w, T, rho, occ, aV, aT, aR):
    """Attenuation3D (ZQ_CUDA_Attenuation3D.h:6-8)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    return check(_lib.lib().Attenuation3D(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32),
                                          _p(T, np.float32), _p(rho, np.float32), _occ(occ), aV, aT, aR, W, H, D),
                 "Attenuation3D")


def solve_open(u, v, w, occ, max_iter):
    """SolveOpenPoissonRedBlackwithOccupy3D_MAC (call site ZQ_SmokeSimulationUtils3D.cpp:2354)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    return check(_lib.lib().SolveOpenPoissonRedBlackwithOccupy3D_MAC(
        _p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ), W, H, D, int(max_iter)), "SolveOpen")


def solve_closed(u, v, w, occ, div_per_volume, max_iter):
    """SolveClosedPoissonRedBlackwithOccupy3D_MAC (call site ZQ_SmokeSimulationUtils3D.cpp:2424)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    return check(_lib.lib().SolveClosedPoissonRedBlackwithOccupy3D_MAC(
        _p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ), float(div_per_volume), W, H, D,
        int(max_iter)), "SolveClosed")


def solve_open_face_ratio(u, v, w, occ, rU, rV, rW, max_iter):
    """SolveOpenPoissonRedBlackwithFaceRatio3D_MAC (call site ZQ_SmokeSimulationUtils3D.cpp:2630):
    the coarse grid's faces projected in place; rU, rV, rW are the face open fractions (face shapes)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    _check_faces(rU, rV, rW, W, H, D)
    return check(_lib.lib().SolveOpenPoissonRedBlackwithFaceRatio3D_MAC(
        _p(u, np.float32), _p(v, np.float32), _p(w, np.float32), _occ(occ), _p(rU, np.float32), _p(rV, np.float32),
        _p(rW, np.float32), W, H, D, int(max_iter)), "SolveOpenFaceRatio")


def solve_octree(u, v, w, leaves, max_iter, div_per_volume=None):
    """SolveOpenOctreePoissonRedBlack3_3D_MAC / SolveClosedOctreePoissonRedBlack3_3D_MAC (call sites
    ZQ_SmokeOctreeGrid3D.cpp:2897, 3000): the faces corrected in place from the leaf flags of the four
    levels (uint8 arrays of shapes (D >> L, H >> L, W >> L)); div_per_volume selects the closed form.
    The index / neighbour lists are not read by the library: empty ones are passed."""
    l0 = leaves[0]
    D, H, W = l0.shape
    _check_faces(u, v, w, W, H, D)
    for L, a in enumerate(leaves):
        if a.shape != (D >> L, H >> L, W >> L):
            raise ValueError("leaf level %d has shape %s" % (L, a.shape))
    ls = [np.ascontiguousarray(a, np.uint8) for a in leaves]
    lp = [a.ctypes.data_as(C.c_void_p) for a in ls]
    empty = [0, None, 0, None] * 8
    fp = (_p(u, np.float32), _p(v, np.float32), _p(w, np.float32))
    if div_per_volume is None:
        ms = _lib.lib().SolveOpenOctreePoissonRedBlack3_3D_MAC(*fp, *lp, W, H, D, int(max_iter), *empty)
    else:
        ms = _lib.lib().SolveClosedOctreePoissonRedBlack3_3D_MAC(*fp, *lp, float(div_per_volume), W, H, D,
                                                                 int(max_iter), *empty)
    return check(ms, "SolveOctree")


def matrix_mul(A, B):
    """ZQ_CUDA_MatrixMul::MatrixMul (called by SubProjection3DCuda, ZQ_SmokeSimulationUtils3D.cpp:13-16):
    returns A @ B for float32 matrices, each entry summed over k in order."""
    if A.ndim != 2 or B.ndim != 2 or A.shape[1] != B.shape[0]:
        raise ValueError("shapes %s and %s do not multiply" % (A.shape, B.shape))
    out = np.zeros((A.shape[0], B.shape[1]), np.float32)
    check(_lib.lib().MatrixMul(_p(A, np.float32), _p(B, np.float32), A.shape[0], A.shape[1], B.shape[1],
                               _p(out, np.float32)), "MatrixMul")
    return out


def solve_ex(u, v, w, occ, open_mask=_lib.OPEN_ALL, closed_mean=False, omega=1.0, iters=100, mixed=False,
             max_outer=0, tol=1e-6, pass_rule=0):
    """ZQ3D_SolvePoissonRedBlack_ex: returns (ms, stats dict)."""
    W, H, D = _dims_from_occ(occ)
    _check_faces(u, v, w, W, H, D)
    o = _lib.SorOpts(open_mask, int(closed_mean), omega, int(iters), int(mixed), int(max_outer), tol,
                     int(pass_rule))
    st = _lib.ProjStats()
    ms = check(_lib.lib().ZQ3D_SolvePoissonRedBlack_ex(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32),
                                                       _occ(occ), W, H, D, C.byref(o), C.byref(st)), "Solve_ex")
    return ms, st.as_dict()


def mac_to_regular4(u, v, w):
    """ZQ3D_MACtoRegular4: cell-centred (u, v, w, 0) float4 grid, shape (D, H, W, 4)."""
    D, H, W1 = u.shape
    W = W1 - 1
    out = np.zeros((D, H, W, 4), np.float32)
    check(_lib.lib().ZQ3D_MACtoRegular4(_p(u, np.float32), _p(v, np.float32), _p(w, np.float32),
                                        _p(out, np.float32), W, H, D), "MACtoRegular4")
    return out
