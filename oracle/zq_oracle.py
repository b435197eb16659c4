"""TEST INFRASTRUCTURE ONLY: numpy bindings of the CPU oracle (liboracle_zq3d.so).

The oracle restates the reference's 3D smoke step on the CPU (see zq_oracle.h for
the per-function reference citations).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; the product never does.

Array conventions: numpy arrays in the reference layout, x fastest.  A cell
field has shape (D, H, W); faces u (D, H, W+1), v (D, H+1, W), w (D+1, H, W);
occupancy is uint8 (D, H, W); padded pressure (D+2, H+2, W+2).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_zq3d.so")

ADV_VEL_INREG_OUTREG, ADV_VEL_INREG_OUTREG_BFECC, ADV_VEL_INMAC_OUTMAC, \
    ADV_VEL_INMAC_OUTMAC_BFECC, ADV_VEL_INREG_OUTMAC, ADV_VEL_INREG_OUTMAC_BFECC = range(6)
ADV_SCA_MAC, ADV_SCA_MAC_BFECC, ADV_SCA_REG, ADV_SCA_REG_BFECC = range(4)
OPEN_XLO, OPEN_XHI, OPEN_YLO, OPEN_YHI, OPEN_ZLO, OPEN_ZHI = 1, 2, 4, 8, 16, 32
ALL_FACES = 63

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class StepParams(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("D", C.c_int),
        ("h", C.c_float), ("dt", C.c_float),
        ("steps", C.c_int),
        ("Tamb", C.c_float), ("beta", C.c_float), ("alpha", C.c_float), ("eps", C.c_float),
        ("aV", C.c_float), ("aT", C.c_float), ("aR", C.c_float),
        ("adv_vel_type", C.c_int), ("adv_sca_type", C.c_int),
        ("open_mask", C.c_int), ("closed_mean", C.c_int),
        ("omega", C.c_float), ("sor_iters", C.c_int), ("mixed", C.c_int), ("max_outer", C.c_int),
        ("tol", C.c_double),
        ("n_box", C.c_int), ("boxes", C.c_void_p),
        ("n_sphere", C.c_int), ("spheres", C.c_void_p),
    ]


def build():
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    import subprocess
    subprocess.check_call(["make", "-B", "-s"], cwd=_HERE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i, f, d, u32 = C.c_int, C.c_float, C.c_double, C.c_uint32
        vp = C.c_void_p
        L.orc_mac_to_cell4.argtypes = [i, i, i, _f32p, _f32p, _f32p, _f32p]
        L.orc_cell4_to_mac.argtypes = [i, i, i, _f32p, _f32p, _f32p, _f32p]
        L.orc_sample3.argtypes = [_f32p, i, i, i, f, f, f]
        L.orc_sample3.restype = f
        L.orc_add_force.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, _f32p, _f32p, f, f, f, f, f, f,
                                    vp, vp, vp, vp]
        L.orc_attenuation.argtypes = [i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p, f, f, f]
        L.orc_advect_velocity.argtypes = [i, i, i, f, i, f, _f32p, _f32p, _f32p, _u8p, i, i]
        L.orc_advect_velocity.restype = i
        L.orc_advect_scalar.argtypes = [i, i, i, f, i, f, _f32p, _f32p, _f32p, _u8p, _f32p, _f32p, _f32p, _f32p, i]
        L.orc_advect_scalar.restype = i
        L.orc_backtrace.argtypes = [i, i, i, f, i, f, _f32p, _f32p, _f32p, _u8p, i, f, f, f, _f32p]
        L.orc_divergence.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, _f32p]
        L.orc_sor_count.argtypes = [i, i, i, _u8p, i, _u8p]
        L.orc_sor.argtypes = [i, i, i, _u8p, _f32p, _f32p, f, i]
        L.orc_project.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, i, f, f, i, vp]
        L.orc_project_mp.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, i, i, f, i, i, d, _f64p, _f64p]
        L.orc_project_mp.restype = i
        L.orc_project_cg64.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, i, i, d, i, vp]
        L.orc_project_cg64.restype = i
        L.orc_max_abs_div.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, i, d]
        L.orc_max_abs_div.restype = d
        L.orc_closed_mean_divergence.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, vp]
        L.orc_closed_mean_divergence.restype = f
        L.orc_mean_div32_f64.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p]
        L.orc_mean_div32_f64.restype = f
        L.orc_init_fine_grid.argtypes = [i, i, i, i, i, _u8p, _f32p, _f32p, _f32p]
        L.orc_reinject_box.argtypes = [i, i, i, _f32p, _f32p, _f32p]
        L.orc_reinject_sphere.argtypes = [i, i, i, f, f, f, f, f, f, _f32p, _f32p]
        L.orc_stamp_sphere.argtypes = [i, i, i, f, f, f, f, f, f, f, i, _u8p, _f32p, _f32p, _f32p]
        L.orc_noise.argtypes = [i, i, i, u32, f, _f32p]
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = i
        L.orc_step.argtypes = [C.POINTER(StepParams), _f32p, _f32p, _f32p, _f32p, _f32p, _u8p, vp, vp, vp]
        L.orc_step.restype = i
        _lib = L
    return _lib


def _dims(occ):
    D, H, W = occ.shape
    return W, H, D


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- field helpers -------------------------------------------------------------
def zeros_faces(W, H, D):
    return (np.zeros((D, H, W + 1), np.float32), np.zeros((D, H + 1, W), np.float32),
            np.zeros((D + 1, H, W), np.float32))


def mac_to_cell4(u, v, w):
    D, H, W1 = u.shape
    W = W1 - 1
    out = np.zeros((D, H, W, 4), np.float32)
    lib().orc_mac_to_cell4(W, H, D, u, v, w, out)
    return out


def cell4_to_mac(vel4):
    D, H, W, _ = vel4.shape
    u, v, w = zeros_faces(W, H, D)
    lib().orc_cell4_to_mac(W, H, D, np.ascontiguousarray(vel4, np.float32), u, v, w)
    return u, v, w


def sample3(g, x, y, z):
    nz, ny, nx = g.shape
    return lib().orc_sample3(np.ascontiguousarray(g, np.float32), nx, ny, nz, x, y, z)


def add_force(u, v, w, occ, T, rho, h, dt, beta, eps, Tamb, alpha=0.0, extra_force=None, debug=False):
    """In-place force stage.  With debug=True returns (vort4, gradn3, force3)."""
    W, H, D = _dims(occ)
    ef = None if extra_force is None else np.ascontiguousarray(extra_force, np.float32)
    dbg = None
    if debug:
        dbg = (np.zeros((D, H, W, 4), np.float32), np.zeros((D, H, W, 3), np.float32),
               np.zeros((D, H, W, 3), np.float32))
    lib().orc_add_force(W, H, D, u, v, w, occ, T, rho, h, dt, beta, alpha, eps, Tamb, _ptr(ef),
                        _ptr(dbg[0]) if dbg else None, _ptr(dbg[1]) if dbg else None,
                        _ptr(dbg[2]) if dbg else None)
    return dbg


def attenuation(u, v, w, T, rho, occ, aV, aT, aR):
    W, H, D = _dims(occ)
    lib().orc_attenuation(W, H, D, u, v, w, T, rho, occ, aV, aT, aR)


def advect_velocity(u, v, w, occ, h, steps, dtt, adv_type=ADV_VEL_INREG_OUTREG, is_open=True):
    W, H, D = _dims(occ)
    rc = lib().orc_advect_velocity(W, H, D, h, steps, dtt, u, v, w, occ, int(is_open), adv_type)
    if rc:
        raise ValueError("unsupported advection mode %d" % adv_type)


def advect_scalar(u, v, w, occ, T, rho, h, steps, dtt, adv_type=ADV_SCA_REG):
    W, H, D = _dims(occ)
    To = np.zeros_like(T)
    Ro = np.zeros_like(rho)
    rc = lib().orc_advect_scalar(W, H, D, h, steps, dtt, u, v, w, occ, T, rho, To, Ro, adv_type)
    if rc:
        raise ValueError("unsupported scalar advection mode %d" % adv_type)
    return To, Ro


def backtrace(u, v, w, occ, h, steps, dtt, pos, mac=False):
    W, H, D = _dims(occ)
    out = np.zeros(3, np.float32)
    lib().orc_backtrace(W, H, D, h, steps, dtt, u, v, w, occ, int(mac), pos[0], pos[1], pos[2], out)
    return out


def divergence(u, v, w, occ):
    W, H, D = _dims(occ)
    out = np.zeros((D, H, W), np.float32)
    lib().orc_divergence(W, H, D, u, v, w, occ, out)
    return out


def sor_count(occ, open_mask):
    W, H, D = _dims(occ)
    out = np.zeros((D, H, W), np.uint8)
    lib().orc_sor_count(W, H, D, occ, open_mask, out)
    return out


def sor(ncount, b, p_pad, omega, iters):
    D, H, W = ncount.shape
    lib().orc_sor(W, H, D, ncount, b, p_pad, omega, iters)


def project(u, v, w, occ, open_mask, mean=0.0, omega=1.8, iters=200, want_p=False):
    W, H, D = _dims(occ)
    p = np.zeros((D + 2, H + 2, W + 2), np.float32) if want_p else None
    lib().orc_project(W, H, D, u, v, w, occ, open_mask, mean, omega, iters, _ptr(p))
    return p


def project_mp(u, v, w, occ, open_mask, subtract_mean=False, omega=1.8, inner=50, max_outer=100,
               tol=1e-6, p64=None):
    W, H, D = _dims(occ)
    if p64 is None:
        p64 = np.zeros((D + 2, H + 2, W + 2), np.float64)
    stats = np.zeros(4, np.float64)
    outer = lib().orc_project_mp(W, H, D, u, v, w, occ, open_mask, int(subtract_mean), omega, inner,
                                 max_outer, tol, p64, stats)
    return outer, stats, p64


def project_cg64(u, v, w, occ, open_mask, subtract_mean=False, tol=1e-12, max_iter=100000, want_p=False):
    W, H, D = _dims(occ)
    p = np.zeros((D + 2, H + 2, W + 2), np.float64) if want_p else None
    it = lib().orc_project_cg64(W, H, D, u, v, w, occ, open_mask, int(subtract_mean), tol, max_iter, _ptr(p))
    return it, p


def max_abs_div(u, v, w, occ, open_mask, mean=0.0):
    W, H, D = _dims(occ)
    return lib().orc_max_abs_div(W, H, D, u, v, w, occ, open_mask, mean)


def closed_mean_divergence(u, v, w, occ):
    W, H, D = _dims(occ)
    cnt = C.c_int(0)
    m = lib().orc_closed_mean_divergence(W, H, D, u, v, w, occ, C.cast(C.pointer(cnt), C.c_void_p))
    return m, cnt.value


def mean_div32_f64(u, v, w, occ):
    W, H, D = _dims(occ)
    return lib().orc_mean_div32_f64(W, H, D, u, v, w, occ)


def init_fine_grid(W, H, D, shell_mask=ALL_FACES, intobj=0):
    occ = np.zeros((D, H, W), np.uint8)
    u, v, w = zeros_faces(W, H, D)
    lib().orc_init_fine_grid(W, H, D, shell_mask, intobj, occ, u, v, w)
    return occ, u, v, w


def reinject_box(T, rho, box):
    D, H, W = T.shape
    lib().orc_reinject_box(W, H, D, np.asarray(box, np.float32), T, rho)


def reinject_sphere(T, rho, cx, cy, cz, r, rho_src, T_src):
    D, H, W = T.shape
    lib().orc_reinject_sphere(W, H, D, cx, cy, cz, r, rho_src, T_src, T, rho)


def stamp_sphere(occ, u, v, w, cx, cy, cz, r, vel, shell_mask=ALL_FACES):
    W, H, D = _dims(occ)
    lib().orc_stamp_sphere(W, H, D, cx, cy, cz, r, vel[0], vel[1], vel[2], shell_mask, occ, u, v, w)


def noise(W, H, D, seed, amp):
    out = np.zeros((D, H, W), np.float32)
    lib().orc_noise(W, H, D, seed & 0xFFFFFFFF, amp, out)
    return out


def step(params, u, v, w, T, rho, occ, extra_force=None, p64=None, boxes=None, spheres=None):
    """One forward step in place.  params: dict of StepParams fields (W,H,D taken from occ)."""
    W, H, D = _dims(occ)
    P = StepParams()
    for k, val in params.items():
        setattr(P, k, val)
    P.W, P.H, P.D = W, H, D
    bx = np.ascontiguousarray(np.asarray(boxes if boxes is not None else [], np.float32).reshape(-1))
    sp = np.ascontiguousarray(np.asarray(spheres if spheres is not None else [], np.float32).reshape(-1))
    P.n_box, P.boxes = len(bx) // 8, bx.ctypes.data if len(bx) else None
    P.n_sphere, P.spheres = len(sp) // 6, sp.ctypes.data if len(sp) else None
    stats = np.zeros(4, np.float64)
    ef = None if extra_force is None else np.ascontiguousarray(extra_force, np.float32)
    rc = lib().orc_step(C.byref(P), u, v, w, T, rho, occ, _ptr(ef), _ptr(p64), _ptr(stats))
    if rc:
        raise ValueError("orc_step failed")
    return stats
