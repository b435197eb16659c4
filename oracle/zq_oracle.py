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
# ZQ3D_ORACLE_LIB may name another build of the same sources (the sanitizer build)
_LIB_PATH = os.environ.get("ZQ3D_ORACLE_LIB") or os.path.join(_HERE, "liboracle_zq3d.so")

ADV_VEL_INREG_OUTREG, ADV_VEL_INREG_OUTREG_BFECC, ADV_VEL_INMAC_OUTMAC, \
    ADV_VEL_INMAC_OUTMAC_BFECC, ADV_VEL_INREG_OUTMAC, ADV_VEL_INREG_OUTMAC_BFECC = range(6)
ADV_SCA_MAC, ADV_SCA_MAC_BFECC, ADV_SCA_REG, ADV_SCA_REG_BFECC = range(4)
OPEN_XLO, OPEN_XHI, OPEN_YLO, OPEN_YHI, OPEN_ZLO, OPEN_ZHI = 1, 2, 4, 8, 16, 32
ALL_FACES = 63

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


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
        ("pass_rule", C.c_int),
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
        _lib = load(_LIB_PATH)
    return _lib


def load(path):
    """Open an oracle build at `path` and declare its signatures (the tests also load builds of
    deliberately altered sources this way, to show that their pins reject them)."""
    L = C.CDLL(path)
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
    L.orc_project_mp.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, i, i, f, i, i, d, _f64p, _f64p, i]
    L.orc_project_mp.restype = i
    L.orc_fitted_pass_length.argtypes = [i, i, d, d, i, d]
    L.orc_fitted_pass_length.restype = i
    L.orc_extrapolate_warm_start.argtypes = [C.c_size_t, _f64p, _f64p, i, d]
    L.orc_sor_gated.argtypes = [i, i, i, _u8p, _f32p, _f32p, f, i, f]
    L.orc_sor_gated.restype = i
    L.orc_gate_threshold.argtypes = [d, f]
    L.orc_gate_threshold.restype = f
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
    L.orc_set_num_threads.argtypes = [i]
    L.orc_set_num_threads.restype = None
    L.orc_step.argtypes = [C.POINTER(StepParams), _f32p, _f32p, _f32p, _f32p, _f32p, _u8p, vp, vp, vp]
    L.orc_step.restype = i
    L.orc_coarse_setup.argtypes = [i, i, i, i, _u8p, _f32p, _f32p, _f32p, _u8p, _u8p]
    L.orc_coarse_inv_diag.argtypes = [i, i, i, i, i, i, _f32p, _f32p, _f32p, i, i]
    L.orc_coarse_inv_diag.restype = f
    L.orc_fine_to_coarse.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_coarse_project.argtypes = [i, i, i, _f32p, _f32p, _f32p, _u8p, _f32p, _f32p, _f32p, i, f, i,
                                     _f32p, _f32p, _f32p, _f32p, vp, i]
    L.orc_coarse_to_fine.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _u8p, _f32p, _f32p, _f32p]
    L.orc_sub_project_regular.argtypes = [i, i, i, i, _f32p, _u8p, _f32p, _f32p, _f32p]
    L.orc_oct_distance.argtypes = [i, i, i, _u8p, _i32p]
    L.orc_oct_leaves.argtypes = [i, i, i, _u8p, _i32p, i, i, i, _u8p, _u8p, _u8p, _u8p]
    L.orc_oct_neighbors.argtypes = [i, i, i, _u8p, _u8p, _u8p, _u8p, i, i, i, i, _i32p]
    L.orc_oct_neighbors.restype = i
    L.orc_oct_divergence.argtypes = [i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_oct_closed_mean.argtypes = [i, i, i, _u8p, _f32p, _f32p, _f32p]
    L.orc_oct_closed_mean.restype = f
    L.orc_oct_solve.argtypes = [i, i, i, _u8p, _u8p, _u8p, _u8p, _f32p, _f32p, _f32p, _f32p, i, i, f,
                                _f32p, _f32p, _f32p, _f32p]
    L.orc_oct_adjust.argtypes = [i, i, i, _u8p, _u8p, _u8p, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                 _f32p]
    L.orc_oct_special_regions.argtypes = [i, i, i, _u8p, _u8p, i, _f32p, _f32p, _f32p, _i32p]
    L.orc_oct_special_regions.restype = i
    return L


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


def sor_gated(ncount, b, p_pad, omega, iters, thresh):
    D, H, W = ncount.shape
    return int(lib().orc_sor_gated(W, H, D, ncount, b, p_pad, omega, iters, thresh))


def extrapolate_warm_start(p64, prev, valid, beta=0.5):
    """In place: p64 = p64 + beta (p64 - prev) (beta as float32, as the device config holds it),
    prev = the old p64 (valid False: prev = p64 only)."""
    assert p64.shape == prev.shape and p64.dtype == prev.dtype == np.float64
    lib().orc_extrapolate_warm_start(p64.size, p64, prev, int(valid), float(np.float32(beta)))


def gate_threshold(tol, omega):
    return float(lib().orc_gate_threshold(tol, omega))


def project(u, v, w, occ, open_mask, mean=0.0, omega=1.8, iters=200, want_p=False):
    W, H, D = _dims(occ)
    p = np.zeros((D + 2, H + 2, W + 2), np.float32) if want_p else None
    lib().orc_project(W, H, D, u, v, w, occ, open_mask, mean, omega, iters, _ptr(p))
    return p


def project_mp(u, v, w, occ, open_mask, subtract_mean=False, omega=1.8, inner=50, max_outer=100,
               tol=1e-6, p64=None, pass_rule=0):
    W, H, D = _dims(occ)
    if p64 is None:
        p64 = np.zeros((D + 2, H + 2, W + 2), np.float64)
    stats = np.zeros(4, np.float64)
    outer = lib().orc_project_mp(W, H, D, u, v, w, occ, open_mask, int(subtract_mean), omega, inner,
                                 max_outer, tol, p64, stats, int(pass_rule))
    return outer, stats, p64


def fitted_pass_length(iters, fit, r, rprev, nprev, tol):
    return int(lib().orc_fitted_pass_length(iters, int(fit), r, rprev, nprev, tol))


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


# ---- coarse + sub-grid projection (zq_oracle_coarse.cpp; ZQ_SmokeSimulationUtils3D.cpp:3688-3795) ----
def coarse_dims(occ, s):
    W, H, D = _dims(occ)
    if s < 1 or W % s or H % s or D % s:
        raise ValueError("block size %d does not divide %dx%dx%d" % (s, W, H, D))
    return W // s, H // s, D // s


def coarse_setup(occ, s):
    """Coarse face open fractions rU, rV, rW, coarse occupancy and per-block regular flags
    (ZQ_InitCoarseGrid3D, ZQ_SmokeSimulationUtils3D.cpp:1666-1860; :3318-3340)."""
    W, H, D = _dims(occ)
    Wc, Hc, Dc = coarse_dims(occ, s)
    rU, rV, rW = zeros_faces(Wc, Hc, Dc)
    cocc = np.zeros((Dc, Hc, Wc), np.uint8)
    regular = np.zeros((Dc, Hc, Wc), np.uint8)
    lib().orc_coarse_setup(W, H, D, s, np.ascontiguousarray(occ), rU, rV, rW, cocc, regular)
    return rU, rV, rW, cocc, regular


def fine_to_coarse(u, v, w, s):
    """ZQ_MapFine2Coarse3D (ZQ_SmokeSimulationUtils3D.cpp:2170-2258)."""
    D, H, W1 = u.shape
    W = W1 - 1
    cu, cv, cw = zeros_faces(W // s, H // s, D // s)
    lib().orc_fine_to_coarse(W, H, D, s, u, v, w, cu, cv, cw)
    return cu, cv, cw


def coarse_project(cu, cv, cw, cocc, rU, rV, rW, iters, omega=1.0, open_mask=ALL_FACES, vol=None, flux=False):
    """Face-fraction red-black SOR on the coarse grid and the coarse delta (see zq_oracle_coarse.cpp);
    vol (closed domains) removes the divergence's mean per unit fluid volume first.
    Returns the padded pressure and du, dv, dw."""
    Dc, Hc, Wc = cocc.shape
    p = np.zeros((Dc + 2, Hc + 2, Wc + 2), np.float32)
    du, dv, dw = zeros_faces(Wc, Hc, Dc)
    vv = None if vol is None else np.ascontiguousarray(vol, np.float32)
    lib().orc_coarse_project(Wc, Hc, Dc, cu, cv, cw, cocc, rU, rV, rW, open_mask, omega, iters, p, du, dv, dw,
                             _ptr(vv), int(flux))
    return p, du, dv, dw


def coarse_volume(occ, s):
    """Fluid volume fraction per coarse cell, 1 - (solid fine cells)/s^3 in float32
    (BuildVolumeUnOccupyRatio, ZQ_SmokeSimulationUtils3D.cpp:1860-1885)."""
    D, H, W = occ.shape
    n = (occ != 0).reshape(D // s, s, H // s, s, W // s, s).sum(axis=(1, 3, 5)).astype(np.float32)
    return np.float32(1) - n / np.float32(s ** 3)


def coarse_inv_diag(rU, rV, rW, i, j, k, open_mask=ALL_FACES, flux=False):
    Dc, Hc, Wc1 = rU.shape
    return lib().orc_coarse_inv_diag(Wc1 - 1, Hc, Dc, i, j, k, rU, rV, rW, open_mask, int(flux))


def coarse_to_fine(du, dv, dw, occ, u, v, w, s):
    """ZQ_MapCoarse2Fine3D (ZQ_SmokeSimulationUtils3D.cpp:3172-3296), in place."""
    W, H, D = _dims(occ)
    lib().orc_coarse_to_fine(W, H, D, s, du, dv, dw, np.ascontiguousarray(occ), u, v, w)


def block_face_index(s):
    """Flat face numbering of one s^3 block in the reference's order: u (sk, sj, si = 0..s), then v
    (sk, sj = 0..s, si), then w (sk = 0..s, sj, si) (ZQ_SmokeSimulationUtils3D.cpp:1159-1161)."""
    nu = s * s * (s + 1)
    U = lambda i, j, k: k * s * (s + 1) + j * (s + 1) + i
    V = lambda i, j, k: nu + k * (s + 1) * s + j * s + i
    Wf = lambda i, j, k: 2 * nu + k * s * s + j * s + i
    return U, V, Wf, 3 * nu


def interior_rows(s):
    """Block face numbers of the interior faces in the order of the operator's rows
    (ZQ_SmokeSimulationUtils3D.cpp:1342-1383)."""
    U, V, Wf, _ = block_face_index(s)
    rows = [U(i, j, k) for k in range(s) for j in range(s) for i in range(1, s)]
    rows += [V(i, j, k) for k in range(s) for j in range(1, s) for i in range(s)]
    rows += [Wf(i, j, k) for k in range(1, s) for j in range(s) for i in range(s)]
    return rows


def sub_matrix(s):
    """The regular-block operator of ZQ_InitSubMatPoisson3D (ZQ_SmokeSimulationUtils3D.cpp:1151-1391):
    y = (I - D [0; pinv(P) V]) x on the interior faces, built in float64 and rounded to float32.
    P is the block's Neumann Laplacian (-count on the diagonal, 1 per neighbour) without the column
    of cell 0; V the divergence (n x F); D the interior-face gradient p_r - p_l (F x n).  The reference
    inverts P with ZQlib's SVD (absent); numpy's SVD pseudo-inverse is the same matrix."""
    n = s ** 3
    U, V, Wf, F = block_face_index(s)
    cell = lambda i, j, k: k * s * s + j * s + i
    Lap = np.zeros((n, n))
    Vm = np.zeros((n, F))
    Dm = np.zeros((F, n))
    for k in range(s):
        for j in range(s):
            for i in range(s):
                c = cell(i, j, k)
                Vm[c, U(i, j, k)] -= 1
                Vm[c, U(i + 1, j, k)] += 1
                Vm[c, V(i, j, k)] -= 1
                Vm[c, V(i, j + 1, k)] += 1
                Vm[c, Wf(i, j, k)] -= 1
                Vm[c, Wf(i, j, k + 1)] += 1
                for (a, b, cc, ok) in ((i - 1, j, k, i > 0), (i + 1, j, k, i < s - 1), (i, j - 1, k, j > 0),
                                       (i, j + 1, k, j < s - 1), (i, j, k - 1, k > 0), (i, j, k + 1, k < s - 1)):
                    if ok:
                        Lap[c, cell(a, b, cc)] = 1
                        Lap[c, c] -= 1
                if i > 0:
                    Dm[U(i, j, k), cell(i - 1, j, k)] = -1
                    Dm[U(i, j, k), c] = 1
                if j > 0:
                    Dm[V(i, j, k), cell(i, j - 1, k)] = -1
                    Dm[V(i, j, k), c] = 1
                if k > 0:
                    Dm[Wf(i, j, k), cell(i, j, k - 1)] = -1
                    Dm[Wf(i, j, k), c] = 1
    P = Lap[:, 1:]
    invPV1 = np.zeros((n, F))
    invPV1[1:] = np.linalg.pinv(P) @ Vm
    full = np.eye(F) - Dm @ invPV1
    return full[interior_rows(s)].astype(np.float32)


def _block_faces(u, v, w, s, bi, bj, bk):
    """Gather one block's faces in the reference's flat order (float32 vector of length F)."""
    x0, y0, z0 = bi * s, bj * s, bk * s
    return np.concatenate([u[z0:z0 + s, y0:y0 + s, x0:x0 + s + 1].reshape(-1),
                           v[z0:z0 + s, y0:y0 + s + 1, x0:x0 + s].reshape(-1),
                           w[z0:z0 + s + 1, y0:y0 + s, x0:x0 + s].reshape(-1)])


def _scatter_interior(u, v, w, s, bi, bj, bk, y):
    """Write the interior faces (operator row order) of one block back."""
    x0, y0, z0 = bi * s, bj * s, bk * s
    nu, nv = s * s * (s - 1), s * s * (s - 1)
    u[z0:z0 + s, y0:y0 + s, x0 + 1:x0 + s] = y[:nu].reshape(s, s, s - 1)
    v[z0:z0 + s, y0 + 1:y0 + s, x0:x0 + s] = y[nu:nu + nv].reshape(s, s - 1, s)
    w[z0 + 1:z0 + s, y0:y0 + s, x0:x0 + s] = y[nu + nv:].reshape(s - 1, s, s)


def sub_project_regular(u, v, w, s, M, regular):
    """Regular blocks: interior faces = M x, one sequential float32 dot product per row (zq_oracle_coarse.cpp)."""
    D, H, W1 = u.shape
    lib().orc_sub_project_regular(W1 - 1, H, D, s, np.ascontiguousarray(M, np.float32),
                                  np.ascontiguousarray(regular, np.uint8), u, v, w)


def sub_solve_block(x, occ_b, s):
    """ZQ_SolveSubGrid3D (ZQ_SmokeSimulationUtils3D.cpp:1393-1521) on one block with solid cells:
    the closed Poisson system of ZQ_InitMatClosedPoisson3D (:603-747; rows all fluid cells, columns
    all but the first, unit face fractions between fluid cells) solved in least squares (the
    reference runs ZQlib's CG on the normal equations to 1e-12, absent: numpy's minimum-norm
    least-squares solution is what it converges to), pressure rounded to float32 with the first
    fluid cell at 0, interior faces between two fluid cells moved by -(p_r - p_l)
    (ZQ_AdjustClosedVelocity3D, :2867-2927), every face then input + delta.  A block with fewer than two
    fluid cells has no system and is returned unchanged (:1492-1497).  Returns the interior faces in
    operator row order."""
    U, V, Wf, F = block_face_index(s)
    occ_b = np.asarray(occ_b).reshape(s, s, s)
    x = np.asarray(x, np.float32)
    fluid = [(i, j, k) for k in range(s) for j in range(s) for i in range(s) if not occ_b[k, j, i]]
    rows = interior_rows(s)
    if len(fluid) < 2:
        return x[rows].copy()
    idx = {c: n for n, c in enumerate(fluid)}
    A = np.zeros((len(fluid), len(fluid)))
    b = np.zeros(len(fluid))
    for (i, j, k), r in idx.items():
        # left to right, as the reference's expression (:2569-2571)
        d = np.float32(x[U(i + 1, j, k)] - x[U(i, j, k)])
        d = np.float32(d + x[V(i, j + 1, k)])
        d = np.float32(d - x[V(i, j, k)])
        d = np.float32(d + x[Wf(i, j, k + 1)])
        d = np.float32(d - x[Wf(i, j, k)])
        b[r] = d
        for nb in ((i, j, k - 1), (i, j, k + 1), (i, j - 1, k), (i, j + 1, k), (i - 1, j, k), (i + 1, j, k)):
            if nb in idx:
                A[r, idx[nb]] = 1.0
                A[r, r] -= 1.0
    sol = np.linalg.lstsq(A[:, 1:], b, rcond=None)[0]
    p = np.zeros((s, s, s), np.float32)
    for (i, j, k), r in idx.items():
        p[k, j, i] = 0.0 if r == 0 else np.float32(sol[r - 1])
    # output = input + delta on every face, delta = 0 unless both cells are fluid (:1507-1512)
    delta = np.zeros(F, np.float32)
    for k in range(s):
        for j in range(s):
            for i in range(s):
                if occ_b[k, j, i]:
                    continue
                if i > 0 and not occ_b[k, j, i - 1]:
                    delta[U(i, j, k)] = -(p[k, j, i] - p[k, j, i - 1])
                if j > 0 and not occ_b[k, j - 1, i]:
                    delta[V(i, j, k)] = -(p[k, j, i] - p[k, j - 1, i])
                if k > 0 and not occ_b[k - 1, j, i]:
                    delta[Wf(i, j, k)] = -(p[k, j, i] - p[k - 1, j, i])
    return (x + delta)[rows]


def sub_project_irregular(u, v, w, occ, s, regular):
    """Blocks with a solid cell, each through sub_solve_block, in place."""
    Dc, Hc, Wc = regular.shape
    for bk in range(Dc):
        for bj in range(Hc):
            for bi in range(Wc):
                if regular[bk, bj, bi]:
                    continue
                x = _block_faces(u, v, w, s, bi, bj, bk)
                ob = occ[bk * s:(bk + 1) * s, bj * s:(bj + 1) * s, bi * s:(bi + 1) * s]
                _scatter_interior(u, v, w, s, bi, bj, bk, sub_solve_block(x, ob, s))


def apply_guide(du, dv, dw, cu, cv, cw, coeff, guide):
    """ZQ_ApplyGuidingVelocity3D (ZQ_SmokeSimulationUtils3D.cpp:3139-3170): delta += coeff*(G - U) on
    every coarse face, float32."""
    c = np.float32(coeff)
    for d, u, g in zip((du, dv, dw), (cu, cv, cw), guide):
        d += c * (np.asarray(g, np.float32) - u)


def coarse_projection(u, v, w, occ, s, iters, omega=1.0, open_mask=ALL_FACES, M=None, guide=None, guide_coeff=0.0,
                      closed_mean=False, flux=False):
    """ZQ_TotalProjection3D for COARSE_OPEN_POISSON_SOR (ZQ_SmokeSimulationUtils3D.cpp:3688-3793), in
    place: fine -> coarse, coarse face-fraction solve, guiding (guide = coarse faces gu, gv, gw),
    coarse -> fine, sub-projection (s > 1)."""
    rU, rV, rW, cocc, regular = coarse_setup(occ, s)
    cu, cv, cw = fine_to_coarse(u, v, w, s)
    p, du, dv, dw = coarse_project(cu, cv, cw, cocc, rU, rV, rW, iters, omega, open_mask,
                                   coarse_volume(occ, s) if closed_mean else None, flux)
    if guide is not None:
        apply_guide(du, dv, dw, cu, cv, cw, guide_coeff, guide)
    coarse_to_fine(du, dv, dw, occ, u, v, w, s)
    if s > 1:
        sub_project_regular(u, v, w, s, sub_matrix(s) if M is None else M, regular)
        sub_project_irregular(u, v, w, occ, s, regular)
    return p


def coarse_step(params, u, v, w, T, rho, occ, s, boxes=None, guide=None, guide_coeff=0.0, open_mask=ALL_FACES,
                closed_mean=False, flux=False):
    """One step in the order of orc_step (ZQ_SmokeSimulation3D.cpp:156-243) with the coarse + sub-grid
    projection in place of the full-grid solve (COARSE_*_POISSON_SOR): open on all six sides, FP32
    SOR with params' omega and sor_iters on the coarse grid."""
    P = dict(params)
    W, H, D = _dims(occ)
    add_force(u, v, w, occ, T, rho, P["h"], P["dt"], P["beta"], P["eps"], P["Tamb"], alpha=P.get("alpha", 0.0))
    attenuation(u, v, w, T, rho, occ, P["aV"], P["aT"], P["aR"])
    dtt = float(np.float32(P["dt"]) / np.float32(P["steps"]))
    advect_velocity(u, v, w, occ, P["h"], P["steps"], dtt, P["adv_vel_type"], True)
    coarse_projection(u, v, w, occ, s, P["sor_iters"], P["omega"], open_mask, guide=guide, guide_coeff=guide_coeff,
                      closed_mean=closed_mean, flux=flux)
    To, Ro = advect_scalar(u, v, w, occ, T, rho, P["h"], P["steps"], dtt, P["adv_sca_type"])
    T[...] = To
    rho[...] = Ro
    for b in boxes or []:
        reinject_box(T, rho, b)


# ---- octree projection (OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR) ------------------
OCTREE_THRESH = (2, 4, 16)  # octree_thresh1..3 defaults (ZQ_SmokeSimulationUtils3D.cpp:52-54)


def _level_shape(W, H, D, L):
    return (D >> L, H >> L, W >> L)


def octree_distance(occ):
    W, H, D = _dims(occ)
    dist = np.zeros((D, H, W), np.int32)
    lib().orc_oct_distance(W, H, D, np.ascontiguousarray(occ, np.uint8), dist)
    return dist


def octree_leaves(occ, thresh=OCTREE_THRESH):
    """Leaf flags (l0, l1, l2, l3) of _buildOctreeInfo (ZQ_SmokeOctreeGrid3D.cpp:5328-5609)."""
    W, H, D = _dims(occ)
    occ = np.ascontiguousarray(occ, np.uint8)
    dist = octree_distance(occ)
    ls = [np.zeros(_level_shape(W, H, D, L), np.uint8) for L in range(4)]
    lib().orc_oct_leaves(W, H, D, occ, dist, int(thresh[0]), int(thresh[1]), int(thresh[2]), *ls)
    return ls


def octree_neighbors(leaves, L, x, y, z):
    """The neighbour list of the level-L cell (x, y, z): rows (level, x, y, z) in the reference's order."""
    D, H, W = leaves[0].shape
    out = np.zeros((6 * 4 ** L, 4), np.int32)
    n = lib().orc_oct_neighbors(W, H, D, *leaves, L, x, y, z, out)
    return out[:n]


def octree_divergence(u, v, w):
    D, H, W1 = u.shape
    W = W1 - 1
    ds = [np.zeros(_level_shape(W, H, D, L), np.float32) for L in range(4)]
    lib().orc_oct_divergence(W, H, D, u, v, w, *ds)
    return ds


def octree_solve(leaves, divs, iters, closed=False, mean=0.0):
    D, H, W = leaves[0].shape
    ps = [np.zeros(_level_shape(W, H, D, L), np.float32) for L in range(4)]
    lib().orc_oct_solve(W, H, D, *leaves, *divs, int(iters), 1 if closed else 0, float(mean), *ps)
    return ps


def octree_adjust(leaves, ps, u, v, w):
    D, H, W = leaves[0].shape
    lib().orc_oct_adjust(W, H, D, *leaves, *ps, u, v, w)


def octree_special_regions(u, v, w, leaves, iters):
    """The special-region projections of octree guiding, in place (orc_oct_special_regions).  Returns
    (number of regions, labels); raises where the reference's labelling does not end."""
    D, H, W = leaves[0].shape
    labels = np.zeros((D, H, W), np.int32)
    n = lib().orc_oct_special_regions(W, H, D, leaves[0], leaves[1], int(iters), u, v, w, labels)
    if n < 0:
        raise RuntimeError("the reference's special-region labelling does not end on these leaves")
    return n, labels


def octree_sub_project(u, v, w, leaves, mats=None):
    """SubProjection (ZQ_SmokeOctreeGrid3D.cpp:840-1245): every level-L leaf's interior faces (L >= 1)
    through the regular-block operator of ZQ_InitSubMatPoisson3D with s = 2^L (:4994-5032).  At level 1
    the reference's write-back of the w rows never advances its row (:1125-1135), so every interior w
    face of a level-1 leaf takes the first w row, the face (sk = 1, sj = 0, si = 0); kept."""
    for L in (1, 2, 3):
        s = 1 << L
        if not leaves[L].any():
            continue
        M = sub_matrix(s) if mats is None else mats[L]
        sub_project_regular(u, v, w, s, M, leaves[L])
        if L == 1:
            for bk, bj, bi in zip(*np.nonzero(leaves[1])):
                z, y, x = 2 * bk + 1, 2 * bj, 2 * bi
                w[z, y:y + 2, x:x + 2] = w[z, y, x]


def octree_projection(u, v, w, occ, iters, thresh=OCTREE_THRESH, closed=False, mats=None, guiding=False):
    """One octree projection in place (ZQ_SmokeSimulation3D.cpp:478-513): build the leaves, red-black
    solve over the levels (iters = maxIter), face correction, with guiding the special-region
    projections (ApplyGuidingVelocity), sub-projection of the leaf blocks.  Returns the leaves and the
    pressures."""
    leaves = octree_leaves(occ, thresh)
    W, H, D = _dims(occ)
    mean = lib().orc_oct_closed_mean(W, H, D, np.ascontiguousarray(occ, np.uint8), u, v, w) if closed else 0.0
    ps = octree_solve(leaves, octree_divergence(u, v, w), iters, closed, mean)
    octree_adjust(leaves, ps, u, v, w)
    if guiding:
        octree_special_regions(u, v, w, leaves, iters)
    octree_sub_project(u, v, w, leaves, mats)
    return leaves, ps


def octree_step(params, u, v, w, T, rho, occ, boxes=None, thresh=OCTREE_THRESH, closed=False, mats=None,
                guiding=False):
    """One step in the order of orc_step (ZQ_SmokeSimulation3D.cpp:156-243) with the octree projection
    (OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR, params' sor_iters passes; guiding: with the
    special-region projections of useGuidingControl) in place of the full-grid solve."""
    P = dict(params)
    add_force(u, v, w, occ, T, rho, P["h"], P["dt"], P["beta"], P["eps"], P["Tamb"], alpha=P.get("alpha", 0.0))
    attenuation(u, v, w, T, rho, occ, P["aV"], P["aT"], P["aR"])
    dtt = float(np.float32(P["dt"]) / np.float32(P["steps"]))
    advect_velocity(u, v, w, occ, P["h"], P["steps"], dtt, P["adv_vel_type"], True)
    octree_projection(u, v, w, occ, P["sor_iters"], thresh, closed, mats, guiding)
    To, Ro = advect_scalar(u, v, w, occ, T, rho, P["h"], P["steps"], dtt, P["adv_sca_type"])
    T[...] = To
    rho[...] = Ro
    for b in boxes or []:
        reinject_box(T, rho, b)
