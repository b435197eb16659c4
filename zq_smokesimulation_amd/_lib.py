"""ctypes binding of libzq_smoke3d.so (C ABI declared in include/zq_smoke3d.h).

The library is the only compute path: if it is missing or was built for another
machine, importing this module raises.  Nothing here falls back to a CPU
implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# ZQ3D_LIB may name another build of the same sources (the sanitizer build of the host mirror)
LIB_PATH = os.environ.get("ZQ3D_LIB") or os.path.join(LIB_DIR, "libzq_smoke3d.so")

# enum AdvectVelocityType / AdvectScalarType (ZQ_CUDA_Advection3D.h:7-21)
ADV_VEL_INREG_OUTREG, ADV_VEL_INREG_OUTREG_BFECC, ADV_VEL_INMAC_OUTMAC, \
    ADV_VEL_INMAC_OUTMAC_BFECC, ADV_VEL_INREG_OUTMAC, ADV_VEL_INREG_OUTMAC_BFECC = range(6)
ADV_SCA_MAC, ADV_SCA_MAC_BFECC, ADV_SCA_REG, ADV_SCA_REG_BFECC = range(4)
ADD_FORCE_ENTIRE, ADD_FORCE_SLICE = range(2)

OPEN_XLO, OPEN_XHI, OPEN_YLO, OPEN_YHI, OPEN_ZLO, OPEN_ZHI, OPEN_ALL = 1, 2, 4, 8, 16, 32, 63

FIELD_U, FIELD_V, FIELD_W, FIELD_T, FIELD_RHO, FIELD_OCC, FIELD_P64, FIELD_VEL4 = range(8)
OK, ERR_ARG, ERR_HIP, ERR_STATE = range(4)


class SorOpts(C.Structure):
    _fields_ = [
        ("open_mask", C.c_int),
        ("closed_mean", C.c_int),
        ("omega", C.c_float),
        ("iters", C.c_int),
        ("mixed_precision", C.c_int),
        ("max_outer", C.c_int),
        ("tol", C.c_double),
        ("pass_rule", C.c_int),
    ]


class ProjStats(C.Structure):
    _fields_ = [
        ("outer", C.c_int),
        ("sweeps", C.c_int),
        ("max_residual", C.c_double),
        ("max_div", C.c_double),
        ("mean", C.c_double),
        ("converged", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Config(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int),
        ("width", C.c_int), ("height", C.c_int), ("depth", C.c_int),
        ("voxel_size", C.c_float),
        ("step_time", C.c_float),
        ("steps", C.c_int),
        ("Tamb", C.c_float),
        ("buoyancy", C.c_float),
        ("gravity", C.c_float),
        ("confine", C.c_float),
        ("vel_atten", C.c_float), ("temp_atten", C.c_float), ("dens_atten", C.c_float),
        ("adv_vel_type", C.c_int),
        ("adv_sca_type", C.c_int),
        ("sor", SorOpts),
        ("warm_start", C.c_int),
        ("sub_size", C.c_int),
        ("coarse_flux", C.c_int),
        ("octree", C.c_int),
        ("octree_thresh", C.c_int * 3),
        ("warm_extrapolate", C.c_float),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("ms_force", C.c_double),
        ("ms_advect_velocity", C.c_double),
        ("ms_project", C.c_double),
        ("ms_sor", C.c_double),
        ("ms_advect_scalar", C.c_double),
        ("ms_reinject", C.c_double),
        ("sor_launches", C.c_longlong),
        ("steps", C.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TUNE_MAX = 32


class TuneReport(C.Structure):
    _fields_ = [
        ("count", C.c_int),
        ("best", C.c_int),
        ("ty", C.c_int * TUNE_MAX),
        ("zc", C.c_int * TUNE_MAX),
        ("us_per_iter", C.c_float * TUNE_MAX),
    ]

    def as_dict(self):
        n = self.count
        return {"best": self.best,
                "candidates": [{"ty": self.ty[i], "zc": self.zc[i], "us_per_iter": round(self.us_per_iter[i], 3)}
                               for i in range(n)]}


class ZQ3DError(RuntimeError):
    pass


_lib = None


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libzq_smoke3d.so not found at %s: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            % LIB_PATH)
    # One HIP runtime per process: torch ships its own libamdhip64 (same SONAME, libamdhip64.so.7).
    # Loaded first, it also serves this library, so torch tensors and streams and these kernels
    # share one runtime; loaded second, it would bring up a second HIP/HSA runtime beside ours.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    f, i, u, vp, b = C.c_float, C.c_int, C.c_uint, C.c_void_p, C.c_bool
    L.ZQ_Cuda_Prepare_Advection.argtypes = [u, u, u, f, u, f]
    L.ZQ_Cuda_Prepare_Advection.restype = None
    L.Advect_Velocity.argtypes = [vp, vp, vp, vp, b, i]
    L.Advect_Velocity.restype = f
    L.Advect_Scalar.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i]
    L.Advect_Scalar.restype = f
    L.AddForce3D.argtypes = [vp, vp, vp, vp, vp, f, f, f, f, f, i, i, i, i]
    L.AddForce3D.restype = f
    L.Attenuation3D.argtypes = [vp, vp, vp, vp, vp, vp, f, f, f, i, i, i]
    L.Attenuation3D.restype = f
    L.SolveOpenPoissonRedBlackwithOccupy3D_MAC.argtypes = [vp, vp, vp, vp, i, i, i, i]
    L.SolveOpenPoissonRedBlackwithOccupy3D_MAC.restype = f
    L.SolveClosedPoissonRedBlackwithOccupy3D_MAC.argtypes = [vp, vp, vp, vp, f, i, i, i, i]
    L.SolveClosedPoissonRedBlackwithOccupy3D_MAC.restype = f
    L.ZQ3D_SolvePoissonRedBlack_ex.argtypes = [vp, vp, vp, vp, i, i, i, C.POINTER(SorOpts), C.POINTER(ProjStats)]
    L.ZQ3D_SolvePoissonRedBlack_ex.restype = f
    L.ZQ3D_SolvePoissonRedBlack_noalloc.argtypes = [vp, vp, vp, vp, i, i, i, C.POINTER(SorOpts), C.POINTER(ProjStats),
                                                    vp]
    L.ZQ3D_SolvePoissonRedBlack_noalloc.restype = i
    L.ZQ3D_MACtoRegular4.argtypes = [vp, vp, vp, vp, i, i, i]
    L.ZQ3D_MACtoRegular4.restype = f
    L.SolveOpenPoissonRedBlackwithFaceRatio3D_MAC.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, i, i]
    L.SolveOpenPoissonRedBlackwithFaceRatio3D_MAC.restype = f
    L.MatrixMul.argtypes = [vp, vp, i, i, i, vp]
    L.MatrixMul.restype = f
    lists = [i, vp, i, vp] * 8
    L.SolveOpenOctreePoissonRedBlack3_3D_MAC.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, i, i] + lists
    L.SolveOpenOctreePoissonRedBlack3_3D_MAC.restype = f
    L.SolveClosedOctreePoissonRedBlack3_3D_MAC.argtypes = [vp, vp, vp, vp, vp, vp, vp, f, i, i, i, i] + lists
    L.SolveClosedOctreePoissonRedBlack3_3D_MAC.restype = f
    L.zq3d_config_default.argtypes = [C.POINTER(Config), i, i, i]
    L.zq3d_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.zq3d_set_guide.argtypes = [vp, f, vp, vp, vp]
    L.zq3d_set_coarse_sweep.argtypes = [vp, i, i]
    L.zq3d_get_coarse_sweep.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
    L.zq3d_destroy.argtypes = [vp]
    L.zq3d_destroy.restype = None
    L.zq3d_set_stream.argtypes = [vp, vp]
    L.zq3d_synchronize.argtypes = [vp]
    L.zq3d_upload.argtypes = [vp, i, vp, C.c_size_t]
    L.zq3d_download.argtypes = [vp, i, vp, C.c_size_t]
    L.zq3d_device_ptr.argtypes = [vp, i, C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.zq3d_set_box_sources.argtypes = [vp, i, vp]
    L.zq3d_set_sphere_sources.argtypes = [vp, i, vp]
    L.zq3d_stamp_sphere_obstacle.argtypes = [vp, f, f, f, f, f, f, f, i]
    L.zq3d_add_force_attenuate.argtypes = [vp, vp]
    L.zq3d_advect_velocity.argtypes = [vp]
    L.zq3d_project.argtypes = [vp, C.POINTER(ProjStats)]
    L.zq3d_advect_scalar.argtypes = [vp]
    L.zq3d_reinject.argtypes = [vp]
    L.zq3d_step.argtypes = [vp, vp, C.POINTER(ProjStats)]
    L.zq3d_mac_to_regular.argtypes = [vp]
    L.zq3d_set_profiling.argtypes = [vp, i]
    L.zq3d_get_timing.argtypes = [vp, C.POINTER(Timing)]
    L.zq3d_autotune_sor.argtypes = [vp, i, C.POINTER(TuneReport)]
    L.zq3d_set_sor_shape.argtypes = [vp, i, i]
    L.zq3d_sor_kernel.argtypes = [vp]
    L.zq3d_sor_kernel.restype = C.c_char_p
    L.zq3d_last_error.argtypes = []
    L.zq3d_last_error.restype = C.c_char_p
    L.zq3d_release_workspace.argtypes = []
    L.zq3d_release_workspace.restype = None
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = _bind_ensemble(_bind_adjoint(_bind_scene(_load())))
    return _lib


def last_error():
    return lib().zq3d_last_error().decode()


def check(rc, what=""):
    """Raise ZQ3DError for a nonzero status (Part 3) or a negative time (Part 1/2)."""
    if isinstance(rc, float):
        if rc < 0:
            raise ZQ3DError("%s failed: %s" % (what, last_error()))
        return rc
    if rc != OK:
        raise ZQ3DError("%s failed (code %d): %s" % (what, rc, last_error()))
    return rc


MAX_SOURCES = 16


class Scene(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int),
        ("config", Config),
        ("frame_count", C.c_int),
        ("solver", C.c_int),
        ("wall_faces", C.c_int),
        ("internal_object", C.c_int),
        ("use_moving_object", C.c_int),
        ("export_type", C.c_int),
        ("device_resident", C.c_int),
        ("sub_size", C.c_int), ("max_iter", C.c_int), ("inner_iter", C.c_int), ("steps", C.c_int),
        ("vel_atten_coeff", C.c_float), ("temp_atten_coeff", C.c_float), ("dens_atten_coeff", C.c_float),
        ("adv_vel_type_para", C.c_int),
        ("adv_sca_type_para", C.c_int),
        ("add_force_type", C.c_int),
        ("n_box_sources", C.c_int),
        ("box_sources", (C.c_float * 8) * MAX_SOURCES),
        ("n_sphere_sources", C.c_int),
        ("sphere_sources", (C.c_float * 6) * MAX_SOURCES),
        ("temp_noise", C.c_float),
        ("noise_seed", C.c_uint),
        ("moving_object_file", C.c_char * 200),
    ]


def _bind_scene(L):
    vp = C.c_void_p
    L.zq3d_scene_load.argtypes = [C.c_char_p, C.POINTER(Scene)]
    L.zq3d_scene_init_occupancy.argtypes = [C.c_char_p, vp, C.c_size_t]
    L.zq3d_scene_apply_moving_objects.argtypes = [C.c_char_p, C.c_int, vp, vp, vp, vp]
    L.zq3d_di3_read.argtypes = [C.c_char_p, C.POINTER(C.c_int * 4), vp, C.c_size_t]
    L.ZQ3D_AddForce3D_ex.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
    L.ZQ3D_AddForce3D_ex.restype = C.c_float
    return L


class AdjStats(C.Structure):
    _fields_ = [
        ("forward_sweeps", C.c_longlong),
        ("adjoint_sweeps", C.c_longlong),
        ("recomputed_steps", C.c_int),
        ("max_div", C.c_double),
        ("ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind_adjoint(L):
    vp = C.c_void_p
    L.zq3d_adj_create.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.zq3d_adj_destroy.argtypes = [vp]
    L.zq3d_adj_destroy.restype = None
    L.zq3d_adj_set_stream.argtypes = [vp, vp]
    L.zq3d_adj_set_sources.argtypes = [vp, C.c_int, vp, C.c_int, vp]
    L.zq3d_adj_set_adjoint_tol.argtypes = [vp, C.c_double]
    L.zq3d_adj_set_pressure_store.argtypes = [vp, C.c_int]
    L.zq3d_adj_set_sor_shape.argtypes = [vp, C.c_int, C.c_int]
    L.zq3d_adj_set_initial.argtypes = [vp, C.c_int, vp]
    L.zq3d_adj_forward.argtypes = [vp, vp, C.POINTER(AdjStats), vp]
    L.zq3d_adj_backward.argtypes = [vp, vp, vp, vp, C.POINTER(AdjStats)]
    L.zq3d_adj_get.argtypes = [vp, C.c_int, C.c_int, vp]
    L.zq3d_adj_generation.argtypes = [vp]
    L.zq3d_adj_generation.restype = C.c_longlong
    L.zq3d_adj_get_tape.argtypes = [vp, C.c_int, C.c_int, vp]
    L.zq3d_adj_evaluate.argtypes = [vp, C.c_int, vp, C.POINTER(AdjStats), vp]
    return L


ENS_TUNE_MAX = 24


class EnsTune(C.Structure):
    _fields_ = [
        ("count", C.c_int),
        ("kind", C.c_int * ENS_TUNE_MAX),
        ("tbx", C.c_int * ENS_TUNE_MAX), ("tby", C.c_int * ENS_TUNE_MAX), ("blocks", C.c_int * ENS_TUNE_MAX),
        ("us_per_iter", C.c_float * ENS_TUNE_MAX),
        ("best_kind", C.c_int), ("best_tbx", C.c_int), ("best_tby", C.c_int), ("best_sweep_blocks", C.c_int),
        ("best_us_per_iter", C.c_float),
        ("best_blocks", C.c_int),
        ("probe_ms", C.c_float),
    ]

    def as_dict(self):
        n = self.count
        kinds = ("half_sweeps", "split_one_launch")
        return {"candidates": [{"kind": kinds[self.kind[i]], "tbx": self.tbx[i], "tby": self.tby[i],
                                "blocks": self.blocks[i], "us_per_iter": round(self.us_per_iter[i], 3)}
                               for i in range(n)],
                "best": {"kind": kinds[self.best_kind], "tbx": self.best_tbx, "tby": self.best_tby,
                         "sweep_blocks": self.best_sweep_blocks,
                         "us_per_iter": round(self.best_us_per_iter, 3), "blocks": self.best_blocks,
                         "probe_ms": round(self.probe_ms, 4)}}


def _bind_ensemble(L):
    vp = C.c_void_p
    L.zq3d_ens_create.argtypes = [C.c_int, C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.zq3d_ens_destroy.argtypes = [vp]
    L.zq3d_ens_destroy.restype = None
    L.zq3d_ens_set_sources.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, vp]
    L.zq3d_ens_upload.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    L.zq3d_ens_download.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    L.zq3d_ens_step.argtypes = [vp, C.c_int, C.c_int]
    L.zq3d_ens_synchronize.argtypes = [vp]
    L.zq3d_ens_cells.argtypes = [vp]
    L.zq3d_ens_cells.restype = C.c_longlong
    L.zq3d_ens_stream.argtypes = [vp, C.POINTER(C.c_void_p)]
    L.zq3d_ens_autotune.argtypes = [vp, C.c_int, C.POINTER(EnsTune)]
    L.zq3d_ens_set_sweep.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    return L
