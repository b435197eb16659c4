"""CPU: the C-ABI library loads and exports every function include/zq_smoke3d.h declares (no GPU
calls), and the oracle library loads."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
        name = m.group(1)
        prev = text[:m.start()].rstrip()
        if name in ("if", "sizeof", "defined") or prev.endswith("#define"):
            continue
        # a declaration: preceded by a return type token
        if re.search(r"(float|void|int|char\s*\*|const char\s*\*)\s*\**\s*$", prev):
            names.add(name)
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_functions(os.path.join(ROOT, "include", "zq_smoke3d.h"))
    for ref in ("ZQ_Cuda_Prepare_Advection", "Advect_Velocity", "Advect_Scalar", "AddForce3D", "Attenuation3D",
                "SolveOpenPoissonRedBlackwithOccupy3D_MAC", "SolveClosedPoissonRedBlackwithOccupy3D_MAC",
                "SolveOpenPoissonRedBlackwithFaceRatio3D_MAC", "MatrixMul", "SolveOpenOctreePoissonRedBlack3_3D_MAC",
                "SolveClosedOctreePoissonRedBlack3_3D_MAC"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    import zq_smokesimulation_amd._lib as L
    lib = C.CDLL(L.LIB_PATH)
    names = declared_functions(os.path.join(ROOT, "include", "zq_smoke3d.h"))
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_size_matches_library():
    """zq3d_config_default runs on the host only; it must agree with the ctypes layout."""
    import zq_smokesimulation_amd._lib as L
    cfg = L.Config()
    assert L.lib().zq3d_config_default(C.byref(cfg), 64, 32, 16) == 0
    assert cfg.struct_size == C.sizeof(L.Config)
    assert (cfg.width, cfg.height, cfg.depth) == (64, 32, 16)
    assert cfg.steps == 20 and cfg.Tamb == 10.0 and cfg.sor.open_mask == 63
    assert cfg.octree == 0 and list(cfg.octree_thresh) == [2, 4, 16]
    assert abs(cfg.voxel_size - 1.0 / 64) < 1e-9


def test_bad_arguments_fail_without_gpu():
    """Argument validation happens before any device call."""
    import zq_smokesimulation_amd._lib as L
    cfg = L.Config()
    L.lib().zq3d_config_default(C.byref(cfg), 1, 8, 8)
    h = C.c_void_p()
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert "dimensions" in L.last_error()
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.adv_vel_type = 7
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.sub_size = 3  # does not divide 8
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert "sub_size" in L.last_error()
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.sub_size = 4
    cfg.sor.mixed_precision = 1
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.coarse_flux = 1  # without a coarse grid
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert "coarse_flux" in L.last_error()
    L.lib().zq3d_config_default(C.byref(cfg), 12, 8, 8)
    cfg.octree = 1  # 12 is no multiple of 8
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert "octree" in L.last_error()
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.octree = 1
    cfg.sub_size = 2
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.octree = 1
    cfg.octree_thresh = (1, -1, 4)
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    lists = [0, None, 0, None] * 8
    assert L.lib().SolveOpenOctreePoissonRedBlack3_3D_MAC(None, None, None, None, None, None, None, 12, 8, 8, 5,
                                                          *lists) < 0
    assert "multiples of 8" in L.last_error()
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    cfg.sor.omega = 2.5
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert L.lib().Attenuation3D(None, None, None, None, None, None, 1.0, 1.0, 1.0, 8, 8, 8) < 0


def test_oracle_loads():
    import zq_oracle
    assert zq_oracle.lib() is not None


def test_noalloc_rejects_grids_beyond_its_scratch_without_gpu():
    """The size check runs before any device call: a grid whose scratch exceeds the static array is
    refused with the bytes it would need."""
    import numpy as np
    import zq_smokesimulation_amd._lib as L
    host = np.zeros(8, np.float32)
    p = host.ctypes.data_as(C.c_void_p)
    opts = L.SorOpts(open_mask=0, closed_mean=1, omega=1.8, iters=40, mixed_precision=1, max_outer=100, tol=1e-6)
    rc = L.lib().ZQ3D_SolvePoissonRedBlack_noalloc(p, p, p, p, 128, 128, 128, C.byref(opts), None, None)
    assert rc == L.ERR_ARG
    msg = L.last_error()
    assert "bytes of scratch" in msg and str(16 << 20) in msg
    opts.mixed_precision = 0
    rc = L.lib().ZQ3D_SolvePoissonRedBlack_noalloc(p, p, p, p, 128, 100, 100, C.byref(opts), None, None)
    assert rc == L.ERR_ARG
    rc = L.lib().ZQ3D_SolvePoissonRedBlack_noalloc(p, p, p, p, 64, 64, 64, None, None, None)
    assert rc == L.ERR_ARG and "null options" in L.last_error()


def test_mixed_precision_needs_refinement_passes_without_gpu():
    """The default config allows refinement passes (maxOuter's default, 1000); a mixed-precision config
    with max_outer = 0, which would never project, is refused before any device call."""
    import zq_smokesimulation_amd._lib as L
    cfg = L.Config()
    L.lib().zq3d_config_default(C.byref(cfg), 8, 8, 8)
    assert cfg.sor.max_outer == 1000
    cfg.sor.mixed_precision = 1
    cfg.sor.max_outer = 0
    h = C.c_void_p()
    assert L.lib().zq3d_create(C.byref(cfg), C.byref(h)) == L.ERR_ARG
    assert "max_outer" in L.last_error()
    assert L.lib().zq3d_adj_create(C.byref(cfg), 4, 4, 2, C.byref(h)) == L.ERR_ARG
