"""CPU: the host mirror of the reference's set-up path (ZQ_SmokeSimulationPara3D::LoadFromFile,
ZQ_InitFineGrid3D, LoadMovingObjects3D + ZQ_ApplyMovingObject3D) on the reference's own config
and moving-object files, which are kept under tests/golden/configs as data fixtures.

Expected values are read off the fixture files by hand or derived in closed form from the
reference's formulas (ZQ_SmokeSimulationUtils3D.cpp:1523-1664, 1933-2046)."""
import math
import os

import numpy as np
import pytest

import zq_oracle as O
from zq_smokesimulation_amd import scene

CFG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs")


def cfg(name):
    return os.path.join(CFG, name)


SOURCE = "source 0.5 0.1 0.5 0.1 0.05 0.1 0.8 20\n"


def write_cfg(tmp_path, text, name="c.txt"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def write_scene(tmp_path, text):
    """A config needs at least one smoke source or LoadFromFile rejects it."""
    return write_cfg(tmp_path, SOURCE + text)


def test_preview64_values():
    s = scene.load_scene(cfg("expsimu3D_preview64.txt"))
    c = s.config
    assert (c.width, c.height, c.depth) == (64, 64, 64)
    assert s.sub_size == 2
    # "step 30" is not a key LoadFromFile reads; steps keeps its default of 20
    assert s.steps == 20 and c.steps == 20
    assert c.step_time == pytest.approx(0.03)
    assert s.frame_count == 200 and s.max_iter == 600 and s.inner_iter == 20
    assert c.Tamb == pytest.approx(1.0) and c.confine == pytest.approx(30.0)
    assert s.vel_atten_coeff == pytest.approx(10.0)
    assert c.vel_atten == pytest.approx(math.exp(-0.03 * 10.0), rel=1e-6)
    assert c.temp_atten == 1.0 and c.dens_atten == 1.0
    assert scene.SOLVERS[s.solver] == "COARSE_OPEN_POISSON_SOR"
    assert s.internal_object == 2  # box
    assert s.use_moving_object == 0
    assert s.adv_vel_type_para == 3  # inReg_outReg_bfecc (matched case-insensitively)
    assert s.adv_sca_type_para == 2  # Reg
    assert s.export_type == 0  # di3
    assert s.n_box_sources == 1 and s.n_sphere_sources == 0
    np.testing.assert_allclose(list(s.box_sources[0]), [0.5, 0.1, 0.5, 0.1, 0.05, 0.1, 0.8, 20.0], rtol=1e-7)
    # an open solver: every domain face is open, the closed mean is off
    assert c.sor.open_mask == O.ALL_FACES and c.sor.closed_mean == 0
    assert c.sor.iters == 600


@pytest.mark.parametrize("name, res, intobj, solver, force_slice", [
    ("expsimu3D_fine128.txt", (128, 128, 128), 1, "OPEN_OCTREE_POISSON", 0),
    ("expsimu3D_fine192.txt", (192, 288, 192), 1, "COARSE_OPEN_POISSON_SOR", 1),
    ("expsimu3D_fine256.txt", (256, 256, 256), 1, "COARSE_OPEN_POISSON_SOR", 1),
    ("smokesim3d_fine192.txt", (128, 256, 128), 1, "COARSE_OPEN_POISSON_SOR", 1),
    ("expsimu3D_preview128.txt", (128, 128, 128), 0, "OPEN_OCTREE_POISSON", 0),
])
def test_fixture_headline_values(name, res, intobj, solver, force_slice):
    s = scene.load_scene(cfg(name))
    assert (s.config.width, s.config.height, s.config.depth) == res
    assert s.internal_object == intobj
    assert scene.SOLVERS[s.solver] == solver
    assert s.add_force_type == force_slice


def test_smokesim_fine192_sub_size():
    assert scene.load_scene(cfg("smokesim3d_fine192.txt")).sub_size == 4


def test_preview128_moving_object_file():
    s = scene.load_scene(cfg("expsimu3D_preview128.txt"))
    assert s.use_moving_object == 1
    assert s.moving_object_file.decode() == "mvobjs3D_preview128.txt"


def test_unknown_adv_vel_type_rejected():
    # inReg_outMAC2 is not one of the six names LoadFromFile accepts
    with pytest.raises(ValueError):
        scene.load_scene(cfg("smokesim3d_preview64.txt"))


def test_missing_file_rejected(tmp_path):
    with pytest.raises(ValueError):
        scene.load_scene(str(tmp_path / "absent.txt"))


def test_unknown_solver_falls_back_to_open_flux(tmp_path):
    p = write_scene(tmp_path, "resolution 16 16 16\nglobalGridSolver NOT_A_SOLVER\n")
    assert scene.SOLVERS[scene.load_scene(p).solver] == "OPEN_FLUX"


def test_closed_solver_sets_closed_mean(tmp_path):
    p = write_scene(tmp_path, "resolution 16 12 8\nglobalGridSolver TOTAL_CLOSED_POISSON_SOR\n")
    s = scene.load_scene(p)
    assert s.config.sor.open_mask == 0 and s.config.sor.closed_mean == 1
    assert s.wall_faces == O.ALL_FACES


@pytest.mark.parametrize("solver, sub, closed", [("TOTAL_OPEN_FLUX_SOR", 1, 0), ("TOTAL_CLOSED_FLUX_SOR", 1, 1),
                                                 ("COARSE_OPEN_FLUX_SOR", 4, 0), ("COARSE_CLOSED_FLUX_SOR", 4, 1)])
def test_flux_solvers_map_to_the_coarse_flux_form(tmp_path, solver, sub, closed):
    # (maxIter-1)/innerIter+1 passes of innerIter iterations: 40, 15 -> 3 passes, 45 iterations
    p = write_scene(tmp_path, "resolution 16 12 8\nsubSize 4\nmaxIter 40\ninnerIter 15\nglobalGridSolver %s\n" % solver)
    c = scene.load_scene(p).config
    assert (c.sub_size, c.coarse_flux, c.sor.iters, c.sor.mixed_precision) == (sub, 1, 45, 0)
    assert (c.sor.open_mask, c.sor.closed_mean) == ((0, 1) if closed else (O.ALL_FACES, 0))


@pytest.mark.parametrize("solver, closed", [("OPEN_OCTREE_POISSON_SOR", 0), ("CLOSED_OCTREE_POISSON_SOR", 1)])
def test_octree_solvers_map_to_the_octree_projection(tmp_path, solver, closed):
    p = write_scene(tmp_path, "resolution 16 24 8\nmaxIter 33\noctree_thresh1 3\noctree_thresh3 9\n"
                    "globalGridSolver %s\n" % solver)
    s = scene.load_scene(p)
    c = s.config
    assert (c.octree, list(c.octree_thresh), c.sor.iters, c.sub_size, c.sor.mixed_precision) == (1, [3, 4, 9], 33, 0, 0)
    assert c.sor.closed_mean == closed
    # the closed solver's domain is walled by a shell of solid cells (its occupancy), not by open_mask
    assert s.wall_faces == (O.ALL_FACES if closed else 0)


def test_octree_pcg_solvers_stay_unmapped(tmp_path):
    for solver in ("OPEN_OCTREE_POISSON", "CLOSED_OCTREE_FLUX"):
        c = scene.load_scene(write_scene(tmp_path, "resolution 16 16 16\nglobalGridSolver %s\n" % solver)).config
        assert c.octree == 0 and c.sub_size == 0


def test_coarse_poisson_solvers_keep_max_iter(tmp_path):
    for solver in ("COARSE_OPEN_POISSON_SOR", "COARSE_CLOSED_POISSON_SOR"):
        p = write_scene(tmp_path, "resolution 16 12 8\nsubSize 4\nmaxIter 40\ninnerIter 15\nglobalGridSolver %s\n" % solver)
        c = scene.load_scene(p).config
        assert (c.sub_size, c.coarse_flux, c.sor.iters, c.sor.open_mask) == (4, 0, 40, O.ALL_FACES)


def test_box_object_closed_form():
    # |k+.5-D/2| < D/20, |j+.5-H/2| < H/20, |i+.5-W/2| < W/10 with integer D/20 etc.: at 64^3
    # k, j in 29..34 and i in 26..37
    occ = scene.initial_occupancy(cfg("expsimu3D_preview64.txt"))
    want = np.zeros((64, 64, 64), np.uint8)
    want[29:35, 29:35, 26:38] = 1
    np.testing.assert_array_equal(occ, want)


def test_sphere_object_brute_force():
    occ = scene.initial_occupancy(cfg("expsimu3D_fine128.txt"))
    c = np.arange(128) + 0.5 - 64
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    np.testing.assert_array_equal(occ, (r2 < 128 * 128 / 400.0).astype(np.uint8))


@pytest.mark.parametrize("solver, shell", [("TOTAL_CLOSED_POISSON_SOR", O.ALL_FACES), ("TOTAL_OPEN_POISSON_SOR", 0)])
@pytest.mark.parametrize("obj, code", [("null", 0), ("sphere", 1), ("box", 2), ("oval", 3), ("special1", 4)])
def test_init_occupancy_matches_oracle(tmp_path, solver, shell, obj, code):
    p = write_scene(tmp_path, "resolution 40 36 28\nintObj %s\nglobalGridSolver %s\n" % (obj, solver))
    occ = scene.initial_occupancy(p)
    want, _, _, _ = O.init_fine_grid(40, 36, 28, shell_mask=shell, intobj=code)
    np.testing.assert_array_equal(occ, want)


def test_moving_object_stamp(tmp_path):
    # mvobjs3D_preview64.txt: one 5x5x1 object, centre (2,2,0), at (32,32,32), 10 zero velocities
    mv = os.path.join(CFG, "mvobjs3D_preview64.txt")
    p = write_scene(tmp_path, "resolution 64 64 64\nuseMovingObject true\nmovingObjectFile %s\n"
                  "globalGridSolver TOTAL_OPEN_POISSON_SOR\n" % mv)
    W = H = D = 64
    u, v, w = O.zeros_faces(W, H, D)
    u += 7.0
    occ = scene.apply_moving_objects(p, 3, u, v, w)
    mask = np.array([[0, 0, 1, 0, 0], [0, 0, 1, 0, 0], [0, 0, 1, 0, 0], [0, 1, 0, 1, 0], [1, 0, 0, 0, 1]], np.uint8)
    want = np.zeros((D, H, W), np.uint8)
    want[32, 30:35, 30:35] = mask
    np.testing.assert_array_equal(occ, want)
    # zero object velocity is written on both faces of every solid voxel; other faces keep theirs
    touched = np.zeros_like(u, bool)
    kk, jj, ii = np.nonzero(want)
    touched[kk, jj, ii] = True
    touched[kk, jj, ii + 1] = True
    assert np.all(u[touched] == 0.0) and np.all(u[~touched] == 7.0)


def test_moving_object_motion(tmp_path):
    # a 2x1x1 object moving +1 voxel in x per step for two steps, then at rest
    mv = write_cfg(tmp_path, "1\n2 1 1\n0 0 0\n4 5 6\n2\n1 0 0\n1 0 0\n1 1\n", "obj.txt")
    p = write_scene(tmp_path, "resolution 16 12 8\nstepTime 0.5\nuseMovingObject true\n"
                  "movingObjectFile %s\nglobalGridSolver TOTAL_CLOSED_POISSON_SOR\n" % mv)
    shell, _, _, _ = O.init_fine_grid(16, 12, 8, shell_mask=O.ALL_FACES)
    for frames, x0, vel in ((0, 4, 0.0), (1, 5, 0.0), (2, 6, 0.0), (5, 6, 0.0)):
        u, v, w = O.zeros_faces(16, 12, 8)
        occ = scene.apply_moving_objects(p, frames, u, v, w)
        want = shell.copy()
        want[6, 5, x0:x0 + 2] = 1
        np.testing.assert_array_equal(occ, want, err_msg="frames=%d" % frames)
        # 1 voxel/step on the faces is (int)(1 * h/dt) = (int)((1/16)/0.5) = 0
        assert np.all(u[6, 5, x0:x0 + 3] == vel)


def test_moving_object_velocity_scale(tmp_path):
    # the voxel size is 1/width; 3 voxels per step at h/dt = (1/8)/0.03125 = 4 -> 12 on the faces
    mv = write_cfg(tmp_path, "1\n1 1 1\n0 0 0\n3 3 3\n1\n3 -2 1\n1\n", "obj.txt")
    p = write_scene(tmp_path, "resolution 8 8 8\nstepTime 0.03125\nuseMovingObject true\n"
                  "movingObjectFile %s\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\n" % mv)
    u, v, w = O.zeros_faces(8, 8, 8)
    occ = scene.apply_moving_objects(p, 0, u, v, w)
    assert occ.sum() == 1 and occ[3, 3, 3] == 1
    assert u[3, 3, 3] == 12 and u[3, 3, 4] == 12
    assert v[3, 3, 3] == -8 and v[3, 4, 3] == -8
    assert w[3, 3, 3] == 4 and w[4, 3, 3] == 4
    assert np.count_nonzero(u) == 2 and np.count_nonzero(v) == 2 and np.count_nonzero(w) == 2


def test_truncated_moving_object_file_rejected(tmp_path):
    mv = write_cfg(tmp_path, "1\n2 2 1\n0 0 0\n4 4 4\n1\n0 0 0\n1 1\n", "obj.txt")
    p = write_scene(tmp_path, "resolution 8 8 8\nuseMovingObject true\nmovingObjectFile %s\n" % mv)
    u, v, w = O.zeros_faces(8, 8, 8)
    with pytest.raises(ValueError):
        scene.apply_moving_objects(p, 0, u, v, w)


def test_no_source_rejected(tmp_path):
    with pytest.raises(ValueError):
        scene.load_scene(write_cfg(tmp_path, "resolution 16 16 16\n"))
