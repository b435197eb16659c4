"""GPU: the C++ driver (ZQ_SmokeSimulation3D + the command line of ZQ_SmokeSimulation3D/main.cpp)
run on small configs, frame by frame, against the oracle's step; and ZQ3D_AddForce3D_ex with its
density term and extra force field against the oracle's force stage.

The command line runs as a child process and writes frame<N>.di3 files; density is read back from
them.  Open domains are compared bit for bit; closed ones to a stated tolerance (the mean
divergence is summed in another order on the device)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import zq_oracle as O
from util import faces

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "zq_smokesimulation_amd", "lib", "ZQ_SmokeSimulation3D")


def read_di3(path):
    raw = open(path, "rb").read()
    assert raw[:5] == b"float"
    W, H, D, ch = np.frombuffer(raw[16:32], np.int32)
    assert ch == 1
    return np.frombuffer(raw[33:], np.float32).reshape(D, H, W)


def run_cli(tmp_path, text, frames):
    cfg = tmp_path / "scene.txt"
    cfg.write_text(text)
    out = tmp_path / "out"
    out.mkdir()
    print("cli", cfg, flush=True)
    r = subprocess.run([CLI, str(cfg), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return [read_di3(str(out / ("frame%d.di3" % f))) for f in range(frames)]


def oracle_frames(path, frames, moving=False, sub_size=0, guides=None, guide_coeff=0.0):
    """The oracle's step driven by the parameters the config parses to (sub_size > 0: the coarse +
    sub-grid projection of the COARSE_*_POISSON_SOR solvers)."""
    from zq_smokesimulation_amd import scene
    s = scene.load_scene(path)
    c = s.config
    W, H, D = c.width, c.height, c.depth
    p = dict(h=c.voxel_size, dt=c.step_time, steps=c.steps, Tamb=c.Tamb, beta=c.buoyancy, alpha=c.gravity,
             eps=c.confine, aV=c.vel_atten, aT=c.temp_atten, aR=c.dens_atten, adv_vel_type=c.adv_vel_type,
             adv_sca_type=c.adv_sca_type, open_mask=c.sor.open_mask, closed_mean=c.sor.closed_mean,
             omega=c.sor.omega, sor_iters=c.sor.iters, mixed=0, max_outer=0, tol=1e-6)
    occ = scene.initial_occupancy(path)
    u, v, w = O.zeros_faces(W, H, D)
    T = np.zeros((D, H, W), np.float32)
    R = np.zeros_like(T)
    boxes = [list(s.box_sources[i]) for i in range(s.n_box_sources)]
    out = []
    for f in range(frames):
        if moving:
            occ = scene.apply_moving_objects(path, f, u, v, w)
        if c.octree:
            O.octree_step(p, u, v, w, T, R, occ, boxes=boxes, thresh=tuple(c.octree_thresh),
                          closed=bool(c.sor.closed_mean), guiding=guides is not None)
        elif c.coarse_flux:
            # the *_FLUX_SOR solvers: the flux form on the coarse grid (s = 1 for TOTAL_*)
            O.coarse_step(p, u, v, w, T, R, occ, c.sub_size, boxes=boxes, open_mask=c.sor.open_mask,
                          closed_mean=bool(c.sor.closed_mean), flux=True)
        elif sub_size:
            O.coarse_step(p, u, v, w, T, R, occ, sub_size, boxes=boxes,
                          guide=guides[f] if guides else None, guide_coeff=guide_coeff)
        else:
            O.step(p, u, v, w, T, R, occ, boxes=boxes)
        out.append(R.copy())
    return out


BASE = ("resolution 32 24 16\nsource 0.5 0.15 0.5 0.12 0.1 0.12 0.8 20\nTamb 0\nconfineCoeff 2\n"
        "stepTime 0.03\nvelAttenCoeff 1\ndensityAttenCoeff 0.5\nmaxIter 40\nadvVelType inReg_outReg\n"
        "advScaType Reg\nexportType di3\n")


@pytest.mark.parametrize("resident", ["false", "true"])
@pytest.mark.parametrize("obj", ["sphere", "box"])
def test_cli_open_bitexact(tmp_path, resident, obj):
    text = BASE + "frameCount 3\nintObj %s\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\ndeviceResident %s\n" % (
        obj, resident)
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3)
    for f in range(3):
        np.testing.assert_array_equal(got[f], want[f], err_msg="frame %d" % f)
    assert want[2].max() > 0.1


@pytest.mark.parametrize("resident", ["false", "true"])
def test_cli_closed(tmp_path, resident):
    text = BASE + "frameCount 3\nintObj sphere\nglobalGridSolver TOTAL_CLOSED_POISSON_SOR\ndeviceResident %s\n" % (
        resident)
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3)
    for f in range(3):
        scale = max(1.0, float(np.abs(want[f]).max()))
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-5 * scale, err_msg="frame %d" % f)


@pytest.mark.parametrize("resident", ["false", "true"])
@pytest.mark.parametrize("solver", ["COARSE_OPEN_POISSON_SOR", "COARSE_CLOSED_POISSON_SOR"])
def test_cli_coarse_projection(tmp_path, resident, solver):
    """The reference's default solver: the block interiors go through a matrix-free solve on the device
    and a dense operator in the oracle, so frames agree to a tolerance (1e-4 of the density scale)."""
    text = BASE + "frameCount 3\nintObj sphere\nsubSize 4\nglobalGridSolver %s\ndeviceResident %s\n" % (
        solver, resident)
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3, sub_size=4)
    for f in range(3):
        scale = max(1.0, float(np.abs(want[f]).max()))
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-4 * scale, err_msg="frame %d" % f)
    assert want[2].max() > 0.1


@pytest.mark.parametrize("resident", ["false", "true"])
@pytest.mark.parametrize("solver", ["TOTAL_OPEN_FLUX_SOR", "TOTAL_CLOSED_FLUX_SOR", "COARSE_OPEN_FLUX_SOR",
                                    "COARSE_CLOSED_FLUX_SOR"])
def test_cli_flux_solvers(tmp_path, resident, solver):
    """The flux solvers: (maxIter-1)/innerIter+1 passes of innerIter iterations (40, 15 -> 45).
    TOTAL_OPEN_FLUX_SOR (s = 1, open) is bit-exact; block interiors (COARSE_*) and the closed
    mean (summed in another order on the device) agree to 1e-4 of the density scale."""
    from zq_smokesimulation_amd import scene
    text = BASE + "frameCount 3\nintObj sphere\nsubSize 4\ninnerIter 15\nglobalGridSolver %s\ndeviceResident %s\n" % (
        solver, resident)
    got = run_cli(tmp_path, text, 3)
    c = scene.load_scene(str(tmp_path / "scene.txt")).config
    assert c.coarse_flux == 1 and c.sor.iters == 45 and c.sub_size == (1 if solver.startswith("TOTAL") else 4)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3)
    for f in range(3):
        if solver == "TOTAL_OPEN_FLUX_SOR":
            np.testing.assert_array_equal(got[f], want[f], err_msg="frame %d" % f)
        else:
            scale = max(1.0, float(np.abs(want[f]).max()))
            np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-4 * scale, err_msg="frame %d" % f)
    assert want[2].max() > 0.1


def test_cli_random_reinjection_takes_the_staged_path(tmp_path):
    """randDensity scales each source cell's density by rand() % 256 / 255 (ZQ_SmokeSimulationUtils3D.cpp:3842-3882):
    with deviceResident true the driver still runs it on the host grid, so the source holds
    differing densities in [0, 0.8] rather than one value."""
    text = BASE + "frameCount 1\nintObj sphere\nrandDensity true\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\n" \
        "deviceResident true\n"
    got = run_cli(tmp_path, text, 1)[0]
    src = got[got > 0]
    assert src.size > 100
    assert src.max() <= 0.8 + 1e-6 and len(np.unique(src)) > 20


@pytest.mark.parametrize("resident", ["false", "true"])
@pytest.mark.parametrize("solver", ["OPEN_OCTREE_POISSON_SOR", "CLOSED_OCTREE_POISSON_SOR"])
def test_cli_octree_solvers(tmp_path, resident, solver):
    """The octree solvers (leaves from octree_thresh1..3, maxIter passes): leaf interiors go through a
    matrix-free solve on the device and a dense operator in the oracle, and the closed mean is summed
    in another order, so frames agree to 1e-4 of the density scale."""
    from zq_smokesimulation_amd import scene
    text = BASE + ("frameCount 3\nintObj sphere\noctree_thresh1 1\noctree_thresh2 2\noctree_thresh3 4\n"
                   "globalGridSolver %s\ndeviceResident %s\n" % (solver, resident))
    got = run_cli(tmp_path, text, 3)
    c = scene.load_scene(str(tmp_path / "scene.txt")).config
    assert c.octree == 1 and list(c.octree_thresh) == [1, 2, 4] and c.sor.iters == 40
    want = oracle_frames(str(tmp_path / "scene.txt"), 3)
    for f in range(3):
        scale = max(1.0, float(np.abs(want[f]).max()))
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-4 * scale, err_msg="frame %d" % f)
    assert want[2].max() > 0.1


@pytest.mark.parametrize("resident", ["false", "true"])
@pytest.mark.parametrize("adv", ["inReg_outReg_BFECC", "inMAC_outMAC_BFECC", "inReg_outMAC_BFECC"])
def test_cli_bfecc_advection_open(tmp_path, resident, adv):
    text = BASE.replace("inReg_outReg", adv) + \
        "frameCount 2\nintObj sphere\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\ndeviceResident %s\n" % resident
    got = run_cli(tmp_path, text, 2)
    want = oracle_frames(str(tmp_path / "scene.txt"), 2)
    for f in range(2):
        np.testing.assert_array_equal(got[f], want[f], err_msg="frame %d" % f)


@pytest.mark.parametrize("resident", ["false", "true"])
def test_cli_guiding_control(tmp_path, resident):
    """useGuidingControl: frame f blends <guidingFold>/<f>.guidedat (coarse resolution) into the coarse delta."""
    s = 4
    W, H, D = 32, 24, 16
    rng = np.random.default_rng(5)
    gdir = tmp_path / "guide"
    gdir.mkdir()
    guides = []
    for f in range(3):
        gu, gv, gw = faces(rng, W // s, H // s, D // s, 0.2)
        extra = np.zeros(2 * (W // s) * (H // s) * (D // s), np.float32)  # density, temperature
        (gdir / ("%d.guidedat" % f)).write_bytes(np.concatenate([gu.ravel(), gv.ravel(), gw.ravel(), extra]).tobytes())
        guides.append((gu, gv, gw))
    text = BASE + ("frameCount 3\nintObj sphere\nsubSize %d\nglobalGridSolver COARSE_OPEN_POISSON_SOR\n"
                   "useGuidingControl true\nguidingCoeff 0.3\nguidingFold %s\ndeviceResident %s\n" % (s, gdir, resident))
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3, sub_size=s, guides=guides, guide_coeff=0.3)
    for f in range(3):
        scale = max(1.0, float(np.abs(want[f]).max()))
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-4 * scale, err_msg="frame %d" % f)
    # the guide moved the flow: the unguided run differs
    plain = oracle_frames(str(tmp_path / "scene.txt"), 3, sub_size=s)
    assert np.abs(plain[2] - want[2]).max() > 1e-3


def test_cli_mac_advection_open(tmp_path):
    text = BASE.replace("inReg_outReg", "inMAC_outMAC") + \
        "frameCount 2\nintObj null\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\ndeviceResident false\n"
    got = run_cli(tmp_path, text, 2)
    want = oracle_frames(str(tmp_path / "scene.txt"), 2)
    for f in range(2):
        np.testing.assert_array_equal(got[f], want[f], err_msg="frame %d" % f)


def test_cli_moving_object(tmp_path):
    mv = tmp_path / "obj.txt"
    # a 3x3x3 block moving one voxel in +x per frame for two frames
    mv.write_text("1\n3 3 3\n1 1 1\n10 12 8\n2\n1 0 0\n1 0 0\n" + " ".join(["1"] * 27) + "\n")
    text = BASE + ("frameCount 3\nintObj null\nglobalGridSolver TOTAL_OPEN_POISSON_SOR\nuseMovingObject true\n"
                   "movingObjectFile %s\n" % mv)
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3, moving=True)
    for f in range(3):
        np.testing.assert_array_equal(got[f], want[f], err_msg="frame %d" % f)


@pytest.mark.parametrize("resident", ["false", "true"])
def test_cli_octree_guiding(tmp_path, resident):
    """useGuidingControl with an octree solver: the guide files are loaded per frame as the reference
    does, and the projection adds the special-region step (its CUDA branch reads no guide values)."""
    s = 4
    W, H, D = 32, 24, 16
    gdir = tmp_path / "guide"
    gdir.mkdir()
    for f in range(3):
        n = (W // s + 1) * (H // s) * (D // s) + (W // s) * (H // s + 1) * (D // s) + \
            (W // s) * (H // s) * (D // s + 1) + 2 * (W // s) * (H // s) * (D // s)
        (gdir / ("%d.guidedat" % f)).write_bytes(np.zeros(n, np.float32).tobytes())
    text = BASE + ("frameCount 3\nintObj sphere\noctree_thresh1 1\noctree_thresh2 2\noctree_thresh3 4\n"
                   "subSize %d\nglobalGridSolver OPEN_OCTREE_POISSON_SOR\nuseGuidingControl true\n"
                   "guidingCoeff 0.3\nguidingFold %s\ndeviceResident %s\n" % (s, gdir, resident))
    got = run_cli(tmp_path, text, 3)
    want = oracle_frames(str(tmp_path / "scene.txt"), 3, guides=[None] * 3)
    for f in range(3):
        scale = max(1.0, float(np.abs(want[f]).max()))
        np.testing.assert_allclose(got[f], want[f], rtol=0, atol=1e-4 * scale, err_msg="frame %d" % f)
    plain = oracle_frames(str(tmp_path / "scene.txt"), 3)
    assert np.abs(plain[2] - want[2]).max() > 1e-4


def test_cli_rejects_unbuilt_solver(tmp_path):
    cfg = tmp_path / "scene.txt"
    cfg.write_text(BASE + "globalGridSolver OPEN_OCTREE_POISSON\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([CLI, str(cfg), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not built" in r.stdout


@pytest.mark.parametrize("alpha, with_extra", [(0.0, True), (0.07, False), (0.07, True)])
def test_add_force_ex(alpha, with_extra):
    from zq_smokesimulation_amd import _lib
    rng = np.random.default_rng(3)
    W, H, D = 21, 18, 13
    occ = (rng.random((D, H, W)) < 0.1).astype(np.uint8)
    u = rng.standard_normal((D, H, W + 1)).astype(np.float32)
    v = rng.standard_normal((D, H + 1, W)).astype(np.float32)
    w = rng.standard_normal((D + 1, H, W)).astype(np.float32)
    T = rng.random((D, H, W)).astype(np.float32)
    R = rng.random((D, H, W)).astype(np.float32)
    ef = rng.standard_normal((D, H, W, 3)).astype(np.float32) if with_extra else None
    h, dt, beta, eps, Tamb = 1.0 / W, 0.03, 0.7, 4.0, 0.2
    go = [a.copy() for a in (u, v, w)]
    O.add_force(u, v, w, occ, T, R, h, dt, beta, eps, Tamb, alpha=alpha, extra_force=ef)
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    print("add_force_ex", alpha, with_extra, flush=True)
    ms = _lib.lib().ZQ3D_AddForce3D_ex(ptr(go[0]), ptr(go[1]), ptr(go[2]), ptr(occ), ptr(T), ptr(R), ptr(ef),
                                       h, dt, beta, alpha, eps, Tamb, W, H, D)
    assert ms >= 0
    for a, b, n in zip(go, (u, v, w), "uvw"):
        np.testing.assert_array_equal(a, b, err_msg=n)
