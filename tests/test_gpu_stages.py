"""GPU parity of each stage through the C ABI, against the CPU oracle on the same seeded inputs.

Bar: bit-exact.  The kernels and the oracle perform the same IEEE FP32 operations in the same
order (both built with -ffp-contract=off, correctly rounded divide/sqrt), and masks/indices are
integer logic, so any difference is a bug.  Grids are non-cubic and small (the oracle runs
them in well under a second).
"""
import numpy as np
import pytest

import zq_oracle as O
from util import cells, copies, faces, occupancy

pytestmark = pytest.mark.gpu

SHAPES = [(24, 20, 16), (17, 9, 13), (2, 2, 2)]


@pytest.fixture(scope="module")
def zs():
    import zq_smokesimulation_amd as z
    return z.stages


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32)) if a.dtype == np.float32 else np.flatnonzero(a != b)
    assert bad.size == 0, "%s: %d of %d differ, first at %d: %r vs %r" % (
        what, bad.size, a.size, bad[0], a.reshape(-1)[bad[0]], b.reshape(-1)[bad[0]])


@pytest.mark.parametrize("shape", SHAPES)
def test_attenuation(zs, rng, shape):
    W, H, D = shape
    print("attenuation", shape, flush=True)
    u, v, w = faces(rng, W, H, D)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.2, shell=False)
    ref = copies(u, v, w, T, R)
    O.attenuation(*ref, occ, 0.97, 0.9, 0.8)
    zs.attenuation(u, v, w, T, R, occ, 0.97, 0.9, 0.8)
    for a, b, n in zip((u, v, w, T, R), ref, "uvwTR"):
        assert_same(a, b, "atten " + n)


@pytest.mark.parametrize("shape", SHAPES)
def test_add_force(zs, rng, shape):
    W, H, D = shape
    print("add_force", shape, flush=True)
    u, v, w = faces(rng, W, H, D, 0.3)
    T = cells(rng, W, H, D, -1, 2)
    R = cells(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.15, shell=False)
    ref = copies(u, v, w)
    O.add_force(*ref, occ, T, R, 1.0 / W, 0.02, 0.7, 3.0, 0.25)
    zs.add_force(u, v, w, occ, T, 1.0 / W, 0.02, 0.7, 3.0, 0.25)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "force " + n)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_mac_to_regular(zs, rng, shape):
    W, H, D = shape
    u, v, w = faces(rng, W, H, D)
    assert_same(zs.mac_to_regular4(u, v, w), O.mac_to_cell4(u, v, w), "vel4")


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", SHAPES)
def test_advect_velocity(zs, rng, shape, mode):
    W, H, D = shape
    print("advect_velocity", shape, mode, flush=True)
    h = 1.0 / W
    steps, dt = 7, 0.05
    u, v, w = faces(rng, W, H, D, 0.4)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    ref = copies(u, v, w)
    O.advect_velocity(*ref, occ, h, steps, dt / steps, mode, True)
    zs.prepare_advection(W, H, D, h, steps, dt / steps)
    zs.advect_velocity(u, v, w, occ, mode, True)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "advect mode %d %s" % (mode, n))


@pytest.mark.parametrize("mode", [0, 1, 3, 5])
def test_advect_velocity_closed_apply(zs, rng, mode):
    """Closed apply; for the MAC BFECC mode this includes the corrected field the reference leaves on
    the faces its final apply skips."""
    W, H, D = 20, 18, 12
    h = 1.0 / W
    print("advect closed", mode, flush=True)
    u, v, w = faces(rng, W, H, D, 0.4)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    ref = copies(u, v, w)
    O.advect_velocity(*ref, occ, h, 5, 0.01, mode, False)
    zs.prepare_advection(W, H, D, h, 5, 0.01)
    zs.advect_velocity(u, v, w, occ, mode, False)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "advect closed %d %s" % (mode, n))


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("shape", SHAPES)
def test_advect_scalar(zs, rng, shape, mode):
    W, H, D = shape
    print("advect_scalar", shape, mode, flush=True)
    h = 1.0 / W
    u, v, w = faces(rng, W, H, D, 0.5)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    To, Ro = O.advect_scalar(u, v, w, occ, T, R, h, 20, 0.03 / 20, mode)
    zs.prepare_advection(W, H, D, h, 20, 0.03 / 20)
    T2, R2 = np.zeros_like(T), np.zeros_like(R)
    zs.advect_scalar(u, v, w, occ, T, R, T2, R2, mode)
    assert_same(T2, To, "T")
    assert_same(R2, Ro, "rho")


def test_advect_scalar_bfecc_zeroes(zs, rng):
    """The reference's Advect_Scalar has no BFECC case and returns zero fields."""
    W, H, D = 8, 8, 8
    u, v, w = faces(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    zs.prepare_advection(W, H, D, 1.0 / W, 4, 0.01)
    for mode in (1, 3):
        T2, R2 = np.ones_like(T), np.ones_like(R)
        zs.advect_scalar(u, v, w, occ, T, R, T2, R2, mode)
        assert not T2.any() and not R2.any()


def test_advect_unknown_velocity_mode_rejected(zs, rng):
    import zq_smokesimulation_amd as z
    W, H, D = 8, 8, 8
    u, v, w = faces(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    zs.prepare_advection(W, H, D, 1.0 / W, 4, 0.01)
    with pytest.raises(z.ZQ3DError):
        zs.advect_velocity(u, v, w, occ, 6, True)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_sor_open_reference_api(zs, rng, shape):
    W, H, D = shape
    print("sor_open", shape, flush=True)
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    ref = copies(u, v, w)
    O.project(*ref, occ, O.ALL_FACES, 0.0, 1.0, 37)
    zs.solve_open(u, v, w, occ, 37)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "open SOR " + n)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_sor_closed_reference_api(zs, rng, shape):
    W, H, D = shape
    print("sor_closed", shape, flush=True)
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=True)
    mean, _ = O.closed_mean_divergence(u, v, w, occ)
    ref = copies(u, v, w)
    O.project(*ref, occ, 0, mean, 1.0, 41)
    zs.solve_closed(u, v, w, occ, mean, 41)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "closed SOR " + n)


@pytest.mark.parametrize("open_mask", [0, O.OPEN_YHI, O.OPEN_XLO | O.OPEN_ZHI, O.ALL_FACES])
def test_sor_ex_omega(zs, rng, open_mask):
    W, H, D = 22, 19, 15
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    ref = copies(u, v, w)
    O.project(*ref, occ, open_mask, 0.0, 1.8, 60)
    zs.solve_ex(u, v, w, occ, open_mask=open_mask, omega=1.8, iters=60)
    for a, b, n in zip((u, v, w), ref, "uvw"):
        assert_same(a, b, "SOR omega=1.8 mask=%d %s" % (open_mask, n))


def test_sor_mixed_precision(zs, rng):
    """FP64 refinement: same outer-iteration sequence as the oracle; final divergence below tol."""
    W, H, D = 20, 24, 16
    # velocities of 0.05: the faces' FP32 rounding (a few 1e-9) stays well below tol
    u, v, w = faces(rng, W, H, D, scale=0.05)
    occ = occupancy(rng, W, H, D, 0.1, shell=True)
    tol = 1e-7
    ref = copies(u, v, w)
    outer, st_o, _ = O.project_mp(*ref, occ, 0, True, 1.7, 40, 400, tol)
    _, st = zs.solve_ex(u, v, w, occ, open_mask=0, closed_mean=True, omega=1.7, iters=40, mixed=True,
                        max_outer=400, tol=tol)
    assert st["outer"] == outer
    assert st["max_residual"] <= tol
    # the passes go on until the stored (FP32-rounded) faces meet the bound, not only the residual
    assert st["max_div"] <= tol and st["converged"] == 1 and st_o[2] <= tol
    # the FP64 mean is summed in a different order on the device: agreement to rounding
    for a, b in zip((u, v, w), ref):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)


@pytest.mark.parametrize("rule", [1, 2])
@pytest.mark.parametrize("open_mask", [O.OPEN_YHI, 63])
def test_sor_mixed_precision_pass_rules(zs, rng, open_mask, rule):
    """Fitted (pass_rule 1) and gated (2) refinement passes: the device chooses the oracle's pass
    lengths from the same values, so passes, total sweeps and the corrected faces agree bit for bit
    (open domains: no FP64 mean summed in another order), and the bound is met."""
    W, H, D = 36, 28, 24
    print("mixed precision, pass rule", rule, "open mask", open_mask, flush=True)
    u, v, w = faces(rng, W, H, D, scale=0.05)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    tol = 1e-7
    iters = 40 if rule == 1 else 2000
    sweeps = {}
    for r in (0, rule):
        n = 40 if r == 0 else iters
        ref, got = copies(u, v, w), copies(u, v, w)
        outer, st_o, _ = O.project_mp(*ref, occ, open_mask, False, 1.9, n, 400, tol, pass_rule=r)
        _, st = zs.solve_ex(*got, occ, open_mask=open_mask, closed_mean=False, omega=1.9, iters=n, mixed=True,
                            max_outer=400, tol=tol, pass_rule=r)
        assert st["outer"] == outer and st["sweeps"] == int(st_o[0]), (r, st, outer, st_o)
        assert st["converged"] == 1 and st["max_div"] <= tol
        for a, b, name in zip(got, ref, "uvw"):
            assert_same(a, b, "pass rule %d %s" % (r, name))
        sweeps[r] = st["sweeps"]
    assert sweeps[0] % 40 == 0
    print("sweeps:", sweeps, flush=True)


def test_sor_mixed_precision_stops_at_the_rounding_floor(zs, rng):
    """Velocities of order 1 stored in FP32 carry rounding of a few 1e-7 per face, so a bound of 1e-7
    on the stored faces' divergence is out of reach: the passes stop once they no longer shrink it
    (far before max_outer), the result says the bound was not met, and it is the oracle's."""
    W, H, D = 20, 24, 16
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=True)
    ref = copies(u, v, w)
    outer, st_o, _ = O.project_mp(*ref, occ, 0, True, 1.7, 40, 400, 1e-7)
    _, st = zs.solve_ex(u, v, w, occ, open_mask=0, closed_mean=True, omega=1.7, iters=40, mixed=True,
                        max_outer=400, tol=1e-7)
    assert st["outer"] == outer < 100
    assert st["converged"] == 0 and st["max_residual"] <= 1e-7 and 1e-7 < st["max_div"] < 2e-6
    assert st_o[2] > 1e-7


def test_sor_mixed_precision_reports_an_unmet_bound(zs, rng):
    """With too few passes for the tolerance the projection still returns, and says so."""
    W, H, D = 20, 24, 16
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=True)
    _, st = zs.solve_ex(u, v, w, occ, open_mask=0, closed_mean=True, omega=1.7, iters=5, mixed=True,
                        max_outer=2, tol=1e-9)
    assert st["outer"] == 2 and st["converged"] == 0 and st["max_div"] > 1e-9
    _, st = zs.solve_ex(u, v, w, occ, open_mask=0, closed_mean=True, omega=1.7, iters=40, mixed=False)
    assert st["converged"] == -1
