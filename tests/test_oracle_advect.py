"""Pins the oracle's sampling and backtrace (ZQ_CUDA_Advection3D.cu) without retyping it:
scipy's map_coordinates (order 1, mode 'nearest' = trilinear with clamp-to-edge) for the filter,
closed forms for translation, and hand-worked backtrace end points."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

import zq_oracle as O
from util import cells, faces


def test_trilinear_matches_scipy(rng):
    g = rng.standard_normal((7, 5, 9)).astype(np.float32)  # (nz, ny, nx)
    pts = rng.uniform(-3, 12, size=(400, 3)).astype(np.float32)
    ours = np.array([O.sample3(g, x, y, z) for x, y, z in pts])
    ref = map_coordinates(g.astype(np.float64), [pts[:, 2], pts[:, 1], pts[:, 0]], order=1, mode="nearest")
    np.testing.assert_allclose(ours, ref, rtol=0, atol=2e-6)


def test_trilinear_exact_at_texels(rng):
    g = rng.standard_normal((4, 6, 5)).astype(np.float32)
    for (k, j, i) in [(0, 0, 0), (3, 5, 4), (2, 3, 1)]:
        assert O.sample3(g, float(i), float(j), float(k)) == g[k, j, i]


def _uniform(W, H, D, c):
    u = np.full((D, H, W + 1), c[0], np.float32)
    v = np.full((D, H + 1, W), c[1], np.float32)
    w = np.full((D + 1, H, W), c[2], np.float32)
    return u, v, w


def test_zero_velocity_is_identity(rng):
    W, H, D = 10, 8, 6
    u, v, w = np.zeros((D, H, W + 1), np.float32), np.zeros((D, H + 1, W), np.float32), np.zeros((D + 1, H, W), np.float32)
    occ = (rng.random((D, H, W)) < 0.2).astype(np.uint8)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    for mode in (O.ADV_SCA_REG, O.ADV_SCA_MAC):
        To, Ro = O.advect_scalar(u, v, w, occ, T, R, 0.1, 20, 0.001, mode)
        assert np.array_equal(To, T) and np.array_equal(Ro, R)


def test_single_substep_is_identity(rng):
    """steps == 1: the loop accepts the start point and stops, so nothing moves (the off-by-one)."""
    W, H, D = 9, 7, 5
    u, v, w = faces(rng, W, H, D, 2.0)
    occ = np.zeros((D, H, W), np.uint8)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    To, Ro = O.advect_scalar(u, v, w, occ, T, R, 1.0 / W, 1, 0.05, O.ADV_SCA_REG)
    assert np.array_equal(To, T)
    u2, v2, w2 = u.copy(), v.copy(), w.copy()
    O.advect_velocity(u2, v2, w2, occ, 1.0 / W, 1, 0.05, O.ADV_VEL_INMAC_OUTMAC)
    assert np.array_equal(u2, u) and np.array_equal(v2, v) and np.array_equal(w2, w)


@pytest.mark.parametrize("mode", [O.ADV_SCA_REG, O.ADV_SCA_MAC])
def test_uniform_flow_translates_linear_field(mode):
    """T = a + b.x advected by a uniform velocity c moves by (steps-1)*c*dtt/h cells."""
    W, H, D = 40, 36, 32
    h, steps, dtt = 1.0 / W, 11, 0.004
    c = np.array([0.3, -0.2, 0.15], np.float32)
    u, v, w = _uniform(W, H, D, c)
    occ = np.zeros((D, H, W), np.uint8)
    z, y, x = np.meshgrid(np.arange(D) + 0.5, np.arange(H) + 0.5, np.arange(W) + 0.5, indexing="ij")
    b = np.array([0.02, 0.05, -0.03])
    T = (1.0 + b[0] * x + b[1] * y + b[2] * z).astype(np.float32)
    To, _ = O.advect_scalar(u, v, w, occ, T, T.copy(), h, steps, dtt, mode)
    shift = (steps - 1) * c.astype(np.float64) * dtt / h
    expect = 1.0 + b[0] * (x - shift[0]) + b[1] * (y - shift[1]) + b[2] * (z - shift[2])
    inner = (slice(4, -4), slice(4, -4), slice(4, -4))
    np.testing.assert_allclose(To[inner], expect[inner], rtol=0, atol=2e-5)


def test_backtrace_stops_before_solid():
    """+x flow of 0.3 cells per substep from x = 3.5; cells 0 and 1 solid.
    Accepted points 3.5, 3.2, 2.9, 2.6, 2.3, ~2.0; the next (~1.7) is in cell 1 -> stop."""
    W, H, D = 8, 4, 4
    h, dtt = 1.0 / W, 0.01
    cvel = 0.3 * h / dtt
    u, v, w = _uniform(W, H, D, (cvel, 0.0, 0.0))
    occ = np.zeros((D, H, W), np.uint8)
    occ[:, :, :2] = 1
    last = O.backtrace(u, v, w, occ, h, 20, dtt, (3.5, 2.5, 2.5))
    assert abs(last[0] - 2.0) < 1e-5 and last[1] == 2.5 and last[2] == 2.5
    # the end point is the last accepted position, never inside the solid
    assert int(np.floor(last[0])) >= 2


def test_backtrace_stops_at_domain_edge():
    """-y flow toward y = H: the walk leaves [0,H] and the last inside point is returned."""
    W, H, D = 6, 6, 6
    h, dtt = 1.0 / W, 0.01
    cvel = -0.4 * h / dtt
    u, v, w = _uniform(W, H, D, (0.0, cvel, 0.0))
    occ = np.zeros((D, H, W), np.uint8)
    last = O.backtrace(u, v, w, occ, h, 50, dtt, (3.5, 4.5, 3.5))
    # 4.5, 4.9, 5.3, 5.7 inside; 6.1 outside
    assert abs(last[1] - 5.7) < 1e-5


def test_backtrace_substep_count():
    """With no obstacles, `steps` passes apply steps-1 displacements."""
    W = H = D = 64
    h, dtt = 1.0 / W, 0.001
    cvel = 0.25 * h / dtt
    u, v, w = _uniform(W, H, D, (cvel, 0.0, 0.0))
    occ = np.zeros((D, H, W), np.uint8)
    for steps in (1, 2, 5, 20):
        last = O.backtrace(u, v, w, occ, h, steps, dtt, (40.5, 10.5, 10.5))
        assert abs(last[0] - (40.5 - 0.25 * (steps - 1))) < 1e-4


def test_velocity_advection_uniform_flow_unchanged():
    """A uniform velocity field is a fixed point of every velocity advection mode."""
    W, H, D = 12, 10, 8
    c = (0.7, -0.3, 0.2)
    occ = np.zeros((D, H, W), np.uint8)
    for mode in (O.ADV_VEL_INREG_OUTREG, O.ADV_VEL_INMAC_OUTMAC, O.ADV_VEL_INREG_OUTMAC):
        u, v, w = _uniform(W, H, D, c)
        O.advect_velocity(u, v, w, occ, 1.0 / W, 20, 0.002, mode)
        for a, val in zip((u, v, w), c):
            assert np.array_equal(a, np.full_like(a, np.float32(val)))


def test_apply_mask_keeps_solid_faces(rng):
    W, H, D = 10, 9, 8
    u, v, w = faces(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    occ[4, 4, 4] = 1
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    O.advect_velocity(u, v, w, occ, 1.0 / W, 10, 0.01, O.ADV_VEL_INREG_OUTREG)
    assert u[4, 4, 4] == u0[4, 4, 4] and u[4, 4, 5] == u0[4, 4, 5]
    assert v[4, 4, 4] == v0[4, 4, 4] and v[4, 5, 4] == v0[4, 5, 4]
    assert w[4, 4, 4] == w0[4, 4, 4] and w[5, 4, 4] == w0[5, 4, 4]
    assert not np.array_equal(u, u0)


def test_mac_cell_roundtrip_linear():
    """Face averages of a field linear in x reproduce the cell-centre values exactly."""
    W, H, D = 6, 5, 4
    xs = np.arange(W + 1, dtype=np.float32) * 0.25
    u = np.broadcast_to(xs, (D, H, W + 1)).astype(np.float32).copy()
    v = np.zeros((D, H + 1, W), np.float32)
    w = np.zeros((D + 1, H, W), np.float32)
    vel4 = O.mac_to_cell4(u, v, w)
    assert np.array_equal(vel4[..., 0], np.broadcast_to((np.arange(W) + 0.5) * 0.25, (D, H, W)).astype(np.float32))
    u2, _, _ = O.cell4_to_mac(vel4)
    assert np.array_equal(u2[:, :, 1:-1], u[:, :, 1:-1])
    assert np.array_equal(u2[:, :, 0], vel4[:, :, 0, 0]) and np.array_equal(u2[:, :, -1], vel4[:, :, -1, 0])


# ---------------------------------------------------------------- BFECC velocity modes
BFECC = [O.ADV_VEL_INREG_OUTREG_BFECC, O.ADV_VEL_INMAC_OUTMAC_BFECC, O.ADV_VEL_INREG_OUTMAC_BFECC]
PLAIN_OF = {O.ADV_VEL_INREG_OUTREG_BFECC: O.ADV_VEL_INREG_OUTREG, O.ADV_VEL_INMAC_OUTMAC_BFECC: O.ADV_VEL_INMAC_OUTMAC,
            O.ADV_VEL_INREG_OUTMAC_BFECC: O.ADV_VEL_INREG_OUTMAC}


@pytest.mark.parametrize("mode", BFECC)
def test_bfecc_zero_and_uniform_velocity_are_fixed_points(rng, mode):
    """Zero velocity: nothing moves; a uniform velocity is a steady solution, so every face keeps it
    (1.5 c - 0.5 c = c: the increment's weights sum to one)."""
    W, H, D = 9, 7, 6
    occ = (rng.random((D, H, W)) < 0.15).astype(np.uint8)
    for c in ((0.0, 0.0, 0.0), (0.7, -0.4, 0.25)):
        u, v, w = _uniform(W, H, D, c)
        u2, v2, w2 = u.copy(), v.copy(), w.copy()
        O.advect_velocity(u2, v2, w2, occ, 1.0 / W, 6, 0.05, mode, is_open=True)
        np.testing.assert_allclose(u2, u, rtol=0, atol=1e-6)
        np.testing.assert_allclose(v2, v, rtol=0, atol=1e-6)
        np.testing.assert_allclose(w2, w, rtol=0, atol=1e-6)


@pytest.mark.parametrize("mode", BFECC)
def test_bfecc_is_more_accurate_than_the_plain_mode(mode):
    """u = U0 + eps sin(2 pi x / L) along x (v = w = 0): one step is a translation of the ripple by
    (steps-1) U0 dtt / h cells up to O(eps^2).  Linear interpolation damps the ripple; BFECC's error
    compensation must leave a clearly smaller error than the plain mode (a sign slip in the back
    advection or in the increment makes it larger).  The regular-grid pair also averages faces to
    cells and back, which scales the ripple by cos(pi/L)^2 in both modes; that is its target."""
    W, H, D = 64, 6, 6
    L, U0, eps = 8.0, 1.0, 0.01
    h, steps, dtt = 1.0 / W, 5, 0.37 / (4 * W)   # shift 0.37 cells
    shift = (steps - 1) * U0 * dtt / h
    x = np.arange(W + 1, dtype=np.float64)
    prof = U0 + eps * np.sin(2 * np.pi * x / L)
    damp = np.cos(np.pi / L) ** 2 if mode == O.ADV_VEL_INREG_OUTREG_BFECC else 1.0
    want = U0 + eps * damp * np.sin(2 * np.pi * (x - shift) / L)
    occ = np.zeros((D, H, W), np.uint8)
    errs = {}
    for m in (mode, PLAIN_OF[mode]):
        u = np.broadcast_to(prof.astype(np.float32), (D, H, W + 1)).copy()
        v = np.zeros((D, H + 1, W), np.float32)
        w = np.zeros((D + 1, H, W), np.float32)
        O.advect_velocity(u, v, w, occ, h, steps, dtt, m, is_open=True)
        errs[m] = np.abs(u[:, :, 16:48] - want[16:48]).max()
    assert errs[mode] < 0.5 * errs[PLAIN_OF[mode]], errs


def test_mac_bfecc_leaves_the_corrected_field_on_skipped_faces(rng):
    """The MAC variant overwrites the caller's faces with phi~ = 1.5 phi - 0.5 phi^ before the final
    apply (ZQ_CUDA_Advection3D.cu:1301-1307), so closed domain faces (never applied) change, while the
    other modes leave them alone."""
    W, H, D = 10, 8, 7
    u, v, w = faces(rng, W, H, D, 0.5)
    occ = np.zeros((D, H, W), np.uint8)
    out = {}
    for m in (O.ADV_VEL_INMAC_OUTMAC_BFECC, O.ADV_VEL_INMAC_OUTMAC, O.ADV_VEL_INREG_OUTMAC_BFECC):
        u2, v2, w2 = u.copy(), v.copy(), w.copy()
        O.advect_velocity(u2, v2, w2, occ, 1.0 / W, 6, 0.02, m, is_open=False)
        out[m] = u2
    assert not np.array_equal(out[O.ADV_VEL_INMAC_OUTMAC_BFECC][:, :, 0], u[:, :, 0])
    assert np.array_equal(out[O.ADV_VEL_INMAC_OUTMAC][:, :, 0], u[:, :, 0])
    assert np.array_equal(out[O.ADV_VEL_INREG_OUTMAC_BFECC][:, :, 0], u[:, :, 0])
