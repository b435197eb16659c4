"""Pins the oracle's force and attenuation stages (ZQ_CUDA_AddForce3D.cu, ZQ_CUDA_Attenuation3D.cu)
with closed forms and invariants: buoyancy only at rest, the curl of known fields, confinement
perpendicular to the vorticity, unit gradient directions, and the face masks."""
import numpy as np
import pytest

import zq_oracle as O
from util import cells, faces


def _zero(W, H, D):
    return O.zeros_faces(W, H, D)


def test_buoyancy_only_at_rest():
    """u = 0, T uniform: v faces between fluid cells gain dt*beta*(T - Tamb); nothing else moves."""
    W, H, D = 9, 8, 7
    u, v, w = _zero(W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    occ[3, 3, 3] = 1
    T = np.full((D, H, W), 2.5, np.float32)
    R = np.zeros_like(T)
    beta, dt, Tamb = 0.5, 0.02, 0.5
    O.add_force(u, v, w, occ, T, R, 1.0 / W, dt, beta, 7.0, Tamb)
    assert not u.any() and not w.any()
    expect = np.float32(dt) * np.float32(beta * 2.0)
    # faces touching the solid cell stay zero, all others got the buoyancy increment
    touched = np.zeros_like(v, bool)
    touched[3, 3, 3] = touched[3, 4, 3] = True
    assert np.all(v[touched] == 0)
    assert np.all(v[~touched] == expect)


def test_density_term_extension():
    W, H, D = 6, 6, 6
    u, v, w = _zero(W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    T = np.zeros((D, H, W), np.float32)
    R = np.full((D, H, W), 2.0, np.float32)
    O.add_force(u, v, w, occ, T, R, 1.0 / W, 0.5, 0.0, 0.0, 0.0, alpha=0.25)
    assert np.all(v == np.float32(-0.25))


def _rotation(W, H, D, omega):
    """Solid-body rotation about z through the domain centre, sampled at face centres."""
    xc, yc = W / 2.0, H / 2.0
    u = np.zeros((D, H, W + 1), np.float32)
    v = np.zeros((D, H + 1, W), np.float32)
    w = np.zeros((D + 1, H, W), np.float32)
    y = np.arange(H) + 0.5
    x = np.arange(W) + 0.5
    u[:] = (-omega * (y - yc))[None, :, None]
    v[:] = (omega * (x - xc))[None, None, :]
    return u, v, w


def test_solid_rotation_vorticity_and_no_confinement():
    W, H, D = 16, 16, 12
    Om = 0.5
    h = 1.0 / W
    u, v, w = _rotation(W, H, D, Om)
    occ = np.zeros((D, H, W), np.uint8)
    T = np.zeros((D, H, W), np.float32)
    vort, gradn, force = O.add_force(u, v, w, occ, T, T.copy(), h, 0.01, 0.0, 4.0, 0.0, debug=True)
    interior = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(vort[interior][..., 2], 2 * Om / h, rtol=1e-5)
    assert np.all(vort[interior][..., 0] == 0) and np.all(vort[interior][..., 1] == 0)
    # border cells carry no vorticity (the reference zero-fills them)
    assert np.all(vort[0] == 0) and np.all(vort[:, :, 0] == 0)
    # |w| is constant two cells in from the border, so its gradient and the confinement vanish there
    deep = (slice(2, -2), slice(2, -2), slice(2, -2))
    assert np.all(gradn[deep] == 0)
    assert np.all(force[deep] == 0)


def test_shear_vorticity():
    """u = s*y gives w_z = -s, w_x = w_y = 0 (central differences over 2h in units of 1/h)."""
    W, H, D = 10, 12, 8
    s, h = 0.3, 1.0 / 10
    u = np.zeros((D, H, W + 1), np.float32)
    u[:] = (s * (np.arange(H) + 0.5))[None, :, None]
    v = np.zeros((D, H + 1, W), np.float32)
    w = np.zeros((D + 1, H, W), np.float32)
    occ = np.zeros((D, H, W), np.uint8)
    T = np.zeros((D, H, W), np.float32)
    vort, _, _ = O.add_force(u, v, w, occ, T, T.copy(), h, 0.01, 0.0, 0.0, 0.0, debug=True)
    inner = vort[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner[..., 2], -s / h, rtol=1e-5)
    np.testing.assert_allclose(inner[..., 3], s / h, rtol=1e-5)


def test_confinement_is_perpendicular_to_vorticity(rng):
    W, H, D = 14, 12, 10
    u, v, w = faces(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    T = np.zeros((D, H, W), np.float32)
    vort, gradn, force = O.add_force(u, v, w, occ, T, T.copy(), 1.0 / W, 0.01, 0.0, 2.0, 0.0, debug=True)
    dot = (force * vort[..., :3]).sum(-1)
    scale = np.linalg.norm(force, axis=-1) * np.linalg.norm(vort[..., :3], axis=-1) + 1e-30
    assert np.all(np.abs(dot) <= 1e-5 * scale)
    n = np.linalg.norm(gradn, axis=-1)
    assert np.all((np.abs(n - 1) < 1e-5) | (n == 0))


def test_face_masks_match_geometry(rng):
    """One solid cell: the force and attenuation leave exactly its six faces alone; domain faces of
    fluid cells are updated (the reference's 'inside cell' rule)."""
    W, H, D = 8, 7, 6
    for stage in ("force", "atten"):
        u, v, w = faces(rng, W, H, D)
        occ = np.zeros((D, H, W), np.uint8)
        occ[2, 3, 4] = 1
        T, R = cells(rng, W, H, D, 1, 2), cells(rng, W, H, D)
        u0, v0, w0 = u.copy(), v.copy(), w.copy()
        if stage == "force":
            O.add_force(u, v, w, occ, T, R, 1.0 / W, 0.1, 1.0, 0.0, 0.0)
            changed = [v != v0]
            fixed = [(v, v0, [(2, 3, 4), (2, 4, 4)])]
        else:
            O.attenuation(u, v, w, T, R, occ, 0.5, 1.0, 1.0)
            changed = [u != u0, v != v0, w != w0]
            fixed = [(u, u0, [(2, 3, 4), (2, 3, 5)]), (v, v0, [(2, 3, 4), (2, 4, 4)]),
                     (w, w0, [(2, 3, 4), (3, 3, 4)])]
        for a, a0, idx in fixed:
            for t in idx:
                assert a[t] == a0[t]
        total = sum(int(c.sum()) for c in changed)
        nfaces = sum(a.size for a in ((v,) if stage == "force" else (u, v, w)))
        assert total == nfaces - (2 if stage == "force" else 6)


def test_attenuation_closed_form(rng):
    W, H, D = 5, 6, 7
    u, v, w = faces(rng, W, H, D)
    T, R = cells(rng, W, H, D), cells(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    a = [x.copy() for x in (u, v, w, T, R)]
    O.attenuation(u, v, w, T, R, occ, 0.9, 0.8, 0.7)
    for x, x0, f in zip((u, v, w, T, R), a, (0.9, 0.9, 0.9, 0.8, 0.7)):
        assert np.array_equal(x, x0 * np.float32(f))


def _quadratic_field(W, H, D, comp, axis, a):
    """Velocity component `comp` (0 u, 1 v, 2 w) = a * (c + 1/2)^2 on every face, c the cell index along
    `axis` (0 x, 1 y, 2 z; axis != comp), so the averaged cell velocity is exactly a*(c + 1/2)^2."""
    u, v, w = O.zeros_faces(W, H, D)
    f = (u, v, w)[comp]
    c = np.arange(f.shape[2 - axis], dtype=np.float64) + 0.5
    shape = [1, 1, 1]
    shape[2 - axis] = -1
    f[...] = (a * c * c).reshape(shape).astype(np.float32)
    return u, v, w


_CASES = [(comp, axis) for comp in range(3) for axis in range(3) if comp != axis]


@pytest.mark.parametrize("comp,axis", _CASES)
def test_confinement_closed_form(comp, axis):
    """V = a (x_A + 1/2)^2 e_B (A = axis, B = comp) has curl w = a (2 x_A + 1)/h * eps_ABC e_C on interior
    cells (central differences over 2h are exact on a quadratic), so |w| grows along e_A and
    N = grad|w| / |grad|w|| = e_A two or more cells from the border.  Then f = eps h N x w
    = -eps a (2 x_A + 1) e_B: the confinement force opposes the velocity component.  The order
    w x N, or a central difference of the wrong sign, flips f.  On the border the one-sided
    differences (ZQ_CUDA_AddForce3D.cu:64-111) give N = +e_A at x_A = 0 and -e_A at the far end;
    one cell in, at x_A = 1 and along the other axes, N mixes e_A with the border's drop to |w| = 0
    and f = eps h (N x w) takes both of its terms (:161-184)."""
    W, H, D = 11, 10, 9
    dims = (W, H, D)
    a, eps = 0.25, 0.75
    h = 1.0 / 16.0
    u, v, w = _quadratic_field(W, H, D, comp, axis, a)
    occ = np.zeros((D, H, W), np.uint8)
    T = np.zeros((D, H, W), np.float32)
    vort, gradn, force = O.add_force(u, v, w, occ, T, T.copy(), h, 0.01, 0.0, eps, 0.0, debug=True)
    C = 3 - comp - axis
    lc = np.zeros((3, 3, 3))
    lc[0, 1, 2] = lc[1, 2, 0] = lc[2, 0, 1] = 1
    lc[0, 2, 1] = lc[2, 1, 0] = lc[1, 0, 2] = -1
    sgn = lc[axis, comp, C]

    def idx(fixed_axis_index, lo=2):
        """Index (k, j, i) tuple: `axis` coordinate as given, the other two in [lo, dim-1-lo]."""
        sl = [slice(lo, dims[ax] - lo) for ax in range(3)]
        sl[axis] = fixed_axis_index
        return tuple(sl[::-1])

    for c in range(1, dims[axis] - 1):
        wmag = a * (2 * c + 1) / h
        got = vort[idx(c, 1)]
        np.testing.assert_allclose(got[..., C], sgn * wmag, rtol=1e-6)
        assert np.all(got[..., comp] == 0) and np.all(got[..., axis] == 0)
        np.testing.assert_allclose(got[..., 3], wmag, rtol=1e-6)
    # deep interior: N = e_A, f = -eps a (2c+1) e_B; at c = dims - 2 the next cell is the border
    # (|w| = 0), so the central difference turns negative: N = -e_A, f = +eps a (2c+1) e_B
    for c in range(1, dims[axis] - 1):
        s = 1.0 if c < dims[axis] - 2 else -1.0
        n = gradn[idx(c)]
        f = force[idx(c)]
        np.testing.assert_allclose(n[..., axis], s, rtol=1e-6)
        assert np.all(n[..., comp] == 0) and np.all(n[..., C] == 0)
        np.testing.assert_allclose(f[..., comp], -s * eps * a * (2 * c + 1), rtol=1e-5)
        assert np.all(f[..., axis] == 0) and np.all(f[..., C] == 0)
    # one-sided ends along A: |w| = 0 on the border cells, so N = +e_A at c = 0, -e_A at the far end
    np.testing.assert_allclose(gradn[idx(0)][..., axis], 1.0, rtol=1e-6)
    np.testing.assert_allclose(gradn[idx(dims[axis] - 1)][..., axis], -1.0, rtol=1e-6)
    assert np.all(force[idx(0)] == 0) and np.all(force[idx(dims[axis] - 1)] == 0)
    # one cell in from a border across B (the velocity's own axis): |w| drops to 0 on the border
    # plane, so the central difference along B is +/-|w|/2 while along A it is 2a/h
    for side, cb in ((+1, 1), (-1, dims[comp] - 2)):
        for c in range(2, dims[axis] - 2):
            sl = [slice(2, dims[ax] - 2) for ax in range(3)]
            sl[axis] = c
            sl[comp] = cb
            key = tuple(sl[::-1])
            wmag = a * (2 * c + 1) / h
            gA, gB = 2 * a / h, side * 0.5 * wmag
            nA, nB = gA / np.hypot(gA, gB), gB / np.hypot(gA, gB)
            np.testing.assert_allclose(gradn[key][..., axis], nA, rtol=1e-5)
            np.testing.assert_allclose(gradn[key][..., comp], nB, rtol=1e-5)
            # f = eps h (N x w e_C * sgn): (e_A x e_C) = -sgn e_B, (e_B x e_C) = sgn e_A
            wC = sgn * wmag
            np.testing.assert_allclose(force[key][..., comp], eps * h * nA * wC * (-sgn), rtol=1e-5)
            np.testing.assert_allclose(force[key][..., axis], eps * h * nB * wC * sgn, rtol=1e-5)
