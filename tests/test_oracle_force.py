"""Pins the oracle's force and attenuation stages (ZQ_CUDA_AddForce3D.cu, ZQ_CUDA_Attenuation3D.cu)
with closed forms and invariants: buoyancy only at rest, the curl of known fields, confinement
perpendicular to the vorticity, unit gradient directions, and the face masks."""
import numpy as np

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
