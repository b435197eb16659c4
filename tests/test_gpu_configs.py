"""GPU: the plain-forward workloads at their full sizes against the oracle.

Closed box at 64^3 -- closed walls, no obstacle, zero initial velocity, sphere source (r 0.08 at
(0.5, 0.1, 0.5), density 1, temperature 1), seeded +-0.01 temperature noise, alpha 0.05, beta 0.5,
eps 0.3, dt 0.02, red-black SOR with omega 1.8 for 200 iterations, 100 steps.  The oracle's
closed-domain mean folds the same per-tile FP64 partials in the same order as the device, so all
fields agree bit for bit after 100 steps.

Moving sphere at 128^3 -- a solid sphere (radius 0.1) moving on a seeded straight path, open top face,
walls elsewhere, mixed-precision projection to max|div| <= 1e-6.  Every step: the device reaches
the tolerance, the obstacle's faces carry its velocity; the first steps agree with the oracle's to
1e-5 of the field scale.
"""
import numpy as np
import pytest

import zq_oracle as O

pytestmark = pytest.mark.gpu

SRC = [[0.5, 0.1, 0.5, 0.08, 1.0, 1.0]]


def params(W, **kw):
    p = dict(h=1.0 / W, dt=0.02, steps=20, Tamb=0.0, beta=0.5, alpha=0.05, eps=0.3, aV=1.0, aT=1.0, aR=1.0,
             adv_vel_type=0, adv_sca_type=2, open_mask=0, closed_mean=1, omega=1.8, sor_iters=200, mixed=0,
             max_outer=0, tol=1e-6)
    p.update(kw)
    return p


def sim_for(W, p):
    import zq_smokesimulation_amd as z
    return z.SmokeSim3D(W, W, W, voxel_size=p["h"], step_time=p["dt"], steps=p["steps"], Tamb=p["Tamb"],
                        buoyancy=p["beta"], gravity=p["alpha"], confine=p["eps"], vel_atten=p["aV"],
                        temp_atten=p["aT"], dens_atten=p["aR"], adv_vel_type=p["adv_vel_type"],
                        adv_sca_type=p["adv_sca_type"], sor_open_mask=p["open_mask"],
                        sor_closed_mean=p["closed_mean"], sor_omega=p["omega"], sor_iters=p["sor_iters"],
                        sor_mixed_precision=p["mixed"], sor_max_outer=p["max_outer"], sor_tol=p["tol"],
                        warm_start=1)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def test_closed_box_64_100_steps():
    W = 64
    p = params(W)
    occ, u, v, w = O.init_fine_grid(W, W, W, O.ALL_FACES, 0)
    T = O.noise(W, W, W, 0, 0.01)
    R = np.zeros_like(T)
    sim = sim_for(W, p)
    sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
    sim.set_sphere_sources(SRC)
    for s in range(100):
        O.step(p, u, v, w, T, R, occ, spheres=SRC)
        sim.step(stats=False)
        if s % 20 == 19:
            print("closed box 64 step", s + 1, flush=True)
    g = sim.get_state()
    sim.close()
    err_rho = rel_l2(g["rho"], R)
    err_vel = max(rel_l2(g[k], a) for k, a in (("u", u), ("v", v), ("w", w)))
    print("closed box 64 after 100 steps: density rel L2 %.3g, velocity rel L2 %.3g, mass %.6g vs %.6g"
          % (err_rho, err_vel, g["rho"].sum(), R.sum()), flush=True)
    assert R.max() > 0.5 and v.max() > 0.0  # the plume rose
    assert err_rho == 0.0 and err_vel == 0.0
    for k, a in (("u", u), ("v", v), ("w", w), ("T", T), ("rho", R)):
        assert np.array_equal(g[k].view(np.uint32), a.view(np.uint32)), k


def test_moving_sphere_128_open_top():
    W = 128
    shell = O.ALL_FACES & ~O.OPEN_YHI
    p = params(W, open_mask=O.OPEN_YHI, closed_mean=0, mixed=1, sor_iters=40, max_outer=2000, tol=1e-6)
    rng = np.random.default_rng(11)
    start = rng.uniform(0.3, 0.7, size=3)
    speed = rng.uniform(-0.5, 0.5, size=3)  # domain widths per second
    occ, u, v, w = O.init_fine_grid(W, W, W, shell, 0)
    T = O.noise(W, W, W, 0, 0.01)
    R = np.zeros_like(T)
    sim = sim_for(W, p)
    sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
    sim.set_sphere_sources(SRC)
    p64 = np.zeros((W + 2, W + 2, W + 2))
    for s in range(12):
        c = start + speed * p["dt"] * s
        vel = tuple(float(x) for x in speed)
        sim.stamp_sphere_obstacle(float(c[0]), float(c[1]), float(c[2]), 0.1, vel, shell)
        st = sim.step()
        assert st["max_div"] <= 1e-6 and st["max_residual"] <= 1e-6, (s, st)
        if s < 3:
            O.stamp_sphere(occ, u, v, w, float(c[0]), float(c[1]), float(c[2]), 0.1, vel, shell)
            O.step(p, u, v, w, T, R, occ, p64=p64, spheres=SRC)
            g = sim.get_state()
            for name, a in zip(("u", "v", "w", "T", "rho"), (u, v, w, T, R)):
                scale = max(float(np.abs(a).max()), 1e-6)
                err = float(np.abs(g[name] - a).max())
                assert err <= 1e-5 * scale, "step %d %s: max err %g (scale %g)" % (s, name, err, scale)
        print("moving sphere 128 step", s, st["outer"], st["max_div"], flush=True)
    g = sim.get_state()
    sim.close()
    # every face of an obstacle voxel carries the obstacle's velocity (ZQ_ApplyMovingObject3D)
    solid = g["occ"].astype(bool)
    solid[0] = solid[:, 0] = solid[:, :, 0] = solid[:, :, -1] = solid[-1] = False
    kk, jj, ii = np.nonzero(solid)
    assert kk.size > 1000
    assert np.allclose(g["u"][kk, jj, ii], np.float32(speed[0])) and np.allclose(g["u"][kk, jj, ii + 1], np.float32(speed[0]))
    assert np.allclose(g["v"][kk, jj, ii], np.float32(speed[1])) and np.allclose(g["w"][kk, jj, ii], np.float32(speed[2]))
