"""GPU parity of whole steps of the device-resident simulation against the oracle's step.

Bit-exact where no reduction is involved (open domains); with a closed domain the mean
divergence is an FP64 sum whose order differs on the device, so fields are compared to a stated
tolerance.  The mixed-precision runs also check the post-projection divergence bound.
"""
import numpy as np
import pytest

import zq_oracle as O

pytestmark = pytest.mark.gpu


def oracle_params(W, dt=0.02, **kw):
    p = dict(h=1.0 / W, dt=dt, steps=20, Tamb=0.0, beta=0.5, alpha=0.05, eps=0.3, aV=1.0, aT=1.0, aR=1.0,
             adv_vel_type=0, adv_sca_type=2, open_mask=0, closed_mean=1, omega=1.8, sor_iters=200, mixed=0,
             max_outer=0, tol=1e-6)
    p.update(kw)
    return p


def make_sim(W, H, D, p, sphere_sources):
    import zq_smokesimulation_amd as z
    sim = z.SmokeSim3D(W, H, D, voxel_size=p["h"], step_time=p["dt"], steps=p["steps"], Tamb=p["Tamb"],
                       buoyancy=p["beta"], gravity=p["alpha"], confine=p["eps"], vel_atten=p["aV"],
                       temp_atten=p["aT"], dens_atten=p["aR"], adv_vel_type=p["adv_vel_type"],
                       adv_sca_type=p["adv_sca_type"], sor_open_mask=p["open_mask"],
                       sor_closed_mean=p["closed_mean"], sor_omega=p["omega"], sor_iters=p["sor_iters"],
                       sor_mixed_precision=p["mixed"], sor_max_outer=p["max_outer"], sor_tol=p["tol"],
                       sor_pass_rule=p.get("pass_rule", 0), warm_start=p.get("warm_start", 1))
    sim.set_sphere_sources(sphere_sources)
    return sim


def initial_state(W, H, D, shell):
    occ, u, v, w = O.init_fine_grid(W, H, D, shell, 0)
    T = O.noise(W, H, D, 7, 0.01)
    R = np.zeros_like(T)
    return occ, u, v, w, T, R


def run_both(W, H, D, p, nsteps, shell, sphere_sources):
    occ, u, v, w, T, R = initial_state(W, H, D, shell)
    sim = make_sim(W, H, D, p, sphere_sources)
    sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
    p64 = np.zeros((D + 2, H + 2, W + 2)) if p["mixed"] else None
    stats_o, stats_g = [], []
    for s in range(nsteps):
        print("step", s, flush=True)
        stats_o.append(O.step(p, u, v, w, T, R, occ, p64=p64, spheres=sphere_sources))
        stats_g.append(sim.step())
    g = sim.get_state()
    sim.close()
    return (u, v, w, T, R), g, stats_o, stats_g


SRC = [[0.5, 0.2, 0.5, 0.12, 1.0, 1.0]]


def test_step_open_bit_exact():
    """Open domain (p = 0 on all faces), omega 1.8: no reductions, bit-exact for 6 steps."""
    W, H, D = 24, 28, 20
    p = oracle_params(W, open_mask=O.ALL_FACES, closed_mean=0, sor_iters=60)
    ref, g, _, _ = run_both(W, H, D, p, 6, 0, SRC)
    for name, a in zip(("u", "v", "w", "T", "rho"), ref):
        b = g[name]
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, "%s: %d cells differ (max |d| %g)" % (name, bad.size, np.abs(a - b).max())


@pytest.mark.parametrize("mode", [1, 3, 5])
def test_step_open_bfecc_bit_exact(mode):
    """Whole open-domain steps with each BFECC velocity mode (extra device scratch in the simulation)."""
    W, H, D = 20, 16, 18
    p = oracle_params(W, open_mask=O.ALL_FACES, closed_mean=0, sor_iters=40, adv_vel_type=mode)
    ref, g, _, _ = run_both(W, H, D, p, 3, 0, SRC)
    for name, a in zip(("u", "v", "w", "T", "rho"), ref):
        b = g[name]
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, "mode %d %s: %d cells differ (max |d| %g)" % (mode, name, bad.size, np.abs(a - b).max())


def test_step_closed_box_32():
    """Closed shell, alpha/beta buoyancy, confinement, 200 SOR iterations, at 32^3."""
    W = H = D = 32
    p = oracle_params(W)
    ref, g, _, stats_g = run_both(W, H, D, p, 5, O.ALL_FACES, [[0.5, 0.1, 0.5, 0.08, 1.0, 1.0]])
    for name, a in zip(("u", "v", "w", "T", "rho"), ref):
        b = g[name]
        scale = max(float(np.abs(a).max()), 1e-6)
        err = float(np.abs(a - b).max())
        assert err <= 1e-5 * scale, "%s: max err %g (scale %g)" % (name, err, scale)
    assert np.isfinite(stats_g[-1]["max_div"])


def test_step_mixed_precision_moving_sphere():
    """Open top, walls elsewhere, moving solid sphere, mixed-precision SOR to max|div| <= 1e-6."""
    W = H = D = 32
    shell = O.ALL_FACES & ~O.OPEN_YHI
    p = oracle_params(W, open_mask=O.OPEN_YHI, closed_mean=0, mixed=1, sor_iters=30, max_outer=200, tol=5e-7)
    occ, u, v, w, T, R = initial_state(W, H, D, shell)
    sim = make_sim(W, H, D, p, SRC)
    sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
    p64 = np.zeros((D + 2, H + 2, W + 2))
    vel = (0.3, 0.0, 0.1)
    for s in range(4):
        print("mp step", s, flush=True)
        c = (0.3 + 0.02 * s, 0.5, 0.4 + 0.01 * s)
        O.stamp_sphere(occ, u, v, w, c[0], c[1], c[2], 0.1, vel, shell)
        sim.stamp_sphere_obstacle(c[0], c[1], c[2], 0.1, vel, shell)
        st_o = O.step(p, u, v, w, T, R, occ, p64=p64, spheres=SRC)
        st_g = sim.step()
        assert st_g["max_residual"] <= 5e-7
        assert st_g["max_div"] <= 1e-6, st_g
        assert st_g["outer"] == int(st_o[0]) // 30
    g = sim.get_state()
    assert np.array_equal(g["occ"], occ)
    for name, a in zip(("u", "v", "w", "T", "rho"), (u, v, w, T, R)):
        b = g[name]
        scale = max(float(np.abs(a).max()), 1e-6)
        err = float(np.abs(a - b).max())
        assert err <= 1e-5 * scale, "%s: max err %g (scale %g)" % (name, err, scale)


def test_every_sweep_tile_shape_is_bit_identical():
    """The autotuner may pick any fused-sweep tile shape: all give the oracle's iteration exactly."""
    import zq_smokesimulation_amd as z
    W, H, D = 36, 30, 40
    rng = np.random.default_rng(3)
    from util import faces, occupancy
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    ref = [a.copy() for a in (u, v, w)]
    O.project(*ref, occ, O.OPEN_YHI, 0.0, 1.8, 17)
    sim = z.SmokeSim3D(W, H, D, sor_open_mask=O.OPEN_YHI, sor_closed_mean=0, sor_omega=1.8, sor_iters=17)
    rep = sim.autotune_sor(3)
    assert rep["best"] >= 0 and len(rep["candidates"]) >= 6
    for c in rep["candidates"]:
        print("shape", c, flush=True)
        sim.set_sor_shape(c["ty"], c["zc"])
        sim.set_state(u=u, v=v, w=w, occ=occ)
        sim.project()
        g = sim.get_state()
        for name, a in zip("uvw", ref):
            assert np.array_equal(g[name].view(np.uint32), a.view(np.uint32)), (c, name)
    sim.close()


@pytest.mark.parametrize("ty", [-4, -8, -16, -32, -116, -132])
def test_one_launch_iteration_multi_tile(ty):
    """The one-launch red-black iteration on colour-split storage (one or, ty <= -100, two iterations
    per launch) works on tiles of 64 slots (128 cells) in x; at W = 150 a row spans three tiles, the
    last partial, and the chunks of zc planes leave a partial last chunk.  Every (rows, planes)
    shape reproduces the oracle's SOR bit for bit, for an odd iteration count (two-iteration launches
    plus one single) and an even one."""
    import zq_smokesimulation_amd as z
    from util import faces, occupancy
    W, H, D = 150, 22, 19
    rng = np.random.default_rng(5)
    u, v, w = faces(rng, W, H, D)
    occ = occupancy(rng, W, H, D, 0.1, shell=False)
    for iters in (17, 12):
        ref = [a.copy() for a in (u, v, w)]
        O.project(*ref, occ, O.OPEN_YHI | O.OPEN_XLO, 0.0, 1.8, iters)
        sim = z.SmokeSim3D(W, H, D, sor_open_mask=O.OPEN_YHI | O.OPEN_XLO, sor_closed_mean=0, sor_omega=1.8,
                           sor_iters=iters)
        for zc in (8, 16, 32):
            print("shape", ty, zc, "iters", iters, flush=True)
            sim.set_sor_shape(ty, zc)
            assert sim.sor_kernel() == ("k_sor_split_rb2" if ty <= -100 else "k_sor_split_rb")
            sim.set_state(u=u, v=v, w=w, occ=occ)
            sim.project()
            g = sim.get_state()
            for name, a in zip("uvw", ref):
                assert np.array_equal(g[name].view(np.uint32), a.view(np.uint32)), (ty, zc, iters, name)
        sim.close()


def test_one_launch_iteration_mixed_precision_same_bits():
    """Mixed-precision steps (FP64 residual, FP32 correction sweeps, several refinement passes) give
    the same bits -- fields and FP64 pressure -- with the interleaved fused sweeps as with
    single-colour and one-launch sweeps on colour-split storage (the correction kept in split
    storage between passes), and reach the divergence bound."""
    import zq_smokesimulation_amd._lib as L
    W, H, D = 96, 64, 48
    p = oracle_params(W, mixed=1, sor_iters=15, max_outer=400, tol=1e-6, open_mask=O.OPEN_YHI, closed_mean=0)
    shell = O.ALL_FACES & ~O.OPEN_YHI
    out = []
    for shape in ((0, 0), (-1, 0), (-8, 8)):
        occ, u, v, w, T, R = initial_state(W, H, D, shell)
        sim = make_sim(W, H, D, p, SRC)
        sim.set_sor_shape(*shape)
        sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
        sts = []
        for s in range(3):
            print("shape", shape, "step", s, flush=True)
            st = sim.step()
            assert st["max_div"] <= 1e-6, st
            sts.append(st["outer"])
        g = sim.get_state()
        g["p64"] = sim.download(L.FIELD_P64)
        out.append((g, sts))
        sim.close()
    g0, s0 = out[0]
    assert s0[0] > 1
    for g1, s1 in out[1:]:
        assert s0 == s1
        for name in ("u", "v", "w", "T", "rho"):
            assert np.array_equal(g0[name].view(np.uint32), g1[name].view(np.uint32)), name
        assert np.array_equal(g0["p64"].view(np.uint64), g1["p64"].view(np.uint64)), "p64"


@pytest.mark.parametrize("shape", [(0, 0), (-4, 8), (-16, 16)])
def test_gated_passes_match_the_oracle(shape):
    """Gated refinement passes (pass_rule 2): every pass sweeps until an iteration's largest update
    meets the threshold; the device runs the oracle's iterations, so passes, sweep counts and every
    field agree bit for bit on an open-top box, whatever sweep shape the handle carries (the gated
    iteration uses the one-launch kernel at that shape's rows and planes, or 32 x 16)."""
    import zq_smokesimulation_amd._lib as L
    W, H, D = 96, 64, 48
    shell = O.ALL_FACES & ~O.OPEN_YHI
    p = oracle_params(W, mixed=1, sor_iters=1000, max_outer=400, tol=1e-6, open_mask=O.OPEN_YHI, closed_mean=0,
                      pass_rule=2)
    occ, u, v, w, T, R = initial_state(W, H, D, shell)
    sim = make_sim(W, H, D, p, SRC)
    sim.set_sor_shape(*shape)
    sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
    p64 = np.zeros((D + 2, H + 2, W + 2))
    for s in range(4):
        print("gated, shape", shape, "step", s, flush=True)
        st_o = O.step(p, u, v, w, T, R, occ, p64=p64, spheres=SRC)
        st_g = sim.step()
        assert st_g["sweeps"] == int(st_o[0]), (st_g, st_o)
        assert st_g["max_div"] <= 1e-6 and st_g["converged"] == 1, st_g
        assert st_g["sweeps"] % 1000 != 0  # the gate ended every pass
    g = sim.get_state()
    g["p64"] = sim.download(L.FIELD_P64)
    sim.close()
    for name, a in zip(("u", "v", "w", "T", "rho"), (u, v, w, T, R)):
        assert np.array_equal(g[name].view(np.uint32), a.view(np.uint32)), name
    assert np.array_equal(g["p64"].view(np.uint64), p64.view(np.uint64)), "p64"


def test_extrapolated_warm_start_matches_the_oracle():
    """warm_start 2: each projection starts from 2 p(n-1) - p(n-2) (p(n-1) alone on the first step);
    the oracle applies the same extrapolation before its step, so sweeps, fields and the FP64
    pressure agree bit for bit on an open-top box.  (Whether it saves sweeps depends on the flow:
    on this start-up transient it does not; on the closed-box plume it does.)"""
    import zq_smokesimulation_amd._lib as L
    W, H, D = 96, 64, 48
    shell = O.ALL_FACES & ~O.OPEN_YHI
    sweeps = {}
    for warm in (1, 2):
        p = oracle_params(W, mixed=1, sor_iters=80, max_outer=400, tol=1e-6, open_mask=O.OPEN_YHI, closed_mean=0,
                          pass_rule=1, warm_start=warm)
        occ, u, v, w, T, R = initial_state(W, H, D, shell)
        sim = make_sim(W, H, D, p, SRC)
        sim.set_state(u=u, v=v, w=w, T=T, rho=R, occ=occ)
        p64 = np.zeros((D + 2, H + 2, W + 2))
        prev = np.zeros_like(p64)
        total = []
        for s in range(10):
            print("warm start", warm, "step", s, flush=True)
            if warm == 2 and s > 0:  # step 0 starts from zero, step 1 from p(0) alone
                O.extrapolate_warm_start(p64, prev, s > 1)
            st_o = O.step(p, u, v, w, T, R, occ, p64=p64, spheres=SRC)
            st_g = sim.step()
            assert st_g["sweeps"] == int(st_o[0]) and st_g["converged"] == 1, (st_g, st_o)
            total.append(st_g["sweeps"])
        g = sim.get_state()
        g["p64"] = sim.download(L.FIELD_P64)
        sim.close()
        for name, a in zip(("u", "v", "w", "T", "rho"), (u, v, w, T, R)):
            assert np.array_equal(g[name].view(np.uint32), a.view(np.uint32)), (warm, name)
        assert np.array_equal(g["p64"].view(np.uint64), p64.view(np.uint64)), (warm, "p64")
        sweeps[warm] = total
    print("sweeps per step, previous pressure:", sweeps[1], "extrapolated:", sweeps[2], flush=True)
