"""GPU: ensembles of uneven grids (Part 6 of the C ABI) step every member bit for bit as a
SmokeSim3D of the same configuration -- eagerly, through hipGraph replays (checked after every
replay, with eager work and synchronisation between replays), and after the launch-shape
autotuner has changed the captured step."""
import numpy as np
import pytest

import zq_oracle as O

pytestmark = pytest.mark.gpu

MEMBERS = [
    # (W, H, D), closed, obstacle, alpha, beta, sor iters
    ((33, 40, 28), True, False, 0.05, 0.5, 30),
    ((48, 32, 40), False, False, 0.02, 1.0, 40),
    ((40, 44, 36), True, True, 0.1, 0.2, 25),
    ((32, 32, 32), True, False, 0.07, 0.7, 40),
    ((61, 37, 45), False, True, 0.03, 0.9, 35),
]


def member_args(closed, alpha, beta, iters, W):
    return dict(voxel_size=1.0 / W, step_time=0.02, steps=20, Tamb=0.0, buoyancy=beta, gravity=alpha, confine=0.3,
                vel_atten=0.99, temp_atten=0.99, dens_atten=0.99, adv_vel_type=0, adv_sca_type=2,
                sor_open_mask=0 if closed else O.ALL_FACES, sor_closed_mean=1 if closed else 0, sor_omega=1.8,
                sor_iters=iters, sor_mixed_precision=0)


def initial(W, H, D, closed, obstacle, seed):
    rng = np.random.default_rng(seed)
    occ, u, v, w = O.init_fine_grid(W, H, D, O.ALL_FACES if closed else 0, 0)
    if obstacle:
        kk, jj, ii = np.meshgrid(np.arange(D) + 0.5, np.arange(H) + 0.5, np.arange(W) + 0.5, indexing="ij")
        occ[((kk - 0.5 * D) ** 2 + (jj - 0.6 * H) ** 2 + (ii - 0.45 * W) ** 2) < (0.15 * W) ** 2] = 1
    u[...] = 0.05 * rng.standard_normal(u.shape)
    v[...] = 0.05 * rng.standard_normal(v.shape)
    w[...] = 0.05 * rng.standard_normal(w.shape)
    T = O.noise(W, H, D, seed, 0.01)
    rho = np.zeros((D, H, W), np.float32)
    return dict(u=u, v=v, w=w, T=T, rho=rho, occ=occ)


SPHERE = [[0.5, 0.1, 0.5, 0.08, 1.0, 1.0]]
BOX = [[0.3, 0.15, 0.6, 0.05, 0.05, 0.05, 0.8, 2.0]]


def build():
    import zq_smokesimulation_amd as z
    from zq_smokesimulation_amd.ensemble import SmokeEnsemble, member_config
    cfgs, sims = [], []
    for n, ((W, H, D), closed, obst, a, b, it) in enumerate(MEMBERS):
        args = member_args(closed, a, b, it, W)
        cfgs.append(member_config(W, H, D, **args))
        sim = z.SmokeSim3D(W, H, D, **args)
        st = initial(W, H, D, closed, obst, n)
        sim.set_state(**st)
        sim.set_sphere_sources(SPHERE)
        sim.set_box_sources(BOX if n % 2 == 0 else [])
        sims.append((sim, st))
    ens = SmokeEnsemble(cfgs)
    for n, (sim, st) in enumerate(sims):
        ens.set_state(n, **st)
        ens.set_sources(n, boxes=BOX if n % 2 == 0 else [], spheres=SPHERE)
    return ens, [s for s, _ in sims]


def compare(ens, sims, tag):
    for n, sim in enumerate(sims):
        a = ens.get_state(n)
        b = sim.get_state()
        for k in ("u", "v", "w", "T", "rho"):
            np.testing.assert_array_equal(a[k], b[k], err_msg="%s member %d field %s" % (tag, n, k))


def test_ensemble_matches_single_grids():
    ens, sims = build()
    print("ensemble cells", ens.cells(), flush=True)
    for step in range(2):
        print("eager step", step, flush=True)
        ens.step(1, graph=False)
        for sim in sims:
            sim.step()
        ens.synchronize()
        compare(ens, sims, "eager step %d" % step)
    for step in range(3):
        print("graph replay", step, flush=True)
        ens.step(1, graph=True)
        for sim in sims:
            sim.step()
        ens.synchronize()
        compare(ens, sims, "graph replay %d" % step)
    tune = ens.autotune(5)
    print("autotune", tune["best"], flush=True)
    ens.step(2, graph=True)
    for sim in sims:
        sim.step()
        sim.step()
    ens.synchronize()
    compare(ens, sims, "after autotune")
    assert np.abs(ens.download(0, 4)).max() > 0.5  # density was injected and advected
    for sim in sims:
        sim.close()
    ens.close()


def test_every_sweep_candidate_is_bit_identical():
    """Every sweep candidate of the autotuner (half sweeps, and one-launch iterations on colour-split
    storage with tiles of 32 or 64 slots) steps every member bit for bit as its SmokeSim3D; a member
    150 cells wide spans several tiles per row."""
    import zq_smokesimulation_amd as z
    from zq_smokesimulation_amd.ensemble import SmokeEnsemble, member_config
    members = [((150, 20, 18), True, True, 0.05, 0.5, 23), ((33, 40, 28), False, False, 0.02, 1.0, 30),
               ((61, 37, 45), True, True, 0.03, 0.9, 24)]
    cfgs, sims, states = [], [], []
    for n, ((W, H, D), closed, obst, a, b, it) in enumerate(members):
        args = member_args(closed, a, b, it, W)
        cfgs.append(member_config(W, H, D, **args))
        sim = z.SmokeSim3D(W, H, D, **args)
        st = initial(W, H, D, closed, obst, 10 + n)
        sim.set_state(**st)
        sim.set_sphere_sources(SPHERE)
        sims.append(sim)
        states.append(st)
    ens = SmokeEnsemble(cfgs)
    for n, st in enumerate(states):
        ens.set_state(n, **st)
        ens.set_sources(n, spheres=SPHERE)
    tune = ens.autotune(2)
    kinds = {c["kind"] for c in tune["candidates"]}
    assert kinds == {"half_sweeps", "split_one_launch"}
    for c in tune["candidates"]:
        if c["kind"] == "half_sweeps" and c["blocks"] not in (0, 256):
            continue
        print("candidate", c, flush=True)
        ens.set_sweep(c["kind"], c["tbx"], c["tby"], c["blocks"])
        ens.step(1, graph=True)
        for sim in sims:
            sim.step()
        ens.synchronize()
        compare(ens, sims, "candidate %s" % c)
    for sim in sims:
        sim.close()
    ens.close()


def test_ensemble_rejects_mixed_precision():
    from zq_smokesimulation_amd import ZQ3DError
    from zq_smokesimulation_amd.ensemble import SmokeEnsemble, member_config
    with pytest.raises(ZQ3DError):
        SmokeEnsemble([member_config(16, 16, 16, sor_mixed_precision=1)])
