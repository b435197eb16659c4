"""Writes c1_128_200.json: the oracle's run of the 128^3 moving-obstacle workload, for
tests/test_gpu_workloads.py.

128^3 MAC grid, FP32; walls on five faces and an open top (p = 0 above y = H); a solid sphere of
radius 0.1 moving on a seeded straight path (start and end uniform in [0.3, 0.7]^3, the end reached
after 200 steps), voxelized every step, its faces carrying its velocity; sphere source of radius
0.08 at (0.5, 0.1, 0.5) with density 1 and temperature 1, seeded +-0.01 temperature noise, Tamb 0,
alpha 0.05, beta 0.5, eps 0.3, dt 0.02, 20 advection substeps, no attenuation; mixed-precision
red-black SOR (omega 1.8, 40 FP32 sweeps per pass, FP64 residual) to max|div| <= 1e-6 with a warm
start; 200 steps.

Recorded: per step the refinement passes and max|div|, and after the last step the density's
16^3 sums over 8^3 blocks, its total, and SHA-256 digests of the final u, v, w, T, rho bytes.
The oracle is deterministic (per-plane partials folded in plane order), so these numbers do not
depend on the machine or thread count.

    python tests/golden/make_c1_golden.py [--steps 200]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import zq_oracle as O  # noqa: E402

W = 128
STEPS = 200
SRC = [[0.5, 0.1, 0.5, 0.08, 1.0, 1.0]]
SHELL = O.ALL_FACES & ~O.OPEN_YHI


def params():
    return dict(h=1.0 / W, dt=0.02, steps=20, Tamb=0.0, beta=0.5, alpha=0.05, eps=0.3, aV=1.0, aT=1.0, aR=1.0,
                adv_vel_type=0, adv_sca_type=2, open_mask=O.OPEN_YHI, closed_mean=0, omega=1.8, sor_iters=40,
                mixed=1, max_outer=2000, tol=1e-6)


def path(nsteps=STEPS):
    """Sphere centre (domain units) before step s, and its velocity (domain widths per second)."""
    rng = np.random.default_rng(11)
    start = rng.uniform(0.3, 0.7, size=3)
    end = rng.uniform(0.3, 0.7, size=3)
    vel = (end - start) / (nsteps * params()["dt"])
    return [start + (end - start) * s / nsteps for s in range(nsteps)], tuple(float(x) for x in vel)


def initial_state():
    occ, u, v, w = O.init_fine_grid(W, W, W, SHELL, 0)
    T = O.noise(W, W, W, 0, 0.01)
    R = np.zeros_like(T)
    return occ, u, v, w, T, R


def block_sums(a, b=8):
    D, H, Wd = a.shape
    return a.astype(np.float64).reshape(D // b, b, H // b, b, Wd // b, b).sum(axis=(1, 3, 5))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=STEPS)
    ap.add_argument("--out", default=os.path.join(HERE, "c1_128_200.json"))
    args = ap.parse_args()
    p = params()
    centres, vel = path()
    occ, u, v, w, T, R = initial_state()
    p64 = np.zeros((W + 2, W + 2, W + 2))
    passes, divs = [], []
    t0 = time.time()
    for s in range(args.steps):
        c = centres[s]
        O.stamp_sphere(occ, u, v, w, float(c[0]), float(c[1]), float(c[2]), 0.1, vel, SHELL)
        st = O.step(p, u, v, w, T, R, occ, p64=p64, spheres=SRC)
        passes.append(int(round(st[0] / p["sor_iters"])))
        divs.append(float(st[2]))
        if s % 10 == 9:
            print("step %d: passes %d, max|div| %.3g, %.0f s" % (s + 1, passes[-1], divs[-1], time.time() - t0),
                  flush=True)
    out = {
        "workload": __doc__.split("\n\n")[1].replace("\n", " "),
        "steps": args.steps,
        "passes": passes,
        "max_div": divs,
        "rho_total": float(R.astype(np.float64).sum()),
        "rho_block_sums_8": block_sums(R).round(10).tolist(),
        "sha256": {k: digest(a) for k, a in (("u", u), ("v", v), ("w", w), ("T", T), ("rho", R))},
    }
    with open(args.out, "w") as f:
        json.dump(out, f)
    print("wrote %s (%d steps, %.0f s)" % (args.out, args.steps, time.time() - t0))


if __name__ == "__main__":
    main()
