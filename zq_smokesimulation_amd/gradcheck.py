"""Finite-difference check of the smoke-control gradient (the differentiable-control workload).

    python -m zq_smokesimulation_amd.gradcheck [--size 192] [--steps 16] [--every 4] [--voxels 32]

Grid size^3, closed box, sphere source of radius 0.08 at (0.5, 0.1, 0.5) with density 1 and
temperature 1, seeded +-0.01 temperature noise, alpha 0.05, beta 0.5, eps 0.3, dt 0.02, 20 advection
substeps, mixed-precision projection.  The rollout runs `steps` steps from rest with per-step
cell-centred control forces (zero), checkpointed every `every` steps; the loss is
1/2 |rho_final - target|^2 with a target of 8 seeded Gaussian blobs (sigma 0.05-0.15, amplitude
0.5-1).  The gradient from the hand-written adjoint is compared, in the FP64 check mode, with
central differences (eps 1e-3) at `voxels` seeded (step, cell, component) entries among those
whose gradient is at least 1e-2 of the largest; each difference of losses is formed from the two
final densities.  An entry off by more than 1e-4 is measured again with eps/10 (an FD error from a
trilinear kink shrinks with the step; an adjoint error would not).  Prints one JSON line per entry
and a summary line.
"""
import argparse
import json
import time

import numpy as np
import torch

from .adjoint import SmokeRollout
from .sim import noise_field

SOURCE = [0.5, 0.1, 0.5, 0.08, 1.0, 1.0]


def gaussian_blobs(W, H, D, count, seed):
    """Sum of `count` seeded Gaussian blobs: centres in [0.2, 0.8]^3, sigma in [0.05, 0.15],
    amplitude in [0.5, 1] (domain units: x / W etc.)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid((np.arange(D) + 0.5) / D, (np.arange(H) + 0.5) / H, (np.arange(W) + 0.5) / W,
                          indexing="ij")
    out = np.zeros((D, H, W), np.float32)
    for _ in range(count):
        c = rng.uniform(0.2, 0.8, size=3)
        sig = rng.uniform(0.05, 0.15)
        amp = rng.uniform(0.5, 1.0)
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        out += (amp * np.exp(-r2 / (2 * sig * sig))).astype(np.float32)
    return out


def closed_box(W, H, D):
    occ = np.zeros((D, H, W), np.uint8)
    occ[0] = occ[-1] = 1
    occ[:, 0] = occ[:, -1] = 1
    occ[:, :, 0] = occ[:, :, -1] = 1
    return occ


def make_rollout(n, steps, every, dtype, tol, adj_rtol):
    eng = SmokeRollout(n, n, n, steps, every, dtype=dtype, voxel_size=1.0 / n, step_time=0.02, steps=20, Tamb=0.0,
                       buoyancy=0.5, gravity=0.05, confine=0.3, sor_open_mask=0, sor_closed_mean=1, sor_omega=1.8,
                       sor_iters=40, sor_max_outer=5000, sor_tol=tol, warm_start=1)
    eng.set_adjoint_tol(adj_rtol)
    eng.set_initial(T=torch.from_numpy(noise_field(n, n, n, 0, 0.01)), occ=torch.from_numpy(closed_box(n, n, n)))
    eng.set_sources(spheres=[SOURCE])
    return eng


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--voxels", type=int, default=32)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-12, help="forward projection tolerance (max|r|, FP64 mode)")
    ap.add_argument("--adj-rtol", type=float, default=1e-11, help="adjoint projection tolerance (relative)")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    n = args.size
    eng = make_rollout(n, args.steps, args.every, torch.float64, args.tol, args.adj_rtol)
    target = torch.from_numpy(gaussian_blobs(n, n, n, 8, args.seed).astype(np.float64)).cuda()
    controls = torch.zeros(eng.controls_shape(), dtype=torch.float64, device="cuda")

    def loss(c):
        rho = eng.forward(c)
        return 0.5 * float(((rho - target) ** 2).sum()), rho

    t0 = time.perf_counter()
    L0, rho = loss(controls)
    t1 = time.perf_counter()
    g = eng.backward(controls, rho - target)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd_stats, bwd_stats = eng.last_forward, eng.last_backward
    flat = g.reshape(-1)
    gmax = float(flat.abs().max())
    cand = torch.nonzero(flat.abs() >= 1e-2 * gmax).reshape(-1).cpu().numpy()
    rng = np.random.default_rng(args.seed + 1)
    picks = rng.choice(cand, size=min(args.voxels, len(cand)), replace=False)
    rows = []
    def central(idx, eps):
        pert = torch.zeros_like(flat)
        pert[int(idx)] = eps
        pert = pert.reshape(controls.shape)
        _, rp = loss(controls + pert)
        _, rm = loss(controls - pert)
        # L+ - L- from the densities: 1/2 sum (rp - rm)(rp + rm - 2 target).  Differencing the two
        # losses (~1e5 here) would lose the ~1e-6 difference in the rounding of the sums.
        return 0.5 * float(((rp - rm) * (rp + rm - 2 * target)).sum()) / (2 * eps)

    def rel_err(a, b):
        return abs(a - b) / max(abs(a), abs(b), 1e-300)

    for idx in picks:
        ad = float(flat[int(idx)])
        fd = central(idx, args.eps)
        s, k, j, i, c = np.unravel_index(int(idx), controls.shape)
        row = {"step": int(s), "cell": [int(i), int(j), int(k)], "comp": "xyz"[c], "adjoint": ad, "fd": fd,
               "rel_err": rel_err(fd, ad)}
        if row["rel_err"] > 1e-4:
            # A one-sided kink of the trilinear samples (cell faces crossed inside [-eps, eps], felt
            # most where the density jumps at the source's edge) gives an O(eps) FD error: a tenth
            # of the step separates that from an error of the adjoint, which would not shrink.
            fd2 = central(idx, args.eps / 10)
            row.update({"fd_eps_div10": fd2, "rel_err_eps_div10": rel_err(fd2, ad)})
        rows.append(row)
        print(json.dumps(row), flush=True)
    t3 = time.perf_counter()
    rel = [r["rel_err"] for r in rows]
    rel_refined = [r.get("rel_err_eps_div10", r["rel_err"]) for r in rows]
    print(json.dumps({
        "check": "smoke-control gradient vs central differences (FP64 check mode)",
        "grid": [n, n, n], "steps": args.steps, "checkpoint_every": args.every, "eps": args.eps,
        "voxels": len(rows), "max_rel_err": max(rel), "median_rel_err": float(np.median(rel)),
        "within_1e-4": int(sum(r <= 1e-4 for r in rel)),
        "max_rel_err_refined": max(rel_refined), "within_1e-4_refined": int(sum(r <= 1e-4 for r in rel_refined)),
        "loss": L0, "grad_max_abs": gmax,
        "forward_s": round(t1 - t0, 2), "backward_s": round(t2 - t1, 2), "fd_s": round(t3 - t2, 1),
        "forward_sweeps": fwd_stats["forward_sweeps"], "adjoint_sweeps": bwd_stats["adjoint_sweeps"],
        "max_div": fwd_stats["max_div"], "tol": args.tol, "adjoint_rtol": args.adj_rtol,
    }))
    eng.close()


if __name__ == "__main__":
    main()
