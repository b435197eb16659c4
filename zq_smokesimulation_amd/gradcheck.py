"""Finite-difference check of the smoke-control gradient (the differentiable-control workload).

    python -m zq_smokesimulation_amd.gradcheck [--size 192] [--steps 16] [--every 4] [--voxels 32]

Grid size^3, closed box, sphere source of radius 0.08 at (0.5, 0.1, 0.5) with density 1 and
temperature 1, seeded +-0.01 temperature noise, alpha 0.05, beta 0.5, eps 0.3, dt 0.02, 20 advection
substeps, mixed-precision projection (SOR at the grid's optimal omega).  The rollout runs `steps`
steps from rest with per-step cell-centred control forces (zero), checkpointed every `every` steps;
the loss is
1/2 |rho_final - target|^2 with a target of 8 seeded Gaussian blobs (sigma 0.05-0.15, amplitude
0.5-1).  The gradient from the hand-written adjoint is compared, in the FP64 check mode, with
central differences at step eps.

Protocol, fixed before any difference is taken:
  * candidates: the (step, cell, component) entries whose gradient is at least 1e-2 of the largest,
    visited in a seeded random order;
  * each visited entry is evaluated at +eps and -eps (the rollout restarted from the checkpoint
    before the entry's step), giving the one-sided differences D+ = (L+ - L0)/eps and
    D- = (L0 - L-)/eps and the central one Dc = (D+ + D-)/2;
  * kink test: the loss is piecewise smooth (the backtrace's trilinear samples and the confinement's
    |w| and grad|w| / |grad|w|| have kinks), and a kink inside [-eps, eps] moves Dc by up to half
    the jump it puts between D+ and D-.  An entry whose |D+ - D-| exceeds kink_tol * |Dc| is not a
    probe (it is counted as rejected, and its adjoint value is not consulted); a smooth entry's
    |D+ - D-| = eps |L''| is far below that;
  * the first `voxels` entries that pass are the probes: each must satisfy |Dc - g| <= 1e-4 max(|Dc|, |g|).
Differences of losses are formed from the densities (L+ - L0 = 1/2 sum (r+ - r0)(r+ + r0 - 2 target)):
the loss itself is ~1e5 while a difference is ~1e-6.  Prints one JSON line per visited entry and a
summary line.
"""
import argparse
import json
import math
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
    # the FP64 check mode refines to max|r| <= 1e-12: SOR at the grid's optimal omega
    # 2/(1+sin(pi/n)) needs a small fraction of the passes omega = 1.8 would
    omega = 2.0 / (1.0 + math.sin(math.pi / n))
    eng = SmokeRollout(n, n, n, steps, every, dtype=dtype, voxel_size=1.0 / n, step_time=0.02, steps=20, Tamb=0.0,
                       buoyancy=0.5, gravity=0.05, confine=0.3, sor_open_mask=0, sor_closed_mean=1, sor_omega=omega,
                       sor_iters=40, sor_max_outer=5000, sor_tol=tol, warm_start=1)
    eng.set_adjoint_tol(adj_rtol)
    eng.set_initial(T=torch.from_numpy(noise_field(n, n, n, 0, 0.01)), occ=torch.from_numpy(closed_box(n, n, n)))
    eng.set_sources(spheres=[SOURCE])
    return eng


def run_check(n=192, steps=16, every=4, voxels=32, eps=1e-3, tol=1e-12, adj_rtol=1e-11, seed=0, kink_tol=2e-4,
              max_visits=200, log=None):
    """Runs the protocol above; returns (summary dict, rows)."""
    eng = make_rollout(n, steps, every, torch.float64, tol, adj_rtol)
    target = torch.from_numpy(gaussian_blobs(n, n, n, 8, seed).astype(np.float64)).cuda()
    controls = torch.zeros(eng.controls_shape(), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    rho0 = eng.forward(controls)
    L0 = 0.5 * float(((rho0 - target) ** 2).sum())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g = eng.backward(controls, rho0 - target)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd_stats, bwd_stats = eng.last_forward, eng.last_backward
    flat = g.reshape(-1)
    gmax = float(flat.abs().max())
    cand = torch.nonzero(flat.abs() >= 1e-2 * gmax).reshape(-1).cpu().numpy()
    order = np.random.default_rng(seed + 1).permutation(cand)
    per_step = controls[0].numel()

    def dloss(ra, rb):
        # L(a) - L(b) from the two densities
        return 0.5 * float(((ra - rb) * (ra + rb - 2 * target)).sum())

    def rel_err(a, b):
        return abs(a - b) / max(abs(a), abs(b), 1e-300)

    rows, probes, rejected = [], 0, 0
    for idx in order[:max_visits]:
        if probes == voxels:
            break
        s = int(idx) // per_step
        pert = torch.zeros_like(flat)
        pert[int(idx)] = eps
        pert = pert.reshape(controls.shape)
        rp = eng.evaluate(controls + pert, from_step=s)
        rm = eng.evaluate(controls - pert, from_step=s)
        dp = dloss(rp, rho0) / eps
        dm = dloss(rho0, rm) / eps
        dc = 0.5 * (dp + dm)
        _, k, j, i, c = np.unravel_index(int(idx), controls.shape)
        row = {"step": s, "cell": [int(i), int(j), int(k)], "comp": "xyz"[c], "fd": dc, "d_plus": dp, "d_minus": dm,
               "kink": abs(dp - dm) / max(abs(dc), 1e-300)}
        if row["kink"] > kink_tol:
            row["probe"] = False
            rejected += 1
        else:
            ad = float(flat[int(idx)])
            row.update({"probe": True, "adjoint": ad, "rel_err": rel_err(dc, ad)})
            probes += 1
        rows.append(row)
        if log:
            log(json.dumps(row))
    t3 = time.perf_counter()
    rel = [r["rel_err"] for r in rows if r["probe"]]
    summary = {
        "check": "smoke-control gradient vs central differences (FP64 check mode)",
        "grid": [n, n, n], "steps": steps, "checkpoint_every": every, "eps": eps, "kink_tol": kink_tol,
        "candidates": int(len(cand)), "visited": len(rows), "rejected_kinks": rejected, "probes": len(rel),
        "max_rel_err": max(rel) if rel else None, "median_rel_err": float(np.median(rel)) if rel else None,
        "within_1e-4": int(sum(r <= 1e-4 for r in rel)),
        "loss": L0, "grad_max_abs": gmax,
        "forward_s": round(t1 - t0, 2), "backward_s": round(t2 - t1, 2), "fd_s": round(t3 - t2, 1),
        "forward_sweeps": fwd_stats["forward_sweeps"], "adjoint_sweeps": bwd_stats["adjoint_sweeps"],
        "max_div": fwd_stats["max_div"], "tol": tol, "adjoint_rtol": adj_rtol,
    }
    eng.close()
    return summary, rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--voxels", type=int, default=32)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--kink-tol", type=float, default=2e-4)
    ap.add_argument("--tol", type=float, default=1e-12, help="forward projection tolerance (max|r|, FP64 mode)")
    ap.add_argument("--adj-rtol", type=float, default=1e-11, help="adjoint projection tolerance (relative)")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    summary, _ = run_check(args.size, args.steps, args.every, args.voxels, args.eps, args.tol, args.adj_rtol,
                           args.seed, args.kink_tol, log=lambda s: print(s, flush=True))
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
