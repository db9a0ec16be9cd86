#!/usr/bin/env python3
"""Precision-robustness sweep: f32 engine against the f64 engine on shared geometries.

Counterpart of the reference's bench/bench_precision_sweep.cpp (built twice, f64 and f32,
run on one shared geometry file) and python/analysis/plot_precision.py: for each molecule
size N (atoms cycling H, C, N, O, singly ionised -- the reference's make_molecule), sample
geometries once on the device (radius 4.0, min separation 0.25, the reference sweep's
sampler options), run them through the f64 path at the reference's f64 defaults (rtol 1e-8,
atol 1e-16, pe_stop 1e-9) and through the f32 path at the production tolerances (rtol 1e-4,
atol 1e-7, pe_stop 1e-5), both with the reference's Hairer start and energy redistribution,
and report per N: failure codes of the f32 run (the reference's kException -> reject limit,
kNotConverged -> step cap, kNonFinite -> NaN/Inf), the config-norm |dp|/|p| distribution
against the 1% floor, and its dependence on the initial minimum pair distance (the
difficulty axis).

Usage:
  python coulomb-explosion_amd/precision_sweep.py --atoms 2,5,8,10 --sims 4000
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

CYCLE = [("H", 1.008), ("C", 12.011), ("N", 14.007), ("O", 15.999)]


def molecule(ce, n: int):
    return ce.Molecule([ce.atom_from_amu(*CYCLE[i % 4], 1.0) for i in range(n)])


def sweep_one(ce, n: int, sims: int, seed: int, floor: float = 0.01,
              f32=(1e-4, 1e-7, 1e-5), f64=(1e-8, 1e-16, 1e-9)) -> dict:
    mol = molecule(ce, n)
    runs = {}
    for name, (rtol, atol, pe_stop), prec in (("f64", f64, ce.PRECISION_F64),
                                              ("f32", f32, ce.PRECISION_F32)):
        eng = ce.BatchedRK45(mol, 1.0, sims, rtol, atol)
        eng.sample(seed, sims, 4.0, 0.25)
        runs[name] = eng.run(ce.RunConfig(pe_stop_fraction=pe_stop, precision=prec))
        geo = eng.geometries()
    a, b = runs["f32"], runs["f64"]
    ok = (a.status <= ce.SIM_OK_TMAX) & (b.status <= ce.SIM_OK_TMAX)
    pa, pb = a.momenta.reshape(sims, n, 3), b.momenta.reshape(sims, n, 3)
    cfg = (np.linalg.norm((pa - pb).reshape(sims, -1), axis=1) /
           np.linalg.norm(pb.reshape(sims, -1), axis=1))[ok]
    d = np.linalg.norm(geo[:, :, None, :] - geo[:, None, :, :], axis=-1)
    d[:, np.arange(n), np.arange(n)] = np.inf
    min_sep = d.reshape(sims, -1).min(axis=1)[ok]
    edges = np.quantile(min_sep, [0, 0.25, 0.5, 0.75, 1.0]) if len(min_sep) else np.zeros(5)
    bins = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (min_sep >= lo) & (min_sep <= hi)
        if sel.any():
            bins.append({"min_sep": [float(lo), float(hi)], "median": float(np.median(cfg[sel])),
                         "max": float(cfg[sel].max())})
    return {
        "atoms": n, "sims": sims,
        "f32_failures": {"reject_limit": int(np.sum(a.status == ce.SIM_REJECT_LIMIT)),
                         "step_cap": int(np.sum(a.status == ce.SIM_NOT_CONVERGED)),
                         "nonfinite": int(np.sum(a.status == ce.SIM_NONFINITE)),
                         "sampler": int(np.sum(a.status == ce.SIM_SAMPLER_FAIL))},
        "f64_failures": int(np.sum(b.status > ce.SIM_OK_TMAX)),
        "dp_over_p": {"median": float(np.median(cfg)), "p99": float(np.percentile(cfg, 99)),
                      "max": float(cfg.max()), "over_floor": int(np.sum(cfg > floor))},
        "by_min_sep_quartile": bins,
        "mean_steps": {"f32": float(a.steps.mean()), "f64": float(b.steps.mean())},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--atoms", default="2,3,4,5,6,7,8,9,10")
    ap.add_argument("--sims", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=0xC0FFEE)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from __graft_entry__ import load_package
    ce = load_package()
    for n in [int(v) for v in args.atoms.split(",")]:
        print(json.dumps(sweep_one(ce, n, args.sims, args.seed)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
