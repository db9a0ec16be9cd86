#!/usr/bin/env python3
"""Benchmark: converged Coulomb-explosion sims per second on one MI355X.

Default workload ("ch4", the headline; one "step" = one pass of the hot path over
one batch):
  10M CH4 sims (C 21874.66, 4 x H 1837.15 electron masses, all charges +1),
  positions uniform in a ball of radius 2.5 bohr, pairs >= 1.0 bohr apart
  (device counter-based sampler), atoms at rest; f32 state, DP5(4) rtol 1e-6,
  atol 1e-9, h0 1e-3; stop when every atom's momentum changes by < 1e-7
  (relative) over a step or at t = 1e4 a.u. (FP64 time); per-atom momenta
  written to HBM.  The timed region is sampler + integrator, inputs and
  outputs resident in HBM.

Other workloads (--workload):
  mixed  10M sims in one launch, 25% each N2 (2 x 25526.04), CO2 (C 21874.66,
         2 x O 29156.95), CH4, C2H6 (2 C, 6 H); per-atom charges uniform in
         {1, 2}; R = N^(1/3) + 1.5, min sep 0.9; same tolerances and stop rule.
  n10    10M sims of a 10-atom molecule (2 C, 2 O, 6 H), charges uniform in
         {1, 2, 3}, R 3.5, min sep 0.9 (the plain method on that distribution).
  closure  the learned closure on CH4: an MLP ensemble (40 -> 256 x 3 -> 15,
         SiLU, BF16 MFMA / FP32 accumulation, Adam lr 1e-3, batch 8192, 20
         epochs) trained on 2M engine trajectories (untimed), then 8M fresh sims
         per step: plain DP5(4) to t_cut = 200, prior + ensemble residual where
         the members agree to 1e-5, integration continued elsewhere; timed
         beside the plain method on the same sims, whose results set the error;
         a 50k subset is also compared with the engine's FP64 path.
  closure10  the same for N=10 (2 C, 2 O, 6 H; charges uniform {1,2,3}; R 3.5,
         min sep 0.9): 80 -> 512 x 4 -> 30, 4M training trajectories, 5M sims.

Also reported: the error of the f32 results against the engine's FP64 path at
rtol 1e-10 on a subset, the integrator kernel's achieved FP32 rate against the
FP32 peak (HIP events on the engine stream), and (ch4) the CPU baseline (the
lockstep BatchedRK45<float> restatement, one instance per host thread) timed in
the same run.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged sims/sec on 1 MI355X at ≤1e-5 momentum err: closure vs plain RK45"
C_, O_, H_, N14 = 21874.66, 29156.95, 1837.15, 25526.04
CH4_MASSES = [C_] + [H_] * 4
CH4_CHARGES = [1.0] * 5
RTOL, ATOL, H0, T_MAX, MOM_TOL = 1e-6, 1e-9, 1e-3, 1e4, 1e-7
RADIUS, MIN_SEP = 2.5, 1.0
SEED = 49188291
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
# HBM bytes per sim of one integrator launch, from the FETCH_SIZE / WRITE_SIZE passes
# scaled by the plain-copy calibration (x2.0 read, x1.0 write), 2M-sim CH4 launch:
# 125.0 B read + 232.6 B written (profiles/dp5_persistent_ch4_counters.md).
MEASURED_HBM_BYTES_PER_SIM = {"ch4": 125.0 + 232.6}
MIXED = [([N14, N14], 2), ([C_, O_, O_], 3), (CH4_MASSES, 5), ([C_, C_] + [H_] * 6, 8)]
N10_MASSES = [C_, C_, O_, O_] + [H_] * 6


def flops_per_attempt(n: int) -> int:
    """Algorithmic FLOPs of one attempted DP5(4) step on an n-atom molecule:
    6 force evaluations of 20 per pair + 3n, stage combinations 324n, error
    norm 132n (FMA = 2, sqrt/rsqrt/div = 1)."""
    p = n * (n - 1) // 2
    return 6 * (20 * p + 3 * n) + 324 * n + 132 * n


def flops_per_sim_fixed(n: int) -> int:
    """Per-sim work outside the attempts: one force+PE evaluation at t=0 (fixed
    h0) and the FP64 finalize (~10n)."""
    p = n * (n - 1) // 2
    return (20 * p + 3 * n) + 10 * n


def run_config(ce, precision=0):
    return ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, t_max=T_MAX, momentum_tol=MOM_TOL,
                        h0_mode=ce.H0_FIXED, h0=H0, precision=precision)


def load_autotune(ce):
    spec = importlib.util.spec_from_file_location(
        "coulomb_explosion_amd.autotune", os.path.join(os.path.dirname(ce.__file__), "autotune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Workload:
    """A batch of sims of one or more molecule sizes on one engine."""

    def __init__(self, ce, kind: str, sims: int, rtol=RTOL, atol=ATOL):
        self.ce, self.kind, self.sims = ce, kind, sims
        if kind == "ch4":
            mol = ce.Molecule([ce.Atom("C", C_, 1.0)] + [ce.Atom("H", H_, 1.0)] * 4)
            self.eng = ce.BatchedRK45(mol, 1.0, sims, rtol, atol)
            self.sizes = [(5, sims)]
        else:
            if kind == "mixed":
                per = [sims // 4] * 4
                per[-1] += sims - sum(per)
                sp = [ce.Species(ce.Molecule([ce.Atom("X", m, 1.0) for m in masses]), cnt,
                                 n ** (1 / 3) + 1.5, 0.9, charge_levels=2, max_total_charge=2 * n)
                      for (masses, n), cnt in zip(MIXED, per)]
            elif kind == "n10":
                sp = [ce.Species(ce.Molecule([ce.Atom("X", m, 1.0) for m in N10_MASSES]), sims,
                                 3.5, 0.9, charge_levels=3)]
            else:
                raise ValueError(kind)
            self.eng = ce.MixedBatch(sp, 1.0, rtol, atol)
            self.sizes = [(s.molecule.size(), s.count) for s in sp]

    def sample(self, seed=SEED):
        if self.kind == "ch4":
            self.eng.sample(seed, self.sims, RADIUS, MIN_SEP)
        else:
            self.eng.sample(seed)

    def run(self, cfg):
        self.eng.run(cfg, download=False)

    def momenta_status(self):
        """[(momenta [count, n, 3], status, attempts)] per size."""
        if self.kind == "ch4":
            r = self.eng.download()
            return [(r.momenta.reshape(-1, 5, 3), r.status, r.attempts)]
        return [(r["momenta"], r["status"], r["attempts"]) for r in self.eng.download()]

    def describe(self) -> str:
        return {
            "ch4": "CH4 plain method, 10M sims per step: R 2.5, min sep 1.0, rtol 1e-6, atol "
                   "1e-9, h0 1e-3, stop |dp|/|p| < 1e-7 or t = 1e4 (FP64 t), f32 state",
            "mixed": "mixed N=2/3/5/8 (N2, CO2, CH4, C2H6) 25% each in one persistent launch, "
                     "per-atom charges uniform {1,2}, R = N^(1/3)+1.5, min sep 0.9, rtol 1e-6, "
                     "atol 1e-9, h0 1e-3, stop |dp|/|p| < 1e-7 or t = 1e4, f32 state",
            "n10": "N=10 (2C, 2O, 6H), charges uniform {1,2,3}, R 3.5, min sep 0.9, plain "
                   "method, rtol 1e-6, atol 1e-9, h0 1e-3, stop |dp|/|p| < 1e-7 or t = 1e4",
        }[self.kind]


def accuracy_subset(ce, kind: str, n_sub: int) -> dict:
    """f32 engine vs the engine's FP64 path (rtol 1e-10, atol 1e-14) on n_sub
    sims of the workload's distribution (same seed, same geometries/charges)."""
    w32 = Workload(ce, kind, n_sub)
    w64 = Workload(ce, kind, n_sub, rtol=1e-10, atol=1e-14)
    for w, prec in ((w32, 0), (w64, 1)):
        w.sample()
        w.run(run_config(ce, prec))
    per_atom, cfg_norm, steps64 = [], [], []
    for (p32, _, _), (p64, _, a64) in zip(w32.momenta_status(), w64.momenta_status()):
        num = np.linalg.norm(p32 - p64, axis=-1)
        per_atom.append((num / np.linalg.norm(p64, axis=-1)).max(axis=-1))
        k = p64.shape[0]
        cfg_norm.append(np.linalg.norm((p32 - p64).reshape(k, -1), axis=1) /
                        np.linalg.norm(p64.reshape(k, -1), axis=1))
        steps64.append(a64)
    per_atom = np.concatenate(per_atom)
    cfg_norm = np.concatenate(cfg_norm)
    # The same f32 results against the asymptote itself: the FP64 path run without an
    # early stop to t = 1e10 and without redistribution (examples/verify_subset.py:
    # 74-90 integrates the reference's oracle to t = 1e10), on the first sims.
    n_asym = min(n_sub, 20000)
    wa = Workload(ce, kind, n_asym, rtol=1e-10, atol=1e-14)
    wa.sample()
    wa.run(ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, t_max=1e10, momentum_tol=0.0,
                        h0_mode=ce.H0_FIXED, h0=H0, precision=1, redistribute_energy=False))
    w32s = Workload(ce, kind, n_asym)
    w32s.sample()
    w32s.run(run_config(ce, 0))
    asym_pa, asym_cn = [], []
    for (p32, _, _), (pinf, st, _) in zip(w32s.momenta_status(), wa.momenta_status()):
        asym_pa.append((np.linalg.norm(p32 - pinf, axis=-1) /
                        np.linalg.norm(pinf, axis=-1)).max(axis=-1))
        k = pinf.shape[0]
        asym_cn.append(np.linalg.norm((p32 - pinf).reshape(k, -1), axis=1) /
                       np.linalg.norm(pinf.reshape(k, -1), axis=1))
    asym_pa, asym_cn = np.concatenate(asym_pa), np.concatenate(asym_cn)
    return {
        "subset": n_sub,
        "vs": "engine FP64 path, rtol 1e-10 atol 1e-14, same stop rule and geometries",
        "per_atom_rel_max": float(per_atom.max()),
        "per_atom_rel_p99": float(np.percentile(per_atom, 99)),
        "per_atom_rel_median": float(np.median(per_atom)),
        "config_norm_rel_max": float(cfg_norm.max()),
        "within_1e-5": float(np.mean(per_atom <= 1e-5)),
        "fp64_mean_attempts": float(np.concatenate(steps64).mean()),
        "fp64_kernel_ms": w64.eng.stats()["kernel_ms"],
        "config_norm_within_1e-5": float(np.mean(cfg_norm <= 1e-5)),
        "vs_asymptote": {
            "subset": n_asym,
            "vs": "engine FP64 path, rtol 1e-10, no early stop, t = 1e10, no redistribution",
            "per_atom_rel_max": float(asym_pa.max()),
            "per_atom_rel_median": float(np.median(asym_pa)),
            "per_atom_rel_p99": float(np.percentile(asym_pa, 99)),
            "config_norm_rel_max": float(asym_cn.max()),
            "config_norm_rel_median": float(np.median(asym_cn)),
            "per_atom_within_1e-5": float(np.mean(asym_pa <= 1e-5)),
            "per_atom_within_1e-3": float(np.mean(asym_pa <= 1e-3))},
    }


def cpu_baseline(n_sims: int, reps: int, threads: int) -> dict:
    from oracle import ffi
    o = ffi.Oracle()
    rc, geom = o.sample(SEED, 5, n_sims, RADIUS, MIN_SEP)
    if rc != 0:
        raise RuntimeError("cpu baseline sampler failed")
    cfg = ffi.make_config(rtol=RTOL, atol=ATOL, stop_rule=ffi.STOP_MOMENTUM_OR_TMAX, t_max=T_MAX,
                          momentum_tol=MOM_TOL, h0_mode=ffi.H0_FIXED, h0=H0)
    r = o.cpu_baseline(CH4_MASSES, CH4_CHARGES, geom, cfg, threads, reps)
    one = o.cpu_baseline(CH4_MASSES, CH4_CHARGES, geom[: n_sims // 8], cfg, 1, 1)
    # Fidelity of the port: one core at the reference's own benchmark setup
    # (C 12.011 / H 1.008 amu, R 4.0, min sep 0.25, rtol 1e-4, atol 1e-7, PE stop
    # 1e-5, Hairer start), where the reference's Highway engine is published at
    # about 255k sims/s with 8 f32 lanes.
    amu = 1822.888486209
    rc, g2 = o.sample(SEED, 5, 40000, 4.0, 0.25)
    cfg2 = ffi.make_config(rtol=1e-4, atol=1e-7, stop_rule=ffi.STOP_PE_FRACTION,
                           pe_stop_fraction=1e-5)
    ref_masses = [12.011 * amu] + [1.008 * amu] * 4
    fid = o.cpu_baseline(ref_masses, CH4_CHARGES, g2, cfg2, 1, 3)
    # ... and with the reference's own lane width on that setup (8 f32 lanes, AVX2: the
    # published ~255k sims/s, docs/benchmarks/0008-ch4-dataset-gen.md:61-66)
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln for ln in f if ln.startswith("flags")), "")
    except OSError:
        pass
    fid8 = (o.cpu_baseline(ref_masses, CH4_CHARGES, g2, cfg2, 1, 3, isa="avx2")
            if " avx2" in flags and " fma" in flags else None)
    # Lockstep efficiency of the baseline on this workload, as the reference
    # defines it (bench/bench_explosion.cpp:298-336): scalar steps of the sims /
    # (K x shared-dt iterations); scalar steps from the per-lane form on a subset.
    sub = o.lane_run(CH4_MASSES, CH4_CHARGES, geom[:4096], cfg)
    mean_batch = r["batch_steps"] / max(1, r["sims"] // r["lanes"])
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": r["sims"] / r["median_s"],
        "unit": "sims/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{r['sims']} CH4 sims of the same workload (host mt19937_64 sampler, seed "
                   f"{SEED}), lockstep BatchedRK45<float> restatement at {r['isa']}, one "
                   f"instance per thread, median of {reps} passes"),
        "single_core_sims_per_s": one["sims"] / one["median_s"],
        "mean_batch_steps": mean_batch,
        "lockstep_efficiency": float(sub.steps.mean()) / mean_batch,
        "cpu_model": model,
        "host_threads": os.cpu_count(),
        "single_core_at_reference_setup_sims_per_s": fid["sims"] / fid["median_s"],
        "single_core_at_reference_setup_lanes": fid["lanes"],
        "single_core_at_reference_setup_avx2_k8_sims_per_s":
            fid8["sims"] / fid8["median_s"] if fid8 else None,
    }


def generator_end_to_end(args) -> dict | None:
    """The dataset generator CLI (coulomb-explosion_amd/generate_dataset) on the
    headline workload, 60M sims in batches of --sims, records to /dev/null: the
    rate with every record in (pinned) host memory, copies and the writer thread
    overlapped with the next batch's integration (examples/generate_dataset.cpp:
    190-227 times the same loop with its file writes)."""
    import subprocess
    exe = os.path.join(ROOT, "coulomb-explosion_amd", "generate_dataset")
    if not os.path.exists(exe):
        return None
    total = 6 * args.sims
    cmd = [exe, "--out", "/dev/null", "--sims", str(total), "--batch", str(args.sims),
           "--masses", ",".join(str(m) for m in CH4_MASSES), "--charges", "1,1,1,1,1",
           "--radius", str(RADIUS), "--min-separation", str(MIN_SEP), "--rtol", str(RTOL),
           "--atol", str(ATOL), "--stop-rule", "tmax", "--t-max", str(T_MAX),
           "--momentum-tol", str(MOM_TOL), "--h0", str(H0), "--seed", str(SEED),
           "--progress-every", "1000"]
    print(f"[bench] generator: {total} sims to /dev/null", file=sys.stderr, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    out = {"sims": total, "batch": args.sims, "command": " ".join(os.path.relpath(c, ROOT)
                                                                 if c == exe else c for c in cmd)}
    for ln in r.stdout.splitlines():
        if ln.startswith("done:"):
            out["all_sims_per_s"] = float(ln.split("(")[1].split()[0])
        if ln.startswith("steady:"):
            out["steady_sims_per_s"] = float(ln.split("(")[1].split()[0])
    return out


CLOSURE_SETUPS = {
    # masses, charge levels (0 = all +1), R, min sep, hidden widths, train / eval sims
    "closure": dict(masses=CH4_MASSES, levels=0, radius=RADIUS, min_sep=MIN_SEP,
                    hidden=(256, 256, 256), train=2_000_000, eval=8_000_000, tune="ch4",
                    name="CH4"),
    "closure10": dict(masses=N10_MASSES, levels=3, radius=3.5, min_sep=0.9,
                      hidden=(512, 512, 512, 512), train=4_000_000, eval=5_000_000, tune="n10",
                      name="N=10 (2C, 2O, 6H), charges uniform {1,2,3}"),
}


def run_closure(ce, args) -> dict:
    import coulomb_explosion_amd.closure_train as ct
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    su = CLOSURE_SETUPS[args.workload]
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in su["masses"]])
    nat = mol.size()
    cfg = run_config(ce)
    train_sims = args.train_sims or su["train"]
    n = args.eval_sims or su["eval"]
    levels = su["levels"]
    log(f"[bench] {args.workload}: generating {train_sims} training trajectories")
    data = ct.make_data(ce, mol, train_sims, SEED + 1, cfg, args.t_cut, su["radius"],
                        su["min_sep"], RTOL, ATOL, charge_levels=levels)
    log(f"[bench] {args.workload}: {len(data.final)} pairs in {data.seconds:.1f} s; training "
        f"{args.members} members x {args.epochs} epochs")
    ens = ct.fit(ce, data, members=args.members, hidden=su["hidden"], epochs=args.epochs,
                 batch=args.batch, lr=args.lr, seed=1, log=log)
    clo = ct.attach(ce, mol, ens)
    seed = SEED + 2

    def make_engine(count, rtol, atol):
        if levels:
            return ce.MixedBatch([ce.Species(mol, count, su["radius"], su["min_sep"], levels,
                                             levels * nat)], 1.0, rtol, atol)
        return ce.BatchedRK45(mol, 1.0, count, rtol, atol)

    def sample(e, count):
        if levels:
            e.sample(seed)
        else:
            e.sample(seed, count, su["radius"], su["min_sep"])

    eng = make_engine(n, RTOL, ATOL)
    load_autotune(ce).apply_cached(ce, eng, cfg, key_extra=su["tune"])
    eng.enable_state()

    def plain():
        sample(eng, n)
        eng.run(cfg, download=False)

    def closure():
        sample(eng, n)
        clo.run(eng, cfg, args.t_cut, args.gate_tol, download=False)

    for _ in range(args.warmup):
        plain()
        closure()
    times = {}
    for name, fn in (("plain", plain), ("closure", closure)):
        log(f"[bench] {args.workload}: timing {name} x {args.steps}")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        times[name] = (time.perf_counter() - t0) / args.steps
        if name == "plain":
            ref = ct._flat(eng.download())
            plain_kernel_ms = eng.stats()["kernel_ms"]
    res = ct._flat(eng.download())
    st = clo.stats()
    p_ref = ref["momenta"].reshape(n, nat, 3)
    p_clo = res["momenta"].reshape(n, nat, 3)

    def rel(p, q):
        return (np.linalg.norm(p - q, axis=-1) / np.linalg.norm(q, axis=-1)).max(-1)

    err = rel(p_clo, p_ref)
    acc = res["status"] == ce.SIM_CLOSURE
    ok = (err <= 1e-5) & ((res["status"] <= ce.SIM_OK_TMAX) | acc)
    fp64 = None
    if args.fp64_subset:
        # the first fp64_subset sims again on the engine's FP64 path (rtol 1e-10)
        e64 = make_engine(args.fp64_subset, 1e-10, 1e-14)
        sample(e64, args.fp64_subset)
        r64 = ct._flat(e64.run(run_config(ce, precision=1)))
        k = args.fp64_subset
        p64 = r64["momenta"].reshape(k, nat, 3)
        e_clo, e_pl = rel(p_clo[:k], p64), rel(p_ref[:k], p64)
        fp64 = {"subset": k, "closure_within_1e-5": float(np.mean(e_clo <= 1e-5)),
                "closure_max": float(e_clo.max()), "plain_within_1e-5": float(np.mean(e_pl <= 1e-5)),
                "plain_max": float(e_pl.max())}
    # held-out quality of the network itself, gate aside
    test = ct.make_data(ce, mol, 20000, SEED + 3, cfg, args.t_cut, su["radius"], su["min_sep"],
                        RTOL, ATOL, charge_levels=levels)
    mean, spread = ct.predict(ens, test.features, test.prior)
    cur = test.features[:, :6 * nat].reshape(-1, nat, 6)[:, :, 3:]
    quality = {"held_out_sims": len(test.final),
               "median_err_current_momenta": float(np.median(rel(cur, test.final))),
               "median_err_prior": float(np.median(rel(test.prior, test.final))),
               "median_err_prior_plus_network": float(np.median(rel(mean, test.final))),
               "p90_err_prior_plus_network": float(np.quantile(rel(mean, test.final), 0.9)),
               "median_ensemble_spread": float(np.median(spread))}
    widths = [8 * nat, *su["hidden"], 3 * nat]
    mlp_flops = 2 * sum(a * b for a, b in zip(widths[:-1], widths[1:]))
    # rows the members actually evaluated (the first two decide most sims, the
    # others run on the rest: closure_pregate)
    infer_flops = st["member_rows"] * mlp_flops
    infer_tf = infer_flops / (st["infer_ms"] / 1e3) / 1e12 if st["infer_ms"] else 0.0
    return {
        "n": n,
        "closure_s": times["closure"], "plain_s": times["plain"],
        "within": int(ok.sum()),
        "line_config": {
            "workload": (f"{su['name']} learned closure, {n} fresh sims per step (R "
                         f"{su['radius']}, min sep {su['min_sep']}, rtol 1e-6, atol 1e-9, h0 1e-3, "
                         f"stop |dp|/|p| < 1e-7 or t = 1e4): plain DP5(4) to t_cut "
                         f"{args.t_cut:g}, straight-line impulse prior + {args.members}-member MLP "
                         f"ensemble residual, gate {args.gate_tol:g} on the members' per-atom "
                         f"spread (decided from the first two members where their spread "
                         f"already excludes acceptance), integration continued elsewhere"),
            "sims_per_step": n,
            "method": "closure",
            "plain_sims_per_s": n / times["plain"],
            "closure_sims_per_s_all": n / times["closure"],
            "closure_over_plain": times["plain"] / times["closure"],
            "fallback_rate": st["fallback"] / max(1, st["candidates"]),
            "candidates": st["candidates"],
            "member_rows_per_candidate": st["member_rows"] / max(1, st["candidates"]),
            "accepted_by_gate": int(acc.sum()),
            "err_vs_full_integration": {
                "within_1e-5": float(np.mean(ok)),
                "accepted_max": float(err[acc].max()) if acc.any() else None,
                "accepted_median": float(np.median(err[acc])) if acc.any() else None,
                "all_max": float(err.max())},
            "err_vs_fp64": fp64,
            "phase_ms": {"cut": st["cut_ms"], "features_ensemble_gate": st["infer_ms"],
                         "fallback": st["fallback_ms"], "total_device": st["total_ms"],
                         "plain_kernel": plain_kernel_ms},
            "attempts_per_sim": {"cut": st["cut_attempts"] / n,
                                 "fallback": st["fallback_attempts"] / n},
            "network_quality": quality,
            "training": {"trajectories": len(data.final), "generate_s": data.seconds,
                         "train_s": ens.seconds, "members": args.members, "widths": widths,
                         "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
                         "final_loss": [float(h[-1]) for h in ens.losses]},
        },
        "roofline": {
            "bound": "mfma", "kernel": "fused ensemble inference (mlp_fused_forward), "
                                       "with features and gate; FLOP of the member rows "
                                       "evaluated",
            "achieved": infer_tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": infer_tf / 2500.0,
            "traffic": None, "flops_per_sim_per_member": mlp_flops},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["ch4", "mixed", "n10", "closure", "closure10"],
                    default="ch4")
    ap.add_argument("--train-sims", type=int, default=0, help="closure: 0 = the workload's")
    ap.add_argument("--eval-sims", type=int, default=0, help="closure: 0 = the workload's")
    ap.add_argument("--fp64-subset", type=int, default=50_000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--t-cut", type=float, default=200.0)
    ap.add_argument("--gate-tol", type=float, default=1e-5)
    ap.add_argument("--sims", type=int, default=10_000_000)
    ap.add_argument("--accuracy-subset", type=int, default=100_000)
    ap.add_argument("--cpu-sims", type=int, default=400_000)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-closure", action="store_true",
                    help="ch4: skip the learned-closure leg (CH4 closure, 2M training "
                         "trajectories, 8M sims) reported under the line's 'closure' key")
    ap.add_argument("--refill-threshold", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--block-threads", type=int, default=0)
    ap.add_argument("--mixed-launch", choices=["one", "per-species"], default="one",
                    help="mixed: all sizes in one persistent launch (the workload as stated), "
                         "or one launch per size")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch shape on this workload first (cached)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run the library's plain device copy kernel (hbm_copy) over "
                         "this many MiB, 3 times (HBM counter calibration for rocprofv3 "
                         "--pmc), and exit")
    ap.add_argument("--mlp-probe", default="",
                    help="only run the fused inference kernel: widths (e.g. 40,256,256,256,15), "
                         "--sims rows of random inputs, 3 forward passes (counter runs), and exit")
    args = ap.parse_args()

    if args.mlp_probe:
        from __graft_entry__ import load_package
        ce = load_package()
        widths = [int(v) for v in args.mlp_probe.split(",")]
        net = ce.MLP(widths, seed=1)
        if not net.set_fused(True):
            print("bench.py: these widths do not use the fused kernel", file=sys.stderr)
            return 2
        x = np.random.default_rng(0).normal(size=(args.sims, widths[0])).astype(np.float32)
        for _ in range(3):
            y = net.forward(x)
        print(json.dumps({"mlp_probe": widths, "rows": args.sims, "passes": 3,
                          "mean_abs_out": float(np.abs(y).mean())}))
        return 0

    if args.copy_probe_mb:
        import ctypes
        from __graft_entry__ import load_package
        ce = load_package()
        nbytes = args.copy_probe_mb * (1 << 20)
        ms = ctypes.c_double(0)
        ce._check(ce.library().ce_hbm_copy_calibration(nbytes, 3, ctypes.byref(ms)))
        print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                          "launches": 3, "ms_per_copy": ms.value,
                          "gbs": 2 * nbytes / (ms.value * 1e-3) / 1e9}))
        return 0

    if args.gpus != 1:
        print(f"bench.py: this engine is measured on one GPU; --gpus {args.gpus} is not "
              "supported (simulations are independent, so a multi-GPU job is one index range "
              "per GPU with no communication).", file=sys.stderr)
        return 2

    from __graft_entry__ import load_package
    ce = load_package()
    if ce.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1

    if args.workload in CLOSURE_SETUPS:
        r = run_closure(ce, args)
        line = {"metric": METRIC, "value": r["within"] / r["closure_s"], "unit": "sims/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": r["closure_s"] * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
                "data": f"synthetic (device sampler, train seed {SEED + 1}, eval seed {SEED + 2})",
                "config": r["line_config"], "roofline": r["roofline"], "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return 0

    w = Workload(ce, args.workload, args.sims)
    cfg = run_config(ce)
    autotune = load_autotune(ce)
    if args.autotune:
        w.sample()
        entry = autotune.tune(ce, w.eng, cfg, log=lambda m: print(m, file=sys.stderr, flush=True),
                              key_extra=args.workload)
        tuned = entry["shape"]
    else:
        tuned = autotune.apply_cached(ce, w.eng, cfg, key_extra=args.workload)
    shape = dict(tuned or {})
    for k, v in (("refill_threshold", args.refill_threshold), ("chunk", args.chunk),
                 ("blocks_per_cu", args.blocks_per_cu), ("block_threads", args.block_threads)):
        if v:
            shape[k] = v
    if args.workload == "mixed":
        shape["mixed_launch"] = (ce.MIXED_ONE_LAUNCH if args.mixed_launch == "one"
                                 else ce.MIXED_PER_SPECIES)
    w.eng.set_launch(**shape)

    print(f"[bench] {args.workload}: warmup {args.warmup} x {args.sims} sims", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        w.sample()
        w.run(cfg)
    kernel_ms, sampler_ms, attempts, accepted, launches, slots = [], [], 0, 0, 0, 0
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.sample()
        sampler_ms.append(w.eng.stats()["sampler_ms"])
        w.run(cfg)
        st = w.eng.stats()
        kernel_ms.append(st["kernel_ms"])
        attempts += st["attempted_steps"]
        accepted += st["accepted_steps"]
        slots += st["lane_slots"]
        launches = st["launches"]
    elapsed = time.perf_counter() - t0
    per_size = w.momenta_status()
    n_ok = int(sum(np.sum(s <= ce.SIM_OK_TMAX) for _, s, _ in per_size))
    stats = w.eng.stats()

    value = n_ok * args.steps / elapsed
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    # algorithmic FLOPs of the last launch from its per-sim attempt counts
    flops_launch = sum(float(a.astype(np.float64).sum()) * flops_per_attempt(n) +
                       cnt * flops_per_sim_fixed(n)
                       for (n, cnt), (_, _, a) in zip(w.sizes, per_size))
    achieved_tflops = flops_launch / avg_kernel_s / 1e12
    bytes_launch = sum(cnt * (24 * n + 24 * n + 12 * n + 25) for n, cnt in w.sizes)

    writeback = None
    if args.workload == "ch4":
        # Dataset records of the last batch (examples/generate_dataset.cpp:198-217): the
        # pack kernel reads geometry f64, momenta f32, steps, status (185 B/sim) and
        # writes 125 B/sim packed; then the records are copied to host memory.
        t1 = time.perf_counter()
        rec = w.eng.records()
        wall = time.perf_counter() - t1
        pack_ms = w.eng.stats()["pack_ms"]
        n_rec = len(rec)
        writeback = {"records": n_rec, "record_bytes": int(rec.itemsize),
                     "pack_kernel_ms": pack_ms,
                     "pack_hbm_gbs": n_rec * (185 + rec.itemsize) / (pack_ms * 1e-3) / 1e9,
                     "host_copy_gbs_incl_pcie": n_rec * rec.itemsize / max(wall - pack_ms * 1e-3,
                                                                          1e-9) / 1e9}
        # the same records queued into pinned host memory (the generator's path)
        pinned = ce.PinnedBuffer(n_rec * rec.itemsize)
        del rec
        w.eng.records_async(pinned)
        w.eng.records_wait()
        t1 = time.perf_counter()
        w.eng.records_async(pinned)
        w.eng.records_wait()
        wall = time.perf_counter() - t1
        writeback["pinned_queued_pack_and_copy_ms"] = wall * 1e3
        writeback["pinned_host_copy_gbs_incl_pcie"] = n_rec * 125 / wall / 1e9
        del pinned
        writeback["generator"] = generator_end_to_end(args)

    other = None
    if args.workload == "mixed":
        # the other launch mode on the same workload, for comparison (not the value)
        alt = "per-species" if args.mixed_launch == "one" else "one"
        w.eng.set_launch(**dict(shape, mixed_launch=ce.MIXED_PER_SPECIES if alt == "per-species"
                                else ce.MIXED_ONE_LAUNCH))
        alt_ms = []
        for _ in range(3):
            w.sample()
            w.run(cfg)
            alt_ms.append(w.eng.stats()["kernel_ms"])
        other = {"mode": alt, "kernel_ms_avg": float(np.mean(alt_ms[1:])),
                 "launches": w.eng.stats()["launches"]}

    acc = None
    if args.accuracy_subset > 0:
        print(f"[bench] accuracy subset {args.accuracy_subset}", file=sys.stderr, flush=True)
        acc = accuracy_subset(ce, args.workload, args.accuracy_subset)

    closure = None
    if args.workload == "ch4" and not args.no_closure:
        # The closure half of the metric, on its own workload (CH4 learned closure: the
        # network trained on 2M engine trajectories, then 8M fresh sims per step timed
        # beside the plain method on the same sims); the line's value stays the plain
        # method on the headline workload.
        cargs = argparse.Namespace(**vars(args))
        cargs.workload, cargs.steps, cargs.warmup = "closure", 3, 1
        r = run_closure(ce, cargs)
        closure = dict(r["line_config"])
        closure["sims_per_s_within_1e-5"] = r["within"] / r["closure_s"]
        closure["steps"], closure["warmup"] = cargs.steps, cargs.warmup
        closure["roofline"] = r["roofline"]

    cpu = None
    if not args.no_cpu and args.workload == "ch4":
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline {args.cpu_sims} sims x {args.cpu_reps} on {threads} threads",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_sims, args.cpu_reps, threads)

    kernel_name = ("dp5_persistent<float,5>" if args.workload == "ch4" else
                   ("dp5_mixed<float,2,3,5,8,per-sim charges>" if args.mixed_launch == "one" else
                    "dp5_persistent<float,N,per-sim charges>, N = 2, 3, 5, 8 (one launch each)")
                   if args.workload == "mixed" else
                   "dp5_persistent<float,10,per-sim charges>")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "sims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device counter-based uniform-sphere sampler, seed %d)" % SEED,
        "config": {
            "workload": (w.describe() if not (args.workload == "mixed" and
                                               args.mixed_launch == "per-species")
                         else w.describe().replace("in one persistent launch",
                                                   "in one persistent launch per size")),
            "sims_per_step": args.sims,
            "method": "plain DP5(4), per-lane dt, device work-queue refill",
            "converged_per_step": n_ok,
            "mean_accepted_steps": accepted / (args.sims * args.steps),
            "mean_attempted_steps": attempts / (args.sims * args.steps),
            # attempted lane-steps / (64 x wave loop iterations): the per-lane-dt
            # counterpart of the reference's shared-dt lockstep efficiency
            # (bench/bench_explosion.cpp:298-336), ceiling 1
            "lane_utilization": attempts / slots if slots else None,
            "kernel_ms_avg": avg_kernel_s * 1e3,
            "sampler_ms_avg": float(np.mean(sampler_ms)),
            "launches_per_step": launches,
            **({"mixed_launch": args.mixed_launch, "other_mixed_launch": other}
               if args.workload == "mixed" else {}),
            **({"record_writeback": writeback} if writeback else {}),
            "launch": {"grid": stats["grid_blocks"], "block": stats["block_threads"],
                       "cus": stats["n_cus"], **(tuned or {})},
        },
        "roofline": {
            "bound": "valu",
            "kernel": kernel_name,
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": (MEASURED_HBM_BYTES_PER_SIM[args.workload] * args.sims
                        if args.workload in MEASURED_HBM_BYTES_PER_SIM else None),
            "algorithmic_bytes_per_launch": bytes_launch,
            "hbm_gbs_algorithmic": bytes_launch / avg_kernel_s / 1e9,
            "flops_per_attempt": {str(n): flops_per_attempt(n) for n, _ in w.sizes},
        },
        "accuracy": acc,
        "closure": closure,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
