"""Launch-shape autotuner for the persistent DP5(4) kernels.

The kernel's free parameters are the simulations per block (block_threads: one
simulation per lane), the resident blocks per CU (occupancy actually launched),
the refill threshold (idle lanes a wave waits for before it pulls new
simulations) and the queue chunk (simulations a wave claims per atomic).  The
pair loop and the stage sums are fully unrolled per molecule size at compile
time, so the unroll inside a step is fixed by N.  The other unroll, of the
scheduling loop (2 or 4 attempted steps per finalize / refill check), was
measured on 10M CH4 sims: no faster (14.55 and 14.61 ms against 14.54 at one
step per check), and its run-time switch alone slowed the one-step kernel from
13.82 to 14.54 ms (one more scalar register live through the step), so it is
not a tuned dimension.  The tuner times the integration kernel (HIP events on
the engine stream) on the engine's loaded workload, coordinate by coordinate,
never measuring a shape twice, and caches the best shape per (device,
workload, precision, stop rule) in tuned_launch.json next to this file.
"""
from __future__ import annotations

import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned_launch.json")

SPACE = {
    "refill_threshold": [1, 2, 4, 8, 16],
    "chunk": [64, 256, 1024],
    "blocks_per_cu": [0, 1, 2, 3, 4],
    "block_threads": [64, 128, 256],
}
DEFAULT = {"block_threads": 256, "blocks_per_cu": 0, "refill_threshold": 4, "chunk": 64}


def _key(eng, cfg, key_extra: str) -> str:
    st = eng.stats()
    return f"cus{st['n_cus']}-{key_extra}-p{cfg.precision}-s{cfg.stop_rule}"


def load(key: str) -> dict | None:
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        return json.load(f).get(key)


def store(key: str, entry: dict) -> None:
    data = {}
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            data = json.load(f)
    data[key] = entry
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def time_shape(eng, cfg, shape: dict, reps: int = 2) -> float:
    eng.set_launch(**shape)
    best = float("inf")
    for _ in range(reps):
        eng.run(cfg, download=False)
        best = min(best, eng.stats()["kernel_ms"])
    return best


def tune(ce, eng, cfg, key_extra: str = "", rounds: int = 1, log=print) -> dict:
    """Coordinate descent over SPACE on the engine's loaded sims.  Returns
    {"shape": {...}, "kernel_ms": t, "trials": [...]} and caches it."""
    shape = dict(DEFAULT)
    t0 = time.time()
    eng.run(cfg, download=False)  # warm
    best_ms = time_shape(eng, cfg, shape)
    trials = [dict(shape, kernel_ms=best_ms)]
    tried = {tuple(sorted(shape.items()))}
    for _ in range(rounds):
        for name, values in SPACE.items():
            for v in values:
                cand = dict(shape, **{name: v})
                if tuple(sorted(cand.items())) in tried:  # measured already
                    continue
                tried.add(tuple(sorted(cand.items())))
                try:
                    ms = time_shape(eng, cfg, cand)
                except (ValueError, RuntimeError) as e:  # shape rejected by the engine
                    log(f"[autotune] {cand} rejected: {e}")
                    continue
                trials.append(dict(cand, kernel_ms=ms))
                log(f"[autotune] {cand} -> {ms:.3f} ms")
                if ms < best_ms:
                    best_ms, shape = ms, cand
    eng.set_launch(**shape)
    entry = {"shape": shape, "kernel_ms": best_ms, "seconds": time.time() - t0,
             "trials": trials}
    store(_key(eng, cfg, key_extra), entry)
    return entry


def apply_cached(ce, eng, cfg, key_extra: str = "") -> dict | None:
    """Sets the cached best shape on the engine if there is one."""
    e = load(_key(eng, cfg, key_extra))
    if e:
        eng.set_launch(**e["shape"])
        return e["shape"]
    return None
