"""The launch-shape autotuner (coulomb-explosion_amd/autotune.py).

CPU: the coordinate descent, its trial record and its cache, on a stand-in
engine whose kernel time is a known function of the shape; the committed cache
is consistent with its own trials.  GPU: a launch shape never changes a result
(every sim's bits are the same under every shape the tuner may pick), and the
tuner run on a real engine returns a measured shape from its space.
"""
import importlib.util
import json
import os

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "coulomb-explosion_amd")


def load_autotune():
    spec = importlib.util.spec_from_file_location("ce_autotune_under_test",
                                                  os.path.join(PKG, "autotune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class _Cfg:
    def __init__(self, precision=0, stop_rule=1):
        self.precision, self.stop_rule = precision, stop_rule


class _StandInEngine:
    """set_launch / run / stats with a kernel time that is a separable function of
    the shape, minimal at `best`; shapes in `reject` raise like the engine does."""

    def __init__(self, best, reject=()):
        self.best, self.reject = best, [dict(r) for r in reject]
        self.shape, self.runs = {}, 0

    def set_launch(self, **shape):
        if any(all(shape.get(k) == v for k, v in r.items()) for r in self.reject):
            raise ValueError("shape rejected")
        self.shape = dict(shape)

    def run(self, cfg, download=True):
        self.runs += 1

    def stats(self):
        ms = 10.0
        for k, v in self.best.items():
            ms += abs(np.log2(max(self.shape.get(k, v), 1) / max(v, 1)))
        return {"kernel_ms": ms, "n_cus": 256}


@pytest.fixture()
def autotune(tmp_path, monkeypatch):
    mod = load_autotune()
    monkeypatch.setattr(mod, "CACHE", str(tmp_path / "tuned.json"))
    return mod


def test_coordinate_descent_finds_the_separable_optimum(autotune):
    best = {"refill_threshold": 16, "chunk": 256, "blocks_per_cu": 2, "block_threads": 128}
    eng = _StandInEngine(best)
    entry = autotune.tune(None, eng, _Cfg(), key_extra="standin", log=lambda m: None)
    assert entry["shape"] == best
    assert entry["kernel_ms"] == pytest.approx(10.0)
    # every trial is a recorded shape with its time; the best is the smallest
    assert min(t["kernel_ms"] for t in entry["trials"]) == entry["kernel_ms"]
    assert eng.shape == best  # left on the engine
    # no shape is measured twice (one warm-up run, then 2 reps per trial); here
    # every dimension's default is in its value list, so one pass tries each
    # other value once
    keys = [tuple(sorted((k, v) for k, v in t.items() if k != "kernel_ms"))
            for t in entry["trials"]]
    assert len(keys) == len(set(keys))
    n_tried = sum(len(vals) - 1 for vals in autotune.SPACE.values())
    assert len(entry["trials"]) == 1 + n_tried
    assert eng.runs == 1 + 2 * (1 + n_tried)


def test_rejected_shapes_are_skipped(autotune):
    best = {"refill_threshold": 4, "chunk": 64, "blocks_per_cu": 3, "block_threads": 256}
    eng = _StandInEngine(best, reject=[{"blocks_per_cu": 3}])
    logs = []
    entry = autotune.tune(None, eng, _Cfg(), key_extra="standin", log=logs.append)
    assert entry["shape"]["blocks_per_cu"] != 3
    assert any("rejected" in m for m in logs)
    assert all(t.get("blocks_per_cu") != 3 for t in entry["trials"])


def test_cache_is_keyed_by_workload_precision_and_stop_rule(autotune):
    e1 = _StandInEngine({"refill_threshold": 8, "chunk": 64, "blocks_per_cu": 3,
                         "block_threads": 256})
    s1 = autotune.tune(None, e1, _Cfg(0, 1), key_extra="ch4", log=lambda m: None)["shape"]
    e2 = _StandInEngine({"refill_threshold": 2, "chunk": 1024, "blocks_per_cu": 2,
                         "block_threads": 64})
    s2 = autotune.tune(None, e2, _Cfg(1, 1), key_extra="ch4", log=lambda m: None)["shape"]
    assert s1 == e1.best and s2 == e2.best
    with open(autotune.CACHE) as f:
        data = json.load(f)
    assert set(data) == {"cus256-ch4-p0-s1", "cus256-ch4-p1-s1"}
    fresh = _StandInEngine({})
    assert autotune.apply_cached(None, fresh, _Cfg(0, 1), key_extra="ch4") == s1
    assert fresh.shape == s1
    assert autotune.apply_cached(None, fresh, _Cfg(1, 1), key_extra="ch4") == s2
    assert autotune.apply_cached(None, _StandInEngine({}), _Cfg(0, 0), key_extra="ch4") is None
    assert autotune.apply_cached(None, _StandInEngine({}), _Cfg(0, 1), key_extra="n10") is None


def test_committed_cache_is_consistent():
    mod = load_autotune()
    with open(mod.CACHE) as f:
        data = json.load(f)
    assert data, "no tuned shapes committed"
    for key, entry in data.items():
        shape = entry["shape"]
        for k, v in shape.items():
            assert v in mod.SPACE[k] or v == mod.DEFAULT[k], (key, k, v)
        assert entry["kernel_ms"] == min(t["kernel_ms"] for t in entry["trials"]), key
        best = [t for t in entry["trials"] if t["kernel_ms"] == entry["kernel_ms"]][0]
        assert {k: best[k] for k in shape} == shape, key


# ------------------------------------------------------------------------ GPU ----

SHAPES = [
    {},  # engine defaults
    {"block_threads": 64, "blocks_per_cu": 1, "refill_threshold": 1, "chunk": 64},
    {"block_threads": 128, "blocks_per_cu": 4, "refill_threshold": 16, "chunk": 1024},
    {"block_threads": 256, "blocks_per_cu": 3, "refill_threshold": 8, "chunk": 256},
]


def _results(ce, eng, cfg):
    r = eng.run(cfg)
    fields = ("momenta", "steps", "attempts", "status", "t_final")
    if isinstance(r, list):  # a mixed engine: one dict per species
        return tuple(np.concatenate([np.ravel(x[f]) for x in r]) for f in fields)
    return tuple(np.ravel(getattr(r, f)).copy() for f in fields)


@pytest.mark.gpu
def test_launch_shape_never_changes_a_result(gpu):
    ce = gpu
    import bench
    cfg = bench.run_config(ce)
    w = bench.Workload(ce, "ch4", 20_000)
    w.sample(seed=77)
    ref = None
    for shape in SHAPES:
        w.eng.set_launch(**shape)
        got = _results(ce, w.eng, cfg)
        if ref is None:
            ref = got
            continue
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), shape
    # per-sim charges (mixed sizes, one launch and per species)
    m = bench.Workload(ce, "mixed", 8_000)
    m.sample(seed=78)
    ref = None
    for shape in SHAPES[:3]:
        for ml in (ce.MIXED_ONE_LAUNCH, ce.MIXED_PER_SPECIES):
            m.eng.set_launch(**shape, mixed_launch=ml)
            got = _results(ce, m.eng, cfg)
            if ref is None:
                ref = got
                continue
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), (shape, ml)


@pytest.mark.gpu
def test_tuner_on_the_engine(gpu, tmp_path, monkeypatch):
    ce = gpu
    import bench
    mod = load_autotune()
    monkeypatch.setattr(mod, "CACHE", str(tmp_path / "tuned.json"))
    monkeypatch.setattr(mod, "SPACE", {"refill_threshold": [2, 8], "blocks_per_cu": [0, 3]})
    cfg = bench.run_config(ce)
    w = bench.Workload(ce, "ch4", 200_000)
    w.sample(seed=79)
    entry = mod.tune(ce, w.eng, cfg, key_extra="ch4-test", log=lambda m: None)
    shape = entry["shape"]
    assert shape["refill_threshold"] in (2, 4, 8) and shape["blocks_per_cu"] in (0, 3)
    assert w.eng.launch()["refill_threshold"] == shape["refill_threshold"]
    times = [t["kernel_ms"] for t in entry["trials"]]
    assert entry["kernel_ms"] == min(times) and all(t > 0 for t in times)
    # the chosen shape re-measured: the same order of time as its trial
    again = mod.time_shape(w.eng, cfg, shape, reps=3)
    assert again < 1.5 * entry["kernel_ms"]
    fresh = bench.Workload(ce, "ch4", 1000)
    assert mod.apply_cached(ce, fresh.eng, cfg, key_extra="ch4-test") == shape
