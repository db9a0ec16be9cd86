"""Dataset records (examples/generate_dataset.cpp:198-217) written by the device
pack kernel, and the generator CLI; byte layout as examples/to_parquet.py:34-42
reads it."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import CH4_MASSES, ROOT

pytestmark = pytest.mark.gpu


def test_records_match_results(gpu, oracle):
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] +
                      [ce.Atom("H", m, 1.0) for m in CH4_MASSES[1:]])
    ns = 1000  # not a multiple of the 256-record block: exercises the tail
    eng = ce.BatchedRK45(mol, 1.0, ns, 1e-6, 1e-9)
    eng.sample(seed=3, n_sims=ns, radius=2.5, min_separation=1.0)
    res = eng.run(ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, h0_mode=ce.H0_FIXED))
    rec = eng.records()
    assert rec.dtype.itemsize == 125 and rec.shape == (ns,)
    geom = eng.geometries()
    np.testing.assert_array_equal(rec["init_pos"], geom.reshape(ns, -1).astype(np.float32))
    np.testing.assert_array_equal(rec["final_p"], res.momenta_f32.reshape(ns, -1))
    np.testing.assert_array_equal(rec["final_p"], res.momenta.reshape(ns, -1).astype(np.float32))
    np.testing.assert_array_equal(rec["steps"], res.steps)
    np.testing.assert_array_equal(rec["converged"], res.converged)
    part = eng.records(first=300, n_sims=257)
    np.testing.assert_array_equal(part, rec[300:557])
    # raw bytes: little-endian packed, no padding
    raw = rec.tobytes()
    first = np.frombuffer(raw[:60], dtype="<f4")
    np.testing.assert_array_equal(first, geom[0].reshape(-1).astype(np.float32))
    assert raw[124] == rec["converged"][0]


def test_queued_records_into_pinned_memory(gpu):
    """ce_engine_pack_records_async: the synchronous pack's bytes, with two copies
    in flight (the device buffers alternate) and a third call waiting on the
    first on the device, and across runs (a copy still in flight while the next
    run overwrites the results it was packed from)."""
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] +
                      [ce.Atom("H", m, 1.0) for m in CH4_MASSES[1:]])
    ns = 3000
    eng = ce.BatchedRK45(mol, 1.0, ns, 1e-6, 1e-9)
    cfg = ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, h0_mode=ce.H0_FIXED)
    bufs = [ce.PinnedBuffer(ns * 125) for _ in range(3)]
    eng.sample(seed=4, n_sims=ns, radius=2.5, min_separation=1.0)
    eng.run(cfg, download=False)
    want = eng.records().copy()
    views = [eng.records_async(bufs[0]), eng.records_async(bufs[1], first=1000, n_sims=1500),
             eng.records_async(bufs[2], first=17, n_sims=1)]
    eng.records_wait()
    assert views[0].tobytes() == want.tobytes()
    assert views[1].tobytes() == want[1000:2500].tobytes()
    assert views[2].tobytes() == want[17:18].tobytes()
    # batch 2 runs while batch 1's copy may still be in flight
    v1 = eng.records_async(bufs[0])
    eng.sample(seed=5, n_sims=ns, radius=2.5, min_separation=1.0)
    eng.run(cfg, download=False)
    v2 = eng.records_async(bufs[1])
    eng.records_wait()
    assert v1.tobytes() == want.tobytes()
    assert v2.tobytes() == eng.records().tobytes() and v2.tobytes() != want.tobytes()
    with pytest.raises(ValueError):
        eng.records_async(ce.PinnedBuffer(100))


@pytest.mark.parametrize("graph", [True, False])
def test_generator_cli(gpu, tmp_path, graph):
    """Two full batches (replayed from a captured graph unless --no-graph) and a
    partial one give the records of one 5000-sim run."""
    exe = os.path.join(ROOT, "coulomb-explosion_amd", "generate_dataset")
    out = str(tmp_path / "ch4.bin")
    r = subprocess.run([exe, "--out", out, "--sims", "5000", "--batch", "2048", "--radius", "2.5",
                        "--min-separation", "1.0", "--rtol", "1e-6", "--atol", "1e-9",
                        "--stop-rule", "tmax", "--h0", "1e-3", "--seed", "11"] +
                       ([] if graph else ["--no-graph"]),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    meta = json.load(open(out + ".meta.json"))
    assert meta["n_sims"] == 5000 and meta["n_atoms"] == 5 and meta["n_failed_convergence"] == 0
    assert meta["symbols"] == ["C", "H", "H", "H", "H"]
    from __graft_entry__ import load_package
    ce = load_package()
    rec = np.fromfile(out, dtype=ce.record_dtype(5))
    assert rec.shape == (5000,) and (rec["converged"] == 1).all()
    # batches are keyed by global sim index: the same sims as one 5000-sim run
    masses = [12.011 * 1822.888486209] + [1.008 * 1822.888486209] * 4
    mol = ce.Molecule([ce.Atom("C", masses[0], 1.0)] + [ce.Atom("H", m, 1.0) for m in masses[1:]])
    eng = ce.BatchedRK45(mol, 1.0, 5000, 1e-6, 1e-9)
    eng.sample(seed=11, n_sims=5000, radius=2.5, min_separation=1.0)
    eng.run(ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, h0_mode=ce.H0_FIXED, h0=1e-3))
    np.testing.assert_array_equal(eng.records(), rec)


def test_verify_subset_against_fp64_asymptote(gpu, tmp_path):
    """The verify_subset counterpart on a dataset the generator wrote with the
    reference's production settings (pe_stop rule, redistribution)."""
    exe = os.path.join(ROOT, "coulomb-explosion_amd", "generate_dataset")
    out = str(tmp_path / "ch4.bin")
    r = subprocess.run([exe, "--out", out, "--sims", "3000", "--batch", "1024", "--seed", "5"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    from __graft_entry__ import load_package
    ce = load_package()
    import coulomb_explosion_amd.verify_subset as vs
    res = vs.verify(ce, out, 500)
    assert res["n"] == 500 and res["reference_status_ok"] == 1.0
    # rtol 1e-4 / pe_stop 1e-5 with redistribution: well inside the 1% floor
    assert res["over_floor"] == 0 and res["config_norm"]["max"] < 1e-2
    assert res["config_norm"]["median"] < 1e-3


def test_precision_sweep_small(gpu):
    from __graft_entry__ import load_package
    ce = load_package()
    import coulomb_explosion_amd.precision_sweep as ps
    for n in (2, 5, 8):
        r = ps.sweep_one(ce, n, 1000, 7)
        assert r["f64_failures"] == 0 and sum(r["f32_failures"].values()) == 0
        assert r["dp_over_p"]["over_floor"] == 0 and r["dp_over_p"]["median"] < 1e-3
