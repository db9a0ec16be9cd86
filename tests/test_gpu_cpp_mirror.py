"""The C++ mirror (include/coulomb_gpu.hpp) as the reference's callers use it.

* tests/cpp/generator_loop.cpp is the reference generator's batch loop
  (examples/generate_dataset.cpp:159-227: Sampler::sample(mol, State&),
  BatchedRK45<float>::set_lane_geometry(lane, state.positions), run, the record
  layout); its records are the Python mirror's bytes for the same stream.
* tests/cpp/strategy_api.cpp drives the sampler through the abstract Sampler,
  CoulombForce, make_integrator(RK45)->step and run_to_convergence (engine and
  generic-loop paths); checked against the oracle (pinned to the reference).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, per_atom_rel
from oracle import ffi

pytestmark = pytest.mark.gpu

LIB_DIR = os.path.join(ROOT, "coulomb-explosion_amd")


def build(name, tmp_path):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++20", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe,
                    "-L" + LIB_DIR, "-lcoulomb_explosion", "-Wl,-rpath," + LIB_DIR],
                   check=True, capture_output=True, timeout=120)
    return exe


def test_reference_generator_loop_writes_the_python_mirrors_records(gpu, tmp_path):
    ce = gpu
    exe = build("generator_loop", tmp_path)
    lanes, sims = 64, 150  # three batches, the last one partly written
    out = tmp_path / "ch4.bin"
    r = subprocess.run([exe, str(lanes), str(sims), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=ce.record_dtype(5))
    assert len(got) == sims
    mol = ce.Molecule([ce.atom_from_amu("C", 12.011, 1.0)] +
                      [ce.atom_from_amu("H", 1.008, 1.0)] * 4)
    sampler = ce.UniformSphereSampler(4.0, 0.25, 10000, seed=49188291)
    bi = ce.BatchedRK45(mol, 1.0, lanes, 1e-4, 1e-7)
    want = np.zeros(sims, dtype=ce.record_dtype(5))
    done = 0
    while done < sims:
        geom = sampler.sample(mol, lanes)
        bi.set_geometries(geom)
        res = bi.run(ce.RunConfig(pe_stop_fraction=1e-5))
        k = min(lanes, sims - done)
        want["init_pos"][done:done + k] = geom[:k].reshape(k, -1).astype(np.float32)
        want["final_p"][done:done + k] = res.momenta.reshape(lanes, -1)[:k].astype(np.float32)
        want["steps"][done:done + k] = res.steps[:k]
        want["converged"][done:done + k] = res.converged[:k]
        done += k
    assert got.tobytes() == want.tobytes()
    assert got["converged"].all()


def _parse(text):
    rows = {}
    for line in text.strip().splitlines():
        key, *vals = line.split()
        rows[key] = vals
    return rows


def test_strategy_api_vs_oracle(gpu, oracle, tmp_path):
    exe = build("strategy_api", tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = _parse(r.stdout)
    assert "observer_out_of_order" not in rows
    f = lambda k: np.array([float(v) for v in rows[k]])  # noqa: E731
    masses = np.array([21874.66] + [1837.15] * 4)
    charges = [1.0, 2.0, 1.0, 1.0, 2.0]
    # sampler through the abstract interface: the reference's stream, at rest
    assert rows["sampler_name"] == ["uniform-sphere"]
    rc, ref_geom = oracle.sample(12345, 5, 1, 2.5, 1.0)
    geom = f("geom").reshape(5, 3)
    np.testing.assert_array_equal(geom, ref_geom[0])
    assert not f("vel0").any()
    # CoulombForce (f64 kernels) against the oracle's restatement of system.cpp
    np.testing.assert_allclose(f("acc").reshape(5, 3),
                               oracle.accelerations(masses, charges, geom), rtol=1e-13,
                               atol=1e-16)
    assert abs(f("pe")[0] / oracle.potential_energy(masses, charges, geom) - 1) < 1e-14
    assert f("ke")[0] == 0.0
    # twelve Integrator::step calls (each a resumed one-step launch) = one engine
    # launch stopped after 12 steps, bit for bit; and the oracle's per-lane form
    # stopped there (its f64 rounding differs, and at rtol 1e-8 the controller's
    # step choice amplifies that to ~1e-9)
    assert rows["integrator_name"] == ["dormand-prince-45"]
    ce = gpu
    rule = dict(stop_rule=1, momentum_tol=0.0, t_max=1e300, max_steps=12)
    eng = ce.BatchedRK45(ce.Molecule([ce.Atom("X", m, q) for m, q in zip(masses, charges)]),
                         1.0, 1, 1e-8, 1e-12)
    eng.set_geometries(geom[None])
    eng.enable_trace(12)
    one = eng.run(ce.RunConfig(redistribute_energy=False, precision=ce.PRECISION_F64, **rule))
    np.testing.assert_array_equal(f("dt"), one.trace_h.reshape(-1))
    p_steps = f("step_vel").reshape(1, 5, 3) * masses[None, :, None]
    np.testing.assert_array_equal(p_steps.reshape(-1), one.momenta)
    ref = oracle.lane_run(masses, charges, geom[None],
                          ffi.make_config(rtol=1e-8, atol=1e-12, redistribute=False, **rule),
                          precision="f64", trace_cap=12)
    np.testing.assert_allclose(f("dt"), ref.trace_h[0], rtol=1e-8)
    assert per_atom_rel(p_steps, ref.momenta).max() < 1e-9
    # run_to_convergence against the reference form (scalar restatement, bit-exact
    # against the reference engine), PE rule 1e-6 with redistribution
    sc = oracle.scalar_run(masses, charges, geom[None],
                           ffi.make_config(rtol=1e-8, atol=1e-12, pe_stop_fraction=1e-6),
                           precision="f64")
    runs = {}
    for key in ("run_engine", "run_loop"):
        vals = f(key)
        steps, t_final, e0, pe_f, ke_b, ke_a, s, conv, observed, t_obs = vals
        assert steps == sc.steps[0] and conv == 1 and observed == steps
        assert t_obs == pytest.approx(t_final, rel=1e-14)
        assert t_final == pytest.approx(sc.t_final[0], rel=1e-9)
        assert ke_a == pytest.approx(e0, rel=1e-14) and s == pytest.approx(np.sqrt(e0 / ke_b))
        assert pe_f <= 1e-6 * e0
        p = f(key + "_vel").reshape(1, 5, 3) * masses[None, :, None]
        assert per_atom_rel(p, sc.momenta).max() < 1e-9
        runs[key] = p
    # the one-launch path and the generic loop over Integrator::step agree
    assert per_atom_rel(runs["run_engine"], runs["run_loop"]).max() < 1e-13
    assert rows["vv_throws"] == ["1"]
