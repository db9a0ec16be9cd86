"""The headline workload at its full size (10M CH4 sims in one launch), checked
through properties that do not depend on the batch size: every sim stops by the
rule, momentum is conserved, the redistributed kinetic energy equals the initial
potential energy, the kinetic energy splits between the species as the
equal-momentum estimate says, results do not depend on how the sims are batched (a slice
re-run in a small engine gives the same bits), and a slice agrees with the
oracle's per-lane statement."""
import numpy as np
import pytest

from conftest import CH4_CHARGES, CH4_MASSES, per_atom_rel

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
N_SIMS = 10_000_000
SEED = 49188291


def ch4(ce):
    return ce.Molecule([ce.Atom("X", m, q) for m, q in zip(CH4_MASSES, CH4_CHARGES)])


@pytest.fixture(scope="module")
def full_run(gpu):
    ce = gpu
    print(f"[fullsize] CH4 {N_SIMS} sims, seed {SEED}", flush=True)
    eng = ce.BatchedRK45(ch4(ce), 1.0, N_SIMS, 1e-6, 1e-9)
    eng.sample(SEED, N_SIMS, 2.5, 1.0)
    res = eng.run(ce.RunConfig(**RULE))
    geom = eng.geometries()
    return ce, eng, res, geom


def test_every_sim_stops_by_the_rule(full_run):
    ce, eng, res, _ = full_run
    assert res.status.shape == (N_SIMS,)
    assert np.all((res.status == ce.SIM_OK_MOMENTUM) | (res.status == ce.SIM_OK_TMAX))
    assert np.all(res.attempts >= res.steps) and res.steps.min() >= 10 and res.steps.max() < 200
    np.testing.assert_array_equal(res.t_final[res.status == ce.SIM_OK_TMAX], 1e4)
    st = eng.stats()
    assert st["sims"] == N_SIMS and st["attempted_steps"] == int(res.attempts.sum())


def test_momentum_conserved_energy_redistributed_and_partitioned(full_run):
    ce, _, res, geom = full_run
    p = res.momenta.reshape(N_SIMS, 5, 3)
    scale = np.linalg.norm(p, axis=-1).sum(axis=1)
    assert np.max(np.linalg.norm(p.sum(axis=1), axis=-1) / scale) < 1e-5
    m = np.asarray(CH4_MASSES)
    ke = 0.5 * np.sum(np.sum(p * p, axis=-1) / m, axis=1)
    pe0 = np.zeros(N_SIMS)
    for i in range(5):
        for j in range(i + 1, 5):
            pe0 += 1.0 / np.linalg.norm(geom[:, i] - geom[:, j], axis=-1)
    # redistribution makes KE equal E0 = PE0 (E0 from the f32 potential energy)
    assert np.max(np.abs(ke - pe0) / pe0) < 1e-5
    # energy partition per species against the equal-|p| prediction n_s/m_s
    # normalised (demo/sanity_check.py:110-124 checks it to 1%)
    ke_atom = np.sum(p * p, axis=-1) / (2.0 * m)
    h_share = ke_atom[:, 1:].sum() / ke_atom.sum()
    w_c, w_h = 1.0 / m[0], 4.0 / m[1]
    assert abs(h_share - w_h / (w_c + w_h)) < 0.01


def test_slices_rerun_in_small_engines_give_the_same_bits(full_run):
    ce, _, res, _ = full_run
    p = res.momenta.reshape(N_SIMS, 5, 3)
    for first in (0, 4_321_987, N_SIMS - 3000):
        small = ce.BatchedRK45(ch4(ce), 1.0, 3000, 1e-6, 1e-9)
        small.sample(SEED, 3000, 2.5, 1.0, first_index=first)
        r = small.run(ce.RunConfig(**RULE))
        np.testing.assert_array_equal(r.momenta.reshape(3000, 5, 3), p[first:first + 3000])
        np.testing.assert_array_equal(r.steps, res.steps[first:first + 3000])


def test_slice_vs_oracle_lane(full_run, oracle):
    ce, _, res, geom = full_run
    from oracle import ffi
    sl = slice(7_000_000, 7_001_000)
    ref = oracle.lane_run(CH4_MASSES, CH4_CHARGES, geom[sl],
                          ffi.make_config(rtol=1e-6, atol=1e-9, **RULE))
    assert per_atom_rel(res.momenta.reshape(N_SIMS, 5, 3)[sl], ref.momenta).max() < 1e-5
    np.testing.assert_array_equal(res.status[sl], ref.status)
