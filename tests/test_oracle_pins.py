"""Independent pins for the oracle functions the per-sim-charge and SoA GPU paths
are checked against (CPU only):

* ceo_soa_accelerations_f32 / ceo_soa_potential_energy_f32 (the reference's
  batched_force.hpp:31-110 on its SoA layout p[i*k + lane]) against the
  reference fixture's forces and PE (rotated and translated per lane, so a lane
  or atom index mix-up fails), the two-charge closed form of
  tests/test_system.cpp:28-50, the reference's own f32 CoulombForce
  (oracle/_ref, src/system.cpp:5-49) with unequal charges, and f64 brute force;
* ceo_lane_run_q_* (per-sim charges) against per-molecule runs of the scalar
  restatement, which is bit-exact against the reference engine: f64 with the
  same step counts and <= 1e-11 (rounding), f32 within 1e-5;
* ceo_device_sample_q's charge draws: chi-squared tests of the levels and of
  their independence across atoms, a cap on the total that binds (the
  conditional distribution over totals, by enumeration, and uniformity over the
  admissible charge vectors), an infeasible cap, split invariance, and
  positions independent of the charge draws.
"""
import itertools
import json
import os

import numpy as np
import pytest
from scipy.stats import chi2

from conftest import CH4_MASSES, GOLDEN, per_atom_rel
from oracle import ffi


def _cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)["cases"]


def _rotation(seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    return q


def _to_soa(geoms):
    """[k][n][3] -> px, py, pz of the reference's layout p[i*k + lane]."""
    g = np.asarray(geoms, dtype=np.float32)
    return [np.ascontiguousarray(g[:, :, c].T).reshape(-1) for c in range(3)]


def _from_soa(a, n, k):
    return np.stack([np.asarray(c).reshape(n, k).T for c in a], axis=-1)  # [k][n][3]


def _pair_const(q, kc=1.0):
    q = np.asarray(q, dtype=np.float64)
    return (kc * np.outer(q, q)).astype(np.float32).reshape(-1)


def _brute_f64(masses, charges, x):
    """F64 brute force over all ordered pairs (no Newton-3 bookkeeping)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(masses)
    a = np.zeros((n, 3))
    pe = 0.0
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            r = x[i] - x[j]
            d = np.sqrt(r @ r)
            a[i] += charges[i] * charges[j] * r / d ** 3 / masses[i]
            if i < j:
                pe += charges[i] * charges[j] / d
    return a, pe


@pytest.mark.parametrize("case", range(3))
def test_soa_force_and_pe_vs_reference_fixture(oracle, case):
    """Lane 0 holds the fixture geometry, lanes 1..4 rotated and translated copies:
    lane l's accelerations are R_l F/m of the fixture (forces_init) and every
    lane's PE is potential_energy_init, to f32 rounding."""
    c = _cases()[case]
    m = np.array(c["masses"])
    x0 = np.array(c["positions_init"])
    f0 = np.array(c["forces_init"])
    n, k = len(m), 5
    rots = [np.eye(3)] + [_rotation(s) for s in range(1, k)]
    shifts = [np.zeros(3)] + [np.array([0.3, -1.1, 2.0]) * s for s in range(1, k)]
    geoms = [x0 @ r.T + t for r, t in zip(rots, shifts)]
    px, py, pz = _to_soa(geoms)
    pc = _pair_const(c["charges"])
    im = (1.0 / m).astype(np.float32)
    a = _from_soa(oracle.soa_accelerations(n, k, px, py, pz, pc, im), n, k)
    pe = oracle.soa_potential_energy(n, k, px, py, pz, pc)
    for lane in range(k):
        want = (f0 / m[:, None]) @ rots[lane].T
        scale = np.abs(want).max()
        assert np.abs(a[lane] - want).max() <= 2e-5 * scale, (c["name"], lane)
        assert abs(pe[lane] / c["potential_energy_init"] - 1) < 2e-6, (c["name"], lane)


def test_soa_two_charge_closed_form(oracle):
    """Unit charges and masses at distance 1 (tests/test_system.cpp:28-50):
    PE = 1 and a = -/+ x exactly, in any lane of a batch."""
    k = 3
    geoms = [[[0.0, 0, 0], [1.0, 0, 0]], [[5.0, 5, 5], [5.0, 6, 5]], [[0.0, 0, 0], [0.0, 0, -1]]]
    px, py, pz = _to_soa(geoms)
    a = _from_soa(oracle.soa_accelerations(2, k, px, py, pz, _pair_const([1, 1]),
                                           np.ones(2, np.float32)), 2, k)
    pe = oracle.soa_potential_energy(2, k, px, py, pz, _pair_const([1, 1]))
    np.testing.assert_array_equal(pe, [1.0, 1.0, 1.0])
    for lane, g in enumerate(geoms):
        d = np.array(g[0]) - np.array(g[1])
        np.testing.assert_array_equal(a[lane, 0], d)
        np.testing.assert_array_equal(a[lane, 1], -d)


def test_soa_unequal_charges_vs_reference_f32_engine_and_brute_force(oracle):
    """Unequal charges make the pair constants differ, so a pair-index mistake
    changes the result: the SoA restatement agrees with the reference's own f32
    CoulombForce (src/system.cpp, built in oracle/_ref; its operations are
    ordered differently, so to a few ulps) and with f64 brute force."""
    n, k = 5, 7
    q = [1.0, 2.0, 3.0, 1.0, 2.0]
    masses = np.array(CH4_MASSES)
    geoms, _, fails = oracle.device_sample(4242, 0, k, n, 2.5, 1.0)
    assert fails == 0
    px, py, pz = _to_soa(geoms)
    a = _from_soa(oracle.soa_accelerations(n, k, px, py, pz, _pair_const(q),
                                           (1.0 / masses).astype(np.float32)), n, k)
    pe = oracle.soa_potential_energy(n, k, px, py, pz, _pair_const(q))
    ref32 = ffi.RefEngine("f32") if ffi.RefEngine.available("f32") else None
    for lane in range(k):
        g32 = geoms[lane].astype(np.float32).astype(np.float64)
        a64, pe64 = _brute_f64(masses, q, g32)
        scale = np.linalg.norm(a64, axis=1)[:, None]
        assert (np.abs(a[lane] - a64) <= 2e-6 * scale).all(), lane
        assert abs(pe[lane] / pe64 - 1) < 1e-6
        if ref32 is not None:
            ar = ref32.accelerations(masses, q, g32)
            assert (np.abs(a[lane] - ar) <= 1e-6 * scale).all(), lane
            assert abs(pe[lane] / ref32.potential_energy(masses, q, g32) - 1) < 5e-7


def _q_sims(oracle, n_sims, n, levels, max_total, seed=606):
    geom, q, st, fails = oracle.device_sample_q(seed, 0, n_sims, n, 2.5, 1.0, levels, max_total)
    assert fails == 0
    return geom, q


@pytest.mark.parametrize("rule", [dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1),
                                  dict(stop_rule=0, pe_stop_fraction=1e-6)])
def test_lane_run_q_vs_per_molecule_scalar_runs(oracle, rule):
    """Each sim's own charges, through the per-sim-charge lane form, against a
    scalar run of the molecule with those charges (the scalar restatement is
    bit-exact against the reference engine)."""
    geom, q = _q_sims(oracle, 24, 5, 3, 12)
    assert len({tuple(r) for r in q}) > 10  # the sims really carry different charges
    cfg64 = ffi.make_config(rtol=1e-8, atol=1e-12, **rule)
    cfg32 = ffi.make_config(rtol=1e-6, atol=1e-9, **rule)
    lane64 = oracle.lane_run_q(CH4_MASSES, q, geom, cfg64, precision="f64")
    lane32 = oracle.lane_run_q(CH4_MASSES, q, geom, cfg32, precision="f32")
    for s in range(len(q)):
        sc = oracle.scalar_run(CH4_MASSES, q[s], geom[s:s + 1], cfg64, precision="f64")
        assert lane64.steps[s] == sc.steps[0], s
        assert per_atom_rel(lane64.momenta[s:s + 1], sc.momenta).max() < 1e-11, s
        sc32 = oracle.scalar_run(CH4_MASSES, q[s], geom[s:s + 1], cfg32, precision="f32")
        assert per_atom_rel(lane32.momenta[s:s + 1], sc32.momenta).max() < 1e-5, s


@pytest.mark.skipif(not ffi.RefEngine.available(), reason="oracle/_ref not built")
def test_lane_run_q_vs_live_reference_engine(oracle):
    """The same against the reference engine itself (oracle/_ref), f64."""
    ref = ffi.RefEngine("f64")
    geom, q = _q_sims(oracle, 8, 5, 3, 11, seed=99)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-12, stop_rule=1, t_max=2e3, momentum_tol=1e-7)
    lane = oracle.lane_run_q(CH4_MASSES, q, geom, cfg, precision="f64")
    for s in range(len(q)):
        mom, steps, st, tf = ref.run(CH4_MASSES, q[s], geom[s:s + 1], cfg)
        assert lane.steps[s] == steps[0]
        assert per_atom_rel(lane.momenta[s:s + 1], mom).max() < 1e-12


def _chi2_ok(counts, expected, dof, p=1e-4):
    stat = float(((counts - expected) ** 2 / expected).sum())
    return stat < chi2.ppf(1 - p, dof), stat


@pytest.mark.parametrize("levels", [2, 3])
def test_charge_levels_uniform_and_independent(oracle, levels):
    """No cap: every atom's charge is uniform on {1..L}, independently of the
    other atoms (chi-squared on the levels and on the joint of atoms 0 and 3)."""
    n, ns = 5, 20000
    _, q = _q_sims(oracle, ns, n, levels, 0)
    assert set(np.unique(q)) == set(range(1, levels + 1))
    counts = np.array([(q == v).sum() for v in range(1, levels + 1)])
    ok, stat = _chi2_ok(counts, np.full(levels, q.size / levels), levels - 1)
    assert ok, stat
    joint = np.zeros((levels, levels))
    np.add.at(joint, (q[:, 0].astype(int) - 1, q[:, 3].astype(int) - 1), 1)
    ok, stat = _chi2_ok(joint.reshape(-1), np.full(levels * levels, ns / levels ** 2),
                        levels * levels - 1)
    assert ok, stat


def test_charge_cap_binds(oracle):
    """Levels {1,2,3}, N = 5, total <= 8 (of at most 15): redrawing the whole
    vector until the cap holds makes the vectors uniform over the admissible
    set.  The totals follow the enumerated conditional distribution (a redraw
    rule of total < cap would leave total = 8 empty), and every admissible
    vector appears with equal frequency."""
    n, levels, cap, ns = 5, 3, 8, 30000
    _, q = _q_sims(oracle, ns, n, levels, cap, seed=31)
    tot = q.sum(axis=1)
    assert tot.max() <= cap
    adm = [v for v in itertools.product(range(1, levels + 1), repeat=n) if sum(v) <= cap]
    by_total = {t: sum(1 for v in adm if sum(v) == t) for t in range(n, cap + 1)}
    counts = np.array([(tot == t).sum() for t in by_total])
    expected = np.array([by_total[t] / len(adm) * ns for t in by_total])
    ok, stat = _chi2_ok(counts, expected, len(by_total) - 1)
    assert ok, (stat, counts, expected)
    index = {v: i for i, v in enumerate(adm)}
    vc = np.zeros(len(adm))
    for row in q.astype(int):
        vc[index[tuple(row)]] += 1
    ok, stat = _chi2_ok(vc, np.full(len(adm), ns / len(adm)), len(adm) - 1)
    assert ok, stat


def test_charge_cap_vacuous_and_infeasible(oracle):
    """A cap of 2N on levels {1,2} never binds (same draws as no cap); a cap below
    N cannot be met: every sim fails with SAMPLER_FAIL and zeroed outputs."""
    g0, q0, st0, f0 = oracle.device_sample_q(5, 0, 500, 5, 2.5, 1.0, 2, 0)
    g1, q1, st1, f1 = oracle.device_sample_q(5, 0, 500, 5, 2.5, 1.0, 2, 10)
    np.testing.assert_array_equal(q0, q1)
    np.testing.assert_array_equal(g0, g1)
    g, q, st, fails = oracle.device_sample_q(5, 0, 20, 5, 2.5, 1.0, 3, 4)
    assert fails == 20 and (st == ffi.SAMPLER_FAIL).all()
    assert not q.any() and not g.any()


def test_charges_split_invariant_and_positions_independent(oracle):
    """Charges, like positions, depend only on (seed, global index); drawing them
    leaves the positions of the charge-free sampler unchanged."""
    a_g, a_q, _, _ = oracle.device_sample_q(17, 0, 100, 8, 3.5, 0.9, 3, 20)
    b_g, b_q, _, _ = oracle.device_sample_q(17, 60, 40, 8, 3.5, 0.9, 3, 20)
    np.testing.assert_array_equal(a_q[60:], b_q)
    np.testing.assert_array_equal(a_g[60:], b_g)
    plain, _, _ = oracle.device_sample(17, 0, 100, 8, 3.5, 0.9)
    np.testing.assert_array_equal(plain, a_g)
    c_g, c_q, _, _ = oracle.device_sample_q(18, 0, 100, 8, 3.5, 0.9, 3, 20)
    assert not np.array_equal(a_q, c_q)


def test_f32_bar_admits_conditioning_and_rejects_planted_error(oracle):
    """The GPU tests' f32 bar (conftest.assert_f32_bar) on the oracle's own f32
    per-lane form against f64, N = 10 with charges {1,2,3}: the slice holds a
    sim (1071) beyond 1e-5 per atom whose start is that ill-conditioned, and it
    passes; a 2e-5 error planted on a well-conditioned sim fails."""
    from conftest import assert_f32_bar
    masses = [21874.66] * 2 + [29156.95] * 2 + [1837.15] * 6
    rule = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
    g, q, st, fails = oracle.device_sample_q(49188291, 1000, 300, 10, 3.5, 0.9, 3, 0)
    cfg64 = ffi.make_config(rtol=1e-10, atol=1e-14, **rule)
    p32 = oracle.lane_run_q(masses, q, g, ffi.make_config(rtol=1e-6, atol=1e-9, **rule),
                            precision="f32").momenta
    p64 = oracle.lane_run_q(masses, q, g, cfg64, precision="f64").momenta
    assert per_atom_rel(p32, p64)[71] > 1e-5
    assert_f32_bar(oracle, masses, q, g, p32, p64, cfg64)
    bad = p32.copy()
    bad[5, 2] *= 1 + 2e-5
    with pytest.raises(AssertionError):
        assert_f32_bar(oracle, masses, q, g, bad, p64, cfg64)
