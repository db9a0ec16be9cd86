"""Pins the CPU oracle (CPU only).

Every oracle function is checked against something it does not restate:
the reference's golden fixture, the reference engine compiled from its sources
(oracle/_ref) or the vectors it wrote (tests/golden/ref_*.{json,npz}), closed
forms of the Coulomb problem, and published known-answer vectors.
"""
import json
import os

import numpy as np
import pytest
from scipy.optimize import brentq

from conftest import CH4_CHARGES, CH4_MASSES, GOLDEN, config_rel, per_atom_rel
from oracle import ffi

AMU = 1822.888486209


def load_cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)


# ------------------------------------------------------- reference fixture ----
@pytest.mark.parametrize("precision", ["f64"])
def test_fixture_forces_and_pe(oracle, precision):
    """Tier 1 of tests/test_reference.cpp:83-104,156-167: F = m a within 1e-9,
    PE within 1e-12 relative."""
    doc = load_cases()
    assert doc["coulomb_constant"] == 1.0
    for c in doc["cases"]:
        x = np.array(c["positions_init"])
        m = np.array(c["masses"])
        a = oracle.accelerations(m, c["charges"], x)
        np.testing.assert_allclose(a * m[:, None], np.array(c["forces_init"]), rtol=0, atol=1e-9)
        pe = oracle.potential_energy(m, c["charges"], x)
        assert abs(pe / c["potential_energy_init"] - 1.0) < 1e-12


def test_fixture_asymptotic_momenta(oracle):
    """Tier 2 (tests/test_reference.cpp:119-154): rtol 1e-8, atol 1e-16,
    pe_stop 1e-10, redistribution on; each component within 1e-6 relative or
    1e-6 * max|p| absolute; the same for the per-lane form."""
    doc = load_cases()
    cfg = ffi.make_config(rtol=doc["rk_rtol"], atol=doc["rk_atol"], pe_stop_fraction=1e-10)
    for c in doc["cases"]:
        x = np.array(c["positions_init"])[None]
        ref = np.array(c["momenta_final"])
        scale = np.abs(ref).max()
        for run in (oracle.scalar_run(c["masses"], c["charges"], x, cfg),
                    oracle.lane_run(c["masses"], c["charges"], x, cfg, precision="f64")):
            assert run.status[0] == ffi.OK_PE
            p = run.momenta[0]
            ok = (np.abs(p - ref) <= 1e-6 * np.abs(ref)) | (np.abs(p - ref) <= 1e-6 * scale)
            assert ok.all(), (c["name"], p, ref)


# ------------------------------------------------------------ closed forms ----
def test_two_charges_closed_form(oracle):
    """tests/test_system.cpp:28-50: unit charges at r=1: PE = 1, a = -+1 x."""
    m, q = [1.0, 1.0], [1.0, 1.0]
    x = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    assert oracle.potential_energy(m, q, x) == 1.0
    a = oracle.accelerations(m, q, x)
    np.testing.assert_allclose(a, [[-1, 0, 0], [1, 0, 0]], atol=1e-15)


def two_body_state(q1q2, mu, r0, t):
    """Two bodies released at rest: t(r) = sqrt(mu r0^3/(2k)) [sqrt(x(x-1)) +
    ln(sqrt x + sqrt(x-1))], x = r/r0; relative momentum p = sqrt(2 mu k (1/r0 - 1/r))."""
    k = q1q2

    def t_of(x):
        return np.sqrt(mu * r0 ** 3 / (2 * k)) * (np.sqrt(x * (x - 1)) +
                                                   np.log(np.sqrt(x) + np.sqrt(x - 1)))

    xr = brentq(lambda x: t_of(x) - t, 1.0 + 1e-15, 1e12, xtol=1e-14, rtol=1e-15)
    return np.sqrt(2 * mu * k * (1 / r0 - 1 / (xr * r0)))


@pytest.mark.parametrize("q2,m2", [(1.0, 25526.04), (2.0, 25526.04), (1.0, 1837.15)])
def test_two_body_tmax_stop_closed_form(oracle, q2, m2):
    """The t_max rule without redistribution lands exactly at t = 1e4: the
    relative momentum there is fixed by the two-body time integral."""
    m = [25526.04, m2]
    q = [1.0, q2]
    r0 = 1.3
    x = np.array([[[0.0, 0, 0], [r0, 0, 0]]])
    mu = m[0] * m[1] / (m[0] + m[1])
    p_exact = two_body_state(q[0] * q[1], mu, r0, 1e4)
    cfg = ffi.make_config(rtol=1e-12, atol=1e-16, stop_rule=1, t_max=1e4, momentum_tol=1e-14,
                          redistribute=False)
    for run in (oracle.scalar_run(m, q, x, cfg), oracle.lane_run(m, q, x, cfg, precision="f64")):
        assert run.status[0] == ffi.OK_TMAX and run.t_final[0] == 1e4
        p = run.momenta[0]
        np.testing.assert_allclose(p[:, 1:], 0.0, atol=1e-14)
        assert abs(p[1, 0] + p[0, 0]) < 1e-9 * abs(p[1, 0])  # total momentum stays 0
        assert abs(p[1, 0] / p_exact - 1) < 1e-9


def test_two_body_asymptote_closed_form(oracle):
    """|p_inf| = sqrt(2 mu q1 q2 / r0); with redistribution the PE rule reaches it."""
    m, q, r0 = [1822.888486209] * 2, [1.0, 1.0], 1.0
    mu = m[0] / 2
    cfg = ffi.make_config(rtol=1e-10, atol=1e-16, pe_stop_fraction=1e-8)
    run = oracle.scalar_run(m, q, np.array([[[0.0, 0, 0], [r0, 0, 0]]]), cfg)
    assert abs(run.momenta[0, 1, 0] / np.sqrt(2 * mu / r0) - 1) < 1e-9


def test_symmetric_closed_forms(oracle):
    """Equilateral H3: each |p| = sqrt(2 m E0/3).  CH4 with C at the centre of a
    regular H tetrahedron: p_C = 0, |p_H| = sqrt(m_H E0 / 2)."""
    cfg = ffi.make_config(rtol=1e-10, atol=1e-16, pe_stop_fraction=1e-7)
    mh = 1837.15
    h = np.sqrt(3.0)
    x = np.array([[[0.0, 0, 0], [2.0, 0, 0], [1.0, h, 0]]])
    run = oracle.scalar_run([mh] * 3, [1.0] * 3, x, cfg)
    e0 = 1.5
    np.testing.assert_allclose(np.linalg.norm(run.momenta[0], axis=1), np.sqrt(2 * mh * e0 / 3),
                               rtol=1e-9)
    d = 1.09 / 0.529177  # C-H bond, bohr
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3) * d
    x = np.vstack([[0, 0, 0], tet])[None]
    e0 = 4 / d + 6 / (d * np.sqrt(8 / 3))
    run = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, x, cfg)
    assert np.linalg.norm(run.momenta[0, 0]) < 1e-9
    np.testing.assert_allclose(np.linalg.norm(run.momenta[0, 1:], axis=1),
                               np.sqrt(mh * e0 / 2), rtol=1e-9)


def test_energy_and_momentum_invariants(oracle):
    """tests/test_integrator.cpp:34-82 and demo/sanity_check.py:89-108: energy
    conserved to ~rtol on an unequal-mass pair, sum p = 0, KER = PE0."""
    m, q = [1.0, 2.0], [1.0, 1.0]
    x = np.array([[[0.0, 0, 0], [1.0, 0, 0]]])
    cfg = ffi.make_config(rtol=1e-10, atol=1e-16, stop_rule=1, t_max=1e4, momentum_tol=1e-20,
                          redistribute=False)
    run = oracle.scalar_run(m, q, x, cfg)
    assert abs(run.energy_err[0]) < 1e-7
    rc, geom = oracle.sample(5, 5, 50, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-12, pe_stop_fraction=1e-6)
    run = oracle.lane_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f64")
    p = run.momenta
    assert (np.abs(p.sum(axis=1)).max(axis=1) / np.abs(p).max(axis=(1, 2))).max() < 1e-6
    ke = (0.5 * (p ** 2).sum(axis=2) / np.array(CH4_MASSES)).sum(axis=1)
    pe0 = np.array([oracle.potential_energy(CH4_MASSES, CH4_CHARGES, g) for g in geom])
    np.testing.assert_allclose(ke, pe0, rtol=1e-12)


# ------------------------------------------------ reference engine (golden) ----
def test_sampler_matches_reference_stream(oracle):
    with open(os.path.join(GOLDEN, "ref_sampler.json")) as f:
        doc = json.load(f)
    for s in doc["streams"]:
        n_sims = len(s["positions"])
        rc, pos = oracle.sample(s["seed"], s["n_atoms"], n_sims, s["radius"], s["min_separation"])
        assert rc == 0
        np.testing.assert_array_equal(pos, np.array(s["positions"]))


def test_sampler_errors(oracle):
    """src/sampler.cpp:7-15 (invalid_argument) and :50-54 (runtime_error)."""
    assert oracle.sample(1, 5, 1, -1.0, 0.25)[0] == -1
    assert oracle.sample(1, 5, 1, 1.0, -0.1)[0] == -1
    assert oracle.sample(1, 10, 1, 1.0, 1.5, max_attempts=100)[0] == -2


@pytest.mark.parametrize("rule", ["tmax", "pe"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_scalar_restatement_bit_exact_vs_reference_engine(oracle, rule, precision):
    """The scalar restatement equals the reference's engine (built from its
    sources, vectors in ref_dp45.npz) bit for bit."""
    g = np.load(os.path.join(GOLDEN, "ref_dp45.npz"))
    cfg = (ffi.make_config(rtol=1e-6, atol=1e-9, stop_rule=1, t_max=1e4, momentum_tol=1e-7)
           if rule == "tmax" else ffi.make_config(rtol=1e-6, atol=1e-9, pe_stop_fraction=1e-5))
    run = oracle.scalar_run(g["masses"], g["charges"], g["geom"], cfg, precision=precision)
    np.testing.assert_array_equal(run.momenta, g[f"{rule}_{precision}_momenta"])
    np.testing.assert_array_equal(run.steps, g[f"{rule}_{precision}_steps"])
    np.testing.assert_array_equal(run.t_final, g[f"{rule}_{precision}_t"])


@pytest.mark.skipif(not ffi.RefEngine.available(), reason="oracle/_ref not built")
def test_scalar_restatement_vs_live_reference_build(oracle):
    """Same check on fresh geometries against oracle/_ref itself."""
    ref = ffi.RefEngine("f64")
    rc, geom = oracle.sample(31337, 5, 40, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-7, atol=1e-10, stop_rule=1, t_max=2e3, momentum_tol=1e-7)
    mom, steps, st, tf = ref.run(CH4_MASSES, CH4_CHARGES, geom, cfg)
    run = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, geom, cfg)
    np.testing.assert_array_equal(run.momenta, mom)
    np.testing.assert_array_equal(run.steps, steps)
    rc2, g2 = ref.sample(31337, 5, 40, 2.5, 1.0)
    np.testing.assert_array_equal(g2, geom)


# --------------------------------------------------------- per-lane form ----
def test_lane_coefficient_identities(oracle):
    """The derived tables satisfy the identities that make the per-lane form
    the same method: c_s = sum_l a_sl, (A A) computed from A, sum E = 0, E A."""
    c, aa, b, e, ea = oracle.lane_coefficients()
    a = np.zeros((7, 7))
    a[1, 0] = 1 / 5
    a[2, :2] = [3 / 40, 9 / 40]
    a[3, :3] = [44 / 45, -56 / 15, 32 / 9]
    a[4, :4] = [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729]
    a[5, :5] = [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656]
    a[6, :6] = [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]
    np.testing.assert_allclose(c, a.sum(axis=1), atol=1e-15)
    np.testing.assert_allclose(aa, a @ a, atol=1e-14)
    np.testing.assert_allclose(b, a[6], atol=0)
    assert abs(e.sum()) < 1e-16
    np.testing.assert_allclose(ea, e @ a, atol=1e-16)
    # order conditions of the 5th-order weights: sum b = 1, sum b c = 1/2, sum b c^2 = 1/3
    assert abs(b.sum() - 1) < 1e-15
    assert abs(b @ c - 0.5) < 1e-15
    assert abs(b @ c ** 2 - 1 / 3) < 1e-15


@pytest.mark.parametrize("rule", [dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7),
                                  dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1),
                                  dict(stop_rule=0, pe_stop_fraction=1e-6)])
def test_lane_form_equals_reference_form(oracle, rule):
    """The Runge-Kutta-Nystrom arrangement is algebraically the reference's
    DP5(4): in f64 the two agree to rounding, step for step; in f32 within 1e-5."""
    rc, geom = oracle.sample(2024, 5, 64, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-12, **rule)
    a = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f64", trace_cap=8)
    b = oracle.lane_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f64", trace_cap=8)
    np.testing.assert_array_equal(a.steps, b.steps)
    np.testing.assert_allclose(a.trace_h, b.trace_h, rtol=1e-6)
    assert per_atom_rel(b.momenta, a.momenta).max() < 1e-12
    cfg = ffi.make_config(rtol=1e-6, atol=1e-9, **rule)
    a = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f32")
    b = oracle.lane_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f32")
    assert per_atom_rel(b.momenta, a.momenta).max() < 1e-5


@pytest.mark.parametrize("rule", [dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1),
                                  dict(stop_rule=0, pe_stop_fraction=1e-6)])
def test_lane_uniform_form(oracle, rule):
    """The f32 lane form folds a common pair constant c into 1/m and the PE.
    For c = 1 it gives the table form's bits (per-sim-charge entry, pair
    constant 1 * 1); for c = 9 (all charges 3) it agrees with the reference
    form within 1e-5, and the PE scale enters E0 (pe-stop rule with
    redistribution; energy diagnostic)."""
    rc, geom = oracle.sample(77, 5, 48, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-6, atol=1e-9, **rule)
    a = oracle.lane_run(CH4_MASSES, CH4_CHARGES, geom, cfg, precision="f32")
    b = oracle.lane_run_q(CH4_MASSES, np.ones((48, 5)), geom, cfg, precision="f32")
    np.testing.assert_array_equal(a.momenta, b.momenta)
    np.testing.assert_array_equal(a.steps, b.steps)
    q3 = [3.0] * 5
    u = oracle.lane_run(CH4_MASSES, q3, geom, cfg, precision="f32")
    ref = oracle.scalar_run(CH4_MASSES, q3, geom, cfg, precision="f32")
    assert per_atom_rel(u.momenta, ref.momenta).max() < 1e-5
    t = oracle.lane_run_q(CH4_MASSES, np.full((48, 5), 3.0), geom, cfg, precision="f32")
    assert per_atom_rel(u.momenta, t.momenta).max() < 1e-5
    assert np.abs(u.energy_err).max() < 1e-5


# --------------------------------------------------------- lockstep form ----
def test_batched_gate_vs_scalar(oracle):
    """bench/bench_batched_integrator.cpp:101-126: lockstep f64 vs per-sim scalar,
    config-norm < max(50 rtol, 1e-6), all converged."""
    rc, geom = oracle.sample(0xC0FFEE, 5, 16, 4.0, 0.25)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-16)
    bt = oracle.batched_run(CH4_MASSES, CH4_CHARGES, geom, cfg, lanes=8, precision="f64")
    sc = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, geom, cfg)
    assert (bt.status == ffi.OK_PE).all()
    assert config_rel(bt.momenta, sc.momenta).max() < max(50 * 1e-8, 1e-6)


def test_batched_one_lane_is_scalar_form(oracle):
    """With one lane the lockstep engine is the per-sim engine (up to the FMA
    the batched kernels use): same step counts, momenta to rounding."""
    rc, geom = oracle.sample(99, 5, 6, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-12, stop_rule=1, t_max=1e4, momentum_tol=1e-7)
    bt = oracle.batched_run(CH4_MASSES, CH4_CHARGES, geom, cfg, lanes=1, precision="f64")
    sc = oracle.scalar_run(CH4_MASSES, CH4_CHARGES, geom, cfg)
    np.testing.assert_array_equal(bt.steps, sc.steps)
    assert per_atom_rel(bt.momenta, sc.momenta).max() < 1e-10


def test_cpu_baseline_matches_lockstep_restatement(oracle):
    import ctypes as C
    rc, geom = oracle.sample(49188291, 5, 64, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-6, atol=1e-9, stop_rule=1, t_max=1e4, momentum_tol=1e-7,
                          h0_mode=1, h0=1e-3)
    mh = ffi._MolHold(CH4_MASSES, CH4_CHARGES)
    mom = np.zeros((64, 5, 3))
    f = oracle.lib.ceo_cpu_baseline_momenta
    f.argtypes = [C.c_void_p] * 3 + [C.c_size_t, C.c_void_p]
    lanes = f(C.byref(mh.s), C.byref(cfg), geom.ctypes.data, 64, mom.ctypes.data)
    assert lanes in (4, 8, 16)
    bt = oracle.batched_run(CH4_MASSES, CH4_CHARGES, geom, cfg, lanes=lanes)
    assert per_atom_rel(mom, bt.momenta).max() < 1e-5
    r = oracle.cpu_baseline(CH4_MASSES, CH4_CHARGES, geom, cfg, 2, 1)
    assert r["lanes"] == lanes and r["median_s"] > 0


# ------------------------------------------------------------ samplers ----
def test_philox_known_answers(oracle):
    with open(os.path.join(GOLDEN, "philox_kat.json")) as f:
        doc = json.load(f)
    for v in doc["vectors"]:
        out = oracle.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [format(int(x), "08x") for x in out] == v["out"]


def test_device_sampler_properties(oracle):
    """sampler properties of tests/test_sampler.cpp:40-112 for the counter-based
    sampler, plus uniformity in the ball and split-invariance."""
    n, R, ms = 5, 2.5, 1.0
    pos, st, fails = oracle.device_sample(7, 0, 20000, n, R, ms)
    assert fails == 0 and (st == 0).all()
    r = np.linalg.norm(pos, axis=-1)
    assert r.max() <= R * (1 + 1e-7)
    d = np.linalg.norm(pos[:, :, None, :] - pos[:, None, :, :], axis=-1)
    d[:, np.arange(n), np.arange(n)] = np.inf
    assert d.min() >= ms * (1 - 1e-7)
    # first atom is uniform in the ball: (r/R)^3 ~ U(0,1)
    u = np.sort((r[:, 0] / R) ** 3)
    ks = np.max(np.abs(u - (np.arange(1, u.size + 1) / u.size)))
    assert ks < 1.63 / np.sqrt(u.size)  # KS at the 1% level
    for c in range(3):  # isotropy of the first atom
        assert abs(pos[:, 0, c].mean()) < 4 * R / np.sqrt(5 * u.size)
    # results depend only on (seed, global index): any split gives the same sims
    a, _, _ = oracle.device_sample(7, 0, 100, n, R, ms)
    b, _, _ = oracle.device_sample(7, 60, 40, n, R, ms)
    np.testing.assert_array_equal(a[60:], b)
    c_, _, _ = oracle.device_sample(8, 0, 100, n, R, ms)
    assert not np.array_equal(a, c_)
    # infeasible packing fails per sim instead of aborting
    _, st, fails = oracle.device_sample(1, 0, 10, 10, 1.0, 1.5, max_attempts=50)
    assert fails == 10 and (st == ffi.SAMPLER_FAIL).all()
