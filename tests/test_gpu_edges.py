"""Edge cases through the C-ABI: empty batches, single atoms, closed forms on the
GPU itself, infeasible sampling, determinism and the capacity boundary."""
import numpy as np
import pytest

from conftest import CH4_MASSES

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)


def test_empty_batch(gpu):
    ce = gpu
    eng = ce.BatchedRK45(ce.Molecule([ce.Atom("H", 1837.15, 1.0)] * 3), 1.0, 16, 1e-6, 1e-9)
    eng.sample(1, 0, 2.5, 1.0)
    res = eng.run(ce.RunConfig(**RULE))
    assert res.momenta.size == 0 and res.steps.size == 0 and res.batch_steps == 0
    assert eng.stats()["sims"] == 0 and eng.stats()["attempted_steps"] == 0


def test_single_atom(gpu):
    """No pairs: zero force, zero momentum; the PE rule stops at once (PE0 = 0),
    the momentum rule runs to t_max."""
    ce = gpu
    eng = ce.BatchedRK45(ce.Molecule([ce.Atom("C", 21874.66, 1.0)]), 1.0, 8, 1e-6, 1e-9)
    eng.set_geometries(np.zeros((8, 1, 3)) + 0.5)
    r = eng.run(ce.RunConfig(stop_rule=0, pe_stop_fraction=1e-5))
    assert np.all(r.status == ce.SIM_OK_PE) and np.all(r.steps == 0)
    np.testing.assert_array_equal(r.momenta, 0.0)
    r = eng.run(ce.RunConfig(**RULE))
    assert np.all(r.status == ce.SIM_OK_TMAX) and np.all(r.t_final == 1e4)
    np.testing.assert_array_equal(r.momenta, 0.0)


@pytest.mark.parametrize("precision,tol", [(0, 2e-6), (1, 1e-9)])
def test_two_body_asymptote_closed_form(gpu, precision, tol):
    """|p_inf| = sqrt(2 mu q1 q2 / r0), back to back along the axis (energy
    conservation; the reference's two-proton case, test_reference.cpp:119-154)."""
    ce = gpu
    m1, m2, q1, q2 = 1837.15, 25526.04, 1.0, 2.0
    mol = ce.Molecule([ce.Atom("A", m1, q1), ce.Atom("B", m2, q2)])
    rng = np.random.default_rng(3)
    n = 64
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    r0 = rng.uniform(0.8, 3.0, size=n)
    geo = np.stack([axis * r0[:, None] * 0.5, -axis * r0[:, None] * 0.5], axis=1)
    rtol, atol = (1e-6, 1e-9) if precision == 0 else (1e-10, 1e-14)
    eng = ce.BatchedRK45(mol, 1.0, n, rtol, atol)
    eng.set_geometries(geo)
    res = eng.run(ce.RunConfig(stop_rule=0, pe_stop_fraction=1e-9, precision=precision))
    p = res.momenta.reshape(n, 2, 3)
    mu = m1 * m2 / (m1 + m2)
    pinf = np.sqrt(2 * mu * q1 * q2 / r0)
    np.testing.assert_allclose(np.linalg.norm(p[:, 0], axis=1), pinf, rtol=tol)
    np.testing.assert_allclose(p[:, 0], axis * pinf[:, None], rtol=0, atol=tol * pinf.max() * 4)
    np.testing.assert_allclose(p[:, 0] + p[:, 1], 0.0, atol=tol * pinf.max())


def test_symmetric_ch4_closed_form(gpu):
    """C at the centre of a regular H tetrahedron, unit charges: p_C = 0 and
    |p_H| = sqrt(m_H E0 / 2) with E0 = 4 / r_CH + 6 / r_HH."""
    ce = gpu
    mC, mH = CH4_MASSES[0], CH4_MASSES[1]
    mol = ce.Molecule([ce.Atom("C", mC, 1.0)] + [ce.Atom("H", mH, 1.0)] * 4)
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3.0)
    r_ch = 2.05
    geo = np.concatenate([np.zeros((1, 3)), tet * r_ch])[None].repeat(4, axis=0)
    e0 = 4 / r_ch + 6 / (r_ch * np.sqrt(8.0 / 3.0))
    eng = ce.BatchedRK45(mol, 1.0, 4, 1e-10, 1e-14)
    eng.set_geometries(geo)
    res = eng.run(ce.RunConfig(stop_rule=0, pe_stop_fraction=1e-9, precision=1))
    p = res.momenta.reshape(4, 5, 3)
    assert np.abs(p[:, 0]).max() < 1e-9 * np.sqrt(mH * e0)
    np.testing.assert_allclose(np.linalg.norm(p[:, 1:], axis=-1), np.sqrt(mH * e0 / 2), rtol=1e-9)


def test_infeasible_sampling_is_flagged_not_fatal(gpu):
    """Atoms that cannot be placed (min separation too large for the ball) mark
    the sim SAMPLER_FAIL (the reference throws, sampler.cpp:50-54); the batch runs."""
    ce = gpu
    eng = ce.BatchedRK45(ce.Molecule([ce.Atom("H", 1837.15, 1.0)] * 5), 1.0, 32, 1e-6, 1e-9)
    eng.sample(3, 32, 1.0, 1.9, max_attempts=200)
    res = eng.run(ce.RunConfig(**RULE))
    assert np.all(res.status == ce.SIM_SAMPLER_FAIL) and np.all(res.converged == 0)
    assert np.all(res.steps == 0)


def test_repeat_runs_are_bit_identical_and_fill_capacity(gpu):
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] + [ce.Atom("H", CH4_MASSES[1], 1.0)] * 4)
    n = 5000
    eng = ce.BatchedRK45(mol, 1.0, n, 1e-6, 1e-9)
    eng.sample(99, n, 2.5, 1.0)
    a = eng.run(ce.RunConfig(**RULE))
    eng.set_launch(block_threads=64, blocks_per_cu=1, refill_threshold=1, chunk=64)
    b = eng.run(ce.RunConfig(**RULE))  # other launch shape: per-sim results do not change
    np.testing.assert_array_equal(a.momenta, b.momenta)
    np.testing.assert_array_equal(a.steps, b.steps)
    with pytest.raises(ValueError):
        eng.sample(99, n + 1, 2.5, 1.0)


def test_energy_diagnostic_is_opt_in(gpu):
    """Off by default (the reference's Result has no energy error): asking for it
    is an error; switched on, the same integration also yields the FP64
    diagnostic, and the momenta do not change."""
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] + [ce.Atom("H", CH4_MASSES[1], 1.0)] * 4)
    eng = ce.BatchedRK45(mol, 1.0, 2000, 1e-6, 1e-9)
    eng.sample(5, 2000, 2.5, 1.0)
    a = eng.run(ce.RunConfig(**RULE))
    assert a.energy_err is None
    ee = np.zeros(2000)
    res = ce._Results(None, None, None, None, None, None,
                      ee.ctypes.data_as(ce.C.POINTER(ce.C.c_double)), None, 0, 0)
    with pytest.raises(ValueError):
        ce._check(ce.library().ce_engine_download(eng._h, ce.C.byref(res)))
    eng.enable_energy_diagnostic()
    b = eng.run(ce.RunConfig(**RULE))
    np.testing.assert_array_equal(a.momenta, b.momenta)
    np.testing.assert_array_equal(a.steps, b.steps)
    assert np.all(np.isfinite(b.energy_err)) and 0 < np.abs(b.energy_err).max() < 1e-4
    eng.enable_energy_diagnostic(False)
    assert eng.run(ce.RunConfig(**RULE)).energy_err is None


def test_lane_slot_accounting(gpu):
    """stats()["lane_slots"] is 64 x the waves' loop iterations (32 where a sim
    spans two lanes): a multiple of 64 (32), at least the attempted steps (one
    attempt per running sim per iteration), and for a large batch mostly used
    (refill keeps the lanes busy)."""
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] + [ce.Atom("H", CH4_MASSES[1], 1.0)] * 4)
    eng = ce.BatchedRK45(mol, 1.0, 1000000, 1e-6, 1e-9)
    eng.sample(8, 1000000, 2.5, 1.0)
    eng.run(ce.RunConfig(**RULE), download=False)
    st = eng.stats()
    assert st["lane_slots"] % 64 == 0
    assert st["attempted_steps"] <= st["lane_slots"]
    assert st["attempted_steps"] > 0.5 * st["lane_slots"]
    # C2H6 (N = 8): one sim per lane pair, 32 seats per wave iteration
    mol8 = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] * 2 +
                       [ce.Atom("H", CH4_MASSES[1], 1.0)] * 6)
    eng8 = ce.BatchedRK45(mol8, 1.0, 200000, 1e-6, 1e-9)
    eng8.sample(8, 200000, 3.5, 0.9)
    eng8.run(ce.RunConfig(**RULE), download=False)
    st8 = eng8.stats()
    assert st8["lane_slots"] % 32 == 0
    assert st8["attempted_steps"] <= st8["lane_slots"]
    assert st8["attempted_steps"] > 0.5 * st8["lane_slots"]


def test_argument_checks_fail_cleanly(gpu):
    """Sizes that overflow the buffer arithmetic, a negative momentum tolerance
    and an output the last run did not produce are rejected with
    invalid_argument; the engine keeps working afterwards, and the rejected
    download queued no copy (the other output buffers stay untouched)."""
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] + [ce.Atom("H", CH4_MASSES[1], 1.0)] * 4)
    with pytest.raises(ValueError, match="overflow"):
        ce.BatchedRK45(mol, 1.0, 2**62, 1e-6, 1e-9)
    eng = ce.BatchedRK45(mol, 1.0, 500, 1e-6, 1e-9)
    eng.sample(3, 500, 2.5, 1.0)
    with pytest.raises(ValueError, match="momentum_tol"):
        eng.run(ce.RunConfig(**dict(RULE, momentum_tol=-1e-7)))
    a = eng.run(ce.RunConfig(**RULE))
    steps = np.full(500, 7, np.uint32)
    ee = np.zeros(500)
    P = ce.C.POINTER
    res = ce._Results(None, None, steps.ctypes.data_as(P(ce.C.c_uint32)), None, None, None,
                      ee.ctypes.data_as(P(ce.C.c_double)), None, 0, 0)
    with pytest.raises(ValueError, match="energy_err"):
        ce._check(ce.library().ce_engine_download(eng._h, ce.C.byref(res)))
    assert np.all(steps == 7)
    b = eng.download()
    np.testing.assert_array_equal(a.momenta, b.momenta)
    np.testing.assert_array_equal(a.steps, b.steps)


def test_a_failed_allocation_does_not_fail_later_launches(gpu):
    """HIP keeps a thread's last error until it is read: an engine whose device
    buffers do not fit (2^40 CH4 sims) fails with the HIP error, and the
    launches of engines made afterwards (device sampler, integrator, records,
    the mixed-size kernel and its sampler) still succeed, with the same results
    as before the failure."""
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] + [ce.Atom("H", CH4_MASSES[1], 1.0)] * 4)
    eng = ce.BatchedRK45(mol, 1.0, 800, 1e-6, 1e-9)
    eng.sample(3, 800, 2.5, 1.0)
    before = eng.run(ce.RunConfig(**RULE))
    with pytest.raises(RuntimeError, match="hipMalloc"):
        ce.BatchedRK45(mol, 1.0, 2**40, 1e-6, 1e-9)
    eng2 = ce.BatchedRK45(mol, 1.0, 800, 1e-6, 1e-9)
    eng2.sample(3, 800, 2.5, 1.0)
    after = eng2.run(ce.RunConfig(**RULE))
    np.testing.assert_array_equal(after.momenta, before.momenta)
    assert len(eng2.records(0, 800)) == 800
    with pytest.raises(RuntimeError, match="hipMalloc"):
        ce.BatchedRK45(mol, 1.0, 2**40, 1e-6, 1e-9)
    mixed = ce.MixedBatch([ce.Species(mol, 600, 3.2, 0.9, charge_levels=2, max_total_charge=10)],
                          1.0, 1e-6, 1e-9)
    mixed.sample(5)
    (res,) = mixed.run(ce.RunConfig(**RULE))
    assert (res["status"] <= ce.SIM_OK_TMAX).all()


def test_forced_failures_end_sims_with_their_status(gpu, oracle):
    """Failures end a sim with a status, never the run, as the oracle's lane
    statement has them: max_steps reached (not converged, after exactly
    max_steps steps), a rejection storm (a fixed h0 of 1e3 a.u. rejected past
    max_rejections: limit after max_rejections + 1 attempts), and a non-finite
    state (two coincident atoms: every attempt rejected, then non-finite)."""
    from oracle import ffi
    ce = gpu
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in CH4_MASSES])
    eng = ce.BatchedRK45(mol, 1.0, 64, 1e-6, 1e-9)
    eng.sample(5, 64, 2.5, 1.0)
    geom = eng.geometries()
    coincident = geom.copy()
    coincident[:, 1] = coincident[:, 0]
    cases = [("max_steps", dict(max_steps=5), geom, ce.SIM_NOT_CONVERGED, 5, 5),
             ("storm", dict(h0=1e3, max_rejections=2), geom, ce.SIM_REJECT_LIMIT, 0, 3),
             ("coincident", {}, coincident, ce.SIM_NONFINITE, 0, 101)]
    for name, kw, g, status, steps, attempts in cases:
        eng.set_geometries(g)
        r = eng.run(ce.RunConfig(**dict(RULE, **kw)))
        cfg = ffi.make_config(rtol=1e-6, atol=1e-9, **dict(RULE, **kw))
        o = oracle.lane_run(CH4_MASSES, [1.0] * 5, g, cfg, precision="f32")
        assert np.all(r.status == status), name
        assert np.all(r.steps == steps) and np.all(r.attempts == attempts), name
        assert np.all(r.converged == 0), name
        np.testing.assert_array_equal(r.status, o.status)
        np.testing.assert_array_equal(r.attempts, o.attempts)


@pytest.mark.parametrize("per_sim_charges", [False, True])
@pytest.mark.parametrize("h0", [20.0, 100.0])
def test_first_force_of_a_fresh_sim_under_first_step_rejections(gpu, oracle, h0,
                                                                 per_sim_charges):
    """A fresh CH4 sim takes its first force from the second stage of its first
    attempt (it starts at rest, so that stage sits at x0).  With a fixed h0 large
    enough to be rejected once or more, that attempt is a rejected one, with the
    molecule's charges and with per-sim charges: the f64 engine still takes the
    oracle's steps and attempts and agrees to <= 1e-12, and f32 to <= 1e-5."""
    ce = gpu
    from oracle import ffi
    n = 1500
    rule = dict(RULE, h0=h0)
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in CH4_MASSES])
    for precision, rtol, atol in ((1, 1e-10, 1e-14), (0, 1e-6, 1e-9)):
        if per_sim_charges:
            eng = ce.MixedBatch([ce.Species(mol, n, 2.5, 1.0, charge_levels=2,
                                            max_total_charge=10)], 1.0, rtol, atol)
            eng.sample(seed=61)
            g, q = eng.geometries()[0], eng.charges()[0]
            r = eng.run(ce.RunConfig(precision=precision, **rule))[0]
            mom, steps, att = r["momenta"], r["steps"], r["attempts"]
        else:
            eng = ce.BatchedRK45(mol, 1.0, n, rtol, atol)
            eng.sample(61, n, 2.5, 1.0)
            g, q = eng.geometries(), np.ones((n, 5))
            r = eng.run(ce.RunConfig(precision=precision, **rule))
            mom, steps, att = r.momenta.reshape(n, 5, 3), r.steps, r.attempts
        ref = oracle.lane_run_q(CH4_MASSES, q, g, ffi.make_config(rtol=rtol, atol=atol, **rule),
                                precision="f64" if precision == 1 else "f32")
        assert np.mean(att > steps) > 0.9  # (nearly) every sim rejects its first attempt
        err = (np.linalg.norm(mom - ref.momenta, axis=-1) /
               np.linalg.norm(ref.momenta, axis=-1)).max(axis=-1)
        if precision == 1:
            np.testing.assert_array_equal(steps, ref.steps)
            np.testing.assert_array_equal(att, ref.attempts)
            assert err.max() < 1e-12
        else:
            assert err.max() < 1e-5
