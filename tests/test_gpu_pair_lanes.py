"""The pair-lane integrator (f32, N = 8 with equal or per-sim charges: one sim over
two lanes) on the paths the headline rule does not take: the reference's PE rule
with the Hairer start, a stop at max_steps resumed from the saved state, the
step-size trace, per-sim charges with the energy diagnostic, and sampler
failures.  References: the oracle's per-lane form (f32, <= 1e-5 per atom for
>= 99% of sims; an 8-atom f32 trajectory is only good to ~1e-4 on an odd sim)."""
from dataclasses import replace

import numpy as np
import pytest

from conftest import CH4_MASSES, per_atom_rel

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
C2H6_MASSES = [CH4_MASSES[0]] * 2 + [CH4_MASSES[1]] * 6
R8, SEP8 = 8 ** (1 / 3) + 1.5, 0.9


def c2h6(ce, q=1.0):
    return ce.Molecule([ce.Atom("X", m, q) for m in C2H6_MASSES])


def f32_close(res_momenta, ref_momenta):
    err = per_atom_rel(res_momenta, ref_momenta)
    return np.mean(err < 1e-5) >= 0.99 and err.max() < 1e-4, err.max()


def test_pe_rule_and_hairer_start_vs_oracle(gpu, oracle):
    from oracle import ffi
    ce = gpu
    _, geom = oracle.sample(81, 8, 300, R8, SEP8)
    eng = ce.BatchedRK45(c2h6(ce), 1.0, 300, 1e-6, 1e-9)
    eng.set_geometries(geom)
    # the reference's default threshold: at PE = 1e-9 PE0 a step more or less
    # changes the momenta by far less than the f32 bar (at 1e-4 it would not)
    cfg = dict(stop_rule=0, pe_stop_fraction=1e-9)
    res = eng.run(ce.RunConfig(**cfg))
    ref = oracle.lane_run(C2H6_MASSES, [1.0] * 8, geom,
                          ffi.make_config(rtol=1e-6, atol=1e-9, **cfg), precision="f32")
    ok, worst = f32_close(res.momenta, ref.momenta)
    assert ok, worst
    assert np.mean(res.status == ref.status) > 0.99
    assert np.all(res.status == ce.SIM_OK_PE)


def test_max_steps_stop_then_resume_gives_the_uninterrupted_bits(gpu):
    """A stop at max_steps saves x, v, h, t, E0 exactly (after an accepted step
    nothing else carries over), so resuming reproduces the uninterrupted run."""
    ce = gpu
    eng = ce.BatchedRK45(c2h6(ce), 1.0, 2000, 1e-6, 1e-9)
    eng.sample(17, 2000, R8, SEP8)
    eng.enable_state()
    full = eng.run(ce.RunConfig(**RULE))
    cut = eng.run(ce.RunConfig(**dict(RULE, max_steps=12)))
    assert np.all(cut.status == ce.SIM_NOT_CONVERGED) and np.all(cut.steps == 12)
    res = eng.run(replace(ce.RunConfig(**RULE), resume=True))
    np.testing.assert_array_equal(res.steps, full.steps)
    np.testing.assert_array_equal(res.t_final, full.t_final)
    np.testing.assert_array_equal(res.momenta, full.momenta)


def test_step_trace(gpu):
    """The trace holds each accepted step's size (written by one lane of the
    pair): positive for the first `steps` entries, zero after, summing to t."""
    ce = gpu
    eng = ce.BatchedRK45(c2h6(ce), 1.0, 500, 1e-6, 1e-9)
    eng.sample(23, 500, R8, SEP8)
    eng.enable_trace(96)
    res = eng.run(ce.RunConfig(**RULE))
    tr = res.trace_h.reshape(500, 96)
    assert np.all(res.steps <= 96)
    for s in range(500):
        n = res.steps[s]
        assert np.all(tr[s, :n] > 0) and np.all(tr[s, n:] == 0)
    # the last step lands on t_max exactly; the sum of the sizes agrees to rounding
    np.testing.assert_allclose(tr.sum(axis=1), res.t_final, rtol=1e-6)


def test_per_sim_charges_and_energy_diagnostic_vs_oracle(gpu, oracle):
    from oracle import ffi
    ce = gpu
    sp = [ce.Species(c2h6(ce), 800, R8, SEP8, charge_levels=2, max_total_charge=16)]
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.enable_energy_diagnostic()
    mb.sample(seed=29)
    r = mb.run(ce.RunConfig(**RULE))[0]
    g, q = mb.geometries()[0], mb.charges()[0]
    ref = oracle.lane_run_q(np.array(C2H6_MASSES), q, g,
                            ffi.make_config(rtol=1e-6, atol=1e-9, **RULE), precision="f32")
    ok, worst = f32_close(r["momenta"], ref.momenta)
    assert ok, worst
    assert np.mean(r["status"] == ref.status) > 0.99
    de = np.abs(r["energy_err"] - ref.energy_err)
    assert np.median(de) < 1e-6 and de.max() < 2e-5, (np.median(de), de.max())
    st = mb.stats()
    assert st["lane_slots"] % 32 == 0 and st["attempted_steps"] <= st["lane_slots"]


def test_sampler_failures_end_both_lanes(gpu):
    ce = gpu
    eng = ce.BatchedRK45(c2h6(ce), 1.0, 40, 1e-6, 1e-9)
    eng.sample(3, 40, 1.0, 1.9, max_attempts=200)
    res = eng.run(ce.RunConfig(**RULE))
    assert np.all(res.status == ce.SIM_SAMPLER_FAIL) and np.all(res.steps == 0)
    np.testing.assert_array_equal(res.momenta, 0.0)
