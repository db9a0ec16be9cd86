"""The momentum-or-t_max rule at the time limit, and the oracle's cut-and-resume
statement (ceo_lane_run_cut), on the CPU.

* A step shortened to land on the time limit ends the sim by the limit: the
  momentum-change test is not applied to it.  Before that rule, sims whose last
  step before t_cut = 200 was clamped to a few 1e-5 a.u. passed the test with a
  change of ~1e-8 and stopped "settled" at t = 200, far from their asymptote
  (8 of 8M CH4 sims of the closure workload, config-norm error up to 0.12); they
  are the regression cases here.
* The continuation is pinned against the uninterrupted run: with t_cut placed
  exactly where the uninterrupted run's k-th step ends (a step after which the
  controller grew the step), cutting there and continuing with the step size
  the controller proposed after that step reproduces the uninterrupted run (same
  step count and time, <= 1e-12 in f64), which is itself pinned to the
  reference engine.  Continuing with the clamped step size instead (what the
  reference integrator would keep after a capped step) changes the step
  sequence and fails this.
* A cut a sliver after a step boundary clamps the last step to a sliver; the
  continuation starts from at least the size that step was shortened from, so
  it does not crawl back from the sliver with sliver-sized momentum changes
  (which pass the momentum test: sims stopped "settled" just after the cut).
"""
import numpy as np
import pytest

from conftest import CH4_CHARGES, CH4_MASSES, config_rel, per_atom_rel
from oracle import ffi

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
# CH4 sims (device sampler, seed 49188293, R 2.5, min sep 1.0) whose step before
# t = 200 is clamped to a few 1e-5 a.u.; the last one is an ordinary sim.
CLAMP_SEED = 49188293
CLAMP_SIMS = [427138, 1488243, 1789783, 2173943, 3672532, 5808755, 6845453, 6231997]


def _geoms(oracle):
    return np.concatenate([oracle.device_sample(CLAMP_SEED, i, 1, 5, 2.5, 1.0)[0]
                           for i in CLAMP_SIMS])


@pytest.mark.parametrize("form", ["lane_f32", "lane_f64", "scalar_f32"])
def test_step_landing_on_time_limit_is_not_a_momentum_stop(oracle, form):
    g = _geoms(oracle)
    prec = form.split("_")[1]
    cfg = ffi.make_config(rtol=1e-6, atol=1e-9, **dict(RULE, t_max=200.0))
    run = (oracle.lane_run(CH4_MASSES, CH4_CHARGES, g, cfg, precision=prec, trace_cap=64)
           if form.startswith("lane") else
           oracle.scalar_run(CH4_MASSES, CH4_CHARGES, g, cfg, precision=prec, trace_cap=64))
    assert (run.status == ffi.OK_TMAX).all(), run.status
    assert (run.t_final == 200.0).all()
    last = run.trace_h[np.arange(len(g)), run.steps - 1]
    if form == "lane_f32":
        assert (last[:7] < 1e-3).all()  # the clamped steps are tiny here


def test_cut_and_resume_agrees_with_uninterrupted_on_clamped_sims(oracle):
    g = _geoms(oracle)
    cfg = ffi.make_config(rtol=1e-6, atol=1e-9, **RULE)
    full = oracle.lane_run(CH4_MASSES, CH4_CHARGES, g, cfg, precision="f32")
    cut = oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g, cfg, 200.0, precision="f32")
    assert (cut.status == ffi.OK_TMAX).all() and (cut.t_final == 1e4).all()
    assert config_rel(cut.momenta, full.momenta).max() < 1e-6
    assert per_atom_rel(cut.momenta, full.momenta).max() < 2e-5


def test_cut_at_a_step_boundary_reproduces_uninterrupted_run(oracle):
    g, _, _ = oracle.device_sample(5150, 0, 24, 5, 2.5, 1.0)
    cfg = ffi.make_config(rtol=1e-8, atol=1e-12, **RULE)
    full = oracle.lane_run(CH4_MASSES, CH4_CHARGES, g, cfg, precision="f64", trace_cap=256)
    checked = 0
    for s in range(len(g)):
        if full.attempts[s] != full.steps[s]:
            continue  # a rejected proposal near the cut would be clamped differently
        k = int(full.steps[s]) // 2
        while k < full.steps[s] - 2 and not full.trace_h[s, k] > full.trace_h[s, k - 1]:
            k += 1  # a step after which the controller grew the step (f > 1)
        t_cut = 0.0
        for h in full.trace_h[s, :k]:  # the engine's FP64 time, summed step by step
            t_cut += h
        one = oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g[s:s + 1], cfg, t_cut,
                                  precision="f64")
        assert one.steps[0] == full.steps[s] and one.t_final[0] == full.t_final[s]
        assert per_atom_rel(one.momenta, full.momenta[s:s + 1]).max() < 1e-12
        # a cut inside a step continues on another step sequence: equal to the
        # tolerance only
        mid = oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g[s:s + 1], cfg,
                                  t_cut - 0.4 * full.trace_h[s, k - 1], precision="f64")
        d = config_rel(mid.momenta, full.momenta[s:s + 1]).max()
        assert 1e-14 < d < 1e-6
        checked += 1
    assert checked >= 12


def test_cut_rejects_other_rules(oracle):
    g, _, _ = oracle.device_sample(1, 0, 2, 5, 2.5, 1.0)
    with pytest.raises(RuntimeError):
        oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g, ffi.make_config(), 200.0)
    with pytest.raises(RuntimeError):
        oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g, ffi.make_config(**RULE), 0.0)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_cut_just_after_a_step_boundary_continues_normally(oracle, precision):
    """t_cut a sliver after the end of step k (1e-6 a.u.; 3e-5 in f32, whose
    step cap rounds to ~2e-6 at these step sizes): the cut's last step is clamped
    to the sliver; the continuation starts from the size that step was shortened
    from and ends by the time limit, agreeing with the uninterrupted run to the
    tolerance."""
    g, _, _ = oracle.device_sample(5151, 0, 16, 5, 2.5, 1.0)
    rtol, atol = (1e-6, 1e-9) if precision == "f32" else (1e-8, 1e-12)
    cfg = ffi.make_config(rtol=rtol, atol=atol, **RULE)
    full = oracle.lane_run(CH4_MASSES, CH4_CHARGES, g, cfg, precision=precision, trace_cap=256)
    for s in range(len(g)):
        k = int(full.steps[s]) // 2
        t_cut = 0.0
        for h in full.trace_h[s, :k]:
            t_cut += h
        t_cut += 1e-6 if precision == "f64" else 3e-5
        cut = oracle.lane_run(CH4_MASSES, CH4_CHARGES, g[s:s + 1],
                              ffi.make_config(rtol=rtol, atol=atol, **dict(RULE, t_max=t_cut)),
                              precision=precision, trace_cap=256)
        assert cut.status[0] == ffi.OK_TMAX
        assert cut.trace_h[0, cut.steps[0] - 1] < 1e-4  # the sliver
        one = oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, g[s:s + 1], cfg, t_cut,
                                  precision=precision, trace_cap=256)
        assert one.status[0] == ffi.OK_TMAX and one.t_final[0] == 1e4
        assert one.trace_h[0, cut.steps[0]] > 1.0  # no crawl back from the sliver
        assert config_rel(one.momenta, full.momenta[s:s + 1]).max() < 1e-6
