"""The step trace (ce_engine_enable_trace: the reference's StepObserver hook,
include/coulomb/driver.hpp:12-18 / src/driver.cpp:25, as data) against the
oracle's lane_run(trace_cap=...), one lane per sim (N = 5) and the pair-lane
form (N = 8).

* f64: the same step counts as the oracle's per-lane form and the same step
  sizes to rounding (the controller turns the ~1e-16 rounding differences of the
  embedded error estimate into ~1e-8 relative step-size differences, up to ~1e-4
  on an odd step); the first step is h0 exactly.
* f32: the early error estimates are rounding noise, so no other arithmetic
  takes the same step sequence (nor does the oracle); the trace is checked as
  a record of the steps taken: a rerun stopped after k accepted steps ends at
  t = sum of the first k entries, bit for bit (the FP64 time is that sum, in
  that order), for several k; entries past the sim's last step are 0.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
C_, H_ = 21874.66, 1837.15
FORMS = {"ch4": [C_] + [H_] * 4, "c2h6": [C_] * 2 + [H_] * 6}  # f32 C2H6: pair lanes
CAP = 256


def _engine(ce, masses, n_sims, rtol, atol, seed):
    n = len(masses)
    eng = ce.BatchedRK45(ce.Molecule([ce.Atom("X", m, 1.0) for m in masses]), 1.0, n_sims,
                         rtol, atol)
    eng.sample(seed, n_sims, n ** (1 / 3) + 1.5, 0.9)
    eng.enable_trace(CAP)
    return eng


@pytest.mark.parametrize("form", list(FORMS))
def test_f64_trace_vs_oracle(gpu, oracle, form):
    ce = gpu
    from oracle import ffi
    masses = FORMS[form]
    n, ns = len(masses), 1500
    eng = _engine(ce, masses, ns, 1e-8, 1e-12, 5)
    res = eng.run(ce.RunConfig(precision=ce.PRECISION_F64, **RULE))
    tr = res.trace_h.reshape(ns, CAP)
    ref = oracle.lane_run(masses, [1.0] * n, eng.geometries(),
                          ffi.make_config(rtol=1e-8, atol=1e-12, **RULE), precision="f64",
                          trace_cap=CAP)
    np.testing.assert_array_equal(res.steps, ref.steps)
    assert (res.steps < CAP).all()
    live = np.arange(CAP)[None, :] < res.steps[:, None]
    assert (tr[~live] == 0).all() and (tr[live] > 0).all()
    np.testing.assert_array_equal(tr[:, 0], 1e-3)
    d = np.abs(tr[live] / ref.trace_h[live] - 1)
    assert d.max() < 1e-3 and np.median(d) < 1e-6, (d.max(), np.median(d))
    np.testing.assert_allclose(tr.sum(axis=1), res.t_final, rtol=1e-14)


@pytest.mark.parametrize("form", list(FORMS))
def test_f32_trace_records_the_steps_taken(gpu, form):
    ce = gpu
    masses = FORMS[form]
    ns = 1000
    eng = _engine(ce, masses, ns, 1e-6, 1e-9, 6)
    full = eng.run(ce.RunConfig(**RULE))
    tr = full.trace_h.reshape(ns, CAP).copy()
    assert (full.steps < CAP).all()
    live = np.arange(CAP)[None, :] < full.steps[:, None]
    assert (tr[~live] == 0).all() and (tr[live] > 0).all()
    np.testing.assert_array_equal(tr[:, 0], np.float64(np.float32(1e-3)))
    for k in (1, 2, 7, 16, 25):
        part = eng.run(ce.RunConfig(**dict(RULE, max_steps=k)))
        stopped = part.status == ce.SIM_NOT_CONVERGED
        assert stopped.mean() > 0.9
        t = np.zeros(ns)
        for j in range(k):  # the engine's FP64 time: the accepted steps summed in order
            t = t + tr[:, j]
        np.testing.assert_array_equal(part.t_final[stopped], t[stopped])
        np.testing.assert_array_equal(part.trace_h.reshape(ns, CAP)[stopped, :k], tr[stopped, :k])
    # a shorter trace holds the first entries only
    eng.enable_trace(8)
    short = eng.run(ce.RunConfig(**RULE))
    np.testing.assert_array_equal(short.trace_h.reshape(ns, 8), tr[:, :8])
    np.testing.assert_array_equal(short.momenta, full.momenta)
