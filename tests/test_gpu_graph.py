"""A batch captured as a HIP graph (sampler + counter reset + integrator) gives
the eager path's bits on every replay, with eager work between replays."""
import numpy as np
import pytest

from conftest import CH4_CHARGES, CH4_MASSES

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)


def ch4(ce):
    return ce.Molecule([ce.Atom("X", m, q) for m, q in zip(CH4_MASSES, CH4_CHARGES)])


def eager(ce, seed, n, first, cfg):
    eng = ce.BatchedRK45(ch4(ce), 1.0, n, 1e-6, 1e-9)
    eng.sample(seed, n, 2.5, 1.0, first_index=first)
    return eng.run(cfg), eng.stats()


def same(a, b):
    np.testing.assert_array_equal(a.momenta, b.momenta)
    np.testing.assert_array_equal(a.steps, b.steps)
    np.testing.assert_array_equal(a.attempts, b.attempts)
    np.testing.assert_array_equal(a.status, b.status)


def test_replays_match_eager(gpu):
    ce = gpu
    cfg = ce.RunConfig(**RULE)
    eng = ce.BatchedRK45(ch4(ce), 1.0, 4096, 1e-6, 1e-9)
    eng.capture(cfg, seed=11, n_sims=3000, radius=2.5, min_separation=1.0)
    firsts = [0, 5000, 123456, 0, 5000]
    seen = {}
    for i, first in enumerate(firsts):
        r = eng.replay(first)
        st = eng.stats()
        ref, ref_st = eager(ce, 11, 3000, first, cfg)
        same(r, ref)
        assert st["attempted_steps"] == ref_st["attempted_steps"]
        assert st["sims"] == 3000
        if first in seen:
            same(r, seen[first])
        seen[first] = r
        # eager work on the same engine between replays
        eng.sample(99, 1000 + 500 * i, 2.5, 1.0)
        eng.run(ce.RunConfig(**dict(RULE, t_max=300.0)))


def test_capture_needs_single_species_and_capacity(gpu):
    ce = gpu
    cfg = ce.RunConfig(**RULE)
    eng = ce.BatchedRK45(ch4(ce), 1.0, 100, 1e-6, 1e-9)
    with pytest.raises(ValueError):
        eng.capture(cfg, 1, 101, 2.5, 1.0)
    with pytest.raises(RuntimeError):
        ce._check(ce.library().ce_engine_replay(eng._h, 0, None))
