"""Saved state / resume, sim subsets and the learned closure, through the C-ABI."""
from dataclasses import replace

import numpy as np
import pytest

from conftest import CH4_CHARGES, CH4_MASSES, per_atom_rel
from oracle import closure_ref

pytestmark = pytest.mark.gpu

RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)
T_CUT = 200.0


def ch4(ce):
    return ce.Molecule([ce.Atom("X", m, q) for m, q in zip(CH4_MASSES, CH4_CHARGES)])


def engine(ce, n, seed, state=True):
    eng = ce.BatchedRK45(ch4(ce), 1.0, n, 1e-6, 1e-9)
    eng.sample(seed, n, 2.5, 1.0)
    if state:
        eng.enable_state()
    return eng


def pe(x, q):
    n = len(q)
    return sum(q[i] * q[j] / np.linalg.norm(x[i] - x[j]) for i in range(n) for j in range(i + 1, n))


def test_cut_then_resume_matches_uninterrupted(gpu, oracle):
    ce = gpu
    from oracle import ffi
    eng = engine(ce, 2000, 31)
    cfg = ce.RunConfig(**RULE)
    full = eng.run(cfg)
    cut = eng.run(replace(cfg, t_max=T_CUT, redistribute_energy=False))  # momenta = m v
    assert np.all(cut.status == ce.SIM_OK_TMAX) and np.all(cut.t_final == T_CUT)
    # the cut state: momenta (no redistribution) against the oracle's lane run to t_cut
    geo = eng.geometries()
    ref = oracle.lane_run_q(np.array(CH4_MASSES), np.tile(CH4_CHARGES, (2000, 1)), geo,
                            ffi.make_config(rtol=1e-6, atol=1e-9, stop_rule=1, t_max=T_CUT,
                                            momentum_tol=1e-7, h0_mode=1, h0=1e-3,
                                            redistribute=False))
    assert per_atom_rel(cut.momenta, ref.momenta).max() < 1e-5
    st = eng.state()
    np.testing.assert_allclose(st["vel"] * np.array(CH4_MASSES)[None, :, None],
                               cut.momenta.reshape(-1, 5, 3), rtol=1e-6, atol=0)
    for s in range(0, 2000, 97):  # energy of the saved state = E0 (to the tolerance)
        ke = 0.5 * np.sum(np.array(CH4_MASSES)[:, None] * st["vel"][s] ** 2)
        assert (ke + pe(st["pos"][s], CH4_CHARGES)) == pytest.approx(st["e0"][s], rel=1e-4)
        assert st["e0"][s] == pytest.approx(pe(geo[s], CH4_CHARGES), rel=1e-6)
    assert np.all(st["h"] > 0)
    res = eng.run(replace(cfg, resume=True))
    assert np.all(res.t_final == 1e4)
    assert per_atom_rel(res.momenta, full.momenta.reshape(-1, 5, 3)).max() < 1e-5
    assert np.all(res.steps >= cut.steps)
    assert np.abs(res.steps.astype(int) - full.steps.astype(int)).max() <= 3


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_cut_resume_vs_oracle_cut(gpu, oracle, precision):
    """A cut at t_cut followed by a resume to t_max against the oracle's statement
    of the same continuation (ceo_lane_run_cut: the state at the cut, the FSAL
    force, E0, time and counts carried over, and the step size proposed after the
    step that landed on t_cut).  f64: the same step counts and <= 1e-12; f32:
    <= 1e-5 and the same statuses.  A defect in what the resume restores moves
    the f64 result far beyond 1e-12."""
    ce = gpu
    from oracle import ffi
    p = ce.PRECISION_F64 if precision == "f64" else ce.PRECISION_F32
    rtol, atol = (1e-8, 1e-12) if precision == "f64" else (1e-6, 1e-9)
    eng = ce.BatchedRK45(ch4(ce), 1.0, 3000, rtol, atol)
    eng.sample(77, 3000, 2.5, 1.0)
    eng.enable_state()
    cfg = ce.RunConfig(precision=p, **RULE)
    cut = eng.run(replace(cfg, t_max=T_CUT))
    res = eng.run(replace(cfg, resume=True))
    ref = oracle.lane_run_cut(CH4_MASSES, CH4_CHARGES, eng.geometries(),
                              ffi.make_config(rtol=rtol, atol=atol, **RULE), T_CUT,
                              precision=precision)
    assert (cut.status == ce.SIM_OK_TMAX).mean() > 0.99
    np.testing.assert_array_equal(res.status, ref.status)
    np.testing.assert_array_equal(res.t_final, ref.t_final)
    err = per_atom_rel(res.momenta, ref.momenta)
    if precision == "f64":
        np.testing.assert_array_equal(res.steps, ref.steps)
        np.testing.assert_array_equal(res.attempts, ref.attempts)
        assert err.max() < 1e-12
    else:
        assert err.max() < 1e-5


def test_clamped_step_at_the_cut_is_not_a_momentum_stop(gpu, oracle):
    """CH4 sims whose last step before t_cut is clamped to a few 1e-5 a.u.: the
    cut ends them by the time limit (not as settled), and continuing them gives
    the uninterrupted result (these ended 0.03-0.12 off in the config norm while
    the momentum test was applied to that shortened step)."""
    ce = gpu
    from test_oracle_cut import CLAMP_SEED, CLAMP_SIMS
    # + 6147900: its GPU cut ends with a sliver, and the old continuation (10x the
    # sliver) stopped it "settled" at t = 200.00005
    geom = np.concatenate([oracle.device_sample(CLAMP_SEED, i, 1, 5, 2.5, 1.0)[0]
                           for i in CLAMP_SIMS + [6147900]])
    eng = ce.BatchedRK45(ch4(ce), 1.0, len(geom), 1e-6, 1e-9)
    eng.set_geometries(geom)
    eng.enable_state()
    cfg = ce.RunConfig(**RULE)
    full = eng.run(cfg)
    cut = eng.run(replace(cfg, t_max=T_CUT))
    assert np.all(cut.status == ce.SIM_OK_TMAX) and np.all(cut.t_final == T_CUT)
    res = eng.run(replace(cfg, resume=True))
    assert np.all(res.status == ce.SIM_OK_TMAX)
    from conftest import config_rel
    n = len(geom)
    assert config_rel(res.momenta.reshape(n, 5, 3), full.momenta.reshape(n, 5, 3)).max() < 1e-6


def test_resume_needs_state_of_same_precision(gpu):
    ce = gpu
    eng = engine(ce, 64, 1, state=False)
    with pytest.raises(RuntimeError):
        eng.run(ce.RunConfig(resume=True, **RULE))
    eng.enable_state()
    eng.run(ce.RunConfig(**dict(RULE, t_max=T_CUT)))
    with pytest.raises(RuntimeError):
        eng.run(ce.RunConfig(resume=True, precision=1, **RULE))


def test_run_subset_touches_only_listed_sims(gpu):
    ce = gpu
    eng = engine(ce, 1000, 5)
    cfg = ce.RunConfig(**RULE)
    eng.run(replace(cfg, t_max=T_CUT))
    cut = eng.download()
    everything = eng.run(replace(cfg, resume=True))
    eng.run(replace(cfg, t_max=T_CUT))  # same cut state again (deterministic)
    sub = np.array([3, 10, 11, 500, 999, 0])
    res = eng.run_subset(sub, replace(cfg, resume=True))
    assert eng.stats()["sims"] == len(sub)
    m = res.momenta.reshape(-1, 5, 3)
    np.testing.assert_array_equal(m[sub], everything.momenta.reshape(-1, 5, 3)[sub])
    rest = np.setdiff1d(np.arange(1000), sub)
    np.testing.assert_array_equal(m[rest], cut.momenta.reshape(-1, 5, 3)[rest])
    np.testing.assert_array_equal(res.status[rest], cut.status[rest])


def test_features_and_prior_vs_oracle(gpu):
    ce = gpu
    eng = engine(ce, 500, 8)
    eng.run(ce.RunConfig(**dict(RULE, t_max=T_CUT)))
    st = eng.state()
    clo = ce.Closure(ch4(ce))
    f, pr = clo.features(eng)
    m, q = np.array(CH4_MASSES), np.array(CH4_CHARGES)
    for s in range(0, 500, 23):
        x, v = st["pos"][s], st["vel"][s]
        np.testing.assert_allclose(pr[s], closure_ref.prior(x, v, m, q, st["e0"][s]), rtol=1e-12,
                                   atol=1e-12 * np.abs(pr[s]).max())
        np.testing.assert_allclose(f[s], closure_ref.features(x, v, m, q), rtol=2e-7, atol=1e-6)


@pytest.fixture(scope="module", params=[(128, 128), (256, 256)], ids=["layers", "fused"])
def trained(gpu, request):
    """A small ensemble trained on engine data (the real one is in bench.py).  At
    width 256 inference runs through the fused kernel, and the survivors of the
    early rejection keep their count on the device."""
    ce = gpu
    import coulomb_explosion_amd.closure_train as ct
    cfg = ce.RunConfig(**RULE)
    data = ct.make_data(ce, ch4(ce), 100000, 4242, cfg, T_CUT, 2.5, 1.0, 1e-6, 1e-9)
    ens = ct.fit(ce, data, members=3, hidden=request.param, epochs=20, batch=1024, lr=2e-3)
    fused = [m.set_fused(True) for m in ens.members]
    assert all(fused) == (request.param[0] == 256) and len(set(fused)) == 1
    return ct, cfg, ens


def test_network_improves_on_prior(gpu, trained):
    """What the trained network gains over the method without it, on fresh sims."""
    ce = gpu
    ct, cfg, ens = trained
    data = ct.make_data(ce, ch4(ce), 4000, 777, cfg, T_CUT, 2.5, 1.0, 1e-6, 1e-9)
    mean, rel = ct.predict(ens, data.features, data.prior)
    e_prior = per_atom_rel(data.prior, data.final)
    e_net = per_atom_rel(mean, data.final)
    # 8 epochs of this small set already took the median from 1.57% to 0.86%
    assert np.median(e_net) < 0.7 * np.median(e_prior)
    assert rel.shape == (len(data.final),)


def test_closure_gate_extremes(gpu, trained):
    ce = gpu
    ct, cfg, ens = trained
    clo = ct.attach(ce, ch4(ce), ens)
    eng = engine(ce, 3000, 99)
    ref = eng.run(replace(cfg, t_max=T_CUT))
    feat, prior = clo.features(eng)
    resumed = eng.run(replace(cfg, resume=True))
    # gate 0: every cut sim is continued -> the same as cut + resume
    res0 = clo.run(eng, cfg, T_CUT, 0.0)
    s0 = clo.stats()
    assert s0["candidates"] == 3000 and s0["fallback"] == 3000
    np.testing.assert_array_equal(res0.momenta, resumed.momenta)
    np.testing.assert_array_equal(res0.status, resumed.status)
    # gate wide open: every cut sim takes the ensemble mean
    res1 = clo.run(eng, cfg, T_CUT, 1e30)
    s1 = clo.stats()
    assert s1["fallback"] == 0 and np.all(res1.status == ce.SIM_CLOSURE)
    mean, _ = ct.predict(ens, feat, prior)
    np.testing.assert_allclose(res1.momenta.reshape(-1, 5, 3), mean, rtol=1e-6,
                               atol=1e-6 * np.abs(mean).max())
    assert np.all(ref.steps == res1.steps)


def test_closure_gate_follows_spread(gpu, trained):
    ce = gpu
    ct, cfg, ens = trained
    clo = ct.attach(ce, ch4(ce), ens)
    eng = engine(ce, 3000, 123)
    eng.run(replace(cfg, t_max=T_CUT))
    feat, prior = clo.features(eng)
    mean, rel = ct.predict(ens, feat, prior)
    tol = float(np.quantile(rel, 0.5))
    res = clo.run(eng, cfg, T_CUT, tol)
    acc = res.status == ce.SIM_CLOSURE
    far = np.abs(rel - tol) > 1e-3 * tol  # away from the threshold (rounding)
    np.testing.assert_array_equal(acc[far], (rel <= tol)[far])
    # the accepted sims take the mean of all members (the third one on compact rows)
    np.testing.assert_allclose(res.momenta.reshape(-1, 5, 3)[acc], mean[acc], rtol=1e-6,
                               atol=1e-6 * np.abs(mean).max())
    assert np.all(res.status[~acc] == ce.SIM_OK_TMAX) and np.all(res.t_final[~acc] == 1e4)
    st = clo.stats()
    assert st["fallback"] == np.count_nonzero(~acc)
    m, cand = len(ens.members), st["candidates"]
    if not ens.members[0].set_fused(True):
        # layer-by-layer inference: every member runs on every sim
        assert st["member_rows"] == m * cand
        return
    # the first two members decided part of the sims (early rejection), the third
    # member ran on the rest, and the decisions above are still the full gate's
    assert 2 * cand < st["member_rows"] < m * cand
    # at the closed gate the first two members decide every sim that they disagree on;
    # wide open, every member runs on every sim
    clo.run(eng, cfg, T_CUT, 0.0)
    assert clo.stats()["member_rows"] == 2 * cand
    clo.run(eng, cfg, T_CUT, 1e30)
    assert clo.stats()["member_rows"] == m * cand


def test_closure_gate_across_chunks(gpu, trained):
    """Over 2.2M sims the ensemble runs in three chunks of at most 1M: every
    chunk's survivors, compact rows and counts land on the right sims."""
    ce = gpu
    ct, cfg, ens = trained
    n = 2_200_000
    clo = ct.attach(ce, ch4(ce), ens)
    eng = engine(ce, n, 321)
    cand = eng.run(replace(cfg, t_max=T_CUT)).status == ce.SIM_OK_TMAX
    feat, prior = clo.features(eng)
    mean, rel = ct.predict(ens, feat, prior)
    del feat
    tol = float(np.quantile(rel, 0.3))
    res = clo.run(eng, cfg, T_CUT, tol)
    acc = res.status == ce.SIM_CLOSURE
    far = cand & (np.abs(rel - tol) > 1e-3 * tol)
    np.testing.assert_array_equal(acc[far], (rel <= tol)[far])
    assert not np.any(acc[~cand])
    np.testing.assert_allclose(res.momenta.reshape(-1, 5, 3)[acc], mean[acc], rtol=1e-6,
                               atol=1e-6 * np.abs(mean).max())
    st = clo.stats()
    m, nc = len(ens.members), np.count_nonzero(cand)
    assert st["candidates"] == nc and st["fallback"] == nc - np.count_nonzero(acc)
    if ens.members[0].set_fused(True):
        assert 2 * nc < st["member_rows"] < m * nc
    else:
        assert st["member_rows"] == m * nc


def test_closure_per_sim_charges_n10(gpu):
    """The N=10 mixed-charge path (one-species MixedBatch): features and prior
    use each sim's charges; gate 0 equals cut + resume; a wide-open gate takes
    the (untrained) ensemble everywhere."""
    ce = gpu
    C, O, H = 21874.66, 29156.95, 1837.15
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in [C, C, O, O] + [H] * 6])
    cfg = ce.RunConfig(**RULE)

    def engine():
        e = ce.MixedBatch([ce.Species(mol, 1500, 3.5, 0.9, charge_levels=3)], 1.0, 1e-6, 1e-9)
        e.sample(2024)
        e.enable_state()
        return e

    eng = engine()
    eng.run(replace(cfg, t_max=T_CUT))
    st = eng.state()
    q = eng.charges()[0]
    assert len(np.unique(q)) == 3
    clo = ce.Closure(mol)
    f, pr = clo.features(eng)
    m = mol.masses
    for s in range(0, 1500, 101):
        x, v = st["pos"][s], st["vel"][s]
        np.testing.assert_allclose(pr[s], closure_ref.prior(x, v, m, q[s], st["e0"][s]),
                                   rtol=1e-12, atol=1e-12 * np.abs(pr[s]).max())
        np.testing.assert_allclose(f[s], closure_ref.features(x, v, m, q[s]), rtol=2e-7, atol=1e-6)
    resumed = eng.run(replace(cfg, resume=True))[0]
    nets = [ce.MLP([80, 64, 30], seed=k) for k in range(2)]
    clo.set_normalization(f.mean(0), np.maximum(f.std(0), 1e-6), np.ones(30, np.float32))
    clo.set_members(nets)
    res0 = clo.run(engine(), cfg, T_CUT, 0.0)[0]
    np.testing.assert_array_equal(res0["momenta"], resumed["momenta"])
    res1 = clo.run(engine(), cfg, T_CUT, 1e30)[0]
    assert np.all(res1["status"] == ce.SIM_CLOSURE) and clo.stats()["fallback"] == 0


def test_closure_per_sim_charges_n10_fused(gpu):
    """N=10 with per-sim charges and members of the fused 512-wide shape: the
    early rejection, the column-major phase arrays and the survivors' compact rows
    give the full gate's decisions, and the accepted sims take the members' mean."""
    ce = gpu
    import coulomb_explosion_amd.closure_train as ct
    C, O, H = 21874.66, 29156.95, 1837.15
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in [C, C, O, O] + [H] * 6])
    cfg = ce.RunConfig(**RULE)
    eng = ce.MixedBatch([ce.Species(mol, 3000, 3.5, 0.9, charge_levels=3)], 1.0, 1e-6, 1e-9)
    eng.sample(2025)
    eng.enable_state()
    cand = eng.run(replace(cfg, t_max=T_CUT))[0]["status"] == ce.SIM_OK_TMAX
    clo = ce.Closure(mol)
    f, pr = clo.features(eng)
    nets = [ce.MLP([80, 512, 512, 30], seed=11 + k) for k in range(3)]
    assert all(n.set_fused(True) for n in nets)
    ens = ct.Ensemble(nets, f.mean(0).astype(np.float32),
                      np.maximum(f.std(0), 1e-6).astype(np.float32),
                      np.full(30, 0.5, np.float32), [], 0.0)
    clo = ct.attach(ce, mol, ens)
    mean, rel = ct.predict(ens, f, pr)
    tol = float(np.quantile(rel, 0.5))
    res = clo.run(eng, cfg, T_CUT, tol)[0]
    acc = res["status"] == ce.SIM_CLOSURE
    far = cand & (np.abs(rel - tol) > 1e-3 * tol)
    np.testing.assert_array_equal(acc[far], (rel <= tol)[far])
    assert not np.any(acc[~cand])
    np.testing.assert_allclose(np.asarray(res["momenta"]).reshape(-1, 10, 3)[acc], mean[acc],
                               rtol=1e-6, atol=1e-6 * np.abs(mean).max())
    st = clo.stats()
    nc = np.count_nonzero(cand)
    assert st["candidates"] == nc and st["fallback"] == nc - np.count_nonzero(acc)
    assert 2 * nc < st["member_rows"] < 3 * nc

