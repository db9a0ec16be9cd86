"""Mixed-size launches with per-sim charges (the N2 / CO2 / CH4 / C2H6 workload),
through the C-ABI, against the oracle's per-lane statement with the same
geometries and charges."""
import numpy as np
import pytest

from conftest import assert_f32_bar, per_atom_rel

pytestmark = pytest.mark.gpu

C, O, H, N14 = 21874.66, 29156.95, 1837.15, 25526.04
MIXED = [  # (masses, radius = n^(1/3) + 1.5)
    ([N14, N14], 2 ** (1 / 3) + 1.5),
    ([C, O, O], 3 ** (1 / 3) + 1.5),
    ([C, H, H, H, H], 5 ** (1 / 3) + 1.5),
    ([C, C] + [H] * 6, 8 ** (1 / 3) + 1.5),
]
RULE = dict(stop_rule=1, t_max=1e4, momentum_tol=1e-7, h0_mode=1, h0=1e-3)


def species(ce, count, masses_radii=MIXED, levels=2):
    out = []
    for masses, R in masses_radii:
        mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in masses])
        out.append(ce.Species(mol, count, R, 0.9, charge_levels=levels,
                              max_total_charge=2 * len(masses)))
    return out


def test_species_sampler_bit_exact(gpu, oracle):
    ce = gpu
    sp = species(ce, 700)
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.sample(seed=77, first_index=1000)
    geoms, charges = mb.geometries(), mb.charges()
    start = 1000
    for s, g, q in zip(sp, geoms, charges):
        n = s.molecule.size()
        ref_g, ref_q, st, fails = oracle.device_sample_q(77, start, s.count, n, s.radius, 0.9,
                                                         2, 2 * n)
        assert fails == 0
        np.testing.assert_array_equal(g, ref_g)
        np.testing.assert_array_equal(q, ref_q)
        assert set(np.unique(q)) <= {1.0, 2.0}
        start += s.count


def test_species_sampler_binding_cap(gpu, oracle):
    """A cap on the total charge that binds (levels {1,2,3}, N = 5, total <= 8):
    the device sampler's redraws give the oracle's charges bit for bit, and the
    integration with those charges agrees with the oracle's per-lane form
    (f64: same steps, <= 1e-12)."""
    ce = gpu
    from oracle import ffi
    mol = ce.Molecule([ce.Atom("X", m, 1.0) for m in [C, H, H, H, H]])
    sp = [ce.Species(mol, 3000, 2.5, 1.0, charge_levels=3, max_total_charge=8)]
    mb = ce.MixedBatch(sp, 1.0, 1e-10, 1e-14)
    mb.sample(seed=404, first_index=77)
    g, q = mb.geometries()[0], mb.charges()[0]
    ref_g, ref_q, st, fails = oracle.device_sample_q(404, 77, 3000, 5, 2.5, 1.0, 3, 8)
    assert fails == 0
    np.testing.assert_array_equal(g, ref_g)
    np.testing.assert_array_equal(q, ref_q)
    assert q.sum(axis=1).max() == 8 and (q == 3).any()
    res = mb.run(ce.RunConfig(precision=1, **RULE))[0]
    ref = oracle.lane_run_q(mol.masses, q, g, ffi.make_config(rtol=1e-10, atol=1e-14, **RULE),
                            precision="f64")
    np.testing.assert_array_equal(res["steps"], ref.steps)
    assert per_atom_rel(res["momenta"], ref.momenta).max() < 1e-12


@pytest.mark.parametrize("precision,tol", [(0, 1e-5), (1, 1e-12)])
def test_mixed_launch_vs_oracle(gpu, oracle, precision, tol):
    ce = gpu
    from oracle import ffi
    sp = species(ce, 600)
    rtol, atol = (1e-6, 1e-9) if precision == 0 else (1e-10, 1e-14)
    mb = ce.MixedBatch(sp, 1.0, rtol, atol)
    mb.sample(seed=5)
    res = mb.run(ce.RunConfig(precision=precision, **RULE))
    assert mb.stats()["launches"] == 1  # sizes (2, 3, 5, 8): one persistent launch
    for s, g, q, r in zip(sp, mb.geometries(), mb.charges(), res):
        ref = oracle.lane_run_q(s.molecule.masses, q, g,
                                ffi.make_config(rtol=rtol, atol=atol, **RULE),
                                precision="f32" if precision == 0 else "f64")
        assert per_atom_rel(r["momenta"], ref.momenta).max() < tol
        np.testing.assert_array_equal(r["status"], ref.status)
        if precision == 1:
            np.testing.assert_array_equal(r["steps"], ref.steps)


def test_mixed_equals_single_size_launches(gpu):
    """The one-launch mixed kernel gives the same bits as one engine per size."""
    ce = gpu
    sp = species(ce, 500)
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.sample(seed=9)
    cfg = ce.RunConfig(**RULE)
    res = mb.run(cfg)
    for s, g, q, r in zip(sp, mb.geometries(), mb.charges(), res):
        eng = ce.BatchedRK45(s.molecule, 1.0, s.count, 1e-6, 1e-9)
        eng.set_geometries(g)
        eng.set_charges(q)
        one = eng.run(cfg)
        np.testing.assert_array_equal(one.momenta.reshape(r["momenta"].shape), r["momenta"])
        np.testing.assert_array_equal(one.steps, r["steps"])


def test_other_size_sets_run_per_size(gpu, oracle):
    """Sizes outside the compiled mixed set run one launch per species."""
    ce = gpu
    from oracle import ffi
    n10 = [C, C, O, O] + [H] * 6
    sp = [ce.Species(ce.Molecule([ce.Atom("X", m, 1.0) for m in n10]), 300, 3.5, 0.9, 3),
          ce.Species(ce.Molecule([ce.Atom("X", m, 1.0) for m in [C] + [H] * 4]), 200, 2.5, 1.0)]
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.sample(seed=21)
    res = mb.run(ce.RunConfig(**RULE))
    assert mb.stats()["launches"] == 2
    q10 = mb.charges()[0]
    assert set(np.unique(q10)) <= {1.0, 2.0, 3.0} and len(np.unique(q10)) == 3
    for s, g, q, r in zip(sp, mb.geometries(), mb.charges(), res):
        ref = oracle.lane_run_q(s.molecule.masses, q, g, ffi.make_config(rtol=1e-6, atol=1e-9,
                                                                         **RULE))
        # N = 10 with charges up to 3 holds sims whose conditioning amplifies the
        # rsqrt-vs-1/sqrt rounding past 1e-5 per atom: the f32 bar allows for them
        assert_f32_bar(oracle, s.molecule.masses, q, g, r["momenta"], ref.momenta,
                       ffi.make_config(rtol=1e-10, atol=1e-14, **RULE))


@pytest.mark.parametrize("co2_charges", [[1.0, 1.0, 1.0], [2.0, 1.0, 1.0]])
def test_mixed_fixed_charges_vs_oracle(gpu, oracle, co2_charges):
    """Fixed charges in the one-launch mixed kernel: all +1 runs the
    uniform-pair-constant form for every species; one CO2 with a +2 carbon puts
    the whole launch on the pair table.  Against the oracle's f32 lane form."""
    ce = gpu
    from oracle import ffi
    sp = []
    for (masses, R), q in zip(MIXED, [[1.0] * 2, co2_charges, [1.0] * 5, [1.0] * 8]):
        mol = ce.Molecule([ce.Atom("X", m, c) for m, c in zip(masses, q)])
        sp.append(ce.Species(mol, 500, R, 0.9))
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.sample(seed=31)
    res = mb.run(ce.RunConfig(**RULE))
    assert mb.stats()["launches"] == 1
    for s, g, r in zip(sp, mb.geometries(), res):
        ref = oracle.lane_run(s.molecule.masses, s.molecule.charges, g,
                              ffi.make_config(rtol=1e-6, atol=1e-9, **RULE))
        assert per_atom_rel(r["momenta"], ref.momenta).max() < 1e-5
        np.testing.assert_array_equal(r["status"], ref.status)


def test_mixed_one_launch_and_per_species_launches_agree(gpu):
    """Sizes (2, 3, 5, 8) run as one persistent kernel or as one kernel per species
    (the automatic choice for >= 2^19 sims per species): the same bits per sim."""
    ce = gpu
    sp = species(ce, 1 << 19)
    mb = ce.MixedBatch(sp, 1.0, 1e-6, 1e-9)
    mb.sample(seed=17)
    cfg = ce.RunConfig(**RULE)
    mb.run(cfg, download=False)
    assert mb.stats()["launches"] == 4  # automatic: every species is large
    mb.set_launch(mixed_launch=ce.MIXED_ONE_LAUNCH)
    one = mb.run(cfg)
    assert mb.stats()["launches"] == 1
    mb.set_launch(mixed_launch=ce.MIXED_PER_SPECIES)
    per = mb.run(cfg)
    assert mb.stats()["launches"] == 4
    for a, b in zip(one, per):
        np.testing.assert_array_equal(a["momenta"], b["momenta"])
        np.testing.assert_array_equal(a["steps"], b["steps"])
        np.testing.assert_array_equal(a["status"], b["status"])
    with pytest.raises(ValueError):
        mb.set_launch(mixed_launch=3)
