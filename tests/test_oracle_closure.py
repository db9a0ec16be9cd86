"""The closure prior's oracle, pinned by quadrature and by exact two-body facts."""
import numpy as np
import pytest

from oracle import closure_ref


@pytest.mark.parametrize("seed", range(6))
def test_pair_impulse_closed_form_matches_quadrature(seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=3) * 3
    u = rng.normal(size=3) * 0.02
    if r @ u < 0:  # moving apart or sideways, as after the cut
        u = -u
    np.testing.assert_allclose(closure_ref.pair_impulse(r, u),
                               closure_ref.pair_impulse_quadrature(r, u), rtol=1e-7, atol=1e-12)


def test_radial_pair_impulse():
    """Head-on separation: |J| = 1 / (|r| |u|) along r (1/|r(t)|^2 integrated)."""
    r, u = np.array([2.0, 0, 0]), np.array([0.5, 0, 0])
    np.testing.assert_allclose(closure_ref.pair_impulse(r, u), [1.0, 0, 0], rtol=1e-15)


def test_two_body_prior_is_exact_asymptote():
    """Two bodies from rest: the asymptotic momenta are +-sqrt(2 mu E0) along the
    axis; the rescaled prior has that magnitude and direction at any state on
    the radial trajectory (equal charges, momenta equal and opposite)."""
    m = np.array([3.0, 5.0])
    q = np.array([1.0, 2.0])
    d0 = 1.5
    e0 = q[0] * q[1] / d0
    mu = m[0] * m[1] / m.sum()
    # a state on the way out: separation d, relative speed from energy conservation
    d = 4.0
    ur = np.sqrt(2 * (e0 - q[0] * q[1] / d) / mu)
    axis = np.array([0.6, 0.8, 0.0])
    x = np.array([axis * d * m[1] / m.sum(), -axis * d * m[0] / m.sum()])
    v = np.array([axis * ur * m[1] / m.sum(), -axis * ur * m[0] / m.sum()])
    p = closure_ref.prior(x, v, m, q, e0)
    pinf = np.sqrt(2 * mu * e0)
    np.testing.assert_allclose(p, [axis * pinf, -axis * pinf], rtol=1e-12)


def test_prior_energy_and_momentum():
    rng = np.random.default_rng(7)
    m = np.array([21874.66] + [1837.15] * 4)
    q = np.ones(5)
    x = rng.normal(size=(5, 3)) * 5
    v = rng.normal(size=(5, 3)) * 0.01
    v -= (m[:, None] * v).sum(0) / m.sum()  # zero total momentum
    p = closure_ref.prior(x, v, m, q, e0=2.5)
    assert 0.5 * np.sum(np.sum(p * p, 1) / m) == pytest.approx(2.5, rel=1e-13)
    np.testing.assert_allclose(p.sum(0), 0.0, atol=1e-12)  # pair impulses cancel
    f = closure_ref.features(x, v, m, q)
    assert f.shape == (40,)
    np.testing.assert_allclose((m[:, None] * f[:30].reshape(5, 6)[:, :3]).sum(0), 0, atol=1e-9)


def test_pregate_bound_is_sound_and_tight():
    """The two-member early rejection never rejects an ensemble the full gate would
    accept, and it is tight: just inside the bound, other members exist that make
    the gate accept (all placed at the point that maximises |prior + mean| for the
    least spread)."""
    rng = np.random.default_rng(5)
    n_acc = n_rej = 0
    for trial in range(4000):
        m = int(rng.integers(3, 6))
        tol = 10.0 ** rng.uniform(-5, -1)
        prior = rng.normal(size=(2, 3))
        center = rng.normal(size=(2, 3)) * 0.1
        spread = tol * 10.0 ** rng.uniform(-1, 1) * np.abs(prior).max()
        z = center + spread * rng.normal(size=(m, 2, 3))
        acc = closure_ref.gate_accepts(z, prior, tol)
        rej = closure_ref.pregate_rejects(z[0], z[1], prior, m, tol)
        assert not (acc and rej), (trial, m, tol)
        n_acc += acc
        n_rej += rej
    assert n_acc > 100 and n_rej > 100  # both sides exercised
    # tightness on one atom: members 3..M at w = z12 + u* e, e along prior + z12
    for m, tol in [(3, 1e-3), (4, 1e-5), (5, 3e-2)]:
        c = (m - 1) * tol * tol
        prior = np.array([[1.0, 0.5, -0.2]])
        z12 = np.array([[0.1, 0.0, 0.3]])
        a = np.linalg.norm(prior + z12)
        e = (prior + z12) / a
        k = 2.0 * m / (m - 2)
        bound = c * k / (k - c) * a * a
        for frac, expect_accept in [(0.999, True), (1.001, False)]:
            d12 = frac * bound
            half = np.sqrt(d12 / 2.0) * np.array([[0.0, 1.0, 0.0]])  # |z1 - z2|^2 / 2 = d12
            half = half - (half @ e.T) * e  # orthogonal to e
            half *= np.sqrt(d12 / 2.0) / np.linalg.norm(half)
            z1, z2 = z12 + half, z12 - half
            # the mean of all M members at z12 + u* e (u* = c a / (k - c)): the
            # others all at one point w
            u = c * a / (k - c)
            w = z12 + (m * u / (m - 2)) * e
            z = np.concatenate([z1[None], z2[None], np.repeat(w[None], m - 2, axis=0)])
            assert closure_ref.pregate_rejects(z1, z2, prior, m, tol) == (not expect_accept)
            assert closure_ref.gate_accepts(z, prior, tol) == expect_accept
