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
