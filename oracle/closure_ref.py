"""CPU oracle of the closure's feature and prior kernels -- TEST INFRASTRUCTURE ONLY.

The closure has no counterpart in the reference.  Its prior is the remaining
Coulomb impulse along straight-line trajectories from the cut on,
    J_ij = k q_i q_j int_0^inf r(t)/|r(t)|^3 dt,  r(t) = r + u t,
whose closed form (r/|r| + u/|u|) / (|r||u| + r.u) is checked here against
numerical quadrature, followed by a rescale of the kinetic energy to E0.
"""
from __future__ import annotations

import numpy as np


def pair_impulse(r, u):
    """Closed form of int_0^inf (r + u t) / |r + u t|^3 dt."""
    rn, un = np.linalg.norm(r), np.linalg.norm(u)
    return (r / rn + u / un) / (rn * un + r @ u)


def pair_impulse_quadrature(r, u):
    """The same integral by adaptive quadrature (scipy), for pinning."""
    from scipy.integrate import quad
    out = np.zeros(3)
    for c in range(3):
        out[c] = quad(lambda t: (r[c] + u[c] * t) / np.linalg.norm(r + u * t) ** 3, 0, np.inf,
                      epsabs=1e-14, epsrel=1e-12, limit=500)[0]
    return out


def prior(x, v, masses, charges, e0, k=1.0):
    """Prior asymptotic momenta [n, 3] of one sim from its state (x, v [n, 3])."""
    x, v = np.asarray(x, np.float64), np.asarray(v, np.float64)
    m = np.asarray(masses, np.float64)
    q = np.asarray(charges, np.float64)
    p = m[:, None] * v
    n = len(m)
    for i in range(n):
        for j in range(i + 1, n):
            r, u = x[i] - x[j], v[i] - v[j]
            w = np.linalg.norm(r) * np.linalg.norm(u) + r @ u
            if not (w > 1e-300):
                continue
            J = k * q[i] * q[j] * (r / np.linalg.norm(r) + u / np.linalg.norm(u)) / w
            p[i] += J
            p[j] -= J
    ke = 0.5 * np.sum(np.sum(p * p, axis=1) / m)
    return p * (np.sqrt(e0 / ke) if ke > 0 and e0 > 0 else 1.0)


def features(x, v, masses, charges):
    """Raw network input (8n): per atom position about the centre of mass and
    m v, then the charges and the masses."""
    x, v = np.asarray(x, np.float64), np.asarray(v, np.float64)
    m = np.asarray(masses, np.float64)
    cm = (m[:, None] * x).sum(0) / m.sum()
    per_atom = np.concatenate([x - cm, m[:, None] * v], axis=1).reshape(-1)
    return np.concatenate([per_atom, np.asarray(charges, np.float64), m])
