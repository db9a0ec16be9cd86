"""CPU oracle of the closure's feature and prior kernels -- TEST INFRASTRUCTURE ONLY.

The closure has no counterpart in the reference.  Its prior is the remaining
Coulomb impulse along straight-line trajectories from the cut on,
    J_ij = k q_i q_j int_0^inf r(t)/|r(t)|^3 dt,  r(t) = r + u t,
whose closed form (r/|r| + u/|u|) / (|r||u| + r.u) is checked here against
numerical quadrature, followed by a rescale of the kinetic energy to E0.  Also the
ensemble gate and its early rejection from two members (closure.hip,
closure_gate / closure_pregate).
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


def gate_accepts(z, prior, tol):
    """The ensemble gate: z [M][n][3] scaled member outputs, prior [n][3].  Accept
    when for every atom sum_k |z_k - zbar|^2 / (M - 1) <= tol^2 |prior + zbar|^2."""
    z = np.asarray(z, dtype=np.float64)
    m = z.shape[0]
    zbar = z.mean(axis=0)
    d = ((z - zbar) ** 2).sum(axis=(0, 2))
    p2 = ((prior + zbar) ** 2).sum(axis=1)
    return bool(np.all(d / (m - 1) <= tol * tol * p2))


def pregate_rejects(z1, z2, prior, members, tol):
    """Early rejection from the first two of M = `members` outputs: with
    c = (M-1) tol^2, k = 2M/(M-2), D12 = |z1 - z2|^2 / 2 and A = |prior + (z1+z2)/2|
    per atom, the gate cannot accept whatever the other members output when
    D12 > c k A^2 / (k - c) on some atom (derivation in closure.hip)."""
    c = (members - 1) * tol * tol
    k = 2.0 * members / (members - 2)
    z12 = 0.5 * (np.asarray(z1, np.float64) + np.asarray(z2, np.float64))
    d12 = 0.5 * ((np.asarray(z1, np.float64) - z2) ** 2).sum(axis=1)
    a2 = ((prior + z12) ** 2).sum(axis=1)
    return bool(np.any(d12 > c * k / (k - c) * a2))
