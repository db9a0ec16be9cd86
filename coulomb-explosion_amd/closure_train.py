"""Training data and ensemble training for the learned asymptotic closure.

Data come from the engine itself: a batch of sampled sims is integrated to
t_cut with the state kept, the closure's feature kernel gives the network input
and the physical prior, and the same sims are then continued from the saved
state to the full stop rule; the residual target is (final momenta - prior).
Every member of the ensemble is an MLP trained on the GPU by the engine's own
BF16 MFMA forward / backward kernels with Adam.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, replace

import numpy as np


@dataclass
class ClosureData:
    features: np.ndarray   # [sims, 8n] raw
    prior: np.ndarray      # [sims, n, 3]
    final: np.ndarray      # [sims, n, 3] momenta of the continued integration
    seconds: float


def make_data(ce, mol, n_sims, seed, cfg, t_cut, radius, min_sep, rtol, atol,
              first_index=0, chunk=2_000_000, charge_levels=0,
              max_total_charge=0) -> ClosureData:
    """Engine-generated (state at t_cut, asymptotic momenta) pairs; only sims
    stopped by the cut and finished by the continuation are kept.  With
    charge_levels > 0 every sim draws its own charges (uniform in 1..levels)."""
    t0 = time.time()
    clo = ce.Closure(mol)
    n = mol.size()
    feats, priors, finals = [], [], []
    cap = min(chunk, n_sims)
    eng = None
    for c0 in range(0, n_sims, cap):
        cnt = min(cap, n_sims - c0)
        if charge_levels:
            if eng is None or eng.total != cnt:
                eng = ce.MixedBatch([ce.Species(mol, cnt, radius, min_sep, charge_levels,
                                                max_total_charge or charge_levels * n)],
                                    1.0, rtol, atol)
                eng.enable_state()
            eng.sample(seed, first_index=first_index + c0)
        else:
            if eng is None:
                eng = ce.BatchedRK45(mol, 1.0, cap, rtol, atol)
                eng.enable_state()
            eng.sample(seed, cnt, radius, min_sep, first_index=first_index + c0)
        cut = _flat(eng.run(replace(cfg, t_max=t_cut)))
        f, pr = clo.features(eng)
        full = _flat(eng.run(replace(cfg, resume=True)))
        keep = (cut["status"] == ce.SIM_OK_TMAX) & (full["status"] <= ce.SIM_OK_TMAX)
        feats.append(f[keep])
        priors.append(pr[keep])
        finals.append(full["momenta"].reshape(cnt, n, 3)[keep])
    return ClosureData(np.concatenate(feats), np.concatenate(priors), np.concatenate(finals),
                       time.time() - t0)


def _flat(res) -> dict:
    """momenta / status of a BatchedRK45 Result or a one-species MixedBatch result."""
    if isinstance(res, list):
        return {"momenta": res[0]["momenta"], "status": res[0]["status"]}
    return {"momenta": res.momenta, "status": res.status}


@dataclass
class Ensemble:
    members: list
    in_shift: np.ndarray
    in_scale: np.ndarray
    out_scale: np.ndarray
    losses: list           # per member, mean loss per epoch
    seconds: float


def normalization(data: ClosureData):
    f = data.features.astype(np.float64)
    shift = f.mean(0)
    scale = f.std(0)
    scale[scale < 1e-12 * np.maximum(1.0, np.abs(shift))] = 1.0
    resid = (data.final - data.prior).reshape(len(f), -1)
    out = resid.std(0)
    out[out == 0] = 1.0
    return shift.astype(np.float32), scale.astype(np.float32), out.astype(np.float32)


def fit(ce, data: ClosureData, members=4, hidden=(256, 256, 256), epochs=20, batch=8192,
        lr=1e-3, seed=0, log=None) -> Ensemble:
    t0 = time.time()
    shift, scale, out_scale = normalization(data)
    x = ((data.features - shift) / scale).astype(np.float32)
    y = ((data.final - data.prior).reshape(len(x), -1) / out_scale).astype(np.float32)
    widths = [x.shape[1], *hidden, y.shape[1]]
    nets, losses = [], []
    for k in range(members):
        net = ce.MLP(widths, seed=seed * 1000 + k)
        net.set_dataset(x, y)
        hist = []
        for e in range(epochs):
            hist.append(net.train_epoch(batch, lr, shuffle_seed=seed * 100000 + k * 1000 + e))
        if log:
            log(f"[closure] member {k}: loss {hist[0]:.3e} -> {hist[-1]:.3e}")
        nets.append(net)
        losses.append(hist)
    return Ensemble(nets, shift, scale, out_scale, losses, time.time() - t0)


def attach(ce, mol, ens: Ensemble):
    """A Closure object with the ensemble and its normalisation set."""
    clo = ce.Closure(mol)
    clo.set_normalization(ens.in_shift, ens.in_scale, ens.out_scale)
    clo.set_members(ens.members)
    return clo


def predict(ens: Ensemble, features, prior):
    """Ensemble mean and per-atom relative spread for given raw features (host
    helper for tests and reports; the run itself uses the device gate)."""
    # the device's arithmetic: float32 (f - shift) * (1 / scale)
    inv = (np.float32(1.0) / ens.in_scale.astype(np.float32)).astype(np.float32)
    x = ((np.asarray(features, np.float32) - ens.in_shift.astype(np.float32)) * inv).astype(np.float32)
    ys = np.stack([m.forward(x) for m in ens.members]).astype(np.float64) * ens.out_scale
    mean = prior.reshape(len(x), -1) + ys.mean(0)
    n = prior.shape[1]
    spread = ys.std(0, ddof=1).reshape(len(x), n, 3) if len(ens.members) > 1 else None
    mean = mean.reshape(len(x), n, 3)
    rel = (np.linalg.norm(spread, axis=-1) / np.linalg.norm(mean, axis=-1)).max(-1) \
        if spread is not None else None
    return mean, rel
