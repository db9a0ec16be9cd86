"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the reference's
own scalar engine built from its sources (oracle/_ref/libceref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py for the timed CPU baseline.  The product never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

STOP_PE_FRACTION = 0
STOP_MOMENTUM_OR_TMAX = 1
H0_HAIRER = 0
H0_FIXED = 1
OK_PE, OK_MOMENTUM, OK_TMAX, NOT_CONVERGED, REJECT_LIMIT, NONFINITE, SAMPLER_FAIL = range(7)


class Config(C.Structure):
    _fields_ = [
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("pe_stop_fraction", C.c_double),
        ("max_dt", C.c_double),
        ("t_max", C.c_double),
        ("momentum_tol", C.c_double),
        ("h0", C.c_double),
        ("max_steps", C.c_uint64),
        ("stop_rule", C.c_int32),
        ("redistribute", C.c_int32),
        ("h0_mode", C.c_int32),
        ("max_rejections", C.c_int32),
    ]


class Molecule(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("pad_", C.c_int32),
        ("masses", C.POINTER(C.c_double)),
        ("charges", C.POINTER(C.c_double)),
        ("coulomb_k", C.c_double),
    ]


class Results(C.Structure):
    _fields_ = [
        ("momenta", C.POINTER(C.c_double)),
        ("steps", C.POINTER(C.c_uint32)),
        ("attempts", C.POINTER(C.c_uint32)),
        ("status", C.POINTER(C.c_uint8)),
        ("t_final", C.POINTER(C.c_double)),
        ("energy_err", C.POINTER(C.c_double)),
        ("trace_h", C.POINTER(C.c_double)),
        ("trace_cap", C.c_uint32),
        ("pad_", C.c_uint32),
    ]


def make_config(
    rtol=1e-8,
    atol=1e-16,
    stop_rule=STOP_PE_FRACTION,
    pe_stop_fraction=1e-9,
    redistribute=True,
    max_dt=1e12,
    max_steps=1_000_000,
    t_max=1e4,
    momentum_tol=1e-7,
    h0_mode=H0_HAIRER,
    h0=1e-3,
    max_rejections=100,
) -> Config:
    """Defaults follow the reference's RunConfig / IntegratorOptions
    (include/coulomb/driver.hpp:28-50, include/coulomb/integrator.hpp:44-47)."""
    return Config(rtol, atol, pe_stop_fraction, max_dt, t_max, momentum_tol, h0, max_steps,
                  stop_rule, int(redistribute), h0_mode, max_rejections)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class RunOut:
    momenta: np.ndarray  # [sims, n, 3]
    steps: np.ndarray
    attempts: np.ndarray
    status: np.ndarray
    t_final: np.ndarray
    energy_err: np.ndarray
    trace_h: np.ndarray | None
    batch_steps: int = 0


class _MolHold:
    def __init__(self, masses, charges, k=1.0):
        self.m = np.ascontiguousarray(masses, dtype=np.float64)
        self.q = np.ascontiguousarray(charges, dtype=np.float64)
        self.s = Molecule(len(self.m), 0, _p(self.m, C.c_double), _p(self.q, C.c_double), float(k))


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
    return C.CDLL(path)


class Oracle:
    def __init__(self, path=None):
        self.lib = _load(path or os.path.join(HERE, "liboracle.so"))
        L = self.lib
        L.ceo_potential_energy_f64.restype = C.c_double
        L.ceo_potential_energy_f32.restype = C.c_float
        L.ceo_kinetic_energy_f64.restype = C.c_double
        L.ceo_sample_uniform_sphere.argtypes = [C.c_uint64, C.c_int32, C.c_double, C.c_double,
                                                C.c_uint64, C.c_size_t, C.c_void_p]
        L.ceo_device_sample.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_int32, C.c_double,
                                        C.c_double, C.c_uint32, C.c_void_p, C.c_void_p]
        L.ceo_device_sample.restype = C.c_size_t
        for f in ("ceo_scalar_run_f64", "ceo_scalar_run_f32", "ceo_lane_run_f32",
                  "ceo_lane_run_f64"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        for f in ("ceo_batched_run_f32", "ceo_batched_run_f64"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32,
                                      C.c_void_p, C.c_void_p]
        L.ceo_cpu_baseline.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32,
                                       C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ceo_cpu_baseline_isa.restype = C.c_char_p

    # -- physics ------------------------------------------------------------
    def accelerations(self, masses, charges, pos, k=1.0, dtype=np.float64):
        mh = _MolHold(masses, charges, k)
        x = np.ascontiguousarray(pos, dtype=dtype).reshape(-1)
        a = np.zeros_like(x)
        if dtype == np.float64:
            self.lib.ceo_accelerations_f64(C.byref(mh.s), _p(x, C.c_double), _p(a, C.c_double))
        else:
            self.lib.ceo_accelerations_f32(C.byref(mh.s), _p(x, C.c_float), _p(a, C.c_float))
        return a.reshape(-1, 3)

    def potential_energy(self, masses, charges, pos, k=1.0, dtype=np.float64):
        mh = _MolHold(masses, charges, k)
        x = np.ascontiguousarray(pos, dtype=dtype).reshape(-1)
        if dtype == np.float64:
            return self.lib.ceo_potential_energy_f64(C.byref(mh.s), _p(x, C.c_double))
        return self.lib.ceo_potential_energy_f32(C.byref(mh.s), _p(x, C.c_float))

    def kinetic_energy(self, masses, vel):
        mh = _MolHold(masses, np.ones(len(masses)))
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        return self.lib.ceo_kinetic_energy_f64(C.byref(mh.s), _p(v, C.c_double))

    def soa_accelerations(self, n, k, px, py, pz, pair_const, inv_mass):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (px, py, pz, pair_const, inv_mass)]
        out = [np.zeros(n * k, dtype=np.float32) for _ in range(3)]
        f = self.lib.ceo_soa_accelerations_f32
        f.argtypes = [C.c_size_t, C.c_size_t] + [C.c_void_p] * 8
        f(n, k, *[v.ctypes.data for v in a], *[o.ctypes.data for o in out])
        return out

    def soa_potential_energy(self, n, k, px, py, pz, pair_const):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (px, py, pz, pair_const)]
        pe = np.zeros(k, dtype=np.float32)
        f = self.lib.ceo_soa_potential_energy_f32
        f.argtypes = [C.c_size_t, C.c_size_t] + [C.c_void_p] * 5
        f(n, k, *[v.ctypes.data for v in a], pe.ctypes.data)
        return pe

    # -- samplers -----------------------------------------------------------
    def sample(self, seed, n_atoms, n_sims, radius, min_sep, max_attempts=10000):
        out = np.zeros((n_sims, n_atoms, 3), dtype=np.float64)
        rc = self.lib.ceo_sample_uniform_sphere(seed, n_atoms, radius, min_sep, max_attempts,
                                                n_sims, out.ctypes.data)
        return rc, out

    def device_sample(self, seed, first_index, n_sims, n_atoms, radius, min_sep,
                      max_attempts=10000):
        out = np.zeros((n_sims, n_atoms, 3), dtype=np.float64)
        st = np.zeros(n_sims, dtype=np.uint8)
        fails = self.lib.ceo_device_sample(seed, first_index, n_sims, n_atoms, radius, min_sep,
                                           max_attempts, out.ctypes.data, st.ctypes.data)
        return out, st, int(fails)

    def device_sample_q(self, seed, first_index, n_sims, n_atoms, radius, min_sep, levels,
                        max_total=0, max_attempts=10000):
        f = self.lib.ceo_device_sample_q
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_int32, C.c_double, C.c_double,
                      C.c_uint32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        f.restype = C.c_size_t
        out = np.zeros((n_sims, n_atoms, 3), dtype=np.float64)
        q = np.zeros((n_sims, n_atoms), dtype=np.float64)
        st = np.zeros(n_sims, dtype=np.uint8)
        fails = f(seed, first_index, n_sims, n_atoms, radius, min_sep, max_attempts, levels,
                  max_total, out.ctypes.data, q.ctypes.data, st.ctypes.data)
        return out, q, st, int(fails)

    def philox(self, ctr, key):
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        o = np.zeros(4, dtype=np.uint32)
        self.lib.ceo_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
        return o

    def lane_coefficients(self):
        c, aa, b, e, ea = (np.zeros(7), np.zeros(49), np.zeros(7), np.zeros(7), np.zeros(7))
        self.lib.ceo_lane_coefficients(*[_p(a, C.c_double) for a in (c, aa, b, e, ea)])
        return c, aa.reshape(7, 7), b, e, ea

    # -- integrators ----------------------------------------------------------
    def _run(self, fn, masses, charges, pos, cfg, k=1.0, lanes=None, trace_cap=0):
        mh = _MolHold(masses, charges, k)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = len(mh.m)
        ns = pos.size // (3 * n)
        mom = np.zeros((ns, n, 3))
        steps = np.zeros(ns, np.uint32)
        att = np.zeros(ns, np.uint32)
        st = np.zeros(ns, np.uint8)
        tf = np.zeros(ns)
        ee = np.zeros(ns)
        tr = np.zeros((ns, trace_cap)) if trace_cap else None
        res = Results(_p(mom, C.c_double), _p(steps, C.c_uint32), _p(att, C.c_uint32),
                      _p(st, C.c_uint8), _p(tf, C.c_double), _p(ee, C.c_double),
                      _p(tr, C.c_double) if tr is not None else None, trace_cap, 0)
        bs = C.c_uint64(0)
        if lanes is None:
            rc = fn(C.byref(mh.s), C.byref(cfg), pos.ctypes.data, ns, C.byref(res))
        else:
            rc = fn(C.byref(mh.s), C.byref(cfg), pos.ctypes.data, ns, lanes, C.byref(res),
                    C.byref(bs))
        if rc != 0:
            raise RuntimeError(f"oracle run failed rc={rc}")
        return RunOut(mom, steps, att, st, tf, ee, tr, bs.value)

    def scalar_run(self, masses, charges, pos, cfg, k=1.0, precision="f64", trace_cap=0):
        fn = self.lib.ceo_scalar_run_f64 if precision == "f64" else self.lib.ceo_scalar_run_f32
        return self._run(fn, masses, charges, pos, cfg, k, trace_cap=trace_cap)

    def lane_run(self, masses, charges, pos, cfg, k=1.0, precision="f32", trace_cap=0):
        fn = self.lib.ceo_lane_run_f32 if precision == "f32" else self.lib.ceo_lane_run_f64
        return self._run(fn, masses, charges, pos, cfg, k, trace_cap=trace_cap)

    def lane_run_q(self, masses, charges_per_sim, pos, cfg, k=1.0, precision="f32", trace_cap=0):
        """Per-lane form with per-sim charges [sims, n]."""
        q = np.ascontiguousarray(charges_per_sim, dtype=np.float64)
        fn = self.lib.ceo_lane_run_q_f32 if precision == "f32" else self.lib.ceo_lane_run_q_f64
        fn.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_void_p]

        def call(mol, cfg_, pos_, ns, res):
            return fn(mol, cfg_, pos_, q.ctypes.data, ns, res)
        return self._run(call, masses, np.ones(len(masses)), pos, cfg, k, trace_cap=trace_cap)

    def lane_run_cut(self, masses, charges, pos, cfg, t_cut, k=1.0, precision="f32",
                     charges_per_sim=None, trace_cap=0):
        """Per-lane form stopped at t_cut and continued from there to cfg.t_max
        (the engine's cut + resume)."""
        fn = self.lib.ceo_lane_run_cut
        fn.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_double, C.c_int32, C.c_void_p]
        q = (None if charges_per_sim is None
             else np.ascontiguousarray(charges_per_sim, dtype=np.float64))

        def call(mol, cfg_, pos_, ns, res):
            return fn(mol, cfg_, pos_, None if q is None else q.ctypes.data, ns, t_cut,
                      1 if precision == "f64" else 0, res)
        return self._run(call, masses, charges, pos, cfg, k, trace_cap=trace_cap)

    def batched_run(self, masses, charges, pos, cfg, lanes=8, k=1.0, precision="f32"):
        fn = self.lib.ceo_batched_run_f32 if precision == "f32" else self.lib.ceo_batched_run_f64
        return self._run(fn, masses, charges, pos, cfg, k, lanes=lanes)

    def cpu_baseline(self, masses, charges, pos, cfg, n_threads, reps, k=1.0, isa=None):
        """isa None: the widest build the host runs; "avx512", "avx2" or "sse2" to
        pick one (the caller checks that the host supports it)."""
        mh = _MolHold(masses, charges, k)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        ns = pos.size // (3 * len(mh.m))
        med = C.c_double(0)
        bs = C.c_uint64(0)
        cs = C.c_double(0)
        fn = self.lib.ceo_cpu_baseline if isa is None else getattr(self.lib,
                                                                    "ceo_cpu_baseline_" + isa)
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32, C.c_int32,
                       C.c_void_p, C.c_void_p, C.c_void_p]
        lanes = fn(C.byref(mh.s), C.byref(cfg), pos.ctypes.data, ns, n_threads, reps,
                   C.byref(med), C.byref(bs), C.byref(cs))
        return {"lanes": lanes, "median_s": med.value, "batch_steps": bs.value,
                "checksum": cs.value, "sims": (ns // max(lanes, 1)) * max(lanes, 1),
                "isa": isa or self.lib.ceo_cpu_baseline_isa().decode()}


class RefEngine:
    """The reference's scalar engine (src/*.cpp), built by oracle/Makefile."""

    def __init__(self, precision="f64"):
        self.lib = _load(os.path.join(HERE, "_ref", f"libceref_{precision}.so"))
        L = self.lib
        L.cer_potential_energy.restype = C.c_double
        L.cer_potential_energy.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.cer_accelerations.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                        C.c_void_p]
        L.cer_sample.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                 C.c_size_t, C.c_void_p]
        L.cer_run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                              C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64, C.c_double,
                              C.c_double, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p]

    @staticmethod
    def available(precision="f64"):
        return os.path.exists(os.path.join(HERE, "_ref", f"libceref_{precision}.so"))

    def accelerations(self, masses, charges, pos, k=1.0):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        a = np.zeros_like(x)
        self.lib.cer_accelerations(len(m), m.ctypes.data, q.ctypes.data, k, x.ctypes.data,
                                   a.ctypes.data)
        return a.reshape(-1, 3)

    def potential_energy(self, masses, charges, pos, k=1.0):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        return self.lib.cer_potential_energy(len(m), m.ctypes.data, q.ctypes.data, k,
                                             x.ctypes.data)

    def sample(self, seed, n_atoms, n_sims, radius, min_sep, max_attempts=10000):
        out = np.zeros((n_sims, n_atoms, 3))
        rc = self.lib.cer_sample(seed, n_atoms, radius, min_sep, max_attempts, n_sims,
                                 out.ctypes.data)
        return rc, out

    def run(self, masses, charges, pos, cfg: Config, k=1.0):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = len(m)
        ns = pos.size // (3 * n)
        mom = np.zeros((ns, n, 3))
        steps = np.zeros(ns, np.uint32)
        st = np.zeros(ns, np.uint8)
        tf = np.zeros(ns)
        self.lib.cer_run(n, m.ctypes.data, q.ctypes.data, k, cfg.rtol, cfg.atol, cfg.stop_rule,
                         cfg.pe_stop_fraction, cfg.redistribute, cfg.max_dt, cfg.max_steps,
                         cfg.t_max, cfg.momentum_tol, pos.ctypes.data, ns, mom.ctypes.data,
                         steps.ctypes.data, st.ctypes.data, tf.ctypes.data)
        return mom, steps, st, tf
