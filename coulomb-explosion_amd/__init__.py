"""Host-side mirror of the reference's batched f32 engine, over the engine's C-ABI.

The classes keep the reference's names and argument meaning:

* ``Atom`` / ``atom_from_amu`` / ``Molecule``  -- include/coulomb/molecule.hpp:14-35
* ``RunConfig``                               -- include/coulomb/driver.hpp:28-50 (+ the
  engine's stop rule, t_max, momentum_tol, fixed h0, precision)
* ``UniformSphereSampler``                    -- include/coulomb/sampler.hpp:46-72 (host
  mt19937_64 path of the reference, same stream as the reference for a seed)
* ``BatchedRK45``                             -- include/coulomb/batched_integrator.hpp:62-191
  (``set_lane_geometry``, ``run`` -> ``Result{momenta, steps, converged, batch_steps}``)

Errors follow the reference: ``ValueError`` where it throws std::invalid_argument,
``RuntimeError`` where it throws std::runtime_error.  The library
(libcoulomb_explosion.so, built in this directory) is required: importing this
module fails loudly when it is missing, and nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcoulomb_explosion.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "coulomb_explosion.h")

MAX_ATOMS = 10
AMU_TO_ELECTRON_MASS = 1822.888486209  # CODATA 2018 m_u/m_e (include/coulomb/units.hpp:11)

STOP_PE_FRACTION = 0
STOP_MOMENTUM_OR_TMAX = 1
H0_HAIRER = 0
H0_FIXED = 1
PRECISION_F32 = 0
PRECISION_F64 = 1

SIM_OK_PE, SIM_OK_MOMENTUM, SIM_OK_TMAX, SIM_NOT_CONVERGED = 0, 1, 2, 3
SIM_REJECT_LIMIT, SIM_NONFINITE, SIM_SAMPLER_FAIL, SIM_CLOSURE = 4, 5, 6, 7

CE_OK, CE_ERR_INVALID_ARGUMENT, CE_ERR_RUNTIME, CE_ERR_HIP, CE_ERR_NO_DEVICE = 0, -1, -2, -3, -4


# ------------------------------------------------------------------ C types ----
class _MoleculeDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("pad_", C.c_int32), ("masses", C.POINTER(C.c_double)),
                ("charges", C.POINTER(C.c_double)), ("coulomb_k", C.c_double)]


class _Tolerances(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double)]


class _RunConfig(C.Structure):
    _fields_ = [("pe_stop_fraction", C.c_double), ("max_dt", C.c_double), ("t_max", C.c_double),
                ("momentum_tol", C.c_double), ("h0", C.c_double), ("max_steps", C.c_uint64),
                ("stop_rule", C.c_int32), ("redistribute_energy", C.c_int32),
                ("h0_mode", C.c_int32), ("max_rejections", C.c_int32), ("precision", C.c_int32),
                ("resume", C.c_int32)]


class _Results(C.Structure):
    _fields_ = [("momenta", C.POINTER(C.c_double)), ("momenta_f32", C.POINTER(C.c_float)),
                ("steps", C.POINTER(C.c_uint32)), ("attempts", C.POINTER(C.c_uint32)),
                ("status", C.POINTER(C.c_uint8)), ("t_final", C.POINTER(C.c_double)),
                ("energy_err", C.POINTER(C.c_double)), ("trace_h", C.POINTER(C.c_double)),
                ("trace_cap", C.c_uint32), ("pad_", C.c_uint32)]


class _RunStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("sampler_ms", C.c_double), ("pack_ms", C.c_double),
                ("sims", C.c_uint64),
                ("attempted_steps", C.c_uint64), ("accepted_steps", C.c_uint64),
                ("refills", C.c_uint64), ("lane_slots", C.c_uint64), ("grid_blocks", C.c_int32),
                ("block_threads", C.c_int32), ("n_cus", C.c_int32), ("launches", C.c_int32)]


class _SpeciesDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("charge_levels", C.c_int32),
                ("masses", C.POINTER(C.c_double)), ("charges", C.POINTER(C.c_double)),
                ("radius", C.c_double), ("min_separation", C.c_double), ("count", C.c_uint64),
                ("max_total_charge", C.c_int32), ("pad_", C.c_int32)]


class _ClosureStats(C.Structure):
    _fields_ = [("sims", C.c_uint64), ("candidates", C.c_uint64), ("fallback", C.c_uint64),
                ("cut_attempts", C.c_uint64), ("fallback_attempts", C.c_uint64),
                ("cut_ms", C.c_double), ("infer_ms", C.c_double), ("fallback_ms", C.c_double),
                ("total_ms", C.c_double), ("member_rows", C.c_uint64)]


class _LaunchConfig(C.Structure):
    _fields_ = [("block_threads", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("refill_threshold", C.c_int32), ("chunk", C.c_int32),
                ("mixed_launch", C.c_int32)]


MIXED_AUTO, MIXED_ONE_LAUNCH, MIXED_PER_SPECIES = 0, 1, 2


class _DeviceResults(C.Structure):
    _fields_ = [("momenta", C.c_void_p), ("momenta_f32", C.c_void_p), ("steps", C.c_void_p),
                ("attempts", C.c_void_p), ("status", C.c_void_p), ("t_final", C.c_void_p)]


def _load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run `make -C coulomb-explosion_amd` (or "
            "__graft_entry__.build()).  There is no CPU fallback.")
    lib = C.CDLL(path)
    P, S = C.c_void_p, C.c_size_t
    sigs = {
        "ce_last_error": ([], C.c_char_p),
        "ce_device_count": ([P], C.c_int),
        "ce_run_config_default": ([P], None),
        "ce_engine_create": ([P, P, S, P], C.c_int),
        "ce_engine_destroy": ([P], C.c_int),
        "ce_engine_set_geometries": ([P, P, S], C.c_int),
        "ce_engine_sample": ([P, C.c_uint64, C.c_uint64, S, C.c_double, C.c_double, C.c_uint32],
                             C.c_int),
        "ce_engine_get_geometries": ([P, P, S], C.c_int),
        "ce_engine_run": ([P, P, P], C.c_int),
        "ce_engine_download": ([P, P], C.c_int),
        "ce_engine_stats": ([P, P], C.c_int),
        "ce_engine_set_launch": ([P, P], C.c_int),
        "ce_engine_get_launch": ([P, P], C.c_int),
        "ce_engine_enable_trace": ([P, C.c_uint32], C.c_int),
        "ce_engine_enable_energy_diagnostic": ([P, C.c_int32], C.c_int),
        "ce_engine_enable_state": ([P, C.c_int32], C.c_int),
        "ce_engine_get_state": ([P, P, P, P, P, S], C.c_int),
        "ce_engine_run_subset": ([P, P, P, S, P], C.c_int),
        "ce_engine_capture": ([P, P, C.c_uint64, S, C.c_double, C.c_double, C.c_uint32], C.c_int),
        "ce_engine_replay": ([P, C.c_uint64, P], C.c_int),
        "ce_engine_create_mixed": ([P, C.c_int32, C.c_double, P, P], C.c_int),
        "ce_engine_sample_species": ([P, C.c_uint64, C.c_uint64, C.c_uint32], C.c_int),
        "ce_engine_set_charges": ([P, P, S], C.c_int),
        "ce_engine_get_charges": ([P, P, S], C.c_int),
        "ce_record_bytes": ([C.c_int32], C.c_size_t),
        "ce_engine_pack_records": ([P, C.c_uint64, S, P], C.c_int),
        "ce_host_sampler_create": ([C.c_uint64, C.c_double, C.c_double, C.c_uint64, P], C.c_int),
        "ce_host_sampler_reseed": ([P, C.c_uint64], C.c_int),
        "ce_host_sampler_sample": ([P, C.c_int32, S, P], C.c_int),
        "ce_host_sampler_destroy": ([P], C.c_int),
        "ce_batched_accelerations_f32": ([S, S, P, P, P, P, P, P, P, P], C.c_int),
        "ce_batched_potential_energy_f32": ([S, S, P, P, P, P, P], C.c_int),
        "ce_mlp_create": ([P, C.c_int32, C.c_uint64, P], C.c_int),
        "ce_mlp_destroy": ([P], C.c_int),
        "ce_mlp_param_count": ([P], C.c_int64),
        "ce_mlp_get_params": ([P, P], C.c_int),
        "ce_mlp_set_params": ([P, P], C.c_int),
        "ce_mlp_forward": ([P, P, S, P], C.c_int),
        "ce_mlp_gradients": ([P, P, P, S, P, P], C.c_int),
        "ce_mlp_adam_step": ([P, C.c_float], C.c_int),
        "ce_mlp_set_fused": ([P, C.c_int32], C.c_int),
        "ce_mlp_set_dataset": ([P, P, P, S], C.c_int),
        "ce_mlp_train_epoch": ([P, S, C.c_float, C.c_uint64, P], C.c_int),
        "ce_closure_create": ([C.c_int32, P, P, C.c_double, P], C.c_int),
        "ce_closure_destroy": ([P], C.c_int),
        "ce_closure_set_normalization": ([P, P, P, P], C.c_int),
        "ce_closure_set_members": ([P, P, C.c_int32], C.c_int),
        "ce_closure_features": ([P, P, P, P, S], C.c_int),
        "ce_closure_run": ([P, P, P, C.c_double, C.c_double, P], C.c_int),
        "ce_closure_get_stats": ([P, P], C.c_int),
        "ce_run_device_noalloc": ([P, P, P, P, S, P, P], C.c_int),
        "ce_hbm_copy_calibration": ([S, C.c_int32, P], C.c_int),
        "ce_engine_pack_records_async": ([P, C.c_uint64, S, P], C.c_int),
        "ce_engine_records_wait": ([P], C.c_int),
        "ce_host_alloc": ([S, P], C.c_int),
        "ce_host_free": ([P], C.c_int),
        "ce_engine_set_state": ([P, P, P, P, P, S, C.c_int32], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load_library()


def library() -> C.CDLL:
    return _lib


def _check(rc: int) -> None:
    if rc == CE_OK:
        return
    msg = (_lib.ce_last_error() or b"").decode(errors="replace")
    if rc == CE_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(f"coulomb engine error {rc}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    _lib.ce_device_count(C.byref(n))
    return n.value


def run_device_noalloc(mol: "Molecule", coulomb_k: float, rtol: float, atol: float,
                       cfg: "RunConfig", geom_ptr: int, n_sims: int, out: dict,
                       stream: int = 0) -> None:
    """ce_run_device_noalloc: integrate n_sims sims whose geometry (AoS float64
    [sim][n][3]) is at device address `geom_ptr` into caller-owned device buffers
    `out` (addresses under the keys momenta, momenta_f32, steps, attempts, status,
    t_final), queued on the HIP stream `stream` and not waited for.  No device
    memory is allocated."""
    m = np.ascontiguousarray(mol.masses)
    q = np.ascontiguousarray(mol.charges)
    desc = _MoleculeDesc(mol.size(), 0, m.ctypes.data_as(C.POINTER(C.c_double)),
                         q.ctypes.data_as(C.POINTER(C.c_double)), float(coulomb_k))
    tol = _Tolerances(float(rtol), float(atol))
    res = _DeviceResults(*[out[k] for k, _ in _DeviceResults._fields_])
    c = cfg._c()
    _check(_lib.ce_run_device_noalloc(C.byref(desc), C.byref(tol), C.byref(c), geom_ptr,
                                      int(n_sims), C.byref(res), stream))


# --------------------------------------------------------- reference mirrors ----
@dataclass
class Atom:
    symbol: str
    mass: float    # electron masses
    charge: float


def atom_from_amu(symbol: str, mass_amu: float, charge: float) -> Atom:
    """include/coulomb/molecule.hpp:23-25"""
    return Atom(symbol, mass_amu * AMU_TO_ELECTRON_MASS, charge)


@dataclass
class Molecule:
    atoms: list = field(default_factory=list)

    def size(self) -> int:
        return len(self.atoms)

    @property
    def masses(self) -> np.ndarray:
        return np.array([a.mass for a in self.atoms], dtype=np.float64)

    @property
    def charges(self) -> np.ndarray:
        return np.array([a.charge for a in self.atoms], dtype=np.float64)


@dataclass
class RunConfig:
    """include/coulomb/driver.hpp:28-50 defaults, plus the engine's fields."""
    pe_stop_fraction: float = 1e-9
    redistribute_energy: bool = True
    max_dt: float = 1e12
    max_steps: int = 1_000_000
    stop_rule: int = STOP_PE_FRACTION
    t_max: float = 1e4
    momentum_tol: float = 1e-7
    h0_mode: int = H0_HAIRER
    h0: float = 1e-3
    max_rejections: int = 100
    precision: int = PRECISION_F32
    resume: bool = False  # continue from the saved state (BatchedRK45.enable_state)

    def _c(self) -> _RunConfig:
        return _RunConfig(self.pe_stop_fraction, self.max_dt, self.t_max, self.momentum_tol,
                          self.h0, self.max_steps, self.stop_rule, int(self.redistribute_energy),
                          self.h0_mode, self.max_rejections, self.precision, int(self.resume))


@dataclass
class Result:
    """BatchedRK45::Result (batched_integrator.hpp:64-69) plus per-sim extras."""
    momenta: np.ndarray      # float64 [(lane*n + i)*3 + c], flat like the reference
    steps: np.ndarray        # accepted steps per sim
    converged: np.ndarray    # uint8: 1 unless not-converged / failed
    batch_steps: int         # max accepted steps over the sims (no lockstep here)
    status: np.ndarray
    attempts: np.ndarray
    t_final: np.ndarray
    energy_err: np.ndarray | None  # None unless enable_energy_diagnostic() was on for the run
    momenta_f32: np.ndarray
    trace_h: np.ndarray | None = None


class UniformSphereSampler:
    """UniformSphereSampler(Options, seed) (include/coulomb/sampler.hpp:46-72).

    Same stream as the reference for a seed (std::mt19937_64 +
    uniform_real_distribution, in the engine library's host code); stateful,
    so successive ``sample`` calls continue the stream.
    """

    def __init__(self, radius: float = 4.0, min_separation: float = 0.25,
                 max_attempts_per_atom: int = 10000, seed: int = 0):
        h = C.c_void_p()
        _check(_lib.ce_host_sampler_create(seed, radius, min_separation, max_attempts_per_atom,
                                           C.byref(h)))
        self._h = h
        self.radius = float(radius)
        self.min_separation = float(min_separation)
        self.max_attempts_per_atom = int(max_attempts_per_atom)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.ce_host_sampler_destroy(h)
            self._h = None

    def name(self) -> str:
        return "uniform-sphere"

    def reseed(self, seed: int) -> None:
        _check(_lib.ce_host_sampler_reseed(self._h, seed))

    def sample(self, mol_or_n, n_sims: int = 1) -> np.ndarray:
        """n_sims consecutive draws -> positions [n_sims, n, 3] (atoms at rest)."""
        n = mol_or_n.size() if isinstance(mol_or_n, Molecule) else int(mol_or_n)
        out = np.zeros((n_sims, n, 3), dtype=np.float64)
        _check(_lib.ce_host_sampler_sample(self._h, n, n_sims, out.ctypes.data))
        return out


class BatchedRK45:
    """Drop-in for BatchedRK45<float> (batched_integrator.hpp:62-191).

    ``lanes`` is the batch capacity (the reference's SIMD width; any size here).
    """

    def __init__(self, mol: Molecule, coulomb_k: float, lanes: int, rtol: float, atol: float):
        if mol.size() < 1 or mol.size() > MAX_ATOMS:
            raise ValueError(f"BatchedRK45: molecule must have 1..{MAX_ATOMS} atoms")
        if lanes < 1:
            raise ValueError("BatchedRK45: lanes must be positive")
        self.n = mol.size()
        self.k = int(lanes)
        self._m = np.ascontiguousarray(mol.masses)
        self._q = np.ascontiguousarray(mol.charges)
        desc = _MoleculeDesc(self.n, 0, self._m.ctypes.data_as(C.POINTER(C.c_double)),
                             self._q.ctypes.data_as(C.POINTER(C.c_double)), float(coulomb_k))
        tol = _Tolerances(float(rtol), float(atol))
        h = C.c_void_p()
        _check(_lib.ce_engine_create(C.byref(desc), C.byref(tol), self.k, C.byref(h)))
        self._h = h
        self._geom = np.zeros((self.k, self.n, 3), dtype=np.float64)
        self._loaded = 0
        self._dirty = False
        self._trace_cap = 0
        self._energy_diag = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.ce_engine_destroy(h)
            self._h = None

    def lanes(self) -> int:
        return self.k

    def atoms(self) -> int:
        return self.n

    # -- geometry ---------------------------------------------------------------
    def set_lane_geometry(self, lane: int, positions) -> None:
        """batched_integrator.hpp:112-121: positions (n x 3), velocities at rest."""
        if not 0 <= lane < self.k:
            raise ValueError("lane out of range")
        self._geom[lane] = np.asarray(positions, dtype=np.float64).reshape(self.n, 3)
        self._loaded = self.k
        self._dirty = True

    def set_geometries(self, pos) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.n, 3)
        if pos.shape[0] > self.k:
            raise ValueError("more geometries than lanes")
        self._geom[: pos.shape[0]] = pos
        self._loaded = pos.shape[0]
        _check(_lib.ce_engine_set_geometries(self._h, self._geom.ctypes.data, self._loaded))
        self._dirty = False

    def sample(self, seed: int, n_sims: int, radius: float, min_separation: float,
               first_index: int = 0, max_attempts: int = 10000) -> None:
        """Device-side counter-based uniform-sphere sampler."""
        if n_sims > self.k:
            raise ValueError("n_sims exceeds lanes")
        _check(_lib.ce_engine_sample(self._h, seed, first_index, n_sims, radius, min_separation,
                                     max_attempts))
        self._loaded = int(n_sims)
        self._dirty = False

    def geometries(self) -> np.ndarray:
        out = np.zeros((self._loaded, self.n, 3), dtype=np.float64)
        if self._loaded:
            _check(_lib.ce_engine_get_geometries(self._h, out.ctypes.data, self._loaded))
        return out

    # -- configuration ------------------------------------------------------------
    def set_launch(self, block_threads=0, blocks_per_cu=0, refill_threshold=0, chunk=0,
                   mixed_launch=MIXED_AUTO) -> None:
        lc = _LaunchConfig(block_threads, blocks_per_cu, refill_threshold, chunk, mixed_launch)
        _check(_lib.ce_engine_set_launch(self._h, C.byref(lc)))

    def launch(self) -> dict:
        lc = _LaunchConfig()
        _check(_lib.ce_engine_get_launch(self._h, C.byref(lc)))
        return {f: getattr(lc, f) for f, _ in _LaunchConfig._fields_}

    def enable_trace(self, cap: int) -> None:
        _check(_lib.ce_engine_enable_trace(self._h, cap))
        self._trace_cap = int(cap)

    def enable_energy_diagnostic(self, on: bool = True) -> None:
        """Next runs compute (KE + PE - E0)/E0 per sim in FP64 (Result.energy_err;
        None while off)."""
        _check(_lib.ce_engine_enable_energy_diagnostic(self._h, 1 if on else 0))
        self._energy_diag = bool(on)

    # -- run ----------------------------------------------------------------------
    def run(self, cfg: RunConfig | None = None, download: bool = True) -> Result | None:
        """batched_integrator.hpp:123-191."""
        cfg = cfg or RunConfig()
        if self._dirty:
            _check(_lib.ce_engine_set_geometries(self._h, self._geom.ctypes.data, self._loaded))
            self._dirty = False
        c = cfg._c()
        if not download:
            _check(_lib.ce_engine_run(self._h, C.byref(c), None))
            return None
        _check(_lib.ce_engine_run(self._h, C.byref(c), None))
        return self.download()

    def download(self) -> Result:
        ns, n = self._loaded, self.n
        mom = np.zeros((ns, n, 3))
        mom32 = np.zeros((ns, n, 3), dtype=np.float32)
        steps = np.zeros(ns, np.uint32)
        att = np.zeros(ns, np.uint32)
        st = np.zeros(ns, np.uint8)
        tf = np.zeros(ns)
        ee = np.zeros(ns) if self._energy_diag else None
        tr = np.zeros((ns, self._trace_cap)) if self._trace_cap else None

        def ptr(a, ct):
            return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None

        res = _Results(ptr(mom, C.c_double), ptr(mom32, C.c_float), ptr(steps, C.c_uint32),
                       ptr(att, C.c_uint32), ptr(st, C.c_uint8), ptr(tf, C.c_double),
                       ptr(ee, C.c_double), ptr(tr, C.c_double), self._trace_cap, 0)
        _check(_lib.ce_engine_download(self._h, C.byref(res)))
        conv = ((st <= SIM_OK_TMAX) | (st == SIM_CLOSURE)).astype(np.uint8)
        return Result(mom.reshape(-1), steps, conv, int(steps.max()) if ns else 0, st, att, tf,
                      ee, mom32, tr)

    def run_subset(self, sims, cfg: RunConfig | None = None) -> Result:
        """Runs only the listed sims (indices into the loaded batch); the others
        keep their previous results."""
        cfg = cfg or RunConfig()
        idx = np.ascontiguousarray(sims, dtype=np.uint64)
        c = cfg._c()
        _check(_lib.ce_engine_run_subset(self._h, C.byref(c), idx.ctypes.data, idx.size, None))
        return self.download()

    def capture(self, cfg: RunConfig, seed: int, n_sims: int, radius: float,
                min_separation: float, max_attempts: int = 10000) -> None:
        """Record one batch (device sampling + integration) as a HIP graph."""
        if n_sims > self.k:
            raise ValueError("n_sims exceeds lanes")
        c = cfg._c()
        _check(_lib.ce_engine_capture(self._h, C.byref(c), seed, n_sims, radius, min_separation,
                                      max_attempts))
        self._graph_sims = int(n_sims)

    def replay(self, first_index: int, download: bool = True) -> Result | None:
        """Launch the captured batch for sims [first_index, first_index + n_sims)."""
        _check(_lib.ce_engine_replay(self._h, first_index, None))
        self._loaded = self._graph_sims
        self._dirty = False
        return self.download() if download else None

    def enable_state(self, enable: bool = True) -> None:
        """Keep every sim's integration state at its stop for a later resume."""
        _check(_lib.ce_engine_enable_state(self._h, int(enable)))

    def state(self) -> dict:
        """The saved state: pos, vel [sims, n, 3], h and e0 per sim (float64)."""
        ns = self._loaded
        out = {"pos": np.zeros((ns, self.n, 3)), "vel": np.zeros((ns, self.n, 3)),
               "h": np.zeros(ns), "e0": np.zeros(ns)}
        _check(_lib.ce_engine_get_state(self._h, out["pos"].ctypes.data, out["vel"].ctypes.data,
                                        out["h"].ctypes.data, out["e0"].ctypes.data, ns))
        return out

    def records(self, first: int = 0, n_sims: int | None = None) -> np.ndarray:
        """Packed dataset records (examples/generate_dataset.cpp:198-217) as a
        numpy structured array with the dtype of examples/to_parquet.py:34-42."""
        n_sims = self._loaded - first if n_sims is None else n_sims
        rb = _lib.ce_record_bytes(self.n)
        buf = np.zeros(n_sims * rb, dtype=np.uint8)
        _check(_lib.ce_engine_pack_records(self._h, first, n_sims, buf.ctypes.data))
        return buf.view(record_dtype(self.n))

    def records_async(self, out: "PinnedBuffer", first: int = 0,
                      n_sims: int | None = None) -> np.ndarray:
        """ce_engine_pack_records_async into the pinned buffer `out`: returns its
        record view, valid after records_wait()."""
        n_sims = self._loaded - first if n_sims is None else n_sims
        rb = _lib.ce_record_bytes(self.n)
        if out.nbytes < n_sims * rb:
            raise ValueError("pinned buffer too small for the records")
        _check(_lib.ce_engine_pack_records_async(self._h, first, n_sims, out.ptr))
        return out.array[: n_sims * rb].view(record_dtype(self.n))

    def records_wait(self) -> None:
        _check(_lib.ce_engine_records_wait(self._h))

    def stats(self) -> dict:
        s = _RunStats()
        _check(_lib.ce_engine_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _RunStats._fields_ if f != "pad_"}

    def set_charges(self, charges) -> None:
        """Per-sim charges [n_sims, n] for the loaded geometries (replaces the
        molecule's fixed charges on the next runs)."""
        if self._dirty:
            _check(_lib.ce_engine_set_geometries(self._h, self._geom.ctypes.data, self._loaded))
            self._dirty = False
        q = np.ascontiguousarray(charges, dtype=np.float64).reshape(self._loaded, self.n)
        _check(_lib.ce_engine_set_charges(self._h, q.ctypes.data, self._loaded))


@dataclass
class Species:
    """One molecule type of a mixed-size batch: `count` sims drawn in a ball of
    `radius` with pairs >= `min_separation`; charge_levels > 0 gives every sim
    its own charges uniform in {1..charge_levels} (sum <= max_total_charge if > 0)."""
    molecule: Molecule
    count: int
    radius: float
    min_separation: float
    charge_levels: int = 0
    max_total_charge: int = 0


class MixedBatch:
    """Several molecule sizes in one launch (no reference counterpart: the
    reference batches one fixed molecule).  Results come back per species."""

    def __init__(self, species: list, coulomb_k: float, rtol: float, atol: float):
        self.species = list(species)
        self._keep = []
        descs = (_SpeciesDesc * len(self.species))()
        for d, s in zip(descs, self.species):
            m = np.ascontiguousarray(s.molecule.masses)
            q = np.ascontiguousarray(s.molecule.charges)
            self._keep += [m, q]
            d.n_atoms = s.molecule.size()
            d.charge_levels = s.charge_levels
            d.masses = m.ctypes.data_as(C.POINTER(C.c_double))
            d.charges = q.ctypes.data_as(C.POINTER(C.c_double))
            d.radius = s.radius
            d.min_separation = s.min_separation
            d.count = s.count
            d.max_total_charge = s.max_total_charge
        tol = _Tolerances(float(rtol), float(atol))
        h = C.c_void_p()
        _check(_lib.ce_engine_create_mixed(descs, len(self.species), float(coulomb_k),
                                           C.byref(tol), C.byref(h)))
        self._h = h
        self.total = sum(s.count for s in self.species)
        self._per_sim_charges = any(s.charge_levels > 0 for s in self.species)
        self._energy_diag = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.ce_engine_destroy(h)
            self._h = None

    def _split(self, flat: np.ndarray, width: int) -> list:
        out, o = [], 0
        for s in self.species:
            k = s.count * width * s.molecule.size()
            shape = (s.count, s.molecule.size(), 3) if width == 3 else (s.count, s.molecule.size())
            out.append(flat[o:o + k].reshape(shape))
            o += k
        return out

    def _split_sims(self, a: np.ndarray) -> list:
        out, o = [], 0
        for s in self.species:
            out.append(a[o:o + s.count])
            o += s.count
        return out

    def sample(self, seed: int, first_index: int = 0, max_attempts: int = 10000) -> None:
        _check(_lib.ce_engine_sample_species(self._h, seed, first_index, max_attempts))

    @property
    def _loaded(self) -> int:
        return self.total

    def enable_state(self, enable: bool = True) -> None:
        """Keep every sim's integration state at its stop for a later resume."""
        _check(_lib.ce_engine_enable_state(self._h, int(enable)))

    def state(self):
        """The saved state per species: dict(pos, vel [count, n, 3], h, e0) (a
        single dict for a one-species batch)."""
        n_pos = sum(sp.count * 3 * sp.molecule.size() for sp in self.species)
        pos, vel = np.zeros(n_pos), np.zeros(n_pos)
        h, e0 = np.zeros(self.total), np.zeros(self.total)
        _check(_lib.ce_engine_get_state(self._h, pos.ctypes.data, vel.ctypes.data, h.ctypes.data,
                                        e0.ctypes.data, self.total))
        out = [{"pos": p, "vel": v, "h": a, "e0": b} for p, v, a, b in
               zip(self._split(pos, 3), self._split(vel, 3), self._split_sims(h),
                   self._split_sims(e0))]
        return out[0] if len(out) == 1 else out

    def set_geometries(self, per_species: list) -> None:
        flat = np.concatenate([np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
                               for g in per_species])
        _check(_lib.ce_engine_set_geometries(self._h, flat.ctypes.data, self.total))

    def set_charges(self, per_species: list) -> None:
        flat = np.concatenate([np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
                               for q in per_species])
        _check(_lib.ce_engine_set_charges(self._h, flat.ctypes.data, self.total))

    def geometries(self) -> list:
        n = sum(s.count * 3 * s.molecule.size() for s in self.species)
        flat = np.zeros(n)
        _check(_lib.ce_engine_get_geometries(self._h, flat.ctypes.data, self.total))
        return self._split(flat, 3)

    def charges(self) -> list:
        n = sum(s.count * s.molecule.size() for s in self.species)
        flat = np.zeros(n)
        _check(_lib.ce_engine_get_charges(self._h, flat.ctypes.data, self.total))
        return self._split(flat, 1)

    def set_launch(self, block_threads=0, blocks_per_cu=0, refill_threshold=0, chunk=0,
                   mixed_launch=MIXED_AUTO) -> None:
        lc = _LaunchConfig(block_threads, blocks_per_cu, refill_threshold, chunk, mixed_launch)
        _check(_lib.ce_engine_set_launch(self._h, C.byref(lc)))

    def stats(self) -> dict:
        s = _RunStats()
        _check(_lib.ce_engine_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _RunStats._fields_ if f != "pad_"}

    def run(self, cfg: RunConfig | None = None, download: bool = True):
        c = (cfg or RunConfig())._c()
        _check(_lib.ce_engine_run(self._h, C.byref(c), None))
        return self.download() if download else None

    def enable_energy_diagnostic(self, on: bool = True) -> None:
        """Next runs compute (KE + PE - E0)/E0 per sim in FP64 ("energy_err";
        None while off)."""
        _check(_lib.ce_engine_enable_energy_diagnostic(self._h, 1 if on else 0))
        self._energy_diag = bool(on)

    def download(self) -> list:
        """Per species: dict(momenta [count, n, 3], steps, attempts, status, t_final, energy_err)."""
        n = sum(s.count * 3 * s.molecule.size() for s in self.species)
        mom = np.zeros(n)
        steps = np.zeros(self.total, np.uint32)
        att = np.zeros(self.total, np.uint32)
        st = np.zeros(self.total, np.uint8)
        tf = np.zeros(self.total)
        ee = np.zeros(self.total) if self._energy_diag else None

        def ptr(a, ct):
            return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None

        res = _Results(ptr(mom, C.c_double), None, ptr(steps, C.c_uint32), ptr(att, C.c_uint32),
                       ptr(st, C.c_uint8), ptr(tf, C.c_double), ptr(ee, C.c_double), None, 0, 0)
        _check(_lib.ce_engine_download(self._h, C.byref(res)))
        out = []
        for i, m in enumerate(self._split(mom, 3)):
            sl = [self._split_sims(a)[i] for a in (steps, att, st, tf)]
            out.append({"momenta": m, "steps": sl[0], "attempts": sl[1], "status": sl[2],
                        "t_final": sl[3],
                        "energy_err": self._split_sims(ee)[i] if ee is not None else None})
        return out


class PinnedBuffer:
    """Pinned host memory (ce_host_alloc) with a uint8 numpy view."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(_lib.ce_host_alloc(int(nbytes), C.byref(p)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(self.nbytes,))

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.ce_host_free(self.ptr)
            self.ptr = None


def record_dtype(n_atoms: int) -> np.dtype:
    """The record layout as numpy reads it (examples/to_parquet.py:34-42)."""
    return np.dtype([("init_pos", "<f4", (n_atoms * 3,)), ("final_p", "<f4", (n_atoms * 3,)),
                     ("steps", "<u4"), ("converged", "u1")])


def batched_accelerations(n: int, k: int, px, py, pz, pair_const, inv_mass):
    """batched::accelerations<float> (batched_force.hpp:31-84), same SoA layout."""
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (px, py, pz, pair_const, inv_mass)]
    out = [np.zeros(n * k, dtype=np.float32) for _ in range(3)]
    _check(_lib.ce_batched_accelerations_f32(n, k, *[a.ctypes.data for a in arrs],
                                             *[o.ctypes.data for o in out]))
    return out


def batched_potential_energy(n: int, k: int, px, py, pz, pair_const):
    """batched::potential_energy<float> (batched_force.hpp:89-110)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (px, py, pz, pair_const)]
    pe = np.zeros(k, dtype=np.float32)
    _check(_lib.ce_batched_potential_energy_f32(n, k, *[a.ctypes.data for a in arrs],
                                                pe.ctypes.data))
    return pe


class MLP:
    """The closure network on the GPU: SiLU MLP with BF16 MFMA GEMMs (FP32
    accumulation), FP32 master parameters, MSE loss and Adam."""

    def __init__(self, widths, seed: int = 0):
        self.widths = [int(w) for w in widths]
        w = np.asarray(self.widths, dtype=np.int32)
        h = C.c_void_p()
        _check(_lib.ce_mlp_create(w.ctypes.data, len(w), seed, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.ce_mlp_destroy(self._h)
            self._h = None

    @property
    def n_params(self) -> int:
        return int(_lib.ce_mlp_param_count(self._h))

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        _check(_lib.ce_mlp_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        _check(_lib.ce_mlp_set_params(self._h, p.ctypes.data))

    def _rows(self, x, width):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != width:
            raise ValueError(f"expected [batch][{width}] array, got {x.shape}")
        return x

    def forward(self, x) -> np.ndarray:
        x = self._rows(x, self.widths[0])
        y = np.empty((x.shape[0], self.widths[-1]), dtype=np.float32)
        _check(_lib.ce_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradients(self, x, t):
        x = self._rows(x, self.widths[0])
        t = self._rows(t, self.widths[-1])
        if t.shape[0] != x.shape[0]:
            raise ValueError("x and t differ in batch size")
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        _check(_lib.ce_mlp_gradients(self._h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                     g.ctypes.data, C.byref(loss)))
        return float(loss.value), g

    def set_fused(self, enable: bool = True) -> bool:
        """Inference through the fused whole-network kernel where the shape allows;
        returns whether it is in use."""
        rc = _lib.ce_mlp_set_fused(self._h, int(enable))
        if rc < 0:
            _check(rc)
        return rc == 1

    def adam_step(self, lr: float) -> None:
        _check(_lib.ce_mlp_adam_step(self._h, lr))

    def set_dataset(self, x, t) -> None:
        x = self._rows(x, self.widths[0])
        t = self._rows(t, self.widths[-1])
        if t.shape[0] != x.shape[0]:
            raise ValueError("x and t differ in length")
        _check(_lib.ce_mlp_set_dataset(self._h, x.ctypes.data, t.ctypes.data, x.shape[0]))

    def train_epoch(self, batch: int, lr: float, shuffle_seed: int = 0) -> float:
        loss = C.c_float(0.0)
        _check(_lib.ce_mlp_train_epoch(self._h, batch, lr, shuffle_seed, C.byref(loss)))
        return float(loss.value)


class Closure:
    """Learned asymptotic closure for one molecule (see include/coulomb_explosion.h):
    plain DP5(4) to t_cut, straight-line-impulse prior plus an MLP ensemble's
    residual where the members agree, integration continued elsewhere."""

    def __init__(self, mol: Molecule, coulomb_k: float = 1.0):
        self.n = mol.size()
        self._m = np.ascontiguousarray(mol.masses)
        self._q = np.ascontiguousarray(mol.charges)
        h = C.c_void_p()
        _check(_lib.ce_closure_create(self.n, self._m.ctypes.data, self._q.ctypes.data,
                                      float(coulomb_k), C.byref(h)))
        self._h = h
        self._members = []

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.ce_closure_destroy(self._h)
            self._h = None

    @property
    def n_features(self) -> int:
        return 8 * self.n

    def set_normalization(self, in_shift, in_scale, out_scale) -> None:
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (in_shift, in_scale, out_scale)]
        if a[0].size != 8 * self.n or a[1].size != 8 * self.n or a[2].size != 3 * self.n:
            raise ValueError("normalisation sizes must be 8n, 8n, 3n")
        _check(_lib.ce_closure_set_normalization(self._h, *[v.ctypes.data for v in a]))

    def set_members(self, members) -> None:
        arr = (C.c_void_p * max(1, len(members)))(*[m._h.value for m in members])
        _check(_lib.ce_closure_set_members(self._h, arr, len(members)))
        self._members = list(members)  # keep them alive

    def features(self, eng):
        """Raw features [sims, 8n] and prior momenta [sims, n, 3] of the engine's
        saved state."""
        ns = eng._loaded
        f = np.zeros((ns, 8 * self.n), dtype=np.float32)
        pr = np.zeros((ns, self.n, 3))
        _check(_lib.ce_closure_features(self._h, eng._h, f.ctypes.data, pr.ctypes.data, ns))
        return f, pr

    def run(self, eng, cfg: RunConfig, t_cut: float, gate_tol: float, download: bool = True):
        """eng: a BatchedRK45 or a one-species MixedBatch (per-sim charges)."""
        if getattr(eng, "_dirty", False):
            _check(_lib.ce_engine_set_geometries(eng._h, eng._geom.ctypes.data, eng._loaded))
            eng._dirty = False
        c = cfg._c()
        _check(_lib.ce_closure_run(self._h, eng._h, C.byref(c), t_cut, gate_tol, None))
        return eng.download() if download else None

    def stats(self) -> dict:
        s = _ClosureStats()
        _check(_lib.ce_closure_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ClosureStats._fields_}
