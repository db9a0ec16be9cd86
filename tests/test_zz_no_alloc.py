"""ce_run_device_noalloc (the entry point that allocates no device memory) gives
the same results as the engine path it is a variant of (ce_engine_create +
ce_engine_set_geometries + ce_engine_run), bit for bit, on two calls in a row,
the second with fewer sims.  The caller's device buffers are allocated with the
HIP runtime the library links (tests/hipmem.py)."""
import numpy as np
import pytest

from conftest import CH4_CHARGES, CH4_MASSES

pytestmark = pytest.mark.gpu


def _engine_run(ce, mol, geom, cfg):
    eng = ce.BatchedRK45(mol, 1.0, len(geom), 1e-6, 1e-9)
    for lane in range(len(geom)):
        eng.set_lane_geometry(lane, geom[lane])
    return eng.run(cfg)


def _noalloc_run(ce, mol, geom, cfg):
    from hipmem import DeviceArray, synchronize
    ns, n = geom.shape[0], geom.shape[1]
    g = DeviceArray.from_numpy(geom.astype(np.float64))
    out = {"momenta": DeviceArray((ns, n, 3), np.float64, fill_byte=0xFF),
           "momenta_f32": DeviceArray((ns, n, 3), np.float32, fill_byte=0xFF),
           "steps": DeviceArray(ns, np.uint32, fill_byte=0xFF),
           "attempts": DeviceArray(ns, np.uint32, fill_byte=0xFF),
           "status": DeviceArray(ns, np.uint8, fill_byte=0xFF),
           "t_final": DeviceArray(ns, np.float64, fill_byte=0xFF)}
    synchronize()
    ce.run_device_noalloc(mol, 1.0, 1e-6, 1e-9, cfg, g.ptr, ns,
                          {k: v.ptr for k, v in out.items()}, 0)
    synchronize()
    return {k: v.numpy() for k, v in out.items()}


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_noalloc_equals_engine_run_twice(gpu, oracle, precision):
    ce = gpu
    mol = ce.Molecule([ce.Atom("C", CH4_MASSES[0], 1.0)] +
                      [ce.Atom("H", m, 1.0) for m in CH4_MASSES[1:]])
    prec = ce.PRECISION_F32 if precision == "f32" else ce.PRECISION_F64
    cfg = ce.RunConfig(stop_rule=ce.STOP_MOMENTUM_OR_TMAX, t_max=1e4, momentum_tol=1e-7,
                       h0_mode=ce.H0_FIXED, h0=1e-3, precision=prec)
    for call, (seed, ns) in enumerate([(11, 5000), (12, 700)]):
        geom, st, fails = oracle.device_sample(seed, 0, ns, 5, 2.5, 1.0)
        assert fails == 0
        print(f"noalloc {precision} call {call}: {ns} CH4 sims", flush=True)
        ref = _engine_run(ce, mol, geom, cfg)
        got = _noalloc_run(ce, mol, geom, cfg)
        np.testing.assert_array_equal(got["momenta"].reshape(-1), ref.momenta)
        np.testing.assert_array_equal(got["momenta_f32"].reshape(-1), ref.momenta_f32.reshape(-1))
        np.testing.assert_array_equal(got["steps"], ref.steps)
        np.testing.assert_array_equal(got["attempts"], ref.attempts)
        np.testing.assert_array_equal(got["status"], ref.status)
        np.testing.assert_array_equal(got["t_final"], ref.t_final)
        assert np.isin(ref.status, [ce.SIM_OK_MOMENTUM, ce.SIM_OK_TMAX]).all()


def test_noalloc_rejects_bad_arguments(gpu):
    ce = gpu
    mol = ce.Molecule([ce.Atom("H", 1837.15, 1.0)] * 2)
    bad = {k: 0 for k in ("momenta", "momenta_f32", "steps", "attempts", "status", "t_final")}
    with pytest.raises(ValueError):
        ce.run_device_noalloc(mol, 1.0, 1e-6, 1e-9, ce.RunConfig(), 0, 10, bad)
    with pytest.raises(ValueError):
        ce.run_device_noalloc(mol, 1.0, 1e-6, 1e-9, ce.RunConfig(resume=True), 0, 0, bad)
    ce.run_device_noalloc(mol, 1.0, 1e-6, 1e-9, ce.RunConfig(), 0, 0, bad)  # nothing to do
