"""Writes the golden fixtures under tests/golden/ (run once, in the container
that holds the reference; the outputs are committed as data).

  reference_cases.json  the reference's own fixture, copied as data from
                        tests/reference/reference_cases.json (three cases:
                        forces/PE at t=0 and asymptotic momenta of scipy RK45
                        at rtol 1e-8 / atol 1e-16).
  ref_sampler.json      UniformSphereSampler streams of the reference's sampler
                        (src/sampler.cpp), from the library oracle/Makefile
                        compiles out of the reference sources (oracle/_ref).
  ref_dp45.npz          the reference's scalar engine (f64 and the
                        COULOMB_SINGLE_PRECISION f32 build) on those
                        geometries: DormandPrince45 + run_to_convergence
                        (PE-fraction rule) and the engine's t_max / momentum
                        rule driven through Integrator::step.
  philox_kat.json       Philox4x32-10 known-answer vectors published with
                        Random123 (Salmon et al., SC'11; kat_vectors file).

Usage:  python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ffi  # noqa: E402

REF_FIXTURE = "/root/reference/tests/reference/reference_cases.json"
CH4_MASSES = [21874.66] + [1837.15] * 4
CH4_CHARGES = [1.0] * 5


def main():
    shutil.copyfile(REF_FIXTURE, os.path.join(HERE, "reference_cases.json"))

    r64 = ffi.RefEngine("f64")
    r32 = ffi.RefEngine("f32")
    streams = []
    for seed, n, R, ms, count in ((12345, 5, 2.5, 1.0, 16), (49188291, 5, 4.0, 0.25, 16),
                                  (777, 8, 4.0, 0.25, 8), (42, 3, 2.44, 0.9, 8)):
        rc, pos = r64.sample(seed, n, count, R, ms)
        assert rc == 0
        streams.append({"seed": seed, "n_atoms": n, "radius": R, "min_separation": ms,
                        "positions": pos.tolist()})
    with open(os.path.join(HERE, "ref_sampler.json"), "w") as f:
        json.dump({"description": "UniformSphereSampler streams of the reference sampler",
                   "streams": streams}, f)

    rc, geom = r64.sample(12345, 5, 32, 2.5, 1.0)
    out = {"geom": geom, "masses": np.array(CH4_MASSES), "charges": np.array(CH4_CHARGES)}
    cfgs = {
        "tmax": ffi.make_config(rtol=1e-6, atol=1e-9, stop_rule=1, t_max=1e4, momentum_tol=1e-7),
        "pe": ffi.make_config(rtol=1e-6, atol=1e-9, stop_rule=0, pe_stop_fraction=1e-5),
    }
    for name, cfg in cfgs.items():
        for prec, eng in (("f64", r64), ("f32", r32)):
            mom, steps, st, tf = eng.run(CH4_MASSES, CH4_CHARGES, geom, cfg)
            out[f"{name}_{prec}_momenta"] = mom
            out[f"{name}_{prec}_steps"] = steps
            out[f"{name}_{prec}_status"] = st
            out[f"{name}_{prec}_t"] = tf
    np.savez_compressed(os.path.join(HERE, "ref_dp45.npz"), **out)

    kat = {
        "source": "Random123 kat_vectors (Salmon, Moraes, Dror, Shaw, SC'11), philox4x32 R=10",
        "vectors": [
            {"ctr": ["00000000", "00000000", "00000000", "00000000"],
             "key": ["00000000", "00000000"],
             "out": ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]},
            {"ctr": ["ffffffff", "ffffffff", "ffffffff", "ffffffff"],
             "key": ["ffffffff", "ffffffff"],
             "out": ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]},
            {"ctr": ["243f6a88", "85a308d3", "13198a2e", "03707344"],
             "key": ["a4093822", "299f31d0"],
             "out": ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]},
        ],
    }
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
