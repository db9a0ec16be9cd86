"""The C++ mirror (include/coulomb_gpu.hpp) run as the reference generator's loop
gives the Python mirror's bits for the same sampler stream and settings."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_mirror_run_matches_python(gpu, tmp_path):
    ce = gpu
    exe = str(tmp_path / "mirror_demo")
    lib_dir = os.path.join(ROOT, "coulomb-explosion_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "mirror_demo.cpp"), "-o", exe,
                    "-L" + lib_dir, "-lcoulomb_explosion", "-Wl,-rpath," + lib_dir],
                   check=True, capture_output=True, timeout=120)
    r = subprocess.run([exe, "64"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(v) for v in line.split()] for line in r.stdout.strip().splitlines()])
    mol = ce.Molecule([ce.atom_from_amu("C", 12.011, 1.0)] +
                      [ce.atom_from_amu("H", 1.008, 1.0)] * 4)
    sampler = ce.UniformSphereSampler(4.0, 0.25, 10000, seed=49188291)
    bi = ce.BatchedRK45(mol, 1.0, 64, 1e-4, 1e-7)
    for lane in range(64):
        bi.set_lane_geometry(lane, sampler.sample(mol)[0])
    res = bi.run(ce.RunConfig(pe_stop_fraction=1e-5))
    np.testing.assert_array_equal(rows[:, 0], res.steps)
    np.testing.assert_array_equal(rows[:, 1], res.converged)
    np.testing.assert_array_equal(rows[:, 2:], res.momenta.reshape(64, 15))
