"""The C-ABI library on the CPU: it loads, exports every entry point the header
declares, validates arguments like the reference (invalid_argument ->
CE_ERR_INVALID_ARGUMENT / ValueError), and its host sampler reproduces the
reference's stream.  No device compute is launched here."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "coulomb_explosion.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ce_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(ce):
    lib = ce.library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950(ce):
    data = open(ce.LIB_PATH, "rb").read()
    assert b"gfx950" in data


STRUCTS = {"ce_molecule_desc": "_MoleculeDesc", "ce_tolerances": "_Tolerances",
           "ce_run_config": "_RunConfig", "ce_results": "_Results", "ce_run_stats": "_RunStats",
           "ce_species_desc": "_SpeciesDesc", "ce_closure_stats": "_ClosureStats",
           "ce_launch_config": "_LaunchConfig", "ce_device_results": "_DeviceResults"}


def test_struct_layouts_match_header(ce, tmp_path):
    """ctypes mirrors of the header structs: size and every field's offset as the C
    compiler lays the header out (a small C program built with gcc prints them)."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "coulomb_explosion.h"',
             'int main(void) {']
    for cname, pyname in STRUCTS.items():
        st = getattr(ce, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in st._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln.strip()}
    for cname, pyname in STRUCTS.items():
        st = getattr(ce, pyname)
        assert C.sizeof(st) == got[(cname, "size")], cname
        for f, _ in st._fields_:
            assert getattr(st, f).offset == got[(cname, f)], (cname, f)


def test_run_config_defaults_are_the_reference_defaults(ce):
    c = ce._RunConfig()
    ce.library().ce_run_config_default(C.byref(c))
    assert c.pe_stop_fraction == 1e-9 and c.redistribute_energy == 1
    assert c.max_dt == 1e12 and c.max_steps == 1_000_000 and c.max_rejections == 100
    assert c.h0_mode == ce.H0_HAIRER and c.stop_rule == ce.STOP_PE_FRACTION
    py = ce.RunConfig()
    assert (py.pe_stop_fraction, py.max_dt, py.max_steps) == (1e-9, 1e12, 1_000_000)


def test_engine_create_validates_before_touching_a_device(ce):
    lib = ce.library()
    m = np.array([1837.15, 1837.15])
    q = np.array([1.0, 1.0])
    h = C.c_void_p()

    def create(n, masses, rtol=1e-6, atol=1e-9, cap=10):
        desc = ce._MoleculeDesc(n, 0, masses.ctypes.data_as(C.POINTER(C.c_double)),
                                q.ctypes.data_as(C.POINTER(C.c_double)), 1.0)
        tol = ce._Tolerances(rtol, atol)
        return lib.ce_engine_create(C.byref(desc), C.byref(tol), cap, C.byref(h))

    assert create(0, m) == ce.CE_ERR_INVALID_ARGUMENT
    assert create(11, m) == ce.CE_ERR_INVALID_ARGUMENT
    assert create(2, np.array([1.0, -1.0])) == ce.CE_ERR_INVALID_ARGUMENT
    assert create(2, m, rtol=0.0) == ce.CE_ERR_INVALID_ARGUMENT
    assert create(2, m, cap=0) == ce.CE_ERR_INVALID_ARGUMENT
    assert b"capacity" in lib.ce_last_error()
    with pytest.raises(ValueError):
        ce.BatchedRK45(ce.Molecule([]), 1.0, 4, 1e-6, 1e-9)
    if ce.device_count() == 0:
        assert create(2, m) == ce.CE_ERR_NO_DEVICE
        with pytest.raises(RuntimeError):
            ce.BatchedRK45(ce.Molecule([ce.Atom("H", 1837.15, 1.0)] * 2), 1.0, 4, 1e-6, 1e-9)


def test_null_arguments_are_rejected(ce):
    lib = ce.library()
    assert lib.ce_engine_run(None, None, None) == ce.CE_ERR_INVALID_ARGUMENT
    assert lib.ce_engine_stats(None, None) == ce.CE_ERR_INVALID_ARGUMENT
    assert lib.ce_engine_set_geometries(None, None, 0) == ce.CE_ERR_INVALID_ARGUMENT
    assert lib.ce_host_sampler_sample(None, 5, 1, None) == ce.CE_ERR_INVALID_ARGUMENT
    assert lib.ce_batched_accelerations_f32(0, 1, *([None] * 8)) == ce.CE_ERR_INVALID_ARGUMENT
    assert lib.ce_run_device_noalloc(*([None] * 4), 1, None, None) == ce.CE_ERR_INVALID_ARGUMENT


def test_host_sampler_reproduces_reference_stream(ce):
    with open(os.path.join(GOLDEN, "ref_sampler.json")) as f:
        doc = json.load(f)
    for s in doc["streams"]:
        smp = ce.UniformSphereSampler(s["radius"], s["min_separation"], seed=s["seed"])
        ref = np.array(s["positions"])
        # stateful: two calls continue the stream exactly like one
        a = smp.sample(s["n_atoms"], 3)
        b = smp.sample(s["n_atoms"], len(ref) - 3)
        np.testing.assert_array_equal(np.concatenate([a, b]), ref)
        smp.reseed(s["seed"])
        np.testing.assert_array_equal(smp.sample(s["n_atoms"], 2), ref[:2])


def test_host_sampler_errors_follow_reference(ce):
    with pytest.raises(ValueError):
        ce.UniformSphereSampler(-1.0, 0.25, seed=1)
    with pytest.raises(ValueError):
        ce.UniformSphereSampler(1.0, -0.5, seed=1)
    smp = ce.UniformSphereSampler(1.0, 1.5, max_attempts_per_atom=100, seed=1)
    with pytest.raises(RuntimeError, match="could not place"):
        smp.sample(10, 1)


def test_product_does_not_import_oracle():
    """The product package never loads the oracle (it is test infrastructure)."""
    pkg = os.path.join(ROOT, "coulomb-explosion_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.replace("oracle path", ""), fn


@pytest.mark.parametrize("name,std,defines", [("generator_loop", "c++20", []),
                                               ("strategy_api", "c++20", []),
                                               ("strategy_api", "c++20", ["-DCOULOMB_SINGLE_PRECISION"])])
def test_cpp_mirror_compiles_and_links(tmp_path, name, std, defines):
    """include/coulomb_gpu.hpp builds reference-style callers (the generator's
    batch loop; the sampler / force / integrator / driver API, also with the
    reference's single-precision Real) and links the library."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    exe = str(tmp_path / name)
    r = subprocess.run(["g++", "-std=" + std, "-O1", "-Wall", "-Werror", *defines,
                        "-I" + os.path.join(ROOT, "include"), src, "-o", exe,
                        "-L" + os.path.join(ROOT, "coulomb-explosion_amd"), "-lcoulomb_explosion",
                        "-Wl,-rpath," + os.path.join(ROOT, "coulomb-explosion_amd")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_cpp_mirror_header_is_cxx17(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "coulomb_gpu.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("argv,msg", [
    (["--sims", "abc", "--out", "x.bin"], "bad value for --sims"),
    (["--masses", "1,,2", "--out", "x.bin"], "bad value for --masses"),
    (["--sims", "10"], "--out is required"),
    (["--stop-rule", "energy", "--out", "x.bin"], "--stop-rule must be"),
])
def test_generator_cli_rejects_bad_arguments(tmp_path, argv, msg):
    """The dataset generator exits with status 1 and a message on malformed or
    missing flags, before it touches a device."""
    exe = os.path.join(ROOT, "coulomb-explosion_amd", "generate_dataset")
    if not os.path.exists(exe):
        pytest.skip("generate_dataset not built")
    r = subprocess.run([exe] + argv, cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)
