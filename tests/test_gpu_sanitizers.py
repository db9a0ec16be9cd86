"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer
(leak checks on) on the device paths: engine lifetimes, every copy-in /
copy-out path, saved state and resume, subsets, captured graphs, queued records
into pinned memory, mixed engines, the MLP (tests/cpp/sanitize_driver.cpp
`device`), and the dataset generator's full pipeline (device sampler with graph
replay, host sampler, writer thread, pinned double buffers, sidecar).  The
device code is the product's; only host code is instrumented.

Needs the sanitizer builds in the tree (`make -C tests/cpp asan`); without them
the tests fail rather than skip, so a missing build is seen.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from test_sanitizers import PKG_SAN, SAN, run_clean

pytestmark = pytest.mark.gpu


def _exe(path):
    assert os.path.exists(path), f"{path} not built: run make -C tests/cpp asan"
    return path


def test_device_paths_under_asan_ubsan():
    r = run_clean([_exe(os.path.join(SAN, "sanitize_driver")), "device"], device=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitize driver ok (host and device paths)" in r.stdout


@pytest.mark.parametrize("sampler", ["device", "host"])
def test_generator_pipeline_under_asan_ubsan(ce, tmp_path, sampler):
    """The sanitized generator writes the product generator's bytes."""
    sims, batch = 70_000, 16_000  # four full batches (graph replays) and a partial one
    args = ["--sims", str(sims), "--batch", str(batch), "--sampler", sampler]
    out = tmp_path / "san.bin"
    r = run_clean([_exe(os.path.join(PKG_SAN, "generate_dataset")), "--out", str(out)] + args,
                  device=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    ref = tmp_path / "product.bin"
    r = subprocess.run([os.path.join(ROOT, "coulomb-explosion_amd", "generate_dataset"),
                        "--out", str(ref)] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    rec = np.fromfile(out, dtype=ce.record_dtype(5))
    assert len(rec) == sims and rec["converged"].all() and (rec["steps"] > 0).all()
    assert out.read_bytes() == ref.read_bytes()
    meta = json.loads((tmp_path / "san.bin.meta.json").read_text())
    assert meta["n_sims"] == sims
