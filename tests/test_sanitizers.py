"""Sanitizer builds of the host code (the reference's COULOMB_ENABLE_SANITIZERS
option, CMakeLists.txt:32,57-61, and its `debug` preset), on the CPU:

* ThreadSanitizer over the generator's record writer, driven by the generator's
  double-buffer protocol (tests/cpp/record_writer_stress.cpp), and over the
  oracle's threaded CPU baseline on 1 and 4 threads (tests/cpp/cpu_baseline_tsan.cpp);
* AddressSanitizer + UndefinedBehaviorSanitizer (leak checks on) over the
  library's host paths that need no device: argument checks and error paths of
  the C-ABI, the host sampler (tests/cpp/sanitize_driver.cpp), and the
  generator's argument handling.

The device paths under the same build run in test_gpu_sanitizers.py.  Builds
come from tests/cpp/Makefile (the ASan library from `make -C
coulomb-explosion_amd sanitize`; its first build compiles the device code too and
takes about two minutes).
"""
import os
import subprocess

import pytest

from conftest import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
SAN = os.path.join(CPP, "_san")
PKG_SAN = os.path.join(ROOT, "coulomb-explosion_amd", "_san")
SANITIZER_REPORT = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:",
                    "WARNING: ThreadSanitizer")


def sanitizer_env(device=False):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    if device:
        # No quarantine on runs that use the GPU: at process exit the HIP / HSA
        # runtimes free memory after the device runtime has unloaded, and a
        # quarantine recycle then hands ASan's device allocator a chunk it can no
        # longer release (CHECK "!dev_runtime_unloaded_" in
        # sanitizer_allocator_device.h, inside __cxa_finalize; seen once, with
        # whatever host chunks the run had left in the quarantine).
        env["ASAN_OPTIONS"] += ":quarantine_size_mb=0"
    env["LSAN_OPTIONS"] = "suppressions=" + os.path.join(CPP, "lsan_hip.supp")
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:exitcode=66"
    return env


def make(target, timeout):
    r = subprocess.run(["make", "-C", CPP, "-j8", target], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def run_clean(cmd, device=False, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=sanitizer_env(device), **kw)
    for tag in SANITIZER_REPORT:
        assert tag not in r.stderr, r.stderr[-6000:]
    return r


@pytest.fixture(scope="module")
def tsan_built():
    make("tsan", 300)


@pytest.fixture(scope="module")
def asan_built():
    make("asan", 900)


def test_record_writer_under_tsan(tsan_built, tmp_path):
    r = run_clean([os.path.join(SAN, "record_writer_stress"), str(tmp_path)], timeout=120)
    assert r.returncode == 0 and "record writer ok" in r.stdout, r.stderr[-3000:]


def test_threaded_cpu_baseline_under_tsan(tsan_built):
    r = run_clean([os.path.join(SAN, "cpu_baseline_tsan")], timeout=120)
    assert r.returncode == 0 and "cpu baseline ok" in r.stdout, r.stderr[-3000:]


def test_host_abi_under_asan_ubsan(asan_built):
    r = run_clean([os.path.join(SAN, "sanitize_driver")], timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "sanitize driver ok (host paths)" in r.stdout


@pytest.mark.parametrize("args,message", [
    (["--sims"], "unknown or incomplete argument"),
    (["--out", "x.bin", "--sims", "ten"], "bad value for --sims"),
    (["--out", "x.bin", "--bogus", "1"], "unknown or incomplete argument"),
    (["--out", "x.bin", "--stop-rule", "never"], "--stop-rule must be pe or tmax"),
    (["--out", "x.bin", "--masses", "1,2", "--charges", "1"], "--charges must match --masses"),
    (["--out", "x.bin", "--masses", "1,x", "--charges", "1,1"], "bad value for --masses"),
    (["--sims", "10"], "--out is required"),
])
def test_generator_arguments_under_asan_ubsan(asan_built, tmp_path, args, message):
    r = run_clean([os.path.join(PKG_SAN, "generate_dataset")] + args, timeout=60, cwd=tmp_path)
    assert r.returncode == 1 and message in r.stderr, (r.returncode, r.stderr[-3000:])
