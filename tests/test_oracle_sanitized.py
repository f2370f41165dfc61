"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's sources built with both (oracle/Makefile, liboracle_san.so, host code
only) and a set of the oracle's own tests run against that build in a child
process.  An out-of-bounds access, a use after free or undefined behaviour
ends the child with a sanitizer report and fails this test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = ["test_oracle_pins.py", "test_oracle_eos.py", "test_oracle_axes.py", "test_oracle_llf_interface.py",
         "test_oracle_godunov.py", "test_oracle_levelset.py", "test_oracle_rnormal.py", "test_oracle_shyue98.py"]


def _runtime(name):
    try:
        out = subprocess.run(["g++", "-print-file-name=" + name], capture_output=True, text=True, timeout=60)
    except (OSError, subprocess.SubprocessError):
        return None
    path = out.stdout.strip()
    return path if out.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def _sanitized_env():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no AddressSanitizer runtime for g++ here")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "liboracle_san.so"], check=True, timeout=900)
    lib = os.path.join(ROOT, "oracle", "liboracle_san.so")
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + ((" " + env["LD_PRELOAD"]) if env.get("LD_PRELOAD") else "")
    # Python itself is not instrumented: no leak report at exit, and the
    # runtime need not come first among the process's libraries
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["M2C_ORACLE_LIB"] = lib
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "tests")])
    return env


def test_oracle_tests_under_asan_and_ubsan():
    env = _sanitized_env()
    files = [os.path.join(ROOT, "tests", f) for f in TESTS if os.path.exists(os.path.join(ROOT, "tests", f))]
    # the two longest decks (80 s and 12 s unsanitized, several times that
    # under the sanitizers) are left to the plain run
    skip = "not SlottedDisk and not shyue98_first_time_steps"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-k", skip] + files,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    report = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report


def test_the_sanitizer_reports_an_overrun():
    """The harness itself: a call that makes the oracle write past the end of
    its output array (3 fluxes asked for, room for 1) is reported."""
    env = _sanitized_env()
    code = ("import numpy as np, ctypes as C, oracle_lib as ol\n"
            "from m2c_amd import abi\n"
            "L = ol.lib()\n"
            "assert L._name.endswith('liboracle_san.so'), L._name\n"
            "m = abi.materials(abi.material('SG', gamma=1.4))\n"
            "s = abi.default_scheme()\n"
            "V = np.ones((3, 5)); ids = np.zeros(3, np.int32); F = np.zeros((1, 5))\n"
            "L.orc_flux_batch(m, 1, C.byref(s), 3, 0, ol.dp(V), ol.dp(V), ol.ip(ids), ol.dp(F))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr, r.stderr[-3000:]
