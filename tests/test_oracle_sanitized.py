"""The oracle's own checks, run against other builds of the same source
(CPU only):

* the whole oracle suite against the AddressSanitizer + UndefinedBehavior-
  Sanitizer build (`make -C oracle asan`), libasan preloaded into the test
  process, aborting on the first finding;
* a build with a planted mistake (the f64 walk taking its visibility from
  the f32 angle, `-DTVSO_PLANTED_F32_VISIBILITY`), which the constructed
  flip test must fail.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
GCC = shutil.which("gcc")

pytestmark = pytest.mark.skipif(GCC is None, reason="gcc not installed")


def _libasan() -> str | None:
    path = subprocess.run([GCC, "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def _run_oracle_tests(lib: str, extra_env: dict, *pytest_args: str):
    env = dict(os.environ)
    env.update(extra_env)
    env["TVS_ORACLE_LIB"] = lib
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                           os.path.join(ROOT, "tests", "test_oracle.py"), *pytest_args],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


def test_oracle_suite_under_asan_and_ubsan():
    asan = _libasan()
    if asan is None:
        pytest.skip("libasan not available")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    r = _run_oracle_tests(os.path.join(ORACLE, "libtvs_oracle_asan.so"),
                          {"LD_PRELOAD": asan,
                           "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out


def test_planted_f32_visibility_in_the_f64_walk_is_caught(tmp_path):
    lib = str(tmp_path / "libtvs_oracle_planted.so")
    subprocess.run([GCC, "-O2", "-std=c11", "-D_DEFAULT_SOURCE", "-ffp-contract=off",
                    "-fno-fast-math", "-fPIC", "-fopenmp", "-DTVSO_PLANTED_F32_VISIBILITY",
                    "-shared", "-o", lib, os.path.join(ORACLE, "tvs_oracle.c"), "-lm"],
                   check=True)
    r = _run_oracle_tests(lib, {}, "-k", "f64_constructed_visibility_flip")
    out = r.stdout + r.stderr
    assert r.returncode == 1 and "1 failed" in out, out[-4000:]
    # The same tests pass on the faithful build.
    r = _run_oracle_tests(os.path.join(ORACLE, "libtvs_oracle.so"), {}, "-k",
                          "f64_constructed_visibility_flip")
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
