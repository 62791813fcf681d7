"""The host code (scene reader/builder, residency manager, feedback update,
legacy files, C-ABI argument checks) and the CPU oracle run the CPU test
files under AddressSanitizer and UndefinedBehaviorSanitizer without a
report: both libraries are rebuilt with -fsanitize=address,undefined
(`make sanitize`) and loaded through SVO_HIP_LIBRARY / SVO_ORACLE_LIBRARY
into a python process with libasan preloaded."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["test_format.py", "test_builder.py", "test_residency.py", "test_residency_model.py", "test_feedback.py",
         "test_single_svo.py", "test_capi.py", "test_oracle_cpu.py", "test_view_modes.py", "test_golden.py"]


def test_cpu_suite_under_asan_and_ubsan():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    asan = subprocess.run([cxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan not installed")
    for d in ("libsvo_amd", "oracle"):
        subprocess.run(["make", "-C", os.path.join(ROOT, d), "-j", "8", "sanitize"], check=True, capture_output=True)
    env = dict(os.environ)
    env.update({
        "SVO_HIP_LIBRARY": os.path.join(ROOT, "libsvo_amd", "_san", "libsvo_hip.so"),
        "SVO_ORACLE_LIBRARY": os.path.join(ROOT, "oracle", "_san", "libsvo_oracle.so"),
        "LD_PRELOAD": asan,
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"]
    cmd += [os.path.join(ROOT, "tests", f) for f in FILES]
    run = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (run.stdout + run.stderr)[-4000:]
    assert run.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, tail
