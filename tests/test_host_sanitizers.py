"""Host code of libatmo_hip under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/asan/Makefile builds every source of the library with the sanitizers on
the host side (the device code is compiled as usual) and links the checking
program tests/asan/host_checks.cpp, which runs the mesh builders at every polydeg and
the configuration and connectivity checks of the three create calls on inputs
they must refuse; nothing in it touches a device, so it runs on the CPU.  The
sanitizers abort that program on a memory or undefined-behaviour error.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "asan")


@pytest.mark.slow
@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_host_code_under_sanitizers():
    jobs = str(min(8, os.cpu_count() or 1))
    build = subprocess.run(["make", "-j", jobs, "-C", HERE], capture_output=True, text=True,
                           timeout=900)
    assert build.returncode == 0, build.stdout[-2000:] + build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", HIP_VISIBLE_DEVICES="")
    run = subprocess.run([os.path.join(HERE, "build", "host_checks")], capture_output=True,
                         text=True, timeout=600, env=env)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-6000:]
    assert "host checks passed" in run.stdout
