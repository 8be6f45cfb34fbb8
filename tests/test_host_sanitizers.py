"""The host code of liblpe under AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU): the library is rebuilt with host-only sanitizers
(co-cultivation-metabolic-atlas_amd/csrc/sanitize.sh, device code unchanged) and
the host tests - the C ABI's symbol and argument checks (test_host.py) and the
reduction plan with its on_device = 0 reduce / expand / compress / lift paths
against the oracle (test_reduction_host.py) - run against it in a child process
with the sanitizer runtime preloaded; any report aborts that process."""

import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

SCRIPT = os.path.join(ROOT, "co-cultivation-metabolic-atlas_amd", "csrc", "sanitize.sh")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.slow
def test_host_paths_under_asan_ubsan(tmp_path):
    if not (os.path.exists(HIPCC) and shutil.which("bash")):
        pytest.skip("hipcc is needed to build the sanitizer library")
    lib = str(tmp_path / "liblpe_asan.so")
    subprocess.run(["bash", SCRIPT, lib], check=True, capture_output=True, timeout=900)
    runtime = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"], check=True,
                             capture_output=True, text=True).stdout.strip()
    assert os.path.exists(runtime), runtime
    env = dict(os.environ, LD_PRELOAD=runtime, LPE_LIBRARY=lib,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the child reports which library it mapped, so a silent fallback to the
    # uninstrumented build cannot pass
    probe = ("import sys; sys.path.insert(0, %r); from lpe import _lib; _lib.load(); "
             "print(any(%r in l for l in open('/proc/self/maps')))"
             % (os.path.join(ROOT, "co-cultivation-metabolic-atlas_amd"), lib))
    out = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "True", out.stderr[-2000:]
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                          os.path.join(ROOT, "tests", "test_host.py"),
                          os.path.join(ROOT, "tests", "test_reduction_host.py")],
                         env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, (res.stdout[-3000:], res.stderr[-3000:])
    assert "passed" in res.stdout
