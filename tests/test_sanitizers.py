"""CPU-side sanitizer runs: the C restatement of the oracle and the host code of
the library's C-ABI, each built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on the CPU (no GPU call is made: the library
checks exercise argument validation, workspace arithmetic, the z grid and the
host-side planning of request batches)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, cwd, timeout):
    res = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    return res.stdout


@pytest.mark.skipif(shutil.which('gcc') is None, reason='gcc not found')
def test_oracle_c_restatement_under_asan_ubsan():
    out = _run(['make', '-s', 'sanitize'], os.path.join(ROOT, 'oracle'), 300)
    assert 'sancheck passed' in out


@pytest.mark.skipif(not os.path.exists('/opt/rocm/bin/hipcc'), reason='hipcc not found')
def test_library_host_code_under_asan_ubsan():
    out = _run(['make', '-s', '-f', 'Makefile.sanitize', 'run'], os.path.join(ROOT, 'pytransit_amd', 'csrc'), 600)
    assert 'host checks passed' in out
