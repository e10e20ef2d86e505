"""The CPU-side C/C++ under sanitizers: the oracle (OpenMP ScatterSearch
restatement, search, brute force) with ThreadSanitizer + Archer and with
AddressSanitizer + UndefinedBehaviorSanitizer, and the library's host code
(file formats, graph builder) with ASan + UBSan and TSan.  Each make target
builds its own driver and fails on any sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(directory):
    r = subprocess.run(["make", "-C", os.path.join(ROOT, directory), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/lib/libarcher.so") or shutil.which("gcc") is None,
                    reason="needs LLVM's clang with the Archer OpenMP tool and gcc")
def test_oracle_under_sanitizers():
    out = _make("oracle")
    assert out.count("oracle sanitizer driver: ok") == 2


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_sanitizers():
    out = _make("scattersearch_amd/csrc")
    assert out.count("host sanitizer driver: ok") == 2
