"""The host C++ of the library (initial partitioner, pairwise FM, perfectly balanced post-pass)
and of the oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer (g++, no GPU code)
and run over random graphs with isolated nodes, hubs, node and edge weights (tests/sanitize/
host_check.cpp). A memory or undefined-behaviour error aborts the program."""
import os
import subprocess

import pytest

from conftest import ROOT

SOURCES = [
    "tests/sanitize/host_check.cpp",
    "kahip_amd/csrc/initial_partition.cpp",
    "kahip_amd/csrc/fm_refine.cpp",
    "kahip_amd/csrc/rebalance.cpp",
    "oracle/kahip_oracle.cpp",
]


@pytest.mark.slow
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "kahip_amd", "csrc"), "-I" + os.path.join(ROOT, "oracle"),
           *[os.path.join(ROOT, s) for s in SOURCES], "-o", exe]
    subprocess.check_call(cmd)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=900)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "host_check ok" in out.stdout
    assert "runtime error" not in out.stderr


@pytest.mark.slow
def test_sanitizer_build_reports_a_planted_overflow(tmp_path):
    """The same flags turn a heap overflow into a failing run (the check above is live)."""
    src = tmp_path / "oob.cpp"
    src.write_text("#include <vector>\nint main(int c, char**) { std::vector<int> v(4); "
                   "return v.data()[4 + c]; }\n")
    exe = str(tmp_path / "oob")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           str(src), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), timeout=60)
    assert out.returncode != 0 and "heap-buffer-overflow" in out.stderr
