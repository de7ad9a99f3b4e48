"""The library's host-only logic under sanitizers, on the CPU.

tests/cpp/host_logic_test.cpp drives the serving loop's scheduling
(tipl_amd/csrc/server_core.hpp: admission, slots, tickets, the batcher,
failure isolation, shutdown) through a stub device backend, and the host's
wait for results that kernels store into host memory
(tipl_amd/csrc/host_wait.hpp).  It is built twice with g++: with
AddressSanitizer + UndefinedBehaviorSanitizer (lifetimes of request buffers,
tickets and batched jobs; leaks) and with ThreadSanitizer (races between
clients, workers and launchers)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_logic_test.cpp")


def _build_and_run(tmp_path, flags, env_extra):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_logic_test")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"), *flags, SRC, "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host logic ok" in r.stdout
    return r


def test_host_logic_asan_ubsan(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
                   {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
                    "UBSAN_OPTIONS": "print_stacktrace=1"})


def test_host_logic_tsan(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"],
                   {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"})
