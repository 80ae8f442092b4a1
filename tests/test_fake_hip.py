"""The host runtime (engine.c, serve.c, autotune.c, xgpu_api.c) against a CPU
fake of the HIP layer, built with AddressSanitizer and UBSan.

tests/fake/xg_hip_fake.c implements every function of xgpu_amd/csrc/xg_hip.h
on the host: allocations, copies and kernels run synchronously (kernels as
plain C restatements), stream capture records a replayable graph, and the
n-th call of any category (malloc, copy, launch, event, stream, sync, graph)
can be made to fail.  The scenarios check that every failure comes back as the
reference's XGPU_CUDA_ERROR (reference src/cuda_xengine.cu:94-101, there with
exit()) or an allocation error, that nothing leaks (LeakSanitizer), and that a
retry afterwards is exact; a randomized serving scenario (random stream
shapes, host or device blocks and dump rings, server limits, two submitting
threads) checks every dump.  The same scenarios also run under
ThreadSanitizer.  CPU only.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FAKE = os.path.join(HERE, "fake")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_host_runtime_under_fault_injection():
    b = subprocess.run(["make", "-s", "-C", FAKE], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(FAKE, "test_fault_injection")], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all fault-injection scenarios passed" in r.stdout
    for scenario in ("xgpu.h pipeline", "xgpuInit fails cleanly", "xgpuCudaXengine returns",
                     "xgpuAmdCreate fails cleanly", "spill to int64", "graph steps, batch",
                     "server: exact dumps", "randomized serving"):
        assert "ok " + scenario in r.stdout, r.stdout
    assert "runtime error" not in out  # UBSan


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_host_runtime_under_thread_sanitizer():
    """The same scenarios, the server's threads included, under TSan."""
    b = subprocess.run(["make", "-s", "-C", FAKE, "test_fault_injection_tsan"], capture_output=True,
                       text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    r = subprocess.run([os.path.join(FAKE, "test_fault_injection_tsan")], capture_output=True,
                       text=True, timeout=900,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all fault-injection scenarios passed" in r.stdout
    assert "ThreadSanitizer" not in out


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_trace_switch_logs_device_work(tmp_path):
    """XGPU_TRACE=<file>: one line per launch, capture, replay and dump."""
    b = subprocess.run(["make", "-s", "-C", FAKE], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    log = tmp_path / "trace.log"
    r = subprocess.run([os.path.join(FAKE, "test_fault_injection"), "trace", str(log)],
                       capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = log.read_text().splitlines()
    assert all(l.startswith("[xgpu-trace]") for l in lines)
    kinds = [l.split()[3] for l in lines]
    assert kinds == ["launch", "dump", "capture", "replay", "replay"], lines
    assert "kernel=xg_xengine_rgd<32,8>" in lines[0] and "nf=3 ns=8 nt=64 store" in lines[0]
