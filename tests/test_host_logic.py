"""The host library's logic without a GPU: mpfft.cpp and batch.cpp built
against a stand-in HIP runtime (tests/host_stub: allocations are host
memory, kernel launchers touch exactly the buffer ranges a launch reads or
writes, graph capture records and replays them) under AddressSanitizer and
UBSan; checks the batched form's packed layout, tuning, capture and replay,
every plan form's buffers, and the pass-plan invariants the kernels rely
on (kernels.h) for every size and compiled width."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    if subprocess.run(["which", "g++"], capture_output=True).returncode:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("host") / "host_check"
    src = os.path.join(ROOT, "mpfft_amd", "csrc")
    stub = os.path.join(ROOT, "tests", "host_stub")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", f"-I{stub}", "-I/opt/conda/include",
           f"-I{os.path.join(ROOT, 'include')}", "-x", "c++",
           os.path.join(src, "mpfft.cpp"), os.path.join(src, "batch.cpp"),
           os.path.join(stub, "stub.cpp"), os.path.join(stub, "host_check.cpp"),
           "-x", "none", "/lib/x86_64-linux-gnu/libmpfr.so.6",
           "/lib/x86_64-linux-gnu/libgmp.so.10", "-lpthread", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    return str(exe)


def _run(exe, *args, **env):
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=600,
                          env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0", **env))


def test_host_logic_under_asan(host_check):
    r = _run(host_check)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().endswith("PASS")


def test_legacy_precision_mismatch_reported(host_check):
    # mp_sr24_cp_fft with outi at another precision: recorded
    # (MPFFT_ERR_PREC), printed to stderr, no abort
    r = _run(host_check, "prec")
    assert r.returncode == 0, r.stderr[-2000:]
    assert "last_error -2" in r.stdout
    assert "mpfft: mp_sr24_cp_fft:" in r.stderr


def test_legacy_hip_error_aborts(host_check):
    # a HIP error inside a reference entry point aborts the process after
    # printing the error; with MPFFT_NO_ABORT=1 it is recorded instead
    r = _run(host_check, "hip")
    assert r.returncode == -6, (r.returncode, r.stdout[-1000:], r.stderr[-2000:])
    assert "mpfft: mp_sr24_cp_fft:" in r.stderr
    r = _run(host_check, "hip", MPFFT_NO_ABORT="1")
    assert r.returncode == 0, r.stderr[-2000:]
    assert "last_error -5" in r.stdout


def test_batches_from_threads_under_tsan(tmp_path):
    # the batched form from two host threads at once (each batch also builds
    # its plans on threads), under ThreadSanitizer
    if subprocess.run(["which", "g++"], capture_output=True).returncode:
        pytest.skip("g++ not available")
    exe = tmp_path / "host_check_tsan"
    src = os.path.join(ROOT, "mpfft_amd", "csrc")
    stub = os.path.join(ROOT, "tests", "host_stub")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", f"-I{stub}",
           "-I/opt/conda/include", f"-I{os.path.join(ROOT, 'include')}", "-x", "c++",
           os.path.join(src, "mpfft.cpp"), os.path.join(src, "batch.cpp"),
           os.path.join(stub, "stub.cpp"), os.path.join(stub, "host_check.cpp"),
           "-x", "none", "/lib/x86_64-linux-gnu/libmpfr.so.6",
           "/lib/x86_64-linux-gnu/libgmp.so.10", "-lpthread", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe), "threads"], capture_output=True, text=True, timeout=900,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-5000:]
    assert r.stdout.strip().endswith("PASS")
