"""bench.py's command line and bookkeeping, on the CPU: a multi-GPU request
is refused with a message, a failing side leg is recorded in the JSON line
instead of ending the run, and the roofline's product count of the plain
form equals the reference's own multiplication counter
(serial_mp_fft3.c:123-250, through the oracle pinned to it)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from oracle import oracle as O  # noqa: E402


def test_multi_gpu_request_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_side_leg_records_failure():
    line = {}
    bench.side_leg(line, "wide", lambda: 1 / 0)
    assert "ZeroDivisionError" in line["wide"]["error"]
    bench.side_leg(line, "batched", lambda: {"fft_per_s": 1.0})
    assert line["batched"] == {"fft_per_s": 1.0}
    bench.side_leg(line, None, lambda: {"cpu_baseline": {"value": 2.0}, "cpu_port": {}})
    assert line["cpu_baseline"] == {"value": 2.0} and "cpu_port" in line
    bench.side_leg(line, None, lambda: [][1])
    assert "IndexError" in line["cpu_baseline"]["error"]


@pytest.mark.parametrize("n", [8, 64, 1024, 1 << 16])
def test_plain_product_count_matches_reference_counter(n):
    assert bench.cmul(n) == O.counts(n)[0]
