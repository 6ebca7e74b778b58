"""Host-side bookkeeping of bench.py (no GPU): the dominant kernel of the timed region,
the committed traffic records it reads, and the one-GPU guard."""
from __future__ import annotations

import importlib.util
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_dominant_launch_is_the_depth_with_the_most_time(bench):
    # a 20-step run: two passes of 6 and the lag-writing pass of 8
    d, n, ms = bench.dominant_launch([(6, 2.0), (6, 2.2), (8, 3.0)])
    assert (d, n) == (6, 2) and ms == pytest.approx(2.1)
    # a long run: passes of 8 dominate
    d, n, ms = bench.dominant_launch([(8, 2.5)] * 49 + [(8, 3.0)])
    assert (d, n) == (8, 50) and ms == pytest.approx((49 * 2.5 + 3.0) / 50)
    # single steps only
    assert bench.dominant_launch([(1, 1.6)] * 20)[:2] == (1, 20)


def test_traffic_records_match_the_kernel_they_name(bench):
    for ns, name in ((1, "traffic_trt_f64_8192.json"), (6, "traffic_trt_f64_8192_pass6.json"),
                     (8, "traffic_trt_f64_8192_pass8.json")):
        rec = json.loads((ROOT / "profiles" / name).read_text())
        got = bench.traffic_from_profiles(8192, ns)
        assert got == rec["bytes_per_launch"]
        # 144 B per cell per launch, whatever the depth
        assert rec["algorithmic_bytes_per_launch"] == 144 * 8192 * 8192
        want = "k_fused<1, 0>" if ns == 1 else f"k_multi<1, 0, {ns}>"
        assert want in rec["kernel"]
        assert 0.95 < got / rec["algorithmic_bytes_per_launch"] < 1.2
    assert bench.traffic_from_profiles(8192, 5) is None  # no record: null, not a guess


def test_bytes_per_lattice_update(bench):
    assert bench.BYTES_PER_LU == {"f64": 144, "f64s": 144, "f32s": 72, "f16s": 36}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
