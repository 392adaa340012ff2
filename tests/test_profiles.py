"""The committed rocprofv3 summaries parse, and what they record is consistent:
the HIP-event average of the roofline kernel (bench.py without the profiler)
agrees with rocprofv3's average over the same timed region, and the
counter-measured HBM traffic is within the algorithmic bytes' range."""

import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["irfft_mdot_twostream_1024x2048_J4.txt", "maxpy_norm_twostream_1024x2048_J4.txt",
                                  "jac_rfft_twostream_1024x2048_J4.txt"])
def test_profile_summary(name):
    from bench_profiles import read_summary
    k = read_summary(os.path.join(ROOT, "profiles", name))
    for key in ("kernel", "rocprof_avg_us_timed_region", "bench_avg_us_unprofiled", "traffic_bytes_per_launch",
                "fetch_factor", "write_factor", "traffic_over_algorithmic", "pmc_algorithmic_bytes_per_launch"):
        assert key in k, key
    assert abs(k["bench_avg_us_unprofiled"] - k["rocprof_avg_us_timed_region"]) <= 0.1 * k["rocprof_avg_us_timed_region"]
    assert 1.9 < k["fetch_factor"] < 2.1 and 0.95 < k["write_factor"] < 1.05
    assert k["traffic_bytes_per_launch"] > 0
    # counter bytes against the algorithmic bytes of the same dispatches
    assert 0.9 < k["traffic_over_algorithmic"] < 1.2
    assert "command: cd /tmp && export TMPDIR=/tmp && rocprofv3" in open(os.path.join(ROOT, "profiles", name)).read()


def test_counter_summary_covers_the_step_kernels():
    text = open(os.path.join(ROOT, "profiles", "counters_twostream_1024x2048_J4.txt")).read()
    assert "command: cd /tmp && export TMPDIR=/tmp && rocprofv3 --pmc" in text
    for k in ("maxpy_norm_kernel", "irfft_mdot_kernel", "jac_rfft_kernel", "bracket_kernel<4, 0",
              "tridiag_lu_kernel"):
        assert k in text, k


def test_batched_busy_summary():
    text = open(os.path.join(ROOT, "profiles", "batched_ensemble_256_busy.txt")).read()
    assert "command: cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace" in text
    rows = [ln.split() for ln in text.splitlines() if ln.strip() and ln.split()[0].replace('.', '').isdigit()]
    assert len(rows) > 20
