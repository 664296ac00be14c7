"""bench.py reads two figures of its JSON line from the committed rocprofv3
summaries (HBM traffic and counted FP64 MFMA work of the factorisation).
These CPU tests check that both summaries parse and that the figures agree
with what the counters fix: 32 FP64 FLOP per busy SIMD-cycle on every
factorisation kernel, so that the busy share equals TFLOP/s over the
78.6 TFLOP/s peak."""
import os
import re

import pytest

import bench


@pytest.mark.parametrize('section', ['one column', '256 columns'])
def test_traffic_and_mfma_summaries_parse(section):
    t = bench.factor_traffic(section)
    m = bench.factor_mfma(section)
    assert t is not None and t > 0
    assert m is not None and m['gflop'] > 0
    # busy share x peak = counted rate (both from the same dispatches)
    assert m['busy_frac'] * bench.FP64_PEAK / 1e12 == pytest.approx(m['tflops'], rel=0.02)


def test_mfma_counter_rate_is_the_published_peak():
    text = open(os.path.join(bench.ROOT, bench.MFMA_FILE)).read()
    rates = re.findall(r'^\| `k_\w+<[^|]*\| .* \| ([0-9.]+) \|$', text, flags=re.M)
    assert rates
    for r in rates:
        assert float(r) == pytest.approx(32.0)
    # 32 FLOP per SIMD-cycle x 1024 SIMDs x 2.4 GHz
    assert 32 * 1024 * 2.4e9 == pytest.approx(bench.FP64_PEAK, rel=1e-3)
