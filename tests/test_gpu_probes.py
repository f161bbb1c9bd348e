"""The measurement probes bench.py reports next to the datasheet peaks: a
streaming copy and independent FP64 fma chains.  Plausibility only: positive
and below the MI355X datasheet figures (8 TB/s, 78.6 TFLOP/s FP64)."""
import pytest

import xfluids_amd as xf
from parity import announce

pytestmark = pytest.mark.gpu


def test_fma_probe_is_plausible():
    announce("fma probe")
    tf = xf.fma_probe(3)
    print("FP64 fma rate %.1f TFLOP/s" % tf)
    assert 1.0 < tf < 78.6 * 1.05


def test_copy_probe_is_plausible():
    announce("copy probe")
    ms = xf.copy_probe(1 << 28, 3)
    gbs = 2 * (1 << 28) / (ms * 1e-3) / 1e9
    print("copy %.0f GB/s" % gbs)
    assert 100.0 < gbs < 8000.0 * 1.05
