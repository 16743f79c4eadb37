"""The measured-peak probes the rooflines divide by: the matrix-unit probe
and the stage-pattern HBM probe return rates in a plausible band for
MI355X (published: 78.6 TFLOP/s FP64, 8 TB/s HBM)."""
import pytest

pytestmark = pytest.mark.gpu


def test_mfma_probe_fp64(fm):
    r = fm.probe_mfma_peak("fp64", 0.01)
    assert 20.0 < r < 100.0


def test_stream_probe_stage_pattern(fm):
    # 4 source slabs, 16 output slabs of 2^20 doubles (160 MiB)
    r = fm.probe_stream(4, 16, 1 << 20)
    assert 500.0 < r < 10000.0
    # a write-only pattern
    w = fm.probe_stream(0, 8, 1 << 20)
    assert 500.0 < w < 10000.0


def test_stream_probe_rejects_bad_sizes(fm):
    with pytest.raises(Exception):
        fm.probe_stream(4, 0, 1 << 20)
    with pytest.raises(Exception):
        fm.probe_stream(4096, 4096, 1 << 30)
