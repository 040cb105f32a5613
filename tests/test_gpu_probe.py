"""The bandwidth probe the bench reports beside the roofline (GPU)."""
import pytest

pytestmark = pytest.mark.gpu


def test_hbm_probe_reports_plausible_bandwidth(ctx):
    read, copy = ctx.hbm_probe(1 << 28)
    # MI355X: 8 TB/s peak; streaming kernels reach a large fraction of it
    assert 1000 < read < 8000 and 1000 < copy < 8000, (read, copy)
