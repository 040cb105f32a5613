"""The bandwidth probes the bench reports beside the roofline (GPU)."""
import pytest

pytestmark = pytest.mark.gpu


def test_hbm_probe_reports_plausible_bandwidth(ctx):
    read, copy = ctx.hbm_probe(1 << 28)
    # MI355X: 8 TB/s peak; streaming kernels reach a large fraction of it
    assert 1000 < read < 8000 and 1000 < copy < 8000, (read, copy)


def test_lds_probe_reports_plausible_bandwidth(ctx):
    read, peak = ctx.lds_probe()
    # the peak is 256 bytes per clock per CU (gfx950): about 157 TB/s on 256
    # CUs at 2.4 GHz; the probe reaches a sizeable part of it and never
    # exceeds it (MI355X: about 107 TB/s)
    assert 50000 < peak < 300000, peak
    assert 0.2 * peak < read <= 1.02 * peak, (read, peak)
