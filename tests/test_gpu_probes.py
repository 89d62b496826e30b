"""The measurement helpers (sph_bandwidth_probe, sph_flops_probe) run and
report rates inside the MI355X's published peaks."""

import pytest

from sphere_amd.dem import SphereInvalid, bandwidth_probe, flops_probe

pytestmark = pytest.mark.gpu


def test_bandwidth_probe_within_peak():
    gbs = bandwidth_probe(1 << 28, 5)
    assert 500.0 < gbs < 8000.0


@pytest.mark.parametrize("fp64,peak", [(True, 78.6), (False, 157.3)])
def test_flops_probe_within_peak(fp64, peak):
    tf = flops_probe(fp64, blocks=2048, iters=1024, reps=3)
    assert 1.0 < tf <= peak


def test_probes_refuse_empty_work():
    with pytest.raises(SphereInvalid):
        flops_probe(True, blocks=0, iters=16, reps=1)
    with pytest.raises(SphereInvalid):
        bandwidth_probe(0, 1)
