"""The bandwidth probes the benchmark reports beside stream_collide run on
the device and return physical rates; invalid arguments are rejected before
any device call (that check runs without a GPU)."""
import pytest

import ionsolver_amd as I

PEAK_GBS = 8000.0  # published HBM peak of the device


@pytest.mark.gpu
def test_copy_stream_and_pull_probes_return_rates_below_the_hbm_peak():
    ms, gbs = I.hbm_copy_probe(1 << 26, 2)
    assert ms > 0 and 0 < gbs < 1.05 * PEAK_GBS
    for elem in (2, 4):
        for streams in (19, 45):
            ms, gbs = I.hbm_stream_probe(elem, streams, 1 << 18, 2)
            assert ms > 0 and 0 < gbs < 1.05 * PEAK_GBS
            ms, gbs = I.hbm_pull_probe(elem, streams, 64, 48, 32, 2)
            assert ms > 0 and 0 < gbs < 1.05 * PEAK_GBS


@pytest.mark.parametrize("args", [(3, 19, 8, 8, 8), (2, 20, 8, 8, 8), (2, 19, 0, 8, 8), (4, 45, 2048, 1024, 1024),
                                  (2, 19, 1 << 22, 1 << 22, 1 << 22)])  # 2^66 cells: 4 after a 64-bit wrap
def test_pull_probe_rejects_invalid_arguments(args):
    with pytest.raises(I.IonError, match="ion_hbm_pull_probe"):
        I.hbm_pull_probe(*args, reps=1)
