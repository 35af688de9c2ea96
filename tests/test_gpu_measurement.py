"""Measurement hooks of the test header (nlos_engine_test.h): the calibration copy kernel and
the FP16S sweep's access pattern alone.  They return positive mean times for small sizes
and refuse bad requests."""
from __future__ import annotations

import pytest

import nlos_engine as nlos

pytestmark = pytest.mark.gpu


def test_copy_kernel_times(gpu):
    assert nlos.copy_kernel_ms(1 << 20, 2) > 0.0


@pytest.mark.parametrize("dims", [(64, 8, 4), (100, 3, 5), (1, 1, 1)])
def test_pull_pattern_times(gpu, dims):
    assert nlos.pull_pattern_ms(*dims, reps=2) > 0.0


def test_bad_requests_are_refused(gpu):
    with pytest.raises(nlos.EngineError) as ex:
        nlos.pull_pattern_ms(0, 8, 4)
    assert ex.value.code == 1
    with pytest.raises(nlos.EngineError):
        nlos.copy_kernel_ms(8)
