"""The measurement kernels beside the step: the plain copy (HBM ceiling) and
the step's gathers and stores alone (the ceiling of its access pattern) run
and give positive times; the pull moves the bytes a copy of the same size
moves, so it takes no less than about the copy's time."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f64", "f32", "f16c"])
def test_pull_and_copy_times(gpu, storage):
    from lbmengine_hip import _capi
    n = 64
    elem = {"f64": 8, "f32": 4, "f16c": 2}[storage]
    bytes_one_way = n ** 3 * 19 * elem
    copy = _capi.measure_copy_ms(bytes_one_way, 5)
    pull = _capi.measure_pull_ms(_capi.STORAGE[storage], n, n, n, 5)
    assert math.isfinite(copy) and copy > 0
    assert math.isfinite(pull) and pull > 0
    assert pull > 0.5 * copy
    # a lattice whose rows take several blocks, and an odd one
    assert _capi.measure_pull_ms(_capi.STORAGE[storage], 300, 5, 3, 2) > 0
    assert _capi.measure_pull_ms(_capi.STORAGE[storage], 37, 11, 9, 2) > 0
