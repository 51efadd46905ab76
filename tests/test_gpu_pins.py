"""The hand-computed load-control cases of tests/test_oracle_pins.py on the device
path (gf_assemble through the C ABI): table interpolation, the gravity ramp over
the solution window and Shell_1's doubled self-weight."""
import numpy as np
import pytest

from giraffe_amd import runtime
from test_oracle_pins import TABLE_CASES, assemble_at, flat_plate, gravity_model, table_model

pytestmark = pytest.mark.gpu


def test_table_value_device():
    m, tip = table_model()
    for t, fx in TABLE_CASES:
        r = assemble_at(runtime.DeviceModel, m, t)
        assert r[tip, 0] == pytest.approx(fx, rel=1e-14, abs=1e-14), t
        assert r[tip, 5] == pytest.approx(10.0 * fx, rel=1e-14, abs=1e-13), t
        assert np.abs(np.delete(r.ravel(), [6 * tip, 6 * tip + 5])).max() == 0.0


def test_gravity_ramp_device():
    m, W = gravity_model()
    for t, factor in [(0.5, 0.0), (1.0, 0.25), (1.5, 0.5), (2.0, 0.75), (2.5, 1.0)]:
        r = assemble_at(runtime.DeviceModel, m, t, window=(0.5, 2.5))
        assert r[:, 2].sum() == pytest.approx(factor * W, rel=1e-13, abs=1e-9), t


def test_shell_self_weight_device():
    m, W2 = flat_plate()
    r = assemble_at(runtime.DeviceModel, m, 1.0)
    assert r[:, 2].sum() == pytest.approx(W2, rel=1e-12)
