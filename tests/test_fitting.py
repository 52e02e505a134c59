"""Host logic of the windowed fitting driver: the bounded adaptive-magnitude
update (parameter_fitting_functions.py:60-92), checked through properties the
reference's algorithm guarantees rather than by restating it."""
import math

import numpy as np
import pytest

from diffcp_amd.fitting import magnitude_level, single_parameter_fitting


@pytest.mark.parametrize("a,level", [(1234.0, 3), (1.0, 0), (0.05, -2), (-0.0049, -3), (9.99e5, 5)])
def test_magnitude_level_closed_form(a, level):
    assert magnitude_level(a) == level


def test_zero_gradient_is_a_no_op():
    assert single_parameter_fitting(500.0, 0.0, 10.0, 1000.0, 1.0, 10, 3) == (500.0, 1.0)


def test_updates_stay_inside_the_bounds_and_oppose_the_gradient():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        lo, hi = sorted(rng.uniform(-1e3, 1e4, 2))
        if hi - lo < 1e-3:
            continue
        p = rng.uniform(lo, hi)
        if min(p - lo, hi - p) < 1e-9 * (hi - lo):
            continue
        g = float(rng.choice([-1, 1]) * 10.0 ** rng.uniform(-12, 6))
        lr = float(10.0 ** rng.uniform(-6, 6))
        epochs = int(rng.integers(2, 30))
        i = int(rng.integers(0, epochs))
        p2, lr2 = single_parameter_fitting(p, g, lo, hi, lr, epochs, i)
        assert lo <= p2 <= hi
        assert (p2 - p) * g <= 0          # moves against the gradient (or stays)
        assert lr2 > 0 and math.isfinite(lr2)


def test_first_epoch_step_has_the_order_of_the_remaining_distance():
    # first epoch: the step is rescaled to 10^level((bound - p) / epochs)
    for p, g, lo, hi, epochs in ((100.0, -3e-7, 10.0, 1000.0, 9), (0.1, 5.0, 0.05, 0.45, 4),
                                 (2000.0, 1e-3, 10.0, 5000.0, 20)):
        p2, _ = single_parameter_fitting(p, g, lo, hi, 1.0, epochs, 0)
        bound = hi if g < 0 else lo
        assert magnitude_level(abs(p2 - p)) == magnitude_level(abs(bound - p) / epochs)


class _FakeEngine:
    """Two instances; the last window's raw dL/dE of instance 0 is NaN."""
    B = 2

    def set_params(self, p):
        self.params = np.array(p)

    def reset(self, slot, **kw):
        pass

    def set_targets(self, t):
        pass

    def rollout_grad(self, S, g, K, use_graph=True):
        grads = np.full((2, 6), 1e-3)
        if g == 1:
            grads[0, 0] = np.nan
        return np.array([1.0, 1.0]), grads, 0.0

    def download_state(self, slot):
        return {k: np.zeros((2, 3)) for k in ("vertex_x", "grasper_x", "cube_x")}

    def last_final_slot(self):
        return 0

    def split(self, arr, attr):
        return [arr[:1], arr[1:]]


def test_nan_gradient_of_the_last_window_stops_the_candidate():
    """parameter_fitting_functions.py:148-150 stops on a NaN garment_E.grad of
    the last window, which compute_grad's nan_to_num average would hide."""
    from diffcp_amd.fitting import WindowedFit
    t = np.zeros((4, 3))
    drv = WindowedFit(_FakeEngine(), [[t, t], [t, t]], substeps=10)
    p0 = np.array([[500.0, 0.3, 100.0, 100.0, 0.0, 0.0]] * 2)
    res = drv.parameters_learning(p0, epochs=3)
    h = np.stack(res.history)
    assert (h[:, 0] == p0[0]).all()              # stopped from the first epoch on
    assert (h[1:, 1, :2] != p0[1, :2]).all()     # the other candidate keeps fitting
