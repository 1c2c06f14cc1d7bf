"""Host-side argument handling of the static solver (no GPU calls)."""
import pytest

from fea_amd.solver.static_implicit import predictor_order


@pytest.mark.parametrize('arg,order', [(False, 0), (True, 1), ('none', 0), ('linear', 1), ('quadratic', 2),
                                       ('cubic', 3), (0, 0), (4, 4), (6, 6)])
def test_predictor_order(arg, order):
    assert predictor_order(arg) == order


@pytest.mark.parametrize('arg', ['quartic', 7, -1, 2.0, None])
def test_predictor_order_rejects(arg):
    with pytest.raises(ValueError):
        predictor_order(arg)


def test_auto_predictor_caps_at_six():
    assert predictor_order('auto') == 6
