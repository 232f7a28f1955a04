"""The reference's input-validation tests (fastLPR_py/tests/test_errors.py),
against this package: the same inputs raise the same exception types with the
same message keywords, before any device work (these run without a GPU).  One
deliberate difference: d = 4..10 runs in the reference and raises
NotImplementedError here (the engine covers 1-3 dimensions)."""
import numpy as np
import pytest

import fastlpr_amd
from fastlpr_amd import cv_fastkde, cv_fastlpr, get_hlist


@pytest.fixture
def data1d():
    rs = np.random.RandomState(42)
    x = rs.rand(100, 1)
    return x, np.sin(2 * np.pi * x[:, 0]) + 0.1 * rs.randn(100)


@pytest.fixture
def h1d():
    return get_hlist(5, [0.05, 0.5])


def test_empty_hlist(data1d):
    x, y = data1d
    with pytest.raises((ValueError, IndexError)) as e:
        cv_fastlpr(x, y, h=np.array([]).reshape(0, 1), options={"calc_dof": False})
    assert any(w in str(e.value).lower()
               for w in ["empty", "bandwidth", "shape", "size", "index", "stack", "array"])


def test_x_y_sample_mismatch(h1d):
    rs = np.random.RandomState(42)
    with pytest.raises(ValueError) as e:
        cv_fastlpr(rs.rand(100, 1), rs.randn(50), h=h1d, options={"calc_dof": False})
    m = str(e.value).lower()
    assert "sample" in m or "shape" in m or "mismatch" in m


def test_bandwidth_dimension_mismatch(data1d):
    x, y = data1d
    with pytest.raises(ValueError) as e:
        cv_fastlpr(x, y, h=get_hlist([3, 3], [[0.1, 0.5], [0.1, 0.5]]),
                   options={"calc_dof": False})
    m = str(e.value).lower()
    assert "dimension" in m or "bandwidth" in m


@pytest.mark.parametrize("order", [-1, 3, 4, 10, -10, 5])
def test_invalid_order(data1d, h1d, order):
    x, y = data1d
    with pytest.raises(ValueError) as e:
        cv_fastlpr(x, y, h=h1d, options={"order": order, "calc_dof": False})
    m = str(e.value)
    assert "0" in m or "1" in m or "2" in m


@pytest.mark.parametrize("fill,word", [(np.nan, "NaN"), (np.inf, "Inf")])
def test_all_non_finite_y(h1d, fill, word):
    x = np.random.RandomState(42).rand(50, 1)
    with pytest.raises(ValueError, match=word):
        cv_fastlpr(x, np.full(50, fill), h=h1d, options={"calc_dof": False})


@pytest.mark.parametrize("fill,word", [(np.nan, "NaN"), (np.inf, "Inf")])
def test_kde_all_non_finite_x(fill, word):
    with pytest.raises(ValueError, match=word):
        cv_fastkde(np.full((50, 1), fill))


def test_partial_nan_y(h1d):
    rs = np.random.RandomState(42)
    x = rs.rand(100, 1)
    y = np.sin(2 * np.pi * x[:, 0]) + 0.1 * rs.randn(100)
    y[0:5] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        cv_fastlpr(x, y, h=h1d, options={"calc_dof": False})


@pytest.mark.parametrize("h", [[[-0.1], [-0.5], [-1.0]], [[0.1], [-0.1], [0.2]],
                               [[0.0], [0.1], [0.2]]])
def test_non_positive_bandwidth(data1d, h):
    x, y = data1d
    with pytest.raises(ValueError, match="positive"):
        cv_fastlpr(x, y, h=np.array(h), options={"calc_dof": False})


def test_y_wrong_ndim(h1d):
    rs = np.random.RandomState(42)
    with pytest.raises(ValueError) as e:
        cv_fastlpr(rs.rand(50, 1), rs.randn(50, 2, 3), h=h1d, options={"calc_dof": False})
    m = str(e.value).lower()
    assert "1d" in m or "2d" in m or "dimension" in m


def test_kde_too_few_observations():
    with pytest.raises(ValueError) as e:
        cv_fastkde(np.array([[0.5]]))
    m = str(e.value).lower()
    assert "2" in m or "observation" in m or "sample" in m


def test_four_dimensions_not_on_this_engine():
    rs = np.random.RandomState(42)
    with pytest.raises(NotImplementedError):
        cv_fastlpr(rs.rand(50, 4), rs.randn(50), h=get_hlist([2, 2, 2, 2], [[0.3, 0.6]] * 4),
                   options={"calc_dof": False})


def test_exceeds_max_dimension():
    rs = np.random.RandomState(42)
    with pytest.raises(ValueError):
        cv_fastlpr(rs.rand(50, 11), rs.randn(50), options={"calc_dof": False})


def test_options_are_not_mutated(data1d, h1d):
    """options is copied, never written (utils/__init__.py:18)."""
    x, y = data1d
    opts = {"order": 7}
    with pytest.raises(ValueError):
        cv_fastlpr(x, y, h=h1d, options=opts)
    assert opts == {"order": 7}
    assert fastlpr_amd.__version__
