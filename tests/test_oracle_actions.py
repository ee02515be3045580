"""Improved gauge actions Rectangular<2> (rectangular_action.hpp:29-300) in the oracle
against the reference (tests/golden/actions16.npz): the full staple for the
Luescher-Weisz and Iwasaki coefficients and two Checkerboard4 overrelaxation sweeps
with the Iwasaki action; plus what the mathematics fixes: the unit field's staple,
overrelaxation keeping the rectangular action, and (1, 0) reducing to the Wilson action."""
import os

import numpy as np
import pytest

from conftest import assert_field_matches

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (16, 16, 16, 16)


@pytest.fixture(scope="module")
def actions_golden():
    path = os.path.join(HERE, "golden", "actions16.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def smooth16(lto):
    return lto.stout(lto.haar(SHAPE, 0), SHAPE, 0.1, 3)


@pytest.mark.parametrize("name", ["lw", "iwasaki"])
def test_staple_matches_reference(lto, actions_golden, smooth16, name):
    cp, cr = lto.LUESCHER_WEISZ if name == "lw" else lto.IWASAKI
    assert_field_matches(actions_golden, f"staple_{name}_smooth", lto.staple_action_field(smooth16, SHAPE, cp, cr),
                         1e-13)


def test_or_sweeps_match_reference(lto, actions_golden, smooth16):
    cp, cr = lto.IWASAKI
    U = lto.or_sweep_action(smooth16.copy(), SHAPE, cp, cr, 2, 4)
    assert_field_matches(actions_golden, "or4_iwasaki_smooth_n2", U, 1e-12)


def test_unit_field_and_wilson_limit(lto):
    shape = (4, 4, 4, 8)
    U = lto.identity(shape)
    ST = lto.staple_action_field(U, shape, 2.0, -0.25)
    # 6 plaquette staples and 3 * 6 rectangle staples of the unit matrix
    unit = np.zeros(18)
    unit[[0, 8, 16]] = 1.0
    assert np.array_equal(ST, np.broadcast_to((2.0 * 6 - 0.25 * 18) * unit, ST.shape))
    V = lto.stout(lto.haar(shape, 2), shape, 0.1, 2)
    assert np.array_equal(lto.staple_action_field(V, shape, 1.0, 0.0), lto.staple_field(V, shape))


def test_overrelaxation_keeps_rectangular_action(lto):
    shape = (4, 4, 8, 8)
    cp, cr = lto.DBW2
    U = lto.stout(lto.haar(shape, 3), shape, 0.1, 2)
    before = lto.rect_action_normalized(U, shape, cp, cr)
    lto.or_sweep_action(U, shape, cp, cr, 3, 4)
    assert abs(lto.rect_action_normalized(U, shape, cp, cr) - before) < 1e-12
