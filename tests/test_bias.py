"""Metadynamics bias potential: the library's host implementation (C ABI, no GPU
needed) and the oracle restatement against the reference's MetaBiasPotential
outputs (golden vectors), plus closed forms the reference leaves open
(the non-well-tempered rule is compiled out of the reference build)."""
import numpy as np
import pytest

from conftest import rel_err


@pytest.fixture(scope="module")
def la_bias():
    import lettuce_amd as la
    return la.MetaBiasPotential


def golden_bias(golden):
    a = golden["bias_args"]
    return (a[0], a[1], int(a[2]), a[3], a[4], a[5], a[6]), golden["bias_deposits"], golden["bias_eval"]


def test_library_bias_matches_reference(la_bias, golden):
    (cmin, cmax, bins, w, width, wt, thr), deps, evals = golden_bias(golden)
    b = la_bias(cmin, cmax, bins, w, width, wt, thr, well_tempered=True)
    for d in deps:
        b.UpdatePotentialSymmetric(d)
    ref = golden["bias_values"]
    got = np.array([[b.ReturnPotential(x), b.ReturnDerivative(x)] for x in evals])
    assert rel_err(got[:, 0], ref[:, 0]) <= 1e-13
    assert np.max(np.abs(got[:, 1] - ref[:, 1])) <= 1e-12 * max(1.0, np.max(np.abs(ref[:, 1])))


def test_oracle_bias_matches_reference(lto, golden):
    (cmin, cmax, bins, w, width, wt, thr), deps, evals = golden_bias(golden)
    b = lto.Bias(cmin, cmax, bins, w, width, wt, thr, well_tempered=True)
    for d in deps:
        b.deposit_symmetric(d)
    ref = golden["bias_values"]
    got = np.array([[b.potential(x), b.derivative(x)] for x in evals])
    assert rel_err(got[:, 0], ref[:, 0]) <= 1e-13
    assert np.max(np.abs(got[:, 1] - ref[:, 1])) <= 1e-12 * max(1.0, np.max(np.abs(ref[:, 1])))


def test_constructor_rejects_bad_ranges(la_bias):
    import lettuce_amd as la
    with pytest.raises(la.LettuceError):
        la_bias(1.0, 1.0, 10, 0.1, 0.1)
    with pytest.raises(la.LettuceError):
        la_bias(-1.0, 1.0, 0, 0.1, 0.1)


def test_plain_gaussian_deposit_closed_form(la_bias):
    b = la_bias(-2.0, 2.0, 40, 0.3, 0.25, well_tempered=False)
    b.UpdatePotential(0.37)
    x = -2.0 + 0.1 * np.arange(41)
    assert rel_err(b.grid(), 0.3 * np.exp(-0.5 * ((0.37 - x) / 0.25) ** 2)) <= 1e-14
    # grid points interpolate to themselves, midpoints linearly
    assert abs(b.ReturnPotential(x[13]) - b.grid()[13]) < 1e-15
    assert abs(b.ReturnPotential(0.5 * (x[13] + x[14])) - 0.5 * (b.grid()[13] + b.grid()[14])) < 1e-15


def test_symmetric_deposits_give_even_potential(la_bias):
    # plain rule: exactly even (well-tempered deposits see the grid the +q deposit left)
    b = la_bias(-3.0, 3.0, 60, 0.2, 0.3, 5.0, 100.0, well_tempered=False)
    for q in (0.4, 1.7, -0.9):
        b.UpdatePotentialSymmetric(q)
    for q in np.linspace(-2.9, 2.9, 37):
        assert abs(b.ReturnPotential(q) - b.ReturnPotential(-q)) < 1e-13
        assert abs(b.ReturnDerivative(q) + b.ReturnDerivative(-q)) < 1e-12


def test_linear_grid_derivative_and_walls(la_bias):
    b = la_bias(-1.0, 1.0, 20, 0.1, 0.1, 1.0, 7.0)
    x = -1.0 + 0.1 * np.arange(21)
    b.set_grid(2.0 + 3.0 * x)
    for q in np.linspace(-0.85, 0.85, 11):
        assert abs(b.ReturnDerivative(q) - 3.0) < 1e-12
    # outside [CV_min, CV_max): edge value plus threshold * distance^2
    assert abs(b.ReturnPotential(-1.5) - ((2.0 - 3.0) + 7.0 * 0.25)) < 1e-13
    assert abs(b.ReturnPotential(1.2) - ((2.0 + 3.0) + 7.0 * 0.04)) < 1e-13
    b.SetCV_current(0.25)
    assert b.ReturnCV_current() == 0.25
