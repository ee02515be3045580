"""The CPU oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py from the compiled reference at 16^4).

Tolerances: 1e-12 relative for deterministic single passes;
observables compare the whole vector at 1e-12 relative to its largest entry,
and Q additionally as |dQ| <= 1e-12 * sum|terms| through the vector scale.
"""
import numpy as np
import pytest

from conftest import assert_field_matches, rel_err

SHAPE = (16, 16, 16, 16)


@pytest.fixture(scope="module")
def start_fields(lto):
    U0 = lto.haar(SHAPE, 0)
    Us = lto.stout(U0, SHAPE, 0.1, 3)
    return {"haar": U0, "smooth": Us}


def obs_close(a, b, tol):
    """Compare observable vectors entry-wise, scaled by magnitude class."""
    a = np.asarray(a)
    b = np.asarray(b)
    # plaquette sum, action: relative; Q and timeslices: absolute against a natural scale
    assert abs(a[0] - b[0]) <= tol * abs(b[0]) + 1e-9
    assert abs(a[1] - b[1]) <= tol * abs(b[1])
    for k in (2, 3, 4):   # Q via three routes: absolute, scale ~ sum |q_t|
        scale = max(1.0, np.sum(np.abs(b[7:7 + SHAPE[0]])))
        assert abs(a[k] - b[k]) <= tol * scale, (k, a[k], b[k])
    assert abs(a[5] - b[5]) <= tol * abs(b[5])
    assert abs(a[6] - b[6]) <= tol * abs(b[6])
    nt = SHAPE[0]
    qs = max(1.0, np.max(np.abs(b[7:7 + nt])))
    assert np.max(np.abs(a[7:7 + nt] - b[7:7 + nt])) <= tol * qs
    assert rel_err(a[7 + nt:], b[7 + nt:]) <= tol


def test_staple(lto, golden, start_fields):
    assert_field_matches(golden, "staple_haar", lto.staple_field(start_fields["haar"], SHAPE), 1e-13)


@pytest.mark.parametrize("schedule", [2, 4])
def test_overrelaxation_one_sweep(lto, golden, start_fields, schedule):
    W = start_fields["smooth"].copy()
    lto.or_sweep(W, SHAPE, 1, schedule)
    assert_field_matches(golden, f"or{schedule}_smooth_n1", W, 1e-12)
    obs_close(lto.observables(W, SHAPE, 6.0), golden[f"or{schedule}_smooth_n1_obs"], 1e-12)


@pytest.mark.parametrize("schedule", [2, 4])
def test_overrelaxation_four_sweeps_observables(lto, golden, start_fields, schedule):
    # link values drift at ~1e-11 after 4 sweeps (rounding amplified by the
    # SU(2) projections); the action is conserved and observables agree closely
    W = start_fields["smooth"].copy()
    lto.or_sweep(W, SHAPE, 4, schedule)
    ref = golden[f"or{schedule}_smooth_n4_obs"]
    got = lto.observables(W, SHAPE, 6.0)
    assert abs(got[0] - ref[0]) <= 1e-11 * abs(ref[0])


@pytest.mark.parametrize("name", ["haar", "smooth"])
def test_observables(lto, golden, start_fields, name):
    obs_close(lto.observables(start_fields[name], SHAPE, 6.0), golden[f"obs_{name}"], 1e-12)


@pytest.mark.parametrize("name", ["haar", "smooth"])
def test_stout_levels(lto, golden, start_fields, name):
    W = start_fields[name]
    for lev in range(4):
        W = lto.stout(W, SHAPE, 0.12, 1)
        obs_close(lto.observables(W, SHAPE, 6.0), golden[f"stout_{name}_obs"][lev], 1e-12)
    assert_field_matches(golden, f"stout4_{name}", W, 1e-12)


@pytest.mark.parametrize("integrator", ["rk3", "euler"])
def test_flow(lto, golden, start_fields, integrator):
    W = lto.flow(start_fields["smooth"], SHAPE, 0.01, 10, integrator)
    assert_field_matches(golden, f"flow_{integrator}_smooth", W, 1e-12)
    obs_close(lto.observables(W, SHAPE, 6.0), golden[f"flow_{integrator}_smooth_obs"], 1e-12)


def test_clover_derivative(lto, golden, start_fields):
    assert_field_matches(golden, "clover_deriv_smooth", lto.clover_derivative(start_fields["smooth"], SHAPE), 1e-12)


def test_cv_force(lto, golden, start_fields):
    Q, F = lto.cv_force(start_fields["smooth"], SHAPE, 0.12, 4, 1.0)
    assert abs(Q - golden["cv_force_smooth_Q"][0]) <= 1e-12 * max(1.0, abs(Q))
    assert_field_matches(golden, "cv_force_smooth", F, 1e-11)


def test_cv_force_vs_reference_finite_differences(lto, golden, start_fields):
    # the reference's FD (numerical_derivative.hpp:214, h = 1e-4) against the
    # oracle's analytic force, projected as Re tr(lambda_a i F) (:215)
    _, F = lto.cv_force(start_fields["smooth"], SHAPE, 0.12, 4, 1.0)
    lam = gell_mann()
    flat = F.reshape(-1, 18)
    fd = golden["fd_h1e-4"]
    for li, link in enumerate(golden["fd_links"]):
        M = flat[link].reshape(3, 3, 2)
        Fm = M[..., 0] + 1j * M[..., 1]
        an = np.array([np.real(np.trace(lam[a] @ (1j * Fm))) for a in range(8)])
        scale = max(np.max(np.abs(fd[li])), 1e-12)
        assert np.max(np.abs(an - fd[li])) <= 1e-6 * scale + 1e-12, (li, an, fd[li])


def test_hmc_injected_momenta(lto, golden, start_fields):
    U = start_fields["smooth"].copy()
    P = lto.random_momentum(SHAPE, 5, 0)
    H0 = lto.hamiltonian(U, P, SHAPE, 6.0)
    lto.omf2(U, P, SHAPE, 6.0, 1.0, 10)
    H1 = lto.hamiltonian(U, P, SHAPE, 6.0)
    Href = golden["hmc_omf2_smooth_H"]
    assert abs(H0 - Href[0]) <= 1e-12 * abs(Href[0])
    assert abs((H1 - H0) - (Href[1] - Href[0])) <= 1e-7 * max(1.0, abs(Href[1] - Href[0]))
    assert_field_matches(golden, "hmc_omf2_smooth", U, 1e-9)


def test_metad_trajectory_injected_momenta(lto, golden, start_fields):
    """HMCMetaDKernel OMF2 trajectory (tau = 0.5, 5 steps, 4 stout levels, rho = 0.12) on injected
    momenta with a deposited well-tempered bias, against the reference."""
    a = golden["metad_bias_args"]
    bias = lto.Bias(a[0], a[1], int(a[2]), a[3], a[4], a[5], a[6], well_tempered=True)
    for d in golden["metad_bias_deposits"]:
        bias.deposit_symmetric(d)
    U = start_fields["smooth"].copy()
    P = lto.random_momentum(SHAPE, 5, 0)
    H0, H1, cv0, cv1, samples = lto.metad_omf2(U, P, SHAPE, 6.0, 0.5, 5, 0.12, 4, bias)
    Href = golden["metad_omf2_smooth_H"]
    cvref = golden["metad_omf2_smooth_CV"]
    assert abs(H0 - Href[0]) <= 1e-12 * abs(Href[0])
    assert abs((H1 - H0) - (Href[1] - Href[0])) <= 1e-7 * max(1.0, abs(Href[1] - Href[0]))
    assert abs(cv0 - cvref[0]) <= 1e-11 and abs(cv1 - cvref[1]) <= 1e-9
    assert np.max(np.abs(np.array(samples) - golden["metad_omf2_smooth_samples"])) <= 1e-9
    assert_field_matches(golden, "metad_omf2_smooth", U, 1e-9)
    # path update: the reference deposits every sample symmetrically after the trajectory
    for s in samples:
        bias.deposit_symmetric(s)
    x = a[0] + np.arange(int(a[2]) + 1) * (a[1] - a[0]) / int(a[2])
    got = np.array([bias.potential(v) for v in x])
    assert rel_err(got, golden["metad_omf2_smooth_grid_after"]) <= 1e-9


def gell_mann():
    l = np.zeros((8, 3, 3), dtype=complex)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1] = -1j
    l[1][1, 0] = 1j
    l[2][0, 0] = 1
    l[2][1, 1] = -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2] = -1j
    l[4][2, 0] = 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2] = -1j
    l[6][2, 1] = 1j
    l[7][0, 0] = l[7][1, 1] = 1 / np.sqrt(3)
    l[7][2, 2] = -2 / np.sqrt(3)
    return l


@pytest.mark.parametrize("name", ["haar", "smooth"])
def test_charge_density_sums_to_reference_timeslices(lto, golden, start_fields, name):
    """The per-site charge density adds up, timeslice by timeslice, to the reference's
    TopChargeCloverTimeslice and in total to its TopChargeClover (16^4)."""
    q = lto.topcharge_density(start_fields[name], SHAPE)
    ref = golden[f"obs_{name}"]
    nt = SHAPE[0]
    scale = max(1.0, np.sum(np.abs(ref[7:7 + nt])))
    assert np.max(np.abs(q.sum(axis=(1, 2, 3)) - ref[7:7 + nt])) <= 1e-12 * scale
    assert abs(q.sum() - ref[2]) <= 1e-12 * scale
