"""HMC and MetaD-HMC integrator schedules, on the CPU.

Pins: the kick / drift sequence of every integrator (oracle and library) equals the
one the reference's own operator() issues (tests/golden/schedules.json, made by
tests/golden/make_golden.py --schedule from the compiled reference, step sizes to the
last bit); the multiple-timescale schedules are palindromes whose kicks and drifts
add up to the trajectory length; with no bias force they reduce to the single-scale
OMF4; with a bias force the energy violation falls as eps^2; the CV log line format
equals the reference's LogCVData output."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SCHED = os.path.join(HERE, "golden", "schedules.json")
TINY = (4, 4, 4, 4)

REF_NAMES = {"Leapfrog": "leapfrog", "Leapfrog_slow": "leapfrog_slow", "OMF_2": "omf2", "OMF_2_slow": "omf2_slow",
             "OMF_4": "omf4", "OMF_4_slow": "omf4_slow", "Leapfrog_OMF_4": "leapfrog_omf4",
             "OMF_2_OMF_4_slow": "omf2_omf4_slow", "OMF_2_OMF_4": "omf2_omf4"}
MULTI = ["leapfrog_omf4", "omf2_omf4_slow", "omf2_omf4"]


@pytest.fixture(scope="module")
def sched():
    if not os.path.exists(SCHED):
        pytest.skip("schedules.json missing (python tests/golden/make_golden.py --schedule)")
    with open(SCHED) as f:
        return json.load(f)


def _cases(sched):
    for key, ev in sched["schedules"].items():
        name, tau, n = key.split()
        yield REF_NAMES[name], float(tau), int(n), ev


def _same(got, ref):
    """Kinds and step sizes bit for bit; the submeasurement flag where the reference has
    one (UpdateMomenta)."""
    assert len(got) == len(ref)
    for (k, e, s), (rk, re_, rs) in zip(got, ref):
        assert k == rk and e == re_
        if k == "P":
            assert s == rs


def test_oracle_schedules_equal_reference(lto, sched):
    for name, tau, n, ev in _cases(sched):
        _same(lto.integrator_schedule(name, tau, n), ev)


def test_library_schedules_equal_reference(lto, sched):
    import lettuce_amd as la
    for name, tau, n, ev in _cases(sched):
        got = la.IntegratorSchedule(name, tau, n)
        _same(got, ev)
        assert got == lto.integrator_schedule(name, tau, n)   # the bias-kick sample flags too


@pytest.mark.parametrize("name", MULTI)
@pytest.mark.parametrize("n", [1, 2, 5])
def test_multiscale_schedule_invariants(lto, name, n):
    tau = 0.9
    ev = lto.integrator_schedule(name, tau, n)
    tot = {k: sum(e for kk, e, _ in ev if kk == k) for k in "GMU"}
    # the reference's _slow variant drifts the field by 2 eps per step (full-size inner OMF4 steps)
    t_inner = 2.0 * tau if name == "omf2_omf4_slow" else tau
    assert abs(tot["U"] - t_inner) < 1e-14 and abs(tot["G"] - t_inner) < 1e-14
    assert abs(tot["M"] - tau) < 1e-14
    assert [(k, e) for k, e, _ in ev] == [(k, e) for k, e, _ in reversed(ev)]   # time-reversible
    n_bias = sum(1 for k, _, _ in ev if k == "M")
    assert n_bias == {"leapfrog_omf4": n + 1, "omf2_omf4": 2 * n + 1, "omf2_omf4_slow": 3 * n}[name]
    assert sum(s for k, _, s in ev if k == "M") == n          # one path sample per step
    assert all(k != "P" for k, _, _ in ev)


def _start(lto, seed=5):
    U = lto.stout(lto.haar(TINY, seed), TINY, 0.1, 2)
    P = lto.random_momentum(TINY, 17, 3)
    return U, P


@pytest.mark.parametrize("name,tau_omf4,n_omf4", [("leapfrog_omf4", 0.6, 3), ("omf2_omf4", 0.6, 6),
                                                  ("omf2_omf4_slow", 1.2, 6)])
def test_multiscale_without_bias_is_inner_omf4(lto, name, tau_omf4, n_omf4):
    """No bias force: the outer kicks vanish and the inner OMF4 steps are the single-scale
    OMF4 (the merged kicks of OMF_4 only round differently)."""
    U0, P0 = _start(lto)
    U1, P1 = U0.copy(), P0.copy()
    lto.metad_trajectory(U1, P1, TINY, 6.0, 0.6, 3, 0.12, 1, None, name)
    U2, P2 = U0.copy(), P0.copy()
    lto.metad_trajectory(U2, P2, TINY, 6.0, tau_omf4, n_omf4, 0.12, 1, None, "omf4")
    assert np.max(np.abs(U1 - U2)) < 1e-12 and np.max(np.abs(P1 - P2)) < 1e-12


def _steep_bias(lto, q0):
    b = lto.Bias(-8.0, 8.0, 1600, 3.0, 0.25, 10.0, 1000.0, well_tempered=False)
    b.deposit(q0 + 0.15)
    b.deposit(q0 - 0.4)
    return b


@pytest.mark.parametrize("name", ["leapfrog_omf4", "omf2_omf4", "omf2"])
def test_energy_violation_falls_as_eps_squared(lto, name):
    """With a bias force the outer scale is second order: |dH(eps)| / |dH(eps/2)| ~ 4, as
    for the single-scale OMF2.  (The bias is linear between grid points, so its force
    jumps at bin edges; below dH ~ 1e-5 that, not the step size, sets dH.)"""
    U0, P0 = _start(lto, 8)
    q0 = lto.cv_charge(U0, TINY, 0.12, 1)
    bias = _steep_bias(lto, q0)
    assert abs(bias.derivative(q0)) > 0.5
    dh = []
    for n in (8, 16):
        U, P = U0.copy(), P0.copy()
        H0, H1, _, _, samples = lto.metad_trajectory(U, P, TINY, 6.0, 0.5, n, 0.12, 1, bias, name)
        assert len(samples) == n
        dh.append(abs(H1 - H0))
    ratio = dh[0] / dh[1]
    assert 3.0 < ratio < 5.5, f"{name}: dH {dh}, ratio {ratio}"


def test_omf2_slow_variant_does_not_conserve_the_biased_energy(lto):
    """OMF_2_OMF_4_slow as written: the field moves 2 tau while the bias kicks add up to
    tau, so the biased H is not conserved as eps -> 0 (unlike OMF_2_OMF_4)."""
    U0, P0 = _start(lto, 8)
    bias = _steep_bias(lto, lto.cv_charge(U0, TINY, 0.12, 1))
    dh = {}
    for name in ("omf2_omf4", "omf2_omf4_slow"):
        U, P = U0.copy(), P0.copy()
        H0, H1, _, _, _ = lto.metad_trajectory(U, P, TINY, 6.0, 0.25, 8, 0.12, 1, bias, name)
        dh[name] = abs(H1 - H0)
    assert dh["omf2_omf4_slow"] > 20.0 * dh["omf2_omf4"]


def test_metad_trajectory_matches_metad_omf2(lto):
    """The schedule-driven oracle equals the hand-written OMF2 MetaD trajectory (which the
    reference's golden trajectory pins), up to the order of the oracle's threaded sums."""
    U0, P0 = _start(lto, 9)
    bias = _steep_bias(lto, lto.cv_charge(U0, TINY, 0.12, 2))
    Ua, Pa = U0.copy(), P0.copy()
    ra = lto.metad_omf2(Ua, Pa, TINY, 6.0, 0.4, 3, 0.12, 2, bias)
    Ub, Pb = U0.copy(), P0.copy()
    rb = lto.metad_trajectory(Ub, Pb, TINY, 6.0, 0.4, 3, 0.12, 2, bias, "omf2")
    assert np.max(np.abs(Ua - Ub)) < 1e-13 and np.max(np.abs(Pa - Pb)) < 1e-13
    assert abs((ra[1] - ra[0]) - (rb[1] - rb[0])) < 1e-10
    assert np.allclose(ra[2:4], rb[2:4], rtol=0, atol=1e-13)
    assert np.allclose(ra[4], rb[4], rtol=0, atol=1e-13)


def test_cv_log_format_equals_reference(sched):
    """LogCVData writes with the stream's default format: six significant digits, %g."""
    g = sched["cvlog"]
    assert " ".join(format(v, "g") for v in g["values"]) + "\n" == g["text"]


def test_schedule_argument_checks():
    """Unknown integrator codes and negative step counts are errors; zero steps is an
    empty schedule."""
    import ctypes

    import lettuce_amd as la
    assert la.IntegratorSchedule("omf4", 1.0, 0) == []
    n = ctypes.c_int()
    with pytest.raises(la.LettuceError):
        la._lib.call("lt_integrator_schedule", 3, 1.0, 4, None, None, None, 0, ctypes.byref(n))
    with pytest.raises(la.LettuceError):
        la._lib.call("lt_integrator_schedule", la.OMF2, 1.0, -1, None, None, None, 0, ctypes.byref(n))
