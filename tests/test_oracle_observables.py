"""The oracle's restatement of the rest of the reference's Observables()
(SU3_main.cpp:125-136) against the reference's values (tests/golden/obs16.json),
plus closed forms: the unit field, and gauge invariance under a random local
transformation."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (16, 16, 16, 16)
C_PLAQ, C_RECT = 1.0 + 8.0 / 12.0, -1.0 / 12.0


@pytest.fixture(scope="module")
def obs_golden():
    with open(os.path.join(HERE, "golden", "obs16.json")) as f:
        return json.load(f)


def oracle_values(lto, U, shape):
    chain = lto.new_field(shape)
    w2 = lto.wilson_loop(U, chain, shape, 0, 2, True)
    w4 = lto.wilson_loop(U, chain, shape, 2, 4, False)
    w8 = lto.wilson_loop(U, chain, shape, 4, 8, False)
    p = lto.polyakov_loop(U, shape)
    return [lto.rect_action_normalized(U, shape, C_PLAQ, C_RECT), lto.topcharge_plaquette(U, shape), w2, w4, w8,
            p.real, p.imag]


@pytest.mark.parametrize("name", ["haar", "smooth"])
def test_oracle_matches_reference(lto, obs_golden, name):
    U0 = lto.haar(SHAPE, 0)
    U = U0 if name == "haar" else lto.stout(U0, SHAPE, 0.1, 3)
    got = np.array(oracle_values(lto, U, SHAPE))
    ref = np.array(obs_golden[name])
    # absolute scale of each entry: 1 for actions / loops, |Q| or 1 for the charge
    scale = np.maximum(np.abs(ref), 1.0)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale), (got, ref)


def test_unit_field_closed_forms(lto):
    shape = (4, 6, 4, 8)
    U = lto.identity(shape)
    V = int(np.prod(shape))
    assert lto.rect_sums(U, shape) == (18.0 * V, 36.0 * V)
    assert oracle_values(lto, U, shape) == [0.0, 0.0, 0.0, 0.0, 0.0, 3.0, 0.0]


def gauge_transform(lto, U, shape, seed):
    """U_mu(x) -> G(x) U_mu(x) G(x+mu)^+ with Haar-random G."""
    G = lto.haar(shape, seed)[:, 0, :].reshape(-1, 3, 3, 2)
    G = G[..., 0] + 1j * G[..., 1]
    nt, nx, ny, nz = shape
    idx = np.arange(int(np.prod(shape))).reshape(shape)
    W = U.reshape(-1, 4, 3, 3, 2)
    Wc = W[..., 0] + 1j * W[..., 1]
    out = np.empty_like(Wc)
    for mu in range(4):
        nb = np.roll(idx, -1, axis=mu).reshape(-1)
        out[:, mu] = G @ Wc[:, mu] @ np.conj(np.transpose(G[nb], (0, 2, 1)))
    return np.ascontiguousarray(np.stack([out.real, out.imag], axis=-1).reshape(-1, 4, 18))


def test_gauge_invariance(lto):
    shape = (4, 4, 6, 8)
    U = lto.stout(lto.haar(shape, 4), shape, 0.1, 2)
    a = np.array(oracle_values(lto, U, shape))
    b = np.array(oracle_values(lto, gauge_transform(lto, U, shape, 9), shape))
    assert np.max(np.abs(a - b)) < 1e-12


def parity_reduced(U, n_sample=128):
    from conftest import golden_weights
    nl = U.shape[0] * 4
    rng = np.random.default_rng(1234)
    idx = np.sort(rng.choice(nl, 1024, replace=False))[:n_sample]
    flat = U.reshape(-1, 18)
    w = golden_weights(nl)
    return {"sample": flat[idx], "colsum": U.sum(axis=0), "wsum": (flat * w[:, None]).sum(axis=0)}


def test_parity_update_matches_reference(lto, obs_golden):
    """ParityUpdate (parity_update.hpp:22-47) on the smoothed 16^4 field: the reference's
    sampled links and sums."""
    U = lto.stout(lto.haar(SHAPE, 0), SHAPE, 0.1, 3)
    red = parity_reduced(lto.parity_update(U, SHAPE))
    ref = obs_golden["parity_smooth"]
    for k in ("sample", "colsum", "wsum"):
        r = np.array(ref[k])
        assert np.max(np.abs(red[k] - r)) <= 1e-13 * max(1.0, np.max(np.abs(r))), k


def test_parity_update_invariants(lto):
    """The action stays, the clover charge changes sign on every timeslice, twice is the
    identity.  (The single-plaquette charge is not odd under this reflection: the
    plaquettes at x map to plaquettes anchored elsewhere, as the reference notes.)"""
    shape = (4, 6, 4, 8)
    U = lto.stout(lto.haar(shape, 6), shape, 0.1, 2)
    P = lto.parity_update(U, shape)
    assert abs(lto.plaquette_sum(P, shape) - lto.plaquette_sum(U, shape)) < 1e-10
    assert abs(lto.topcharge(P, shape) + lto.topcharge(U, shape)) < 1e-12
    assert np.max(np.abs(lto.topcharge_timeslice(P, shape) + lto.topcharge_timeslice(U, shape))) < 1e-12
    assert np.array_equal(lto.parity_update(P, shape), U)
