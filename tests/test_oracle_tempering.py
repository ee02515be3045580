"""The oracle's replica-exchange rule: it reduces to the reference's two-stream
tempering rule (LettuceGauge/updates/tempering.hpp:57-72) and satisfies detailed
balance for the joint ensemble prod_r exp(-s_label(r) V(Q_r)).  The acceptance
probability is read off the oracle function itself (threshold of the uniform)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def bias(lto):
    b = lto.Bias(-8.0, 8.0, 800, 0.8, 0.3, 10.0, 1000.0, well_tempered=False)
    for q in (0.1, -0.6, 1.4, -2.2, 0.9):
        b.deposit_symmetric(q)
    return b


def threshold(lto, monkeypatch, labels, strengths, cvs, shapes, bias, k):
    """Largest u for which the oracle accepts the exchange of labels (k, k+1), by bisection."""
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        monkeypatch.setattr(lto, "uniform", lambda *a, **kw: mid)
        mode = 1 if k % 2 == 0 else 2
        new, moves = lto.ladder_exchange(labels, strengths, cvs, shapes, bias, 0, 0, mode)
        acc = dict(moves)[k]
        lo, hi = (mid, hi) if acc else (lo, mid)
    return lo


def test_two_replicas_reduce_to_reference_rule(lto, bias, monkeypatch):
    rng = np.random.default_rng(3)
    shapes = [(4, 4, 4, 4)] * 2
    n_acc = 0
    for _ in range(200):
        cvs = rng.uniform(-3, 3, size=2)
        labels = [int(x) for x in rng.permutation(2)]
        u = float(rng.uniform())
        unbiased, biased = labels.index(0), labels.index(1)
        ref = bool(lto.tempering_accept(bias, cvs[unbiased], cvs[biased], u))
        monkeypatch.setattr(lto, "uniform", lambda *a, **kw: u)
        new, moves = lto.ladder_exchange(labels, [0.0, 1.0], cvs, shapes, bias, 0, 0, 1)
        assert moves == [(0, ref)]
        assert new == (labels[::-1] if ref else labels)
        n_acc += ref
    assert 0 < n_acc < 200


def test_uniform_is_the_accept_stream(lto):
    """uniform(seed, 0xFFFFFFFF, 3, c) is the Philox draw of the Metropolis step: key = seed,
    counter = (0xFFFFFFFF, 3 << 24, c_lo, c_hi), u = (a >> 5, b >> 6) -> [0, 1) with 53 bits."""
    seed, c = 0x123456789, (5 << 32) + 7
    o = lto.philox([0xFFFFFFFF, 3 << 24, 7, 5], [0x23456789, 0x1])
    assert lto.uniform(seed, 0xFFFFFFFF, 3, c) == ((o[0] >> 5) * 2.0 ** 26 + (o[1] >> 6)) * 2.0 ** -53


def test_detailed_balance(lto, bias, monkeypatch):
    rng = np.random.default_rng(4)
    strengths = [0.0, 0.35, 0.6, 1.0]
    shapes = [(4, 4, 4, 4)] * 4

    def weight(labels, cvs):
        return np.exp(-sum(strengths[labels[r]] * bias.potential(cvs[r]) for r in range(4)))

    for _ in range(20):
        cvs = rng.uniform(-2.5, 2.5, size=4)
        a = list(rng.permutation(4))
        k = int(rng.integers(0, 3))
        b = list(a)
        i, j = a.index(k), a.index(k + 1)
        b[i], b[j] = k + 1, k
        p_ab = threshold(lto, monkeypatch, a, strengths, cvs, shapes, bias, k)
        p_ba = threshold(lto, monkeypatch, b, strengths, cvs, shapes, bias, k)
        lhs, rhs = weight(a, cvs) * p_ab, weight(b, cvs) * p_ba
        assert abs(lhs - rhs) <= 1e-12 * max(lhs, rhs)
        assert max(p_ab, p_ba) > 1.0 - 1e-12    # Metropolis: one direction always accepted


def test_pairs_alternate_and_skip_unequal_extents(lto, bias):
    labels = [0, 1, 2, 3, 4]
    cvs = [0.0] * 5
    s = [0.0, 0.25, 0.5, 0.75, 1.0]
    shapes = [(4, 4, 4, 4)] * 5
    _, even = lto.ladder_exchange(labels, s, cvs, shapes, bias, 1, 0, 3)
    _, odd = lto.ladder_exchange(labels, s, cvs, shapes, bias, 1, 1, 3)
    assert [k for k, _ in even] == [0, 2] and [k for k, _ in odd] == [1, 3]
    # equal CVs: delta = 0, always accepted
    assert all(acc for _, acc in even + odd)
    shapes[1] = (6, 6, 6, 6)
    _, moves = lto.ladder_exchange(labels, s, cvs, shapes, bias, 1, 0, 1)
    assert [k for k, _ in moves] == [2]
