"""Replica batches and batched PT-MetaD on the GPU (through the C ABI).

Pins: a batched replica against the reference's golden MetaD-HMC trajectory;
batched == single-lattice path; batched == each replica run alone (bit for bit,
same Philox counters), also for mixed extents; graph replays == eager steps;
device bias deposits == host bias deposits; label exchanges == the oracle's
ladder rule, which reduces to the reference's two-stream rule (tempering.hpp:57-72)
and satisfies detailed balance (tests/test_oracle_tempering.py)."""
import numpy as np
import pytest

from conftest import assert_field_matches, rel_err

pytestmark = pytest.mark.gpu

SMALL = (8, 8, 8, 8)
G16 = (16, 16, 16, 16)


def flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def la():
    import lettuce_amd
    return lettuce_amd


def smooth(lto, shape, seed=0, n=3):
    return lto.stout(lto.haar(shape, seed), shape, 0.1, n)


def golden_bias(la, lto, golden):
    a = golden["metad_bias_args"]
    bias = la.MetaBiasPotential(a[0], a[1], int(a[2]), a[3], a[4], a[5], a[6], well_tempered=True)
    obias = lto.Bias(a[0], a[1], int(a[2]), a[3], a[4], a[5], a[6], well_tempered=True)
    for d in golden["metad_bias_deposits"]:
        bias.UpdatePotentialSymmetric(d)
        obias.deposit_symmetric(d)
    return bias, obias


def plain_bias(la, deposits=(0.3, -1.2, 2.1, 0.05)):
    """Non-well-tempered bias with a few deposits: steep enough to give the CV force weight."""
    b = la.MetaBiasPotential(-8.0, 8.0, 800, 1.5, 0.3, 10.0, 1000.0, well_tempered=False)
    for d in deposits:
        b.UpdatePotentialSymmetric(d)
    return b


def make_batch(la, ctx, shapes, starts, seeds, strengths, bias, n_levels=2, rho=0.12):
    b = la.ReplicaBatch(ctx, shapes, n_levels, rho)
    for r, (U0, s) in enumerate(zip(starts, seeds)):
        b.field(r).upload(U0)
        b.set_rng(r, s, 0)
    b.set_ladder(strengths)
    b.set_bias(bias)
    return b


def test_batch_replica_reproduces_golden_metad(la, lto, golden, gpu_ctx):
    """One replica at full strength = the reference's MetaD-HMC OMF2 trajectory with injected momenta."""
    flush("batch golden 16^4")
    bias, _ = golden_bias(la, lto, golden)
    U0 = lto.stout(lto.haar(G16, 0), G16, 0.1, 3)
    b = make_batch(la, gpu_ctx, [G16], [U0], [5], [1.0], bias, n_levels=4, rho=0.12)
    b.metad_hmc(5, beta=6.0, integrator="omf2", tau=0.5, metropolis=False, update_bias=True)
    res = b.results()
    Href = golden["metad_omf2_smooth_H"]
    assert abs(res["dH"][0] - (Href[1] - Href[0])) <= 1e-7 * max(1.0, abs(Href[1] - Href[0]))
    assert abs(res["cv"][0] - golden["metad_omf2_smooth_CV"][1]) <= 1e-9
    assert_field_matches(golden, "metad_omf2_smooth", b.field(0).download(), 1e-9)
    b.get_bias(bias)
    a = golden["metad_bias_args"]
    x = a[0] + np.arange(int(a[2]) + 1) * (a[1] - a[0]) / int(a[2])
    got = np.array([bias.ReturnPotential(v) for v in x])
    assert rel_err(got, golden["metad_omf2_smooth_grid_after"]) <= 1e-9
    b.close()


@pytest.mark.parametrize("integrator", ["omf2", "leapfrog_omf4", "omf2_omf4"])
def test_batch_matches_single_lattice_path(la, lto, gpu_ctx, integrator):
    """A batch of one at strength 1 and lt_metad_hmc_trajectory agree bit for bit while the
    bias is fixed; with deposits (device exp vs host libm exp) to rounding."""
    flush(f"batch vs single-lattice trajectory {integrator}")
    U0 = smooth(lto, SMALL, 4)
    bias_a, bias_b = plain_bias(la), plain_bias(la)
    b = make_batch(la, gpu_ctx, [SMALL], [U0], [9], [1.0], bias_a)
    U = la.GaugeField(gpu_ctx, SMALL)
    U.upload(U0)
    U.set_rng(9, 0)
    Uc, P, F = la.GaugeField(gpu_ctx, SMALL), la.MatrixField(gpu_ctx, SMALL), la.MatrixField(gpu_ctx, SMALL)
    cv = la.CollectiveVariable(gpu_ctx, SMALL, 2, 0.12)
    k = la.HMCMetaDKernel(U, Uc, P, F, bias_b, cv, integrator, 6.0, 0.5, update_bias=False)
    for step in range(3):
        update = step == 2
        k.update_bias = update
        acc = k(4, metropolis_step=True)
        b.metad_hmc(4, beta=6.0, integrator=integrator, tau=0.5, metropolis=True, update_bias=update)
        res = b.results()
        assert res["dH"][0] == k.delta_H, f"step {step}: {res['dH'][0]!r} vs {k.delta_H!r}"
        assert bool(res["accepted"][0]) == acc
        assert res["cv"][0] == k.cv_value
        assert np.array_equal(b.field(0).download(), U.download())
    b.get_bias(bias_a)
    assert rel_err(bias_a.grid(), bias_b.grid()) <= 1e-14
    assert bias_a.ReturnCV_current() == bias_b.ReturnCV_current()
    cv.close()
    b.close()


def run_alone(la, ctx, shape, U0, seed, strength, bias_grid, steps, graph=False):
    bias = plain_bias(la, ())
    bias.set_grid(bias_grid)
    b = make_batch(la, ctx, [shape], [U0], [seed], [strength], bias)
    b.use_graph(graph)
    out = []
    for _ in range(steps):
        b.metad_hmc(3, beta=6.0, integrator="omf2", tau=0.4, metropolis=True)
        res = b.results()
        out.append((res["dH"][0], res["accepted"][0], res["cv"][0], b.field(0).download()))
    b.close()
    return out


@pytest.mark.parametrize("shapes", [
    [SMALL] * 4,
    [SMALL, (6, 6, 6, 6), (4, 4, 8, 8), (4, 6, 8, 10)],
], ids=["equal", "mixed"])
def test_batch_equals_replicas_run_alone(la, lto, gpu_ctx, shapes):
    flush(f"batch == alone {shapes}")
    starts = [smooth(lto, s, 20 + r, 2) for r, s in enumerate(shapes)]
    seeds = [100 + r for r in range(len(shapes))]
    strengths = np.linspace(0.0, 1.0, len(shapes))
    bias = plain_bias(la)
    grid = bias.grid()
    b = make_batch(la, gpu_ctx, shapes, starts, seeds, strengths, bias)
    b.use_graph(True)
    got = []
    for _ in range(2):
        b.metad_hmc(3, beta=6.0, integrator="omf2", tau=0.4, metropolis=True)
        res = b.results()
        got.append((res, [b.field(r).download() for r in range(len(shapes))]))
    for r, s in enumerate(shapes):
        alone = run_alone(la, gpu_ctx, s, starts[r], seeds[r], strengths[r], grid, 2)
        for step, (res, fields) in enumerate(got):
            dH, acc, cvv, U = alone[step]
            assert res["dH"][r] == dH, f"replica {r} step {step}"
            assert res["accepted"][r] == acc
            assert res["cv"][r] == cvv
            assert np.array_equal(fields[r], U), f"replica {r} step {step}: fields differ"
    b.close()


def test_graph_replays_equal_eager(la, lto, gpu_ctx):
    """Every replay of the captured step equals the eager step (with host work in between)."""
    flush("graph replays vs eager")
    shapes = [SMALL] * 3
    starts = [smooth(lto, SMALL, 40 + r, 2) for r in range(3)]
    grid = plain_bias(la).grid()
    batches = []
    for graph in (False, True):
        bias = plain_bias(la, ())
        bias.set_grid(grid)
        b = make_batch(la, gpu_ctx, shapes, starts, [7, 8, 9], [0.0, 0.5, 1.0], bias)
        b.use_graph(graph)
        batches.append((b, bias))
    for step in range(4):
        outs = []
        for b, _ in batches:
            b.metad_hmc(3, beta=6.0, integrator="omf2", tau=0.4, metropolis=True, update_bias=True,
                        swap="alternate")
            res = b.results()
            outs.append((res, [b.field(r).download() for r in range(3)], b.samples()))
            la.PlaquetteSum(b.field(0))     # eager work between replays
        (r0, f0, s0), (r1, f1, s1) = outs
        for key in ("dH", "accepted", "cv", "labels"):
            assert np.array_equal(r0[key], r1[key]), f"step {step}: {key}"
        assert np.array_equal(s0, s1)
        for a, c in zip(f0, f1):
            assert np.array_equal(a, c), f"step {step}: fields differ"
    (b0, bias0), (b1, bias1) = batches
    b0.get_bias(bias0)
    b1.get_bias(bias1)
    assert np.array_equal(bias0.grid(), bias1.grid())
    assert np.array_equal(b0.swap_stats(), b1.swap_stats())
    b0.close()
    b1.close()


@pytest.mark.parametrize("well_tempered", [False, True])
def test_device_deposits_equal_host_deposits(la, lto, gpu_ctx, well_tempered):
    flush(f"device deposits wt={well_tempered}")
    starts = [smooth(lto, SMALL, 60 + r, 2) for r in range(2)]
    dev_bias = la.MetaBiasPotential(-8.0, 8.0, 800, 0.05, 0.2, 10.0, 1000.0, well_tempered=well_tempered)
    host_bias = la.MetaBiasPotential(-8.0, 8.0, 800, 0.05, 0.2, 10.0, 1000.0, well_tempered=well_tempered)
    for d in (0.4, -0.7):
        dev_bias.UpdatePotentialSymmetric(d)
        host_bias.UpdatePotentialSymmetric(d)
    b = make_batch(la, gpu_ctx, [SMALL] * 2, starts, [1, 2], [0.0, 1.0], dev_bias)
    for _ in range(2):
        b.metad_hmc(4, beta=6.0, integrator="omf2", tau=0.4, metropolis=True, update_bias=True)
        res = b.results()
        top = int(np.argmax(res["labels"]))
        for s in b.samples()[top]:
            host_bias.UpdatePotentialSymmetric(s)
        if not res["accepted"][top]:
            assert np.all(b.samples()[top] == b.samples()[top][0])
    b.get_bias(dev_bias)
    g_dev, g_host = dev_bias.grid(), host_bias.grid()
    # same deposit order per grid point; only exp of the device vs the host's libm differs
    assert np.max(np.abs(g_dev - g_host)) <= 1e-14 * np.max(np.abs(g_host))
    assert dev_bias.ReturnCV_current() == res["cv"][top]
    b.close()


def test_exchanges_follow_oracle_ladder(la, lto, gpu_ctx):
    flush("ladder exchanges vs oracle")
    n = 4
    shapes = [SMALL] * n
    starts = [smooth(lto, SMALL, 80 + r, 1) for r in range(n)]
    strengths = np.array([0.0, 0.3, 0.7, 1.0])
    bias = plain_bias(la, (0.0, 0.5, -0.5, 1.0, -1.0, 0.25))
    b = make_batch(la, gpu_ctx, shapes, starts, [11, 12, 13, 14], strengths, bias)
    b.set_ladder(strengths, None, pt_seed=77, pt_counter=0)
    obias = lto.Bias(-8.0, 8.0, 800, 1.5, 0.3, 10.0, 1000.0, well_tempered=False)
    labels = list(range(n))
    proposed = 0
    for step in range(6):
        b.metad_hmc(3, beta=6.0, integrator="omf2", tau=0.4, metropolis=True, update_bias=True, swap="alternate")
        res = b.results()
        b.get_bias(bias)
        obias.grid = bias.grid()
        labels, moves = lto.ladder_exchange(labels, strengths, res["cv"], shapes, obias, 77, step, 3)
        proposed += len(moves)
        assert list(res["labels"]) == labels, f"step {step}"
    stats = b.swap_stats()
    assert stats[:, 0].sum() == proposed
    b.close()


def test_exchange_skips_unequal_extents(la, lto, gpu_ctx):
    flush("no exchange across extents")
    shapes = [SMALL, (6, 6, 6, 6)]
    starts = [smooth(lto, s, 90 + r, 1) for r, s in enumerate(shapes)]
    b = make_batch(la, gpu_ctx, shapes, starts, [1, 2], [0.0, 1.0], plain_bias(la))
    b.metad_hmc(2, beta=6.0, tau=0.2, swap="even", n_traj=3)
    assert list(b.results()["labels"]) == [0, 1]
    assert b.swap_stats().sum() == 0
    b.close()


def test_batch_views_reject_destroy_and_swap(la, gpu_ctx):
    b = la.ReplicaBatch(gpu_ctx, [SMALL] * 2, 1, 0.1)
    f0, f1 = b.field(0), b.field(1)
    with pytest.raises(la.LettuceError):
        la._lib.call("lt_field_destroy", f0.handle)
    with pytest.raises(la.LettuceError):
        la.Swap(f0, f1)
    f0.SetToHaar(3)
    dev, bad = f0.check()
    assert bad == 0 and dev < 1e-13
    with pytest.raises(la.LettuceError):
        b.metad_hmc(2)   # no bias set
    b.close()


def test_context_outlives_its_handles(la):
    ctx = la.Context(0)
    U = la.GaugeField(ctx, (4, 4, 4, 4))
    b = la.ReplicaBatch(ctx, [(4, 4, 4, 4)], 1, 0.1)
    view = b.field(0)
    with pytest.raises(la.LettuceError):
        la._lib.call("lt_ctx_destroy", ctx.handle)
    b.close()
    with pytest.raises(la.LettuceError):
        view.SetToIdentity()        # the batch is gone
    U.close()
    ctx.close()
