"""Launch-shape tuning on the GPU: every thread-block tile gives bit-identical
results (sweeps, flow, batched MetaD), and the autotuner's JSON cache is written
and reused; the read-rate probe behind the roofline ceilings."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G16 = (16, 16, 16, 16)
SMALL = (8, 8, 8, 8)


def flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def la():
    import lettuce_amd
    return lettuce_amd


def sweep_and_flow(la, ctx, U0, tile=None, tune_cache=None):
    U = la.GaugeField(ctx, G16)
    U.upload(U0)
    U.set_rng(7, 0)
    if tile is not None:
        U.set_tile(*tile)
    if tune_cache is not None:
        U.autotune("sweep", tune_cache)
    la.Iterator.Checkerboard(la.HeatbathKernel(U, 6.0), 1)
    la.Iterator.Checkerboard(la.OverrelaxationSubgroupKernel(U), 2)
    W = la.GaugeField(ctx, G16)
    if tile is not None:
        W.set_tile(*tile)
    if tune_cache is not None:
        W.autotune("flow", tune_cache)
    fl = la.GradientFlowKernel(U, W, "rk3", 0.02)
    fl(3)
    out = (U.download(), W.download(), U.get_tile(), W.get_tile())
    fl.close()
    return out


def test_tiles_and_autotune_give_identical_results(la, lto, gpu_ctx, tmp_path):
    flush("tile invariance 16^4")
    U0 = lto.haar(G16, 3)
    ref = sweep_and_flow(la, gpu_ctx, U0)
    for tile in [(1, 1, 4, 8), (2, 2, 4, 8), (1, 4, 8, 2), (4, 4, 4, 2), (1, 2, 16, 8)]:
        got = sweep_and_flow(la, gpu_ctx, U0, tile=tile)
        assert got[2] == tile
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), f"tile {tile}"
    cache = str(tmp_path / "tiles.json")
    got = sweep_and_flow(la, gpu_ctx, U0, tune_cache=cache)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    entries = json.load(open(cache))
    keys = sorted(entries)
    assert len(keys) == 2 and keys[0].endswith("|16x16x16x16|flow") and keys[1].endswith("|16x16x16x16|sweep")
    assert tuple(entries[keys[1]][:4]) == got[2] and tuple(entries[keys[0]][:4]) == got[3]
    assert len(entries[keys[0]]) == 9 and len(entries[keys[1]]) == 9      # shape, time and site order
    # a cached entry is applied as stored (five numbers: the shape's default site order)
    entries[keys[1]] = [2, 2, 4, 8, 1.0]
    json.dump(entries, open(cache, "w"))
    U = la.GaugeField(gpu_ctx, G16)
    tile, ms = U.autotune("sweep", cache)
    assert tile == (2, 2, 4, 8) and ms == 1.0
    U.close()
    # site orders (t / y chunks, XCD spread) and a linear map as cached entries: same results
    for sweep_entry, flow_entry in [([1, 1, 16, 8, 1.0, 1, 1, 1, 0], [1, 1, 16, 8, 1.0, 1, 1, 1, 0]),
                                    ([1, 2, 4, 8, 1.0, 2, 2, 4, 1], [1, 2, 4, 8, 1.0, 2, 4, 2, 0]),
                                    ([0, 0, 0, 128, 1.0, 1, 1, 1, 0], [0, 0, 0, 256, 1.0, 1, 1, 1, 1])]:
        flush(f"site order {sweep_entry} / {flow_entry}")
        json.dump({keys[1]: sweep_entry, keys[0]: flow_entry}, open(cache, "w"))
        got = sweep_and_flow(la, gpu_ctx, U0, tune_cache=cache)
        assert got[2] == tuple(sweep_entry[:4]) and got[3] == tuple(flow_entry[:4])
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_batch_autotune_keeps_results(la, lto, gpu_ctx, tmp_path):
    flush("batch autotune invariance")
    shapes = [SMALL, SMALL, (4, 4, 8, 8)]
    starts = [lto.stout(lto.haar(s, 50 + r), s, 0.1, 1) for r, s in enumerate(shapes)]
    out = []
    for tune in (False, True):
        b = la.ReplicaBatch(gpu_ctx, shapes, 2, 0.12)
        bias = la.MetaBiasPotential(-8.0, 8.0, 800, 1.5, 0.3, 10.0, 1000.0, well_tempered=False)
        for q in (0.3, -0.8):
            bias.UpdatePotentialSymmetric(q)
        if tune:
            b.autotune(str(tmp_path / "batch.json"))
        for r, U0 in enumerate(starts):
            b.field(r).upload(U0)
            b.set_rng(r, 30 + r, 0)
        b.set_ladder([0.0, 1.0, 0.5])
        b.set_bias(bias)
        b.metad_hmc(3, tau=0.3, update_bias=True)
        res = b.results()
        out.append((res["dH"], res["cv"], [b.field(r).download() for r in range(3)]))
        b.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for a, c in zip(out[0][2], out[1][2]):
        assert np.array_equal(a, c)
    assert len(json.load(open(tmp_path / "batch.json"))) == 2     # one entry per distinct extent


def test_read_rate_probe(la, gpu_ctx):
    """The read-rate probe: an L2-resident buffer reads faster than one that streams from
    HBM, both rates positive and finite; bad footprints are refused."""
    flush("read-rate probe")
    l2 = gpu_ctx.read_rate(1 << 21)
    hbm = gpu_ctx.read_rate(1 << 28)
    assert 0.0 < hbm < l2 < 1e15
    for bad in (3 << 20, 1 << 10, 1 << 31):
        with pytest.raises(la.LettuceError):
            gpu_ctx.read_rate(bad)
