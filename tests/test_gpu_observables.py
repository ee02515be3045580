"""The rest of the reference's Observables() on the GPU -- tree-level Symanzik action,
TopChargePlaquette, Wilson loops L = 2, 4, 8 (one chain workspace, as the reference
chains them), Polyakov loop -- against the reference's values at 16^4
(tests/golden/obs16.json) and the oracle on other extents."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (16, 16, 16, 16)
C_PLAQ, C_RECT = 1.0 + 8.0 / 12.0, -1.0 / 12.0


def flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def la():
    import lettuce_amd
    return lettuce_amd


def gpu_values(la, ctx, U, shape):
    f = la.GaugeField(ctx, shape)
    f.upload(U)
    chain = la.MatrixField(ctx, shape)
    out = [la.RectangularAction(6.0, C_PLAQ, C_RECT).ActionNormalized(f), la.TopChargePlaquette(f),
           la.WilsonLoop(0, 2, True, f, chain), la.WilsonLoop(2, 4, False, f, chain),
           la.WilsonLoop(4, 8, False, f, chain)]
    p = la.PolyakovLoop(f)
    f.close()
    chain.close()
    return np.array(out + [p.real, p.imag])


def oracle_values(lto, U, shape):
    chain = lto.new_field(shape)
    w = [lto.wilson_loop(U, chain, shape, 0, 2, True), lto.wilson_loop(U, chain, shape, 2, 4, False),
         lto.wilson_loop(U, chain, shape, 4, 8, False)]
    p = lto.polyakov_loop(U, shape)
    return np.array([lto.rect_action_normalized(U, shape, C_PLAQ, C_RECT), lto.topcharge_plaquette(U, shape)] + w
                    + [p.real, p.imag])


@pytest.mark.parametrize("name", ["haar", "smooth"])
def test_matches_reference(la, lto, gpu_ctx, name):
    flush(f"observables vs reference 16^4 {name}")
    with open(os.path.join(HERE, "golden", "obs16.json")) as f:
        ref = np.array(json.load(f)[name])
    U0 = lto.haar(SHAPE, 0)
    U = U0 if name == "haar" else lto.stout(U0, SHAPE, 0.1, 3)
    got = gpu_values(la, gpu_ctx, U, SHAPE)
    assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1.0)), (got, ref)


@pytest.mark.parametrize("shape", [(4, 6, 8, 10), (6, 4, 4, 8), (2, 2, 4, 4)])
def test_matches_oracle(la, lto, gpu_ctx, shape):
    flush(f"observables vs oracle {shape}")
    U = lto.stout(lto.haar(shape, 5), shape, 0.1, 2)
    got = gpu_values(la, gpu_ctx, U, shape)
    ref = oracle_values(lto, U, shape)
    assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1.0)), (got, ref)


def test_unit_field(la, lto, gpu_ctx):
    flush("observables unit field")
    shape = (4, 4, 6, 8)
    got = gpu_values(la, gpu_ctx, lto.identity(shape), shape)
    assert np.array_equal(got, np.array([0.0, 0.0, 0.0, 0.0, 0.0, 3.0, 0.0]))


@pytest.mark.parametrize("shape", [(16, 16, 16, 16), (4, 6, 8, 10)])
def test_parity_update(la, lto, gpu_ctx, shape):
    """ParityUpdate on the GPU against the oracle (the oracle matches the reference at
    16^4), the action kept, the clover charge reversed, applied twice the identity."""
    flush(f"parity update {shape}")
    U = lto.stout(lto.haar(shape, 0), shape, 0.1, 3)
    f = la.GaugeField(gpu_ctx, shape)
    f.upload(U)
    tmp = la.GaugeField(gpu_ctx, shape)
    q0 = la.TopChargeClover(f)
    p0 = la.PlaquetteSum(f)
    la.ParityUpdate(f, tmp)
    got = f.download()
    ref = lto.parity_update(U, shape)
    assert np.max(np.abs(got - ref)) < 1e-14
    assert abs(la.TopChargeClover(f) + q0) < 1e-12 and abs(la.PlaquetteSum(f) - p0) < 1e-9
    la.ParityUpdate(f, tmp)
    assert np.max(np.abs(f.download() - U)) < 1e-14
    f.close()
    tmp.close()


@pytest.mark.parametrize("shape,tiles", [((16, 16, 16, 16), [None, (1, 1, 4, 8), (2, 2, 4, 8), (1, 2, 16, 8)]),
                                         ((4, 6, 8, 10), [None]), ((32, 32, 32, 32), [None])])
def test_plaquette_sum(la, lto, gpu_ctx, shape, tiles):
    """PlaquetteSum (plaquette.hpp:26-43) spreads the six planes of a site over its four links
    in the tile order of the link kernels: against the oracle's site-by-site sum to 1e-12
    relative for every tile (the partial sums follow the tile, so bits may differ between
    tiles), against the observables pass to rounding, and exactly 18 V on the unit field
    (every plane term is 3 exactly: G_00 = G_11 = 1, G_01 = G_10 = 0)."""
    flush(f"plaquette sum {shape}")
    U = lto.haar(shape, 11) if shape[0] > 16 else lto.stout(lto.haar(shape, 11), shape, 0.1, 2)
    ref = lto.plaquette_sum(U, shape)
    f = la.GaugeField(gpu_ctx, shape)
    f.upload(U)
    for tile in tiles:
        if tile is not None:
            f.set_tile(*tile)
        got = la.PlaquetteSum(f)
        assert abs(got - ref) <= 1e-12 * abs(ref), (tile, got, ref)
    assert abs(la.Observables(f)["plaquette_sum"] - got) <= 1e-12 * abs(ref)
    f.SetToIdentity()
    assert la.PlaquetteSum(f) == 18.0 * lto.volume(shape)
    f.close()
