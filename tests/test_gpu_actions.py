"""Improved gauge actions Rectangular<2> on the GPU: the full staple and Checkerboard4
overrelaxation against the reference's values (tests/golden/actions16.npz), heatbath
against the oracle with the same Philox stream, the rectangular action kept by
overrelaxation, and the schedule check."""
import os

import numpy as np
import pytest

from conftest import assert_field_matches

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (16, 16, 16, 16)


def flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def la():
    import lettuce_amd
    return lettuce_amd


@pytest.fixture(scope="module")
def actions_golden():
    with np.load(os.path.join(HERE, "golden", "actions16.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def smooth16(lto):
    return lto.stout(lto.haar(SHAPE, 0), SHAPE, 0.1, 3)


def make(la, ctx, shape, U):
    f = la.GaugeField(ctx, shape)
    f.upload(U)
    return f


@pytest.mark.parametrize("name", ["lw", "iwasaki"])
def test_staple_matches_reference(la, gpu_ctx, actions_golden, smooth16, name):
    flush(f"rectangular staple {name}")
    f = make(la, gpu_ctx, SHAPE, smooth16)
    out = la.MatrixField(gpu_ctx, SHAPE)
    act = la.LuescherWeiszAction(6.0) if name == "lw" else la.IwasakiAction(6.0)
    la.StapleAction(f, out, act)
    assert_field_matches(actions_golden, f"staple_{name}_smooth", out.download(), 1e-13)
    f.close()
    out.close()


def test_or_sweeps_match_reference(la, gpu_ctx, actions_golden, smooth16):
    flush("rectangular OR sweeps")
    f = make(la, gpu_ctx, SHAPE, smooth16)
    orx = la.OverrelaxationSubgroupKernel(f, action=la.IwasakiAction(6.0))
    la.Iterator.Checkerboard4(orx, 2)
    assert_field_matches(actions_golden, "or4_iwasaki_smooth_n2", f.download(), 1e-12)
    f.close()


def test_heatbath_matches_oracle_stream(la, lto, gpu_ctx):
    flush("rectangular heatbath")
    shape = (8, 8, 8, 8)
    U = lto.stout(lto.haar(shape, 4), shape, 0.1, 2)
    f = make(la, gpu_ctx, shape, U)
    f.set_rng(31, 2)
    hb = la.HeatbathKernel(f, 4.2, action=la.LuescherWeiszAction(4.2))
    la.Iterator.Checkerboard4(hb, 1)
    assert f.get_rng() == (31, 3)
    ref = U.copy()
    cp, cr = lto.LUESCHER_WEISZ
    lto.hb_sweep_action(ref, shape, 4.2, cp, cr, 1, 4, 31, 2)
    diff = np.max(np.abs(f.download() - ref).reshape(-1, 18), axis=1)
    assert np.max(diff) <= 1e-11, np.max(diff)
    f.close()


def test_overrelaxation_keeps_action_and_schedule_check(la, lto, gpu_ctx):
    flush("rectangular OR invariant")
    shape = (8, 8, 12, 16)
    cp, cr = lto.DBW2
    f = make(la, gpu_ctx, shape, lto.stout(lto.haar(shape, 5), shape, 0.1, 2))
    act = la.DBW2Action(6.0)
    before = act.ActionNormalized(f)
    la.Iterator.Checkerboard4(la.OverrelaxationSubgroupKernel(f, action=act), 3)
    assert abs(act.ActionNormalized(f) - before) < 1e-12
    with pytest.raises(la.LettuceError):
        la.Iterator.Checkerboard(la.OverrelaxationSubgroupKernel(f, action=act), 1)
    f.close()
