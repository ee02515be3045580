import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref16.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def lto():
    import lto as _lto
    _lto.lib()
    return _lto


@pytest.fixture(scope="session")
def golden():
    if not os.path.exists(GOLDEN):
        pytest.skip("golden vectors missing (python tests/golden/make_golden.py)")
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    import lettuce_amd as la
    ctx = la.Context(0)
    yield ctx
    import gc
    gc.collect()     # objects of finished tests release their handles first
    ctx.close()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def golden_weights(nlinks):
    i = np.arange(nlinks, dtype=np.int64)
    return ((i * 7919) % 65521).astype(np.float64) / 65521.0


def reduce_field(F, idx):
    flat = F.reshape(-1, 18)
    w = golden_weights(flat.shape[0])
    return {"sample": flat[idx], "colsum": F.sum(axis=0), "wsum": (flat * w[:, None]).sum(axis=0)}


def assert_field_matches(golden, prefix, F, tol):
    """Compare a full field against the golden sample and checksums of the reference."""
    red = reduce_field(F, golden["sample_index"])
    for k in ("sample", "colsum", "wsum"):
        ref = golden[f"{prefix}_{k}"]
        err = rel_err(red[k], ref)
        assert err <= tol, f"{prefix}_{k}: relative error {err:.3e} > {tol:.1e}"
