"""BMW configuration files packed and unpacked on the GPU (lt_field_save_bmw /
lt_field_load_bmw): bytes equal to what the reference's SaveConfigBMW[Full] wrote
for the same field (tests/golden/io16.json), loads of files the oracle writes
(byte-identical to the reference's), round trips on other extents, and the
reference's failure cases."""
import hashlib
import json
import os

import numpy as np
import pytest

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
def io_golden():
    with open(os.path.join(HERE, "golden", "io16.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def field16(lto):
    return lto.third_row_fma(lto.stout(lto.haar(SHAPE, 0), SHAPE, 0.1, 3))


def make_field(la, ctx, shape, U=None):
    f = la.GaugeField(ctx, shape)
    if U is not None:
        f.upload(U)
    return f


@pytest.mark.parametrize("full", [False, True])
def test_save_equals_reference_file(la, gpu_ctx, io_golden, field16, full, tmp_path):
    flush(f"bmw save 16^4 full={full}")
    f = make_field(la, gpu_ctx, SHAPE, field16)
    path = str(tmp_path / "u.conf")
    cs = f.save_bmw(path, full=full)
    data = open(path, "rb").read()
    ref = io_golden["bmw18" if full else "bmw12"]
    assert data[:data.index(b"\n")].decode() == ref["header_line"]
    assert ref["header_line"].endswith(" " + cs)
    assert len(data) == ref["bytes"]
    assert hashlib.sha256(data).hexdigest() == ref["sha256"]
    assert f.bmw_checksum(full) == int(cs, 16)
    f.close()


@pytest.mark.parametrize("full", [False, True])
def test_load_reference_bytes(la, lto, gpu_ctx, io_golden, field16, full, tmp_path):
    """The oracle's file for field16 is the reference's (same SHA-256); loading it gives
    the field back bit for bit (row 2 included: the file's row 2 is the device's)."""
    flush(f"bmw load 16^4 full={full}")
    data, cs = lto.bmw_file_bytes(field16, SHAPE, full)
    assert hashlib.sha256(data).hexdigest() == io_golden["bmw18" if full else "bmw12"]["sha256"]
    path = tmp_path / "ref.conf"
    path.write_bytes(data)
    f = make_field(la, gpu_ctx, SHAPE)
    ok, dev = f.load_bmw(str(path), full=full)
    assert ok and dev < 1e-12
    assert np.array_equal(f.download(), field16)
    assert la.LoadConfigBMWFull(f, str(path)) if full else la.LoadConfigBMW(f, str(path))
    f.close()


@pytest.mark.parametrize("shape", [(4, 6, 8, 10), (2, 4, 4, 6)])
@pytest.mark.parametrize("full", [False, True])
def test_roundtrip_other_extents(la, lto, gpu_ctx, shape, full, tmp_path):
    flush(f"bmw roundtrip {shape} full={full}")
    U = lto.third_row_fma(lto.haar(shape, 11))
    f = make_field(la, gpu_ctx, shape, U)
    path = str(tmp_path / "r.conf")
    cs = f.save_bmw(path, full=full)
    data, cs_oracle = lto.bmw_file_bytes(U, shape, full)
    assert cs == cs_oracle and open(path, "rb").read() == data
    g = make_field(la, gpu_ctx, shape)
    ok, dev = g.load_bmw(path, full=full)
    assert ok and dev < 1e-12
    assert np.array_equal(g.download(), U)
    f.close()
    g.close()


def test_failure_cases(la, lto, gpu_ctx, field16, tmp_path):
    """A changed mantissa bit is a checksum mismatch (reported, not fatal, as in
    LoadConfigBMW); other extents, a missing file, a foreign header, non-SU(3) links
    and saving over an existing file without overwrite fail."""
    flush("bmw failure cases")
    shape = (4, 4, 4, 8)
    U = lto.third_row_fma(lto.haar(shape, 3))
    f = make_field(la, gpu_ctx, shape, U)
    path = tmp_path / "a.conf"
    f.save_bmw(str(path))
    raw = bytearray(path.read_bytes())
    raw[4096 + 7] ^= 0x01          # last byte of the first big-endian double: lowest mantissa bit
    bad = tmp_path / "flip.conf"
    bad.write_bytes(bytes(raw))
    ok, dev = f.load_bmw(str(bad))
    assert not ok and dev < 1e-12
    assert la.LoadConfigBMW(f, str(bad))                         # still SU(3): the reference returns true
    other = make_field(la, gpu_ctx, (4, 4, 8, 4))
    assert not la.LoadConfigBMW(other, str(path))                # extents differ
    assert not la.LoadConfigBMW(f, str(tmp_path / "missing.conf"))
    assert not la.LoadConfigBMWFull(f, str(path))                # "#BMW" header read as "#18BMW"
    raw = bytearray(path.read_bytes())
    raw[4096 + 8 * 12 * 0 + 0] ^= 0x40                           # exponent bit of the first double
    broken = tmp_path / "broken.conf"
    broken.write_bytes(bytes(raw))
    with pytest.raises(la.LettuceError):
        f.load_bmw(str(broken))
    assert not la.SaveConfigBMW(f, str(path))                    # exists, no overwrite
    assert la.SaveConfigBMW(f, str(path), overwrite=True)
    short = tmp_path / "short.conf"
    short.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(la.LettuceError):
        f.load_bmw(str(short))
    f.close()
    other.close()


def test_checkpoint_resume(la, gpu_ctx, tmp_path):
    """Rotating checkpoints of field and generator state; a run resumed from the newest
    checkpoint continues the same Markov chain bit for bit."""
    flush("checkpoint resume")
    shape = (8, 8, 8, 8)
    U = la.GaugeField(gpu_ctx, shape)
    U.SetToHaar(2)
    U.set_rng(9, 0)
    hb = la.HeatbathKernel(U, 5.9)
    cm = la.CheckpointManager(tmp_path, 2)
    prng = la.PhiloxState(U)
    for _ in range(3):
        hb.sweep(1, 2)
        cm.AlternatingCheckpoints(la.SaveConfigBMWFull, prng, U, "config.conf", "prng_state.txt",
                                  "distribution_state.txt")
    hb.sweep(2, 2)
    expect = U.download()
    V = la.GaugeField(gpu_ctx, shape)
    assert cm.LoadCheckpoint(la.LoadConfigBMWFull, la.PhiloxState(V), V, "config.conf", "prng_state.txt",
                             "distribution_state.txt")
    la.HeatbathKernel(V, 5.9).sweep(2, 2)
    assert np.array_equal(V.download(), expect)
    assert os.path.exists(tmp_path / "backup1" / "config.conf")
    U.close()
    V.close()
