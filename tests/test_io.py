"""Files exchanged with the reference, on the CPU: the oracle's BMW writer and the
library's bias-potential files against what the reference's own writers produced
(tests/golden/io16.json, made by tests/golden/make_golden.py --io from the compiled
reference), the checksum against a brute-force restatement, and the checkpoint
rotation (pure file handling)."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
IO_GOLDEN = os.path.join(HERE, "golden", "io16.json")
SHAPE = (16, 16, 16, 16)


@pytest.fixture(scope="module")
def io_golden():
    if not os.path.exists(IO_GOLDEN):
        pytest.skip("io16.json missing (python tests/golden/make_golden.py --io)")
    with open(IO_GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def field16(lto):
    return lto.third_row_fma(lto.stout(lto.haar(SHAPE, 0), SHAPE, 0.1, 3))


@pytest.mark.parametrize("full", [False, True])
def test_oracle_bmw_file_equals_reference(lto, io_golden, field16, full):
    """bmw_format.hpp:315-402 / 517-600: same header line and the same bytes (SHA-256)."""
    data, cs = lto.bmw_file_bytes(field16, SHAPE, full)
    ref = io_golden["bmw18" if full else "bmw12"]
    assert data[:data.index(b"\n")].decode() == ref["header_line"]
    assert ref["header_line"].endswith(" " + cs)
    assert len(data) == ref["bytes"]
    assert hashlib.sha256(data).hexdigest() == ref["sha256"]


def adler64_loop(words):
    """Checksum_Adler64::Add / Finalize (bmw_format.hpp:55-92) word by word."""
    p = 4294967291
    nblk, bsz = len(words), len(words[0])
    A = B = 0
    for n, block in enumerate(words):
        sa = sb = 0
        r = (nblk - n) * 2 * bsz
        for w in block:
            h, l = int(w) >> 32, int(w) & 0xFFFFFFFF
            sa += h + l
            sb += ((2 + r * h + (r - 1) * l) % (1 << 64)) % p
            r -= 2
        A = (A + sa) % p
        B = (B + sb) % p
    return ((B % p) << 32) + (A + 1) % p


def test_checksum_vectorised_equals_word_loop(lto):
    rng = np.random.default_rng(5)
    W = rng.integers(0, 2**64, size=(37, 48), dtype=np.uint64)
    assert lto.bmw_checksum(W) == adler64_loop(W.tolist())
    W[3, 7] ^= np.uint64(1)   # one flipped bit changes the checksum
    assert lto.bmw_checksum(W) == adler64_loop(W.tolist())


def test_checksum_closed_form_single_word(lto):
    """One block of one word w = (h, l): A = h + l + 1, B = 2 + 2h + l (r = 2)."""
    h, l = 123456789, 987654321
    w = np.array([[(h << 32) | l]], dtype=np.uint64)
    assert lto.bmw_checksum(w) == ((2 + 2 * h + l) << 32) + (h + l + 1)


def _bias_from_golden(la, g):
    cmin, cmax, bins, w, width, wt, thr = g["args"]
    b = la.MetaBiasPotential(cmin, cmax, int(bins), w, width, wt, thr, well_tempered=True)
    for d in g["deposits"]:
        b.UpdatePotentialSymmetric(d)
    return b


def test_bias_files_equal_reference(io_golden, tmp_path):
    """SaveParameters + SavePotential (metadynamics.hpp:304-346) write the reference's text."""
    import lettuce_amd as la
    g = io_golden["bias"]
    b = _bias_from_golden(la, g)
    path = str(tmp_path / "metapotential.txt")
    b.SaveParameters(path, overwrite=True)
    b.SavePotential(path)
    assert open(path).read() == g["text"]
    b.close()


def test_oracle_bias_text_equals_reference(lto, io_golden):
    g = io_golden["bias"]
    cmin, cmax, bins, w, width, wt, thr = g["args"]
    b = lto.Bias(cmin, cmax, int(bins), w, width, wt, thr, well_tempered=True)
    for d in g["deposits"]:
        b.deposit_symmetric(d)
    assert lto.bias_file_text(cmin, cmax, int(bins), w, width, wt, thr, b.grid, 0) == g["text"]


def test_bias_load_takes_last_snapshot(io_golden, tmp_path):
    """LoadPotential (metadynamics.hpp:351-505): parameters from the head, the grid and
    exceeded_count from the last two lines of a file SavePotential appended to."""
    import lettuce_amd as la
    g = io_golden["bias"]
    b = _bias_from_golden(la, g)
    path = str(tmp_path / "pot.txt")
    b.SaveParameters(path, overwrite=True)
    b.SavePotential(path)
    b.UpdatePotential(9.5)        # outside [CV_min, CV_max): exceeded_count 1
    b.SavePotential(path)
    c = la.MetaBiasPotential(-1.0, 1.0, 4, 1.0, 1.0)
    assert c.LoadPotential(path)
    info = c.info()
    assert info["bin_number"] == 80 and info["CV_min"] == -8.0 and info["CV_max"] == 8.0
    assert info["weight"] == 0.05 and info["width"] == 0.2 and info["threshold_weight"] == 1000.0
    assert info["exceeded_count"] == 1
    # the file holds six significant digits
    np.testing.assert_allclose(c.grid(), b.grid(), rtol=5e-6, atol=1e-300)
    assert not c.LoadPotential(str(tmp_path / "missing.txt"))
    bad = tmp_path / "bad.txt"
    bad.write_text("SU(3)_version_0.9\n")
    assert not c.LoadPotential(str(bad))
    b.close()
    c.close()


def test_checkpoint_rotation(tmp_path):
    """AlternatingCheckpoints (checkpoint_manager.hpp:110-148): generation k-1 moves to
    backup k, the newest files are written to the checkpoint directory."""
    import lettuce_amd as la

    class FakeField:
        def __init__(self):
            self.gen = 0
            self.rng = (7, 0)

        def get_rng(self):
            return self.rng

        def set_rng(self, seed, counter=0):
            self.rng = (seed, counter)

    def save(U, path, overwrite):
        assert overwrite
        with open(path, "w") as f:
            f.write(f"gen {U.gen}")

    def load(U, path):
        U.gen = int(open(path).read().split()[1])
        return True

    U = FakeField()
    prng = la.PhiloxState(U)
    cm = la.CheckpointManager(tmp_path, 3)
    assert os.path.isdir(tmp_path / "backup1") and os.path.isdir(tmp_path / "backup2")
    for gen in range(1, 5):
        U.gen = gen
        U.rng = (7, 100 * gen)
        cm.AlternatingCheckpoints(save, prng, U, "config.conf", "prng_state.txt", "distribution_state.txt")
    assert (tmp_path / "config.conf").read_text() == "gen 4"
    assert (tmp_path / "backup1" / "config.conf").read_text() == "gen 3"
    assert (tmp_path / "backup2" / "config.conf").read_text() == "gen 2"
    assert (tmp_path / "backup1" / "prng_state.txt").read_text().split() == ["philox4x32-10", "7", "300"]
    U.gen, U.rng = 0, (0, 0)
    assert cm.LoadCheckpoint(load, prng, U, "config.conf", "prng_state.txt", "distribution_state.txt")
    assert U.gen == 4 and U.rng == (7, 400)
    cm.AlternatingConfigCheckpoints(save, U, "config_temper.conf")
    assert (tmp_path / "config_temper.conf").read_text() == "gen 4"
    cm.CreateCompletedFile(tmp_path)
    assert (tmp_path / "completed_run").exists()


def test_philox_state_of_several_streams(tmp_path):
    """A two-stream tempering run keeps three Philox streams (the unbiased field's, the
    MetaD field's and the swap proposals'); one state file holds all of them."""
    import lettuce_amd as la

    class Holder:
        def __init__(self, seed, counter):
            self.rng = (seed, counter)

        def get_rng(self):
            return self.rng

        def set_rng(self, seed, counter=0):
            self.rng = (seed, counter)

    a, b, c = Holder(5, 12), Holder(6, 3), Holder(77, 2)
    st = la.PhiloxState(a, b, c)
    assert st.SaveState(str(tmp_path / "p.txt"), str(tmp_path / "d.txt"))
    assert not st.SaveState(str(tmp_path / "p.txt"), str(tmp_path / "d.txt"))
    assert (tmp_path / "p.txt").read_text().splitlines() == ["philox4x32-10 5 12", "philox4x32-10 6 3",
                                                             "philox4x32-10 77 2"]
    a.rng = b.rng = c.rng = (0, 0)
    assert st.LoadState(str(tmp_path / "p.txt"))
    assert (a.rng, b.rng, c.rng) == ((5, 12), (6, 3), (77, 2))
    with pytest.raises(ValueError):
        la.PhiloxState(a, b).LoadState(str(tmp_path / "p.txt"))
