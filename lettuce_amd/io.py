"""Configuration files, generator state and rotating checkpoints.

  reference (LettuceGauge/...)                                 here
  SaveConfigBMW / LoadConfigBMW (IO/config_io/bmw_format.hpp)  SaveConfigBMW / LoadConfigBMW
  SaveConfigBMWFull / LoadConfigBMWFull ("#18BMW")             SaveConfigBMWFull / LoadConfigBMWFull
  CheckpointManager (IO/config_io/checkpoint_manager.hpp)      CheckpointManager
  PRNG4D::SaveState / LoadState (the PCG generators)           PhiloxState.SaveState / LoadState

The BMW files are written and read on the device (lt_field_save_bmw /
lt_field_load_bmw); the functions here keep the reference's bool results.  The
random numbers here come from counter-based Philox streams, so a generator state is
two integers (seed, counter) per field; PhiloxState writes them in a one-line text
file and leaves the reference's second file (the normal distribution's cached
value) as a marker, since Box-Muller draws here keep no state between calls.
"""
import os
import shutil

from ._lib import LettuceError

__all__ = ["SaveConfigBMW", "SaveConfigBMWFull", "LoadConfigBMW", "LoadConfigBMWFull", "CheckpointManager",
           "PhiloxState"]


def _save(U, filename, overwrite, full):
    if os.path.exists(filename) and not overwrite:
        return False
    U.save_bmw(filename, full=full, overwrite=True)
    return True


def SaveConfigBMW(U, filename, overwrite=False):
    """bmw_format.hpp:315-402: False if the file exists and overwrite is False."""
    return _save(U, filename, overwrite, False)


def SaveConfigBMWFull(U, filename, overwrite=False):
    """bmw_format.hpp:517-600 ("#18BMW", all three rows)."""
    return _save(U, filename, overwrite, True)


def _load(U, filename, full):
    if not os.path.exists(filename):
        return False
    try:
        U.load_bmw(filename, full=full)
    except LettuceError:
        return False
    return True


def LoadConfigBMW(U, filename):
    """bmw_format.hpp:153-313: False for a missing file, a header that does not match,
    other extents, or links that are not SU(3) within 1e-6.  A checksum mismatch is
    reported by GaugeField.load_bmw, not a failure (as in the reference)."""
    return _load(U, filename, False)


def LoadConfigBMWFull(U, filename):
    """bmw_format.hpp:425-515."""
    return _load(U, filename, True)


class PhiloxState:
    """Generator state of one or more Philox streams: fields (or replicas of a batch),
    whose seed and counter drive their heatbath draws, momenta and Metropolis draws, and
    MetadynamicsTemperingKernel swap proposals -- the counterpart of PRNG4D::SaveState /
    LoadState, which keeps all of a run's draws in one saved generator.  The file holds
    one line "philox4x32-10 <seed> <counter>" per holder, in the order given."""

    TAG = "philox4x32-10"

    def __init__(self, *holders):
        if not holders:
            raise ValueError("PhiloxState needs at least one field or tempering kernel")
        self.holders = holders
        self.field = holders[0]

    def SaveState(self, filename_prng, filename_distribution, overwrite=False):
        for f in (filename_prng, filename_distribution):
            if os.path.exists(f) and not overwrite:
                return False
        lines = []
        for h in self.holders:
            seed, counter = h.get_rng()
            lines.append(f"{self.TAG} {seed} {counter}\n")
        with open(filename_prng, "w") as f:
            f.writelines(lines)
        with open(filename_distribution, "w") as f:
            f.write(f"{self.TAG} stateless\n")
        return True

    def LoadState(self, filename_prng, filename_distribution=None):
        with open(filename_prng) as f:
            rows = [line.split() for line in f if line.strip()]
        if len(rows) != len(self.holders):
            raise ValueError(f"{filename_prng}: {len(rows)} generator states for {len(self.holders)} holders")
        for row in rows:
            if len(row) != 3 or row[0] != self.TAG:
                raise ValueError(f"{filename_prng}: not a {self.TAG} state")
        for h, (_, seed, counter) in zip(self.holders, rows):
            h.set_rng(int(seed), int(counter))
        return True


class CheckpointManager:
    """Rotating checkpoints (checkpoint_manager.hpp:28-159): the newest files live in
    the checkpoint directory, older ones in backup1, backup2, ... (higher = older);
    n_checkpoint_files generations are kept."""

    def __init__(self, checkpoint_directory, n_checkpoint_files, create_subdirectories=True):
        self.checkpoint_directory = str(checkpoint_directory)
        self.n_checkpoint_files = int(n_checkpoint_files)
        if create_subdirectories:
            self.CreateSubdirectories()

    def CheckpointDirectory(self):
        return self.checkpoint_directory

    @staticmethod
    def ReturnBackupSubdirectory(checkpoint_counter):
        return "" if checkpoint_counter == 0 else f"backup{checkpoint_counter}"

    def _dir(self, k):
        return os.path.join(self.checkpoint_directory, self.ReturnBackupSubdirectory(k))

    def CreateSubdirectories(self):
        for k in range(1, self.n_checkpoint_files):
            os.makedirs(self._dir(k), exist_ok=True)

    @staticmethod
    def RotateFiles(current_filepath, old_filepath, copy_file=True):
        if os.path.exists(current_filepath):
            if copy_file:
                shutil.copyfile(current_filepath, old_filepath)
            else:
                os.replace(current_filepath, old_filepath)

    def _rotate(self, names):
        # from the oldest generation down: generation k-1 -> k; the newest one is moved
        # (not copied) so that the save below finds no file in its place
        for k in range(self.n_checkpoint_files - 1, 0, -1):
            for name in names:
                self.RotateFiles(os.path.join(self._dir(k - 1), name), os.path.join(self._dir(k), name), k != 1)

    def AlternatingConfigCheckpoints(self, SaveFunction, U, filename_config):
        self._rotate([filename_config])
        SaveFunction(U, os.path.join(self.checkpoint_directory, filename_config), True)

    def AlternatingCheckpoints(self, SaveFunction, prng, U, filename_config, filename_prng,
                               filename_normal_distribution):
        self._rotate([filename_config, filename_prng, filename_normal_distribution])
        SaveFunction(U, os.path.join(self.checkpoint_directory, filename_config), True)
        prng.SaveState(os.path.join(self.checkpoint_directory, filename_prng),
                       os.path.join(self.checkpoint_directory, filename_normal_distribution), True)

    def LoadCheckpoint(self, LoadFunction, prng, U, filename_config, filename_prng, filename_normal_distribution):
        ok = LoadFunction(U, os.path.join(self.checkpoint_directory, filename_config))
        prng.LoadState(os.path.join(self.checkpoint_directory, filename_prng),
                       os.path.join(self.checkpoint_directory, filename_normal_distribution))
        return ok

    @staticmethod
    def CreateCompletedFile(completed_filepath):
        open(os.path.join(str(completed_filepath), "completed_run"), "w").close()
